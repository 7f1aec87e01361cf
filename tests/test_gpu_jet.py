"""MHD jet inflow boundary (y_JetInflow) on the GPU through the C ABI, against the oracle.

The inflow strip |x_c| < 0.05 of the bottom boundary holds the jet state in its ghost cells
(boundary.cpp:148-172), its ghost faces (constrained_transport.cpp:50-63) and, evenly reflected, the
vector potential (boundary.cpp:279-291); the implicit subsystems treat its ghost as a Dirichlet value
(imex_subsystems.cpp:107-118, 163-165, 202-203, 266-267, 453-475; implicit_operators.cpp:559-561, 772-779).
Outflow (transmissive) everywhere else.  Stage-level: every subsystem of both formulas and the
potential treatment on the oracle's captured working set (both stages); operators and GMRES on random
vectors.  Whole runs: conserved fields after N steps (relative L2 per field <= 1e-9) on the graph and host
schedules, both formulas, PotentialA and UnStaggerCT; the mixed form within 1e-6.
"""
import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import rel_l2, field_errors, make_case, make_case_a, interior

pytestmark = pytest.mark.gpu

SHAPES = [(32, 48), (40, 24)]
_snap = {}


def snapshot(n, m, point, **over):
    key = (n, m, point, tuple(sorted(over.items())))
    if key not in _snap:
        p, u, bx, by, A = make_case_a("jet", n, m, **over)
        o = orc.Oracle(p)
        o.set_state(u, bx, by)
        o.set_potential(A)
        o.step(1)
        st, sc, _ = o.step_capture(point)
        if p.magfield in (im.MAGFIELD_UNSTAGGER, im.MAGFIELD_POTENTIAL_A):
            A, As = o.capture_potential()
        else:
            As = A
        _snap[key] = (p, st, sc, A, As)
    return _snap[key]


SUBS = [("explicit_conv", 0, {}, [0, 1, 2, 3], 1e-14),
        ("mag_update_b", 1, {}, [5, 6, 7], 1e-11),
        ("explicit_press", 2, {}, [1, 2, 3, 4], 1e-14),
        ("pre_subsystem_pressure", 3, {}, [1, 2, 4], 1e-10),
        ("ct_update", 4, {}, [5, 6], 1e-13),
        ("explicit_energy", 2, dict(formula=im.FORMULA_ENERGY), [1, 2, 3, 4], 1e-14),
        ("pre_subsystem_energy", 3, dict(formula=im.FORMULA_ENERGY), [1, 2, 4], 1e-11),
        ("mag_update_a", 1, dict(magfield=im.MAGFIELD_POTENTIAL_A), [5, 6, 7], 1e-11)]


@pytest.mark.parametrize("n,m", SHAPES)
@pytest.mark.parametrize("rks", [0, 1])
@pytest.mark.parametrize("name,point,over,slots,tol", SUBS, ids=[s[0] for s in SUBS])
def test_jet_stage_parity(n, m, rks, name, point, over, slots, tol):
    print(f"[jet] {n}x{m} rks={rks} {name}", flush=True)
    p, st, sc, A, As = snapshot(n, m, 10 * rks + point, **over)
    if name == "mag_update_a":
        ref, A_ref, rc, _ = orc.sub_a(p, sc, st, A, As, name)
    else:
        ref, rc, _ = orc.sub(p, sc, st, name)
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        if name == "mag_update_a":
            s.set_stage_potential(A, As)
        s.sub(name)
        out, _ = s.get_stage()
        A_out = s.get_stage_potential()[0] if name == "mag_update_a" else None
    finally:
        s.close()
    for k in slots:
        e = rel_l2(out["u_con"][..., k], ref["u_con"][..., k])
        assert e <= tol, f"{name} slot {k}: rel L2 {e:.3e} > {tol}"
    if name == "ct_update":
        for f in ("bx", "by"):
            assert rel_l2(out[f], ref[f]) <= 1e-13, f
    if A_out is not None:
        for k in range(3):
            assert rel_l2(A_out[..., k], A_ref[..., k]) <= tol, k


@pytest.mark.parametrize("op,point,over,pc", [(im.OP_P_CARRIER, 3, {}, 1), (im.OP_P_JACOBIAN, 3, {}, 1),
                                             (im.OP_MAG, 1, {}, 0),
                                             (im.OP_ENERGY, 3, dict(formula=im.FORMULA_ENERGY), 1),
                                             (im.OP_POTENTIAL, 1, dict(magfield=im.MAGFIELD_POTENTIAL_A), 0)])
def test_jet_operator_and_gmres_parity(op, point, over, pc):
    print(f"[jet op] 32x48 op={op}", flush=True)
    p, st, sc, A, As = snapshot(32, 48, point, **over)
    rng = np.random.default_rng(31)
    N = p.nx * p.ny
    dim = 3 * N if op in (im.OP_MAG, im.OP_POTENTIAL) else N
    x, b = rng.standard_normal(dim), rng.standard_normal(dim)
    pf = np.abs(rng.standard_normal(N)) + 0.5 if op == im.OP_P_JACOBIAN else None
    ref = orc.apply_op(p, sc, st, op, x, pf, 7.5 if pf is not None else 0.0)
    xr, ir, _ = orc.gmres(p, sc, st, op, b, np.zeros(dim), pc, pf, 7.5 if pf is not None else 0.0)
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        out = s.apply_op(op, x, pf, 7.5 if pf is not None else 0.0)
        xg, ig, _ = s.gmres(op, b, np.zeros(dim), pc, pf, 7.5 if pf is not None else 0.0)
    finally:
        s.close()
    assert rel_l2(out, ref) <= 1e-14, rel_l2(out, ref)
    e = rel_l2(xg, xr)
    print(f"  gmres iters gpu {ig['iters']} oracle {ir['iters']} rel {e:.2e}", flush=True)
    assert e <= 1e-12, e
    assert abs(ig["iters"] - ir["iters"]) <= 1


def run_pair(p, u, bx, by, A, steps):
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.set_potential(A)
    o.step(steps)
    ref = o.get_state()
    A_ref = o.get_potential()
    o.close()
    useA = p.magfield in (im.MAGFIELD_UNSTAGGER, im.MAGFIELD_POTENTIAL_A)
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        if useA:
            s.set_potential(A)
        rc, stats = s.step(steps)
        got = s.get_state()
        A_got = s.get_potential() if useA else None
    finally:
        s.close()
    return ref, A_ref, got, A_got, stats


@pytest.mark.parametrize("n,m,steps,over", [
    (32, 48, 5, dict(schedule=im.SCHED_GRAPH)), (40, 24, 3, dict(schedule=im.SCHED_HOST)),
    (48, 72, 4, dict(schedule=im.SCHED_BUDGET)), (32, 48, 4, dict(formula=im.FORMULA_ENERGY)),
    (32, 48, 3, dict(magfield=im.MAGFIELD_POTENTIAL_A)), (32, 48, 3, dict(magfield=im.MAGFIELD_UNSTAGGER)),
    (32, 48, 3, dict(ct_mode=im.CT_PLAIN)), (32, 48, 3, dict(recon=im.RECON_WENO5, recon_state=im.RECONSTATE_PRIM)),
    (32, 48, 3, dict(ct_mode=im.CT_PLAIN, recon=im.RECON_ENO3, recon_state=im.RECONSTATE_PRIM))])
def test_jet_full_run_parity(n, m, steps, over):
    print(f"[jet run] {n}x{m} steps={steps} {over}", flush=True)
    p, u, bx, by, A = make_case_a("jet", n, m, **over)
    (uc_r, _, bx_r, by_r, sc_r), A_r, (uc, _, bxg, byg, sc), A_g, stats = run_pair(p, u, bx, by, A, steps)
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-9, errs
    assert abs(sc["t"] - sc_r["t"]) <= 1e-12 * abs(sc_r["t"])
    scale = np.linalg.norm(interior(by_r))               # B_x starts at zero: measure both faces against |B_y|
    assert np.linalg.norm(interior(bxg) - interior(bx_r)) <= 1e-9 * scale
    assert np.linalg.norm(interior(byg) - interior(by_r)) <= 1e-9 * scale
    if A_g is not None:
        for k in range(3):
            assert rel_l2(A_g[..., k], A_r[..., k]) <= 1e-9, k


def test_jet_mixed_form_within_1e6():
    print("[jet mixed] 48x72 steps=4", flush=True)
    p, u, bx, by = make_case("jet", 48, 72)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(4)
    uc_r = o.get_state()[0]
    p.precision = im.PREC_MIXED
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        s.step(4)
        uc = s.get_state()[0]
    finally:
        s.close()
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-6, errs
