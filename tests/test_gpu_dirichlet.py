"""Dirichlet ambient boundaries (x_Dirichlet / y_Dirichlet) on the GPU through the C ABI, against the oracle.

Ghost cells hold the ambient state rho_amb, v = 0, p_amb with the boundary cell's B (bound_cond_u,
boundary.cpp:56-76, 128-147; E = p_amb/(gamma-1) + |B|^2/2 in a conservative grid, also inside
TVDReconstruction and for the face states); the pressure solve has strong Dirichlet rows on the boundary
cells (imex_subsystems.cpp:160-161, 263-264; implicit_operators.cpp:549-550, 652-655, 697-699); the energy
operator's Dirichlet ghosts are homogeneous with the ambient energy lifted into the right-hand side
(implicit_operators.cpp:772-779, imex_subsystems.cpp:453-475).  Stage-level: every subsystem of both
formulas and the potential treatments on captured stages of a magnetised blast, both stages; operators and
GMRES on random vectors.  Whole runs: conserved fields after N steps (relative L2 per field <= 1e-9) on both
axes and on one, every schedule, both formulas, PotentialA, UnStaggerCT, plain CT, PrimState; the mixed form
within 1e-6.
"""
import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import rel_l2, field_errors, make_case_a, interior

pytestmark = pytest.mark.gpu

WALLS = [(im.BC_DIRICHLET, im.BC_DIRICHLET), (im.BC_DIRICHLET, im.BC_PERIODIC)]
_snap = {}


def case(n, m, walls, amp=10.0, **over):
    p, u, bx, by, A = make_case_a("blast", n, m, amp=amp, **over)
    p.bc_x, p.bc_y = walls
    p.rho_amb, p.p_amb = 1.0, 0.1
    u = orc.bc(p, u, 4)                 # the primitive ghosts of the ambient
    return p, u, bx, by, A


def snapshot(n, m, walls, point, **over):
    key = (n, m, walls, point, tuple(sorted(over.items())))
    if key not in _snap:
        p, u, bx, by, A = case(n, m, walls, **over)
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
        ("mag_update_a", 1, dict(magfield=im.MAGFIELD_POTENTIAL_A), [5, 6, 7], 1e-11),
        ("unstaggered_ct", 4, dict(magfield=im.MAGFIELD_UNSTAGGER), [5, 6], 1e-13)]


@pytest.mark.parametrize("walls", WALLS, ids=["xy", "x"])
@pytest.mark.parametrize("rks", [0, 1])
@pytest.mark.parametrize("name,point,over,slots,tol", SUBS, ids=[s[0] for s in SUBS])
def test_dirichlet_stage_parity(walls, rks, name, point, over, slots, tol):
    print(f"[dirichlet] 24x20 walls={walls} rks={rks} {name}", flush=True)
    p, st, sc, A, As = snapshot(24, 20, walls, 10 * rks + point, **over)
    useA = name in ("mag_update_a", "unstaggered_ct")
    if useA:
        ref, A_ref, rc, _ = orc.sub_a(p, sc, st, A, As, name)
    else:
        ref, rc, _ = orc.sub(p, sc, st, name)
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        if useA:
            s.set_stage_potential(A, As)
        s.sub(name)
        out, _ = s.get_stage()
        A_out = s.get_stage_potential()[0] if useA else None
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
            assert rel_l2(A_out[..., k], A_ref[..., k]) <= max(tol, 1e-13), k


@pytest.mark.parametrize("walls", WALLS, ids=["xy", "x"])
@pytest.mark.parametrize("op,point,over,pc", [(im.OP_MAG, 1, {}, 0), (im.OP_P_CARRIER, 3, {}, 1),
                                             (im.OP_P_JACOBIAN, 3, {}, 1),
                                             (im.OP_ENERGY, 3, dict(formula=im.FORMULA_ENERGY), 1),
                                             (im.OP_POTENTIAL, 1, dict(magfield=im.MAGFIELD_POTENTIAL_A), 0)])
def test_dirichlet_operator_and_gmres_parity(walls, op, point, over, pc):
    print(f"[dirichlet op] 24x20 walls={walls} op={op}", flush=True)
    p, st, sc, A, As = snapshot(24, 20, walls, point, **over)
    rng = np.random.default_rng(41)
    N = p.nx * p.ny
    dim = 3 * N if op in (im.OP_MAG, im.OP_POTENTIAL) else N
    x, b = rng.standard_normal(dim), rng.standard_normal(dim)
    pf = np.abs(rng.standard_normal(N)) + 0.5 if op == im.OP_P_JACOBIAN else None
    mb = 7.5 if pf is not None else 0.0
    ref = orc.apply_op(p, sc, st, op, x, pf, mb)
    xr, ir, _ = orc.gmres(p, sc, st, op, b, np.zeros(dim), pc, pf, mb)
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        out = s.apply_op(op, x, pf, mb)
        xg, ig, _ = s.gmres(op, b, np.zeros(dim), pc, pf, mb)
    finally:
        s.close()
    assert rel_l2(out, ref) <= 1e-14, rel_l2(out, ref)
    e = rel_l2(xg, xr)
    print(f"  gmres iters gpu {ig['iters']} oracle {ir['iters']} rel {e:.2e}", flush=True)
    assert e <= 1e-12, e
    assert abs(ig["iters"] - ir["iters"]) <= 1


@pytest.mark.parametrize("walls,n,m,steps,over", [
    (WALLS[0], 32, 32, 4, dict(schedule=im.SCHED_GRAPH)), (WALLS[1], 40, 24, 3, dict(schedule=im.SCHED_HOST)),
    (WALLS[0], 48, 40, 3, dict(schedule=im.SCHED_BUDGET)), (WALLS[0], 32, 32, 3, dict(formula=im.FORMULA_ENERGY)),
    (WALLS[0], 32, 32, 3, dict(magfield=im.MAGFIELD_POTENTIAL_A)), (WALLS[1], 32, 32, 3, dict(magfield=im.MAGFIELD_UNSTAGGER)),
    (WALLS[0], 32, 32, 3, dict(ct_mode=im.CT_PLAIN)),
    (WALLS[0], 32, 32, 3, dict(recon=im.RECON_ENO3, recon_state=im.RECONSTATE_PRIM))], ids=lambda v: str(v))
def test_dirichlet_full_run_parity(walls, n, m, steps, over):
    print(f"[dirichlet run] walls={walls} {n}x{m} steps={steps} {over}", flush=True)
    p, u, bx, by, A = case(n, m, walls, **over)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.set_potential(A)
    o.step(steps)
    uc_r, _, bx_r, by_r, sc_r = o.get_state()
    A_r = o.get_potential()
    o.close()
    useA = p.magfield in (im.MAGFIELD_UNSTAGGER, im.MAGFIELD_POTENTIAL_A)
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        if useA:
            s.set_potential(A)
        rc, stats = s.step(steps)
        uc, _, bxg, byg, sc = s.get_state()
        A_g = s.get_potential() if useA else None
    finally:
        s.close()
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-9, errs
    assert abs(sc["t"] - sc_r["t"]) <= 1e-12 * abs(sc_r["t"])
    assert rel_l2(interior(bxg), interior(bx_r)) <= 1e-9 and rel_l2(interior(byg), interior(by_r)) <= 1e-9
    if A_g is not None:
        for k in range(3):
            assert rel_l2(A_g[..., k], A_r[..., k]) <= 1e-9, k


def test_dirichlet_mixed_form_within_1e6():
    print("[dirichlet mixed] 48x48 steps=4", flush=True)
    p, u, bx, by, A = case(48, 48, WALLS[0])
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
