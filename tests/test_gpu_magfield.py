"""Magnetic treatments other than StaggerCT (config.hpp:28, main.cpp:424-434, 497-510) on the GPU
through the C ABI, against the oracle.

UnStaggerCT: MagSubsystem_UpdateB, then Unstaggered_CT_Consistent (constrained_transport.cpp:502-540)
evolves the vector potential A_z.  PotentialA: MagSubsystem_UpdateA (imex_subsystems.cpp:5-74) solves
for A with Magnetic_Subsystem_potential_AX (implicit_operators.cpp:355-503), no CT.  NoDivFree:
MagSubsystem_UpdateB only.  Stage-level: both A sub-steps on the oracle's captured working set and
A / A_s; the potential operator and its GMRES solve on random vectors, periodic and transmissive.
Whole runs: conserved fields, faces and A after N steps (relative L2 <= 1e-9), the cell-centred
div B metric, every schedule, the mixed form (<= 1e-6) and a batch bit for bit equal to separate runs.
"""
import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import rel_l2, field_errors, make_case_a, interior

pytestmark = pytest.mark.gpu

CASES = [("vortex", 40, 24, None), ("ot", 32, 32, None), ("blast", 24, 40, 10.0)]
_snap = {}


def snapshot(case, n, m, amp, mf, point):
    key = (case, n, m, amp, mf, point)
    if key not in _snap:
        p, u, bx, by, A = make_case_a(case, n, m, amp=amp, magfield=mf)
        o = orc.Oracle(p)
        o.set_state(u, bx, by)
        o.set_potential(A)
        o.step(1)
        st, sc, _ = o.step_capture(point)
        A, As = o.capture_potential()
        _snap[key] = (p, st, sc, A, As)
    return _snap[key]


@pytest.mark.parametrize("case,n,m,amp", CASES)
@pytest.mark.parametrize("rks", [0, 1])
@pytest.mark.parametrize("name,mf,point,slots,tol", [
    ("mag_update_a", im.MAGFIELD_POTENTIAL_A, 1, [5, 6, 7], 1e-11),
    ("unstaggered_ct", im.MAGFIELD_UNSTAGGER, 4, [5, 6], 1e-13)])
def test_potential_stage_parity(case, n, m, amp, rks, name, mf, point, slots, tol):
    print(f"[magfield] {case} {n}x{m} rks={rks} {name}", flush=True)
    p, st, sc, A, As = snapshot(case, n, m, amp, mf, 10 * rks + point)
    ref, A_ref, rc, _ = orc.sub_a(p, sc, st, A, As, name)
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        s.set_stage_potential(A, As)
        s.sub(name)
        out, _ = s.get_stage()
        A_out, _ = s.get_stage_potential()
    finally:
        s.close()
    for k in slots:
        e = rel_l2(out["u_con"][..., k], ref["u_con"][..., k])
        assert e <= tol, f"{name} slot {k}: rel L2 {e:.3e} > {tol}"
    for k in range(3):
        e = rel_l2(A_out[..., k], A_ref[..., k])
        assert e <= tol, f"{name} A[{k}]: rel L2 {e:.3e} > {tol}"


@pytest.mark.parametrize("case,n,m,amp", CASES)
def test_potential_operator_and_gmres_parity(case, n, m, amp):
    print(f"[potential op] {case} {n}x{m}", flush=True)
    p, st, sc, A, As = snapshot(case, n, m, amp, im.MAGFIELD_POTENTIAL_A, 1)
    rng = np.random.default_rng(23)
    N = p.nx * p.ny
    x = rng.standard_normal(3 * N)
    b = rng.standard_normal(3 * N)
    ref = orc.apply_op(p, sc, st, im.OP_POTENTIAL, x)
    xr, ir, _ = orc.gmres(p, sc, st, im.OP_POTENTIAL, b, np.zeros(3 * N), 0)
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        out = s.apply_op(im.OP_POTENTIAL, x)
        xg, ig, _ = s.gmres(im.OP_POTENTIAL, b, np.zeros(3 * N), 0)
    finally:
        s.close()
    e = rel_l2(out, ref)
    assert e <= 1e-14, e
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
    st_ref = o.stats()
    o.close()
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        if p.magfield in (im.MAGFIELD_UNSTAGGER, im.MAGFIELD_POTENTIAL_A):
            s.set_potential(A)
        rc, stats = s.step(steps)
        got = s.get_state()
        A_got = s.get_potential() if p.magfield in (im.MAGFIELD_UNSTAGGER, im.MAGFIELD_POTENTIAL_A) else None
    finally:
        s.close()
    return ref, A_ref, st_ref, got, A_got, stats


@pytest.mark.parametrize("mf", [im.MAGFIELD_UNSTAGGER, im.MAGFIELD_POTENTIAL_A, im.MAGFIELD_NODIVFREE])
@pytest.mark.parametrize("case,n,m,steps,amp,sched", [
    ("vortex", 64, 64, 8, None, im.SCHED_GRAPH), ("ot", 64, 64, 5, None, im.SCHED_GRAPH),
    ("blast", 48, 48, 4, 10.0, im.SCHED_GRAPH), ("vortex", 40, 24, 3, None, im.SCHED_HOST),
    ("ot", 48, 40, 3, None, im.SCHED_BUDGET), ("blast", 64, 64, 3, None, im.SCHED_GRAPH)])
def test_magfield_full_run_parity(mf, case, n, m, steps, amp, sched):
    print(f"[magfield run] mf={mf} {case} {n}x{m} steps={steps} schedule={sched}", flush=True)
    p, u, bx, by, A = make_case_a(case, n, m, amp=amp, magfield=mf, schedule=sched)
    (uc_r, _, bx_r, by_r, sc_r), A_r, st_r, (uc, _, bxg, byg, sc), A_g, stats = run_pair(p, u, bx, by, A, steps)
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), "div_b", stats["div_b"], st_r["div_b"], flush=True)
    assert max(errs) <= 1e-9, errs
    assert abs(sc["t"] - sc_r["t"]) <= 1e-12 * abs(sc_r["t"])
    assert np.array_equal(bxg, bx_r) and np.array_equal(byg, by_r)      # the faces are left alone
    assert abs(stats["div_b"] - st_r["div_b"]) <= 1e-9 * max(st_r["div_b"], 1e-300) + 1e-13
    if A_g is not None:
        for k in range(3):
            assert rel_l2(A_g[..., k], A_r[..., k]) <= 1e-9, k


@pytest.mark.parametrize("mf", [im.MAGFIELD_UNSTAGGER, im.MAGFIELD_POTENTIAL_A])
def test_magfield_mixed_form_within_1e6(mf):
    print(f"[magfield mixed] mf={mf} ot 64x64 steps=6", flush=True)
    p, u, bx, by, A = make_case_a("ot", 64, 64, magfield=mf)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.set_potential(A)
    o.step(6)
    uc_r = o.get_state()[0]
    p.precision = im.PREC_MIXED
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        s.set_potential(A)
        s.step(6)
        uc = s.get_state()[0]
    finally:
        s.close()
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-6, errs


def test_magfield_batch_matches_single_runs():
    print("[magfield batch] PotentialA vortices 40x24, 32x32, 24x48", flush=True)
    shapes = [(40, 24), (32, 32), (24, 48)]
    cases = [make_case_a("vortex", n, m, magfield=im.MAGFIELD_POTENTIAL_A) for n, m in shapes]
    b = im.Batch(cases[0][0], shapes)
    try:
        for i, (p, u, bx, by, A) in enumerate(cases):
            b.set_instance(i, u, bx, by)
            b.set_potential_of(i, A)
        b.step(3)
        outs = [b.get_instance(i)[0] for i in range(len(shapes))]
        As = [b.get_potential_of(i) for i in range(len(shapes))]
    finally:
        b.close()
    for i, (p, u, bx, by, A) in enumerate(cases):
        s = im.Solver(p)
        try:
            s.set_state(u, bx, by)
            s.set_potential(A)
            s.step(3)
            assert np.array_equal(interior(outs[i]), interior(s.get_state()[0])), i
            assert np.array_equal(As[i], s.get_potential()), i
        finally:
            s.close()


def test_potential_required_and_rejected():
    p, u, bx, by, A = make_case_a("vortex", 24, magfield=im.MAGFIELD_POTENTIAL_A)
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        with pytest.raises(im.ImexError):
            s.step(1)                      # the vector potential was never set
    finally:
        s.close()
    q, u, bx, by, A = make_case_a("vortex", 24)
    s = im.Solver(q)
    try:
        with pytest.raises(im.ImexError):
            s.set_potential(A)             # StaggerCT does not evolve A
    finally:
        s.close()
