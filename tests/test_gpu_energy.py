"""Energy formula (config.hpp:27, main.cpp:470-486): HIP path through the C ABI against the oracle.

Stage-level: Energy_flux_Update (flux.cpp:165-200) and PreSubsystem_Energy (imex_subsystems.cpp:
415-534) on the oracle's working set captured right before them; Energy_Subsystem_AX
(implicit_operators.cpp:741-790) and its preconditioned GMRES solve on random vectors.  Whole runs:
conserved fields after N steps (relative L2 per field <= 1e-9, div B <= 1e-12) on the three
workloads, the blast at amplitude 10 (at amplitude 100 the energy formula drives the pressure
negative within a few steps, in the oracle as on the GPU), on every schedule and with the RK EOS;
the mixed form within 1e-6; a batch of uneven instances bit for bit equal to separate runs.
"""
import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import rel_l2, field_errors, make_case, interior

pytestmark = pytest.mark.gpu

EN = dict(formula=im.FORMULA_ENERGY)
CASES = [("vortex", 40, 24, None), ("ot", 32, 32, None), ("blast", 24, 40, 10.0)]
_snap = {}


def snapshot(case, n, m, amp, point):
    key = (case, n, m, amp, point)
    if key not in _snap:
        p, u, bx, by = make_case(case, n, m, amp=amp, **EN)
        o = orc.Oracle(p)
        o.set_state(u, bx, by)
        o.step(1)
        st, sc, _ = o.step_capture(point)
        _snap[key] = (p, st, sc)
    return _snap[key]


@pytest.mark.parametrize("case,n,m,amp", CASES)
@pytest.mark.parametrize("rks", [0, 1])
@pytest.mark.parametrize("name,point,slots,tol", [("explicit_energy", 2, [1, 2, 3, 4], 1e-14),
                                                   ("pre_subsystem_energy", 3, [1, 2, 4], 1e-11)])
def test_energy_stage_parity(case, n, m, amp, rks, name, point, slots, tol):
    print(f"[energy] {case} {n}x{m} rks={rks} {name}", flush=True)
    p, st, sc = snapshot(case, n, m, amp, 10 * rks + point)
    ref, rc, _ = orc.sub(p, sc, st, name)
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        s.sub(name)
        out, _ = s.get_stage()
    finally:
        s.close()
    for k in slots:
        e = rel_l2(out["u_con"][..., k], ref["u_con"][..., k])
        assert e <= tol, f"{name} slot {k}: rel L2 {e:.3e} > {tol}"


@pytest.mark.parametrize("case,n,m,amp", CASES)
def test_energy_operator_and_gmres_parity(case, n, m, amp):
    print(f"[energy op] {case} {n}x{m}", flush=True)
    p, st, sc = snapshot(case, n, m, amp, 3)
    rng = np.random.default_rng(17)
    N = p.nx * p.ny
    x = rng.standard_normal(N)
    b = rng.standard_normal(N)
    ref = orc.apply_op(p, sc, st, im.OP_ENERGY, x)
    sols = {pc: orc.gmres(p, sc, st, im.OP_ENERGY, b, np.zeros(N), pc) for pc in (0, 1)}
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        out = s.apply_op(im.OP_ENERGY, x)
        got = {pc: s.gmres(im.OP_ENERGY, b, np.zeros(N), pc) for pc in (0, 1)}
    finally:
        s.close()
    assert rel_l2(out, ref) <= 1e-14, rel_l2(out, ref)
    for pc in (0, 1):
        xr, ir, _ = sols[pc]
        xg, ig, _ = got[pc]
        e = rel_l2(xg, xr)
        print(f"  precond {pc}: iters gpu {ig['iters']} oracle {ir['iters']} rel {e:.2e}", flush=True)
        assert e <= 1e-12, e
        assert abs(ig["iters"] - ir["iters"]) <= 1


def run_pair(p, u, bx, by, steps):
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(steps)
    ref = o.get_state()
    o.close()
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        rc, stats = s.step(steps)
        got = s.get_state()
    finally:
        s.close()
    return ref, got, stats


@pytest.mark.parametrize("case,n,m,steps,amp,sched", [
    ("vortex", 64, 64, 10, None, im.SCHED_GRAPH), ("ot", 64, 64, 6, None, im.SCHED_GRAPH),
    ("blast", 64, 64, 6, 10.0, im.SCHED_GRAPH), ("vortex", 40, 24, 4, None, im.SCHED_HOST),
    ("ot", 48, 40, 4, None, im.SCHED_BUDGET), ("blast", 24, 40, 3, 10.0, im.SCHED_HOST)])
def test_energy_full_run_parity(case, n, m, steps, amp, sched):
    print(f"[energy run] {case} {n}x{m} steps={steps} schedule={sched}", flush=True)
    p, u, bx, by = make_case(case, n, m, amp=amp, schedule=sched, **EN)
    (uc_r, _, bx_r, by_r, sc_r), (uc, _, bxg, byg, sc), stats = run_pair(p, u, bx, by, steps)
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-9, errs
    assert abs(sc["t"] - sc_r["t"]) <= 1e-12 * abs(sc_r["t"])
    assert rel_l2(interior(bxg), interior(bx_r)) <= 1e-9
    assert rel_l2(interior(byg), interior(by_r)) <= 1e-9
    assert stats["div_b"] <= 1e-12, stats["div_b"]
    assert stats["newton_iters"] == 0


def test_energy_full_run_parity_rk_eos():
    print("[energy run] vortex RK EOS 32x32 steps=4", flush=True)
    p, u, bx, by = make_case("vortex", 32, 32, eos=im.EOS_RK, **EN)
    (uc_r, *_), (uc, *_), stats = run_pair(p, u, bx, by, 4)
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-9, errs


@pytest.mark.parametrize("mixed_pressure", [0, 1])
def test_energy_mixed_form_within_1e6(mixed_pressure):
    print(f"[energy mixed] ot 64x64 steps=6 mixed_pressure={mixed_pressure}", flush=True)
    p, u, bx, by = make_case("ot", 64, 64, **EN)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(6)
    uc_r = o.get_state()[0]
    p.precision = im.PREC_MIXED
    p.mixed_pressure = mixed_pressure
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        rc, stats = s.step(6)
        uc = s.get_state()[0]
    finally:
        s.close()
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-6, errs
    assert stats["div_b"] <= 1e-12


def test_energy_batch_matches_single_runs():
    print("[energy batch] vortices 40x24, 32x32, 24x48", flush=True)
    shapes = [(40, 24), (32, 32), (24, 48)]
    cases = [make_case("vortex", n, m, **EN) for n, m in shapes]
    b = im.Batch(cases[0][0], shapes)
    singles = []
    try:
        for i, (p, u, bx, by) in enumerate(cases):
            b.set_instance(i, u, bx, by)
        b.step(4)
        outs = [b.get_instance(i)[0] for i in range(len(shapes))]
    finally:
        b.close()
    for p, u, bx, by in cases:
        s = im.Solver(p)
        try:
            s.set_state(u, bx, by)
            s.step(4)
            singles.append(s.get_state()[0])
        finally:
            s.close()
    for i in range(len(shapes)):
        assert np.array_equal(interior(outs[i]), interior(singles[i])), i


def test_energy_transposes_are_not_offered():
    p, st, sc = snapshot("vortex", 40, 24, None, 3)
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        with pytest.raises(im.ImexError):
            s.apply_op_T(im.OP_ENERGY, np.zeros(p.nx * p.ny))
        with pytest.raises(im.ImexError):
            s.adjoint_forward(1, 1)
    finally:
        s.close()
