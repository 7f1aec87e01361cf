"""Mixed-precision form against the FP64 CPU oracle.

The mixed form applies the magnetic operator (and, with mixed_pressure = 1, the pressure
operator) in FP32 inside GMRES solves whose basis is stored in FP32; dot products, norms,
the Hessenberg/Givens solve and the outer residual of the iterative refinement are FP64,
and so are CT and the Newton residual.  With mixed_flux = 1 it also reconstructs and forms
the explicit fluxes in FP32 (the conservative update stays FP64).
Bound: relative L2 at most 1e-6 per conserved field after N steps, div B at most 1e-12
(CT is unchanged).  The accuracy reached is printed for each case.
"""
import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import rel_l2, field_errors, make_case, interior

pytestmark = pytest.mark.gpu

CASES = [("vortex", 64, 64, 10), ("ot", 64, 64, 8), ("blast", 64, 64, 3), ("vortex", 40, 24, 3)]


def run_gpu(p, u, bx, by, steps):
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        rc, stats = s.step(steps)
        out = s.get_state()
    finally:
        s.close()
    return out, stats


@pytest.mark.parametrize("flux", [0, 1])
@pytest.mark.parametrize("case,n,m,steps", CASES)
def test_mixed_form_within_1e6_of_oracle(case, n, m, steps, flux):
    print(f"[mixed] {case} {n}x{m} steps={steps} mixed_flux={flux}", flush=True)
    p, u, bx, by = make_case(case, n, m)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(steps)
    uc_r, up_r, bx_r, by_r, sc_r = o.get_state()
    p.precision = im.PREC_MIXED
    p.mixed_flux = flux
    (uc, up, bxg, byg, sc), stats = run_gpu(p, u, bx, by, steps)
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs),
          f"refinement sweeps {stats['mixed_refinements']}", flush=True)
    assert max(errs) <= 1e-6, errs
    assert rel_l2(interior(bxg), interior(bx_r)) <= 1e-6
    assert stats["div_b"] <= 1e-12, stats["div_b"]
    assert stats["mixed_refinements"] > 0


@pytest.mark.parametrize("case,n,m,steps", [("ot", 64, 64, 8), ("blast", 48, 40, 3)])
def test_mixed_fp32_pressure_solves_within_1e6(case, n, m, steps):
    """The opt-in FP32 inner solves of the pressure system (mixed_pressure = 1) keep the same bound."""
    print(f"[mixed-pressure] {case} {n}x{m} steps={steps}", flush=True)
    p, u, bx, by = make_case(case, n, m)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(steps)
    uc_r = o.get_state()[0]
    p.precision = im.PREC_MIXED
    p.mixed_pressure = 1
    (uc, up, bxg, byg, sc), stats = run_gpu(p, u, bx, by, steps)
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-6, errs
    assert stats["div_b"] <= 1e-12


@pytest.mark.parametrize("case,n,m,steps,flux", [("vortex", 64, 64, 12, 0), ("ot", 64, 64, 8, 0), ("ot", 64, 64, 8, 1)])
def test_mixed_schedules_bitwise_identical(case, n, m, steps, flux):
    print(f"[mixed-schedule] {case} {n}x{m} steps={steps} mixed_flux={flux}", flush=True)
    outs = []
    for sched in (im.SCHED_HOST, im.SCHED_BUDGET, im.SCHED_GRAPH):
        p, u, bx, by = make_case(case, n, m)
        p.precision = im.PREC_MIXED
        p.mixed_flux = flux
        p.schedule = sched
        (uc, up, bxg, byg, sc), stats = run_gpu(p, u, bx, by, steps)
        print(f"  sched {sched}: budget steps {stats['budget_steps']} reruns {stats['budget_reruns']} "
              f"sweeps {stats['mixed_refinements']} iters {stats['mag_gmres_iters']}/{stats['p_gmres_iters']}", flush=True)
        outs.append((uc, bxg, byg, stats))
    ref = outs[0]
    for o in outs[1:]:
        for a, b in zip(o[:3], ref[:3]):
            assert np.array_equal(a, b)
        for k in ("mag_gmres_iters", "p_gmres_iters", "newton_iters", "mixed_refinements"):
            assert o[3][k] == ref[3][k], k
