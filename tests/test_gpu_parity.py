"""Parity of the HIP path (through the C ABI) against the FP64 CPU oracle.

Stage-level tests capture the oracle's working set right before one subsystem
of a chosen IMEX stage, run that subsystem on both sides and compare every
output array, ghosts included.  Whole-run tests compare the conserved fields
after N steps (relative L2 per field, bound 1e-9) and check div B (<= 1e-12).

Tolerances: explicit / CT kernels repeat the oracle's arithmetic, so they agree
to a few ulps; outputs of GMRES / Newton differ by the solve tolerance
(GMRES stops at 1e-14 relative, dot products are summed in a different order).
"""
import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import rel_l2, field_errors, make_case, interior

pytestmark = pytest.mark.gpu

SUBS = ["explicit_conv", "mag_update_b", "explicit_press", "pre_subsystem_pressure", "ct_update"]
CASES = [("vortex", 40, 24), ("ot", 32, 32), ("blast", 24, 40)]
_snap_cache = {}


def snapshot(case, n, m, point, presteps=1):
    key = (case, n, m, point, presteps)
    if key not in _snap_cache:
        p, u, bx, by = make_case(case, n, m)
        o = orc.Oracle(p)
        o.set_state(u, bx, by)
        o.step(presteps)
        st, sc, _ = o.step_capture(point)
        _snap_cache[key] = (p, st, sc)
    return _snap_cache[key]


def run_gpu_sub(p, st, sc, name):
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        s.sub(name)
        out, _ = s.get_stage()
    finally:
        s.close()
    return out


SLOTS = {"explicit_conv": [0, 1, 2, 3], "mag_update_b": [5, 6, 7], "explicit_press": [1, 2, 3, 4],
         "pre_subsystem_pressure": [1, 2, 4], "ct_update": [5, 6]}
TOL = {"explicit_conv": 1e-14, "mag_update_b": 1e-11, "explicit_press": 1e-14, "pre_subsystem_pressure": 1e-10,
       "ct_update": 1e-13}


@pytest.mark.parametrize("case,n,m", CASES)
@pytest.mark.parametrize("rks", [0, 1])
@pytest.mark.parametrize("sub", range(5))
def test_stage_subsystem_parity(case, n, m, rks, sub):
    name = SUBS[sub]
    print(f"[parity] {case} {n}x{m} rks={rks} {name}", flush=True)
    p, st, sc = snapshot(case, n, m, 10 * rks + sub)
    ref, rc, _ = orc.sub(p, sc, st, name)
    out = run_gpu_sub(p, st, sc, name)
    tol = TOL[name]
    for k in SLOTS[name]:
        e = rel_l2(out["u_con"][..., k], ref["u_con"][..., k])
        assert e <= tol, f"{name} slot {k}: rel L2 {e:.3e} > {tol}"
    if name == "mag_update_b":
        for f in ("fx", "fy"):
            e = rel_l2(out[f], ref[f])
            assert e <= 1e-11, f"{f}: {e:.3e}"
    if name == "ct_update":
        for f in ("bx", "by"):
            e = rel_l2(out[f], ref[f])
            assert e <= 1e-13, f"{f}: {e:.3e}"


@pytest.mark.parametrize("case,n,m", CASES)
@pytest.mark.parametrize("op", [im.OP_MAG, im.OP_P_CARRIER, im.OP_P_JACOBIAN])
def test_operator_parity(case, n, m, op):
    print(f"[op] {case} {n}x{m} op={op}", flush=True)
    point = 1 if op == im.OP_MAG else 3
    p, st, sc = snapshot(case, n, m, point)
    rng = np.random.default_rng(7)
    N = p.nx * p.ny
    x = rng.standard_normal(3 * N if op == im.OP_MAG else N)
    pf = None
    mbig = 0.0
    if op == im.OP_P_JACOBIAN:
        pf = np.abs(rng.standard_normal(N)) + 0.5
        pf[::7] = 0.1 * p.p_floor          # some cells below the floor: slope = Mbig there
        mbig = 123.25
    ref = orc.apply_op(p, sc, st, op, x, pf, mbig)
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        out = s.apply_op(op, x, pf, mbig)
    finally:
        s.close()
    e = rel_l2(out, ref)
    assert e <= 1e-14, f"operator {op}: {e:.3e}"


@pytest.mark.parametrize("case,n,m", CASES)
@pytest.mark.parametrize("op,precond", [(im.OP_MAG, 0), (im.OP_P_CARRIER, 1), (im.OP_P_JACOBIAN, 1)])
def test_gmres_parity(case, n, m, op, precond):
    print(f"[gmres] {case} {n}x{m} op={op}", flush=True)
    point = 1 if op == im.OP_MAG else 3
    p, st, sc = snapshot(case, n, m, point)
    rng = np.random.default_rng(11)
    N = p.nx * p.ny
    dim = 3 * N if op == im.OP_MAG else N
    b = rng.standard_normal(dim)
    x0 = np.zeros(dim)
    pf = np.full(N, 1.0) if op == im.OP_P_JACOBIAN else None
    xr, ir, _ = orc.gmres(p, sc, st, op, b, x0, precond, pf, 0.0)
    s = im.Solver(p)
    try:
        s.set_stage(st, sc)
        xg, ig, _ = s.gmres(op, b, x0, precond, pf, 0.0)
    finally:
        s.close()
    e = rel_l2(xg, xr)
    assert e <= 1e-12, f"gmres x: {e:.3e} (iters gpu {ig['iters']} oracle {ir['iters']})"
    assert abs(ig["iters"] - ir["iters"]) <= 1


@pytest.mark.parametrize("case,n,m,steps", [("vortex", 64, 64, 10), ("ot", 64, 64, 5), ("blast", 64, 64, 3),
                                            ("vortex", 40, 24, 3), ("blast", 24, 40, 2)])
def test_full_run_parity(case, n, m, steps):
    print(f"[run] {case} {n}x{m} steps={steps}", flush=True)
    p, u, bx, by = make_case(case, n, m)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(steps)
    uc_r, up_r, bx_r, by_r, sc_r = o.get_state()
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        rc, stats = s.step(steps)
        uc, up, bxg, byg, sc = s.get_state()
    finally:
        s.close()
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), "dt", sc["dt"], sc_r["dt"], flush=True)
    assert max(errs) <= 1e-9, errs
    assert abs(sc["t"] - sc_r["t"]) <= 1e-12 * abs(sc_r["t"])
    assert rel_l2(interior(bxg), interior(bx_r)) <= 1e-9
    assert rel_l2(interior(byg), interior(by_r)) <= 1e-9
    assert stats["div_b"] <= 1e-12, stats["div_b"]
