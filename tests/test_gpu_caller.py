"""The compiled C++ caller (imex_amd/caller/imex_caller.cpp, built by imex_amd/Makefile) on the GPU.

It is the reference-side program a maintainer would write: it includes include/imex.h, links
libimex_amd.so and hands the library its Grid / Array buffers as the reference holds them.  The
test writes those buffers (and the imex_params / imex_scalars structs, byte for byte as the C ABI
lays them out) into a scratch directory, runs the program as a child process and reads back
what it wrote.

  whole step   imex_create / imex_set_state / imex_step per pass / imex_get_state (main.cpp:334-561):
               conserved fields against the FP64 oracle (<= 1e-9 per field), div B <= 1e-12, and the
               same bits as the ctypes binding's run of the same problem
  one stage    imex_set_stage / one subsystem / imex_get_stage (declarations.hpp:74-96), on the
               oracle's working set right before that subsystem, against the oracle's subsystem
               (the stage-parity tolerances)
"""
import os
import subprocess

import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import field_errors, make_case, rel_l2, interior

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "imex_amd", "imex_caller")


def _caller(*args, timeout=120):
    if not os.path.exists(CALLER):
        pytest.fail("imex_amd/imex_caller is not built (make -C imex_amd)")
    r = subprocess.run([CALLER, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    print(r.stdout.strip(), r.stderr.strip(), flush=True)
    assert r.returncode == 0, f"imex_caller {args} exited {r.returncode}: {r.stderr}"
    return r


def _write(path, a):
    np.ascontiguousarray(a, dtype=np.float64).tofile(path)


def _read(path, shape):
    a = np.fromfile(path, dtype=np.float64)
    assert a.size == int(np.prod(shape)), path
    return a.reshape(shape)


@pytest.mark.parametrize("case,n,m,steps", [("ot", 64, 64, 5), ("blast", 48, 40, 3)])
def test_caller_whole_step(tmp_path, case, n, m, steps):
    print(f"[caller] step {case} {n}x{m} steps={steps}", flush=True)
    p, u, bx, by = make_case(case, n, m)
    (tmp_path / "params.bin").write_bytes(bytes(p))
    _write(tmp_path / "u_pri.bin", u)
    _write(tmp_path / "bx.bin", bx)
    _write(tmp_path / "by.bin", by)
    _caller("step", tmp_path, steps)
    uc = _read(tmp_path / "out_u_con.bin", im.grid_shape(p))
    bxg = _read(tmp_path / "out_bx.bin", im.array_shape(p))
    byg = _read(tmp_path / "out_by.bin", im.array_shape(p))
    sc = im.Scalars.from_buffer_copy((tmp_path / "out_scalars.bin").read_bytes())
    st = im.Stats.from_buffer_copy((tmp_path / "out_stats.bin").read_bytes())
    assert sc.timestep == steps and st.steps == steps

    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(steps)
    uc_r, _, bx_r, by_r, sc_r = o.get_state()
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field vs oracle:", " ".join(f"{e:.2e}" for e in errs), f"div B {st.div_b:.2e}", flush=True)
    assert max(errs) <= 1e-9, errs
    assert rel_l2(interior(bxg), interior(bx_r)) <= 1e-9 and rel_l2(interior(byg), interior(by_r)) <= 1e-9
    assert abs(sc.t - sc_r["t"]) <= 1e-12 * sc_r["t"]
    assert st.div_b <= 1e-12

    s = im.Solver(p)                 # the same library through the ctypes binding: the same bits
    try:
        s.set_state(u, bx, by)
        for _ in range(steps):
            s.step(1)
        uc_p, _, bx_p, by_p, _ = s.get_state()
    finally:
        s.close()
    assert np.array_equal(uc, uc_p) and np.array_equal(bxg, bx_p) and np.array_equal(byg, by_p)


STAGE_CASES = [("explicit_conv", 0, [0, 1, 2, 3], 1e-14), ("mag_update_b", 1, [5, 6, 7], 1e-11),
               ("explicit_press", 2, [1, 2, 3, 4], 1e-14), ("pre_subsystem_pressure", 3, [1, 2, 4], 1e-10),
               ("ct_update", 4, [5, 6], 1e-13)]


@pytest.mark.parametrize("name,point,slots,tol", STAGE_CASES)
def test_caller_stage_drop_in(tmp_path, name, point, slots, tol):
    case, n, m, rks = "ot", 40, 32, 1
    print(f"[caller] stage {name}: {case} {n}x{m} rks={rks}", flush=True)
    p, u, bx, by = make_case(case, n, m)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(1)
    st, sc, _ = o.step_capture(10 * rks + point)
    ref, _, _ = orc.sub(p, sc, st, name)

    scs = im.Scalars()
    for k, v in sc.items():
        setattr(scs, k, v)
    (tmp_path / "params.bin").write_bytes(bytes(p))
    (tmp_path / "scalars.bin").write_bytes(bytes(scs))
    for k, a in st.items():
        _write(tmp_path / f"{k}.bin", a)
    _caller("stage", tmp_path, name)
    out = {k: _read(tmp_path / f"out_{k}.bin", a.shape) for k, a in st.items()}
    for k in slots:
        e = rel_l2(out["u_con"][..., k], ref["u_con"][..., k])
        print(f"  slot {k}: rel L2 {e:.2e}", flush=True)
        assert e <= tol, (k, e)
    if name == "mag_update_b":
        assert rel_l2(out["fx"], ref["fx"]) <= 1e-11 and rel_l2(out["fy"], ref["fy"]) <= 1e-11
    if name == "ct_update":
        assert rel_l2(out["bx"], ref["bx"]) <= 1e-13 and rel_l2(out["by"], ref["by"]) <= 1e-13
