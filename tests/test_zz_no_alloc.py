"""imex_div_b_noalloc (no device allocation) against imex_div_b, through the C ABI.

The entry point takes device face arrays and a device output and keeps its partial sums in the
library's fixed device array; it must give imex_div_b's metric (same partial sums, same summation
order) for consecutive calls, the second on a smaller grid, which reuses the partials of the first.
"""
import ctypes as C

import numpy as np
import pytest

import imex_amd as im
from helpers import make_case

pytestmark = pytest.mark.gpu


def _reference(case, n, m, steps, seed):
    """A context's state and its imex_div_b; with a seed, faces with a real divergence."""
    p, u, bx, by = make_case(case, n, m)
    if seed is not None:
        rng = np.random.default_rng(seed)
        bx = bx + 0.1 * rng.standard_normal(bx.shape)
        by = by + 0.1 * rng.standard_normal(by.shape)
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        if steps:
            s.step(steps)
        _, _, bxs, bys, _ = s.get_state()
        ref = s.div_b()
    finally:
        s.close()
    return p, bxs, bys, ref


class _Hip:
    """The caller's side: device buffers from the HIP runtime itself (the entry point allocates none)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0          # hipMemcpyHostToDevice
        return p

    def download(self, p, n):
        a = np.empty(n, dtype=np.float64)
        assert self.L.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0          # hipMemcpyDeviceToHost
        return a

    def free(self, *ps):
        for p in ps:
            self.L.hipFree(p)


def _noalloc(p, bx, by):
    hip = _Hip()
    dbx, dby = hip.upload(bx), hip.upload(by)
    out = hip.upload(np.array([np.nan]))
    try:
        rc = im.lib().imex_div_b_noalloc(C.byref(p), dbx, dby, out, None)
        assert rc == 0, im.lib().imex_last_error()
        assert hip.L.hipDeviceSynchronize() == 0
        return float(hip.download(out, 1)[0])
    finally:
        hip.free(dbx, dby, out)


def test_noalloc_div_b_matches_imex_div_b_on_consecutive_shrinking_calls():
    cases = [("ot", 96, 80, 0, 3), ("vortex", 40, 24, 0, 4), ("ot", 64, 64, 3, None), ("vortex", 24, 20, 2, None)]
    for case, n, m, steps, seed in cases:
        print(f"[noalloc] {case} {n}x{m} steps={steps} seed={seed}", flush=True)
        p, bx, by, ref = _reference(case, n, m, steps, seed)
        got = _noalloc(p, bx, by)
        print(f"  imex_div_b {ref:.17e}  noalloc {got:.17e}", flush=True)
        assert abs(got - ref) <= 2e-16 * abs(ref) + 1e-300, (got, ref)
        if seed is not None:
            assert ref > 1.0


def test_noalloc_rejects_bad_arguments():
    p, _, _, _ = make_case("vortex", 16, 16)
    hip = _Hip()
    buf = hip.upload(np.zeros(64))
    L = im.lib()
    try:
        q = im.default_params()
        C.memmove(C.byref(q), C.byref(p), C.sizeof(p))
        q.nx = 40000
        assert L.imex_div_b_noalloc(C.byref(q), buf, buf, buf, None) == im.ERR_ARG
        assert b"32768" in L.imex_last_error()
        assert L.imex_div_b_noalloc(C.byref(p), None, buf, buf, None) == im.ERR_ARG
        q.nx = 0
        assert L.imex_div_b_noalloc(C.byref(q), buf, buf, buf, None) == im.ERR_ARG
    finally:
        hip.free(buf)
