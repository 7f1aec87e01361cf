"""The benchmark configurations themselves, on the GPU, against the FP64 oracle.

  Orszag-Tang 512^2, WENO5, 50 steps      FP64 and mixed forms against the oracle (<= 1e-9 / <= 1e-6 per
                                          conserved field), div B <= 1e-12, same time and solver counts
  magnetic blast 256^2, VanLeer, 20 steps FP64 and mixed against the oracle; the nested Newton's positivity
                                          statistics (smallest pressure, floored cells, Cor. 1 warning) equal;
                                          div B at round-off of |B|/dx (|B| = 28 here)
  256 vortex instances, sizes from        every instance: div B <= 1e-12 and mass conserved; a seeded sample
    {64..512}^2, ideal gas and RK, 10 steps  of instances against the oracle (<= 1e-9), RK ones included
  discrete adjoint, vortex 256^2,          32 seeded directions against central differences (eps 1e-6)
    20 steps, checkpoint every 5

The oracle runs its per-cell loops on the host's cores (OpenMP); the 512^2 run takes it about half a
minute.  Reference: src/main.cpp:334-595 (the step loop the oracle restates).
"""
import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import make_case, field_errors, rel_l2, interior

pytestmark = pytest.mark.gpu
G = im.GC


def _gpu_run(p, u, bx, by, steps, A=None):
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        rc, stats = s.step(steps)
        out = s.get_state()
    finally:
        s.close()
    return out, stats, rc


def _oracle_run(p, u, bx, by, steps):
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    rc = o.step(steps)
    return o.get_state(), o.stats(), rc


def test_orszag_tang_512_50_steps_fp64_and_mixed():
    print("[configs] orszag-tang 512x512 weno5 50 steps: oracle", flush=True)
    p, u, bx, by = make_case("ot", 512, 512)
    (uc_r, up_r, bx_r, by_r, sc_r), st_r, rc_r = _oracle_run(p, u, bx, by, 50)
    for prec, bar in ((im.PREC_FP64, 1e-9), (im.PREC_MIXED, 1e-6)):
        print(f"[configs] orszag-tang 512x512 50 steps precision={prec}", flush=True)
        p.precision = prec
        (uc, up, bxg, byg, sc), stats, rc = _gpu_run(p, u, bx, by, 50)
        errs = field_errors(uc, uc_r)
        print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), f"div B {stats['div_b']:.2e}",
              f"t {sc['t']:.12e} / {sc_r['t']:.12e}", flush=True)
        assert max(errs) <= bar, errs
        assert rel_l2(interior(bxg), interior(bx_r)) <= bar and rel_l2(interior(byg), interior(by_r)) <= bar
        assert stats["div_b"] <= 1e-12
        assert abs(sc["t"] - sc_r["t"]) <= (1e-12 if prec == im.PREC_FP64 else 1e-7) * sc_r["t"]
        if prec == im.PREC_FP64:
            assert abs(stats["p_gmres_iters"] - st_r["p_gmres_iters"]) <= 0.02 * st_r["p_gmres_iters"] + 4
            assert abs(stats["mag_gmres_iters"] - st_r["mag_gmres_iters"]) <= 0.02 * st_r["mag_gmres_iters"] + 4


def test_blast_256_20_steps_fp64_and_mixed():
    print("[configs] magnetic blast 256x256 vanleer 20 steps: oracle", flush=True)
    p, u, bx, by = make_case("blast", 256, 256)
    (uc_r, _, bx_r, by_r, sc_r), st_r, rc_r = _oracle_run(p, u, bx, by, 20)
    print(f"  oracle: min p {st_r['min_p']:.6e}, max floored {st_r['max_floored']}, warnings {st_r['warnings']}, "
          f"newton {st_r['newton_iters']}", flush=True)
    for prec, bar in ((im.PREC_FP64, 1e-9), (im.PREC_MIXED, 1e-6)):
        print(f"[configs] magnetic blast 256x256 20 steps precision={prec}", flush=True)
        p.precision = prec
        (uc, _, bxg, byg, sc), stats, rc = _gpu_run(p, u, bx, by, 20)
        errs = field_errors(uc, uc_r)
        print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), f"div B {stats['div_b']:.2e}",
              f"min p {stats['min_p']:.6e} floored {stats['max_floored']} warnings {stats['warnings']}", flush=True)
        assert max(errs) <= bar, errs
        # the metric is round-off of faces whose size is |B|/dx: the bar is 1e-12 or 1e-14 of that scale
        scale = np.abs(uc_r[G:-G, G:-G, 5:7]).max() / ((p.x1 - p.x0) / p.nx) * np.sqrt((p.x1 - p.x0) * (p.y1 - p.y0))
        print(f"  div B {stats['div_b']:.2e} (oracle {orc.div_b(p, bx_r, by_r):.2e}, scale {scale:.2e})", flush=True)
        assert stats["div_b"] <= max(1e-12, 1e-14 * scale)
        if prec == im.PREC_FP64:
            assert stats["warnings"] & im.capi.WARN_POSITIVITY == st_r["warnings"] & im.capi.WARN_POSITIVITY
            assert stats["max_floored"] == st_r["max_floored"]
            assert stats["min_p"] == pytest.approx(st_r["min_p"], rel=1e-6)


BATCH_SIZES = (64, 96, 128, 192, 256, 384, 512)


def _batch_workload(n_inst, seed):
    rng = np.random.default_rng(seed)
    sizes = [(int(rng.choice(BATCH_SIZES)), int(rng.choice(BATCH_SIZES))) for _ in range(n_inst)]
    amps = rng.uniform(0.5, 2.0, size=(n_inst, 3))
    return sizes, amps


@pytest.mark.parametrize("eos", [im.EOS_IDEAL, im.EOS_RK])
def test_batch_256_vortices_10_steps(eos):
    sizes, amps = _batch_workload(256, 2024)
    print(f"[configs] batch of 256 vortices eos={eos}, {sum(a * b for a, b in sizes) / 1e6:.2f} Mcells, 10 steps", flush=True)
    p = im.vortex_params(64, eos=eos)
    b = im.Batch(p, sizes)
    inits = []
    try:
        for i, (nx, ny) in enumerate(sizes):
            q = b.params_of(i)
            u, bx, by = im.init_case(q, im.CASE_VORTEX, amps[i, 0], amps[i, 1], amps[i, 2], 0)
            b.set_instance(i, u, bx, by)
        rc, stats = b.step(10)
        print(f"  stats {stats}", flush=True)
        assert rc >= 0
        # a seeded sample against the oracle: the smallest instances of the batch and two random ones
        order = sorted(range(len(sizes)), key=lambda i: sizes[i][0] * sizes[i][1])
        sample = sorted(set(order[:3] + list(np.random.default_rng(7).choice(order[:64], 2, replace=False))))
        for i in range(len(sizes)):
            q = b.params_of(i)
            u0, bx0, by0 = im.init_case(q, im.CASE_VORTEX, amps[i, 0], amps[i, 1], amps[i, 2], 0)
            uc, up, bxg, byg, sc = b.get_instance(i)
            dx, dy = (q.x1 - q.x0) / q.nx, (q.y1 - q.y0) / q.ny
            m0 = orc.pri2con(q, u0)[G:-G, G:-G, 0].sum()
            assert abs(uc[G:-G, G:-G, 0].sum() - m0) <= 1e-12 * m0, i
            assert b.div_b_of(i) <= 1e-12, (i, b.div_b_of(i))
            if i in sample:
                print(f"  instance {i} {sizes[i]} against the oracle", flush=True)
                (uc_r, _, _, _, sc_r), _, _ = _oracle_run(q, u0, bx0, by0, 10)
                errs = field_errors(uc, uc_r)
                print("    rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
                assert max(errs) <= 1e-9, (i, errs)
                assert abs(sc["t"] - sc_r["t"]) <= 1e-12 * sc_r["t"]
    finally:
        b.close()


def test_adjoint_vortex_256_20_steps_32_directions():
    n, steps, every, ndir, eps = 256, 20, 5, 32, 1e-6
    print(f"[configs] adjoint vortex {n}x{n} {steps} steps checkpoint {every} directions {ndir}", flush=True)
    p, u, bx, by = make_case("vortex", n, n)
    r = np.random.default_rng(1000 + 1)
    u[G:-G, G:-G, :] *= 1 + 1e-3 * r.standard_normal(u[G:-G, G:-G, :].shape)
    bx[G:-G, G:-G] *= 1 + 1e-3 * r.standard_normal(bx[G:-G, G:-G].shape)
    by[G:-G, G:-G] *= 1 + 1e-3 * r.standard_normal(by[G:-G, G:-G].shape)
    dx, dy = (p.x1 - p.x0) / p.nx, (p.y1 - p.y0) / p.ny

    def kinetic(uc):
        ui = uc[G:-G, G:-G]
        rho, m = ui[..., 0], ui[..., 1:4]
        m2 = (m * m).sum(-1)
        g = np.zeros_like(uc)
        g[G:-G, G:-G, 0] = -0.5 * m2 / rho ** 2 * dx * dy
        g[G:-G, G:-G, 1:4] = m / rho[..., None] * dx * dy
        return float((0.5 * m2 / rho).sum() * dx * dy), g

    s = im.Solver(p)
    try:
        def J_of(uu, bbx, bby):
            s.set_state(uu, bbx, bby)
            s.step(steps)
            return kinetic(s.get_state()[0])[0]

        s.set_state(u, bx, by)
        s.adjoint_forward(steps, every)
        J, g = kinetic(s.get_state()[0])
        au, abx, aby, info = s.adjoint_backward(g)
        print(f"  J {J:.12e}; {info}", flush=True)
        rng = np.random.default_rng(1)
        ad, fd = [], []
        for k in range(ndir):
            du = np.zeros_like(u); dbx = np.zeros_like(bx); dby = np.zeros_like(by)
            du[G:-G, G:-G, :] = rng.standard_normal(du[G:-G, G:-G, :].shape) * 1e-2
            dbx[G:-G, G:-G] = rng.standard_normal(dbx[G:-G, G:-G].shape) * 1e-2
            dby[G:-G, G:-G] = rng.standard_normal(dby[G:-G, G:-G].shape) * 1e-2
            f = (J_of(u + eps * du, bx + eps * dbx, by + eps * dby) - J_of(u - eps * du, bx - eps * dbx, by - eps * dby)) / (2 * eps)
            ad.append(float((au * du).sum() + (abx * dbx).sum() + (aby * dby).sum()))
            fd.append(f)
    finally:
        s.close()
    ad, fd = np.array(ad), np.array(fd)
    rel = np.abs(ad - fd) / np.abs(fd)
    print(f"  relative error: median {np.median(rel):.2e}, max {rel.max():.2e}, all directions "
          f"{np.linalg.norm(ad - fd) / np.linalg.norm(fd):.2e}; worst solve residual {info['worst_solve_residual']:.1e}",
          flush=True)
    # central differences at eps = 1e-6 carry about 1e-8 absolute noise from the rounding of J (~100) and
    # the solver tolerances; directions whose derivative is small see it as a larger relative error
    assert np.median(rel) <= 2e-5
    assert np.abs(ad - fd).max() <= 2e-5 * np.abs(fd).max()
    assert np.linalg.norm(ad - fd) <= 2e-5 * np.linalg.norm(fd)
