"""Discrete adjoint of the step against central finite differences.

The library's adjoint (imex_adjoint_forward / imex_adjoint_backward) is wrapped here, and only
here, as a torch.autograd.Function: forward runs the steps and returns the objective J (total
kinetic energy at the final step), backward returns dJ/d(initial primitive state, faces).  Each
check compares <grad J, d> with (J(x0 + eps d) - J(x0 - eps d)) / (2 eps) along seeded random
directions d on the interior cells (ghost cells are refilled from the interior by set_state).
The transposed operators are checked by the dot-product test <y, A x> = <A^T y, x>.

The initial data is the MHD vortex with a seeded relative perturbation of 1e-3 on every interior
value.  The unperturbed vortex is symmetric on its grid, so several CFL maxima (B_max_x,
B_max_y, max|vx|, max|vy|) are attained by two cells with exactly the same value; the step is
not differentiable there (finite differences average the two one-sided derivatives, the
adjoint follows the first cell), which shows up as 1e-4..1e-2 disagreement in every direction.
"""
import numpy as np
import pytest

import imex_amd as im
from helpers import make_case

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu
G = im.GC


def vortex(n, m=None, noise=1e-3, seed=123):
    p, u, bx, by = make_case("vortex", n, m or n)
    if noise:
        r = np.random.default_rng(seed)
        u[G:-G, G:-G, :] *= 1 + noise * r.standard_normal(u[G:-G, G:-G, :].shape)
        bx[G:-G, G:-G] *= 1 + noise * r.standard_normal(bx[G:-G, G:-G].shape)
        by[G:-G, G:-G] *= 1 + noise * r.standard_normal(by[G:-G, G:-G].shape)
    return p, u, bx, by


def kinetic(p, uc):
    """J = sum over interior cells of |rho v|^2 / (2 rho) dx dy, and dJ/du_con."""
    dx = (p.x1 - p.x0) / p.nx
    dy = (p.y1 - p.y0) / p.ny
    ui = uc[G:-G, G:-G]
    rho, m = ui[..., 0], ui[..., 1:4]
    m2 = (m * m).sum(-1)
    J = float((0.5 * m2 / rho).sum() * dx * dy)
    g = np.zeros_like(uc)
    g[G:-G, G:-G, 0] = -0.5 * m2 / rho ** 2 * dx * dy
    g[G:-G, G:-G, 1:4] = m / rho[..., None] * dx * dy
    return J, g


def run_J(p, u, bx, by, steps):
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        s.step(steps)
        uc = s.get_state()[0]
    finally:
        s.close()
    return kinetic(p, uc)[0]


class KineticEnergyAfterSteps(torch.autograd.Function):
    """J(u_pri, bx, by) = kinetic energy after `steps` IMEX-RK2 steps, through the C ABI."""

    @staticmethod
    def forward(ctx, u, bx, by, p, steps, every):
        s = im.Solver(p)
        s.set_state(u.detach().numpy(), bx.detach().numpy(), by.detach().numpy())
        s.adjoint_forward(steps, every)
        uc = s.get_state()[0]
        J, g = kinetic(p, uc)
        ctx.solver, ctx.g = s, g
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        s = ctx.solver
        try:
            au, abx, aby, info = s.adjoint_backward(ctx.g)
        finally:
            s.close()
        ctx.info = info
        sc = float(gJ)
        return (torch.from_numpy(au * sc), torch.from_numpy(abx * sc), torch.from_numpy(aby * sc), None, None, None)


def interior_direction(rng, shape_u, shape_f):
    du = np.zeros(shape_u)
    du[G:-G, G:-G, :] = rng.standard_normal(du[G:-G, G:-G, :].shape)
    du[..., 4] *= 0.5                       # keep the pressure perturbation moderate
    dbx = np.zeros(shape_f); dby = np.zeros(shape_f)
    dbx[G:-G, G:-G] = rng.standard_normal(dbx[G:-G, G:-G].shape)
    dby[G:-G, G:-G] = rng.standard_normal(dby[G:-G, G:-G].shape)
    return du, dbx, dby


def fd_check(n, steps, every, ndir, seed, eps=1e-6, scale=1e-2, m=None, time_stop=None, setup=None):
    p, u, bx, by = vortex(n, m)
    if time_stop is not None:
        p.time_stop = time_stop
    if setup is not None:
        setup(p)
    ut = torch.tensor(u, requires_grad=True)
    bxt = torch.tensor(bx, requires_grad=True)
    byt = torch.tensor(by, requires_grad=True)
    J = KineticEnergyAfterSteps.apply(ut, bxt, byt, p, steps, every)
    gu, gbx, gby = torch.autograd.grad(J, (ut, bxt, byt))
    gu, gbx, gby = gu.numpy(), gbx.numpy(), gby.numpy()
    rng = np.random.default_rng(seed)
    errs = []
    for k in range(ndir):
        du, dbx, dby = interior_direction(rng, u.shape, bx.shape)
        du *= scale; dbx *= scale; dby *= scale
        Jp = run_J(p, u + eps * du, bx + eps * dbx, by + eps * dby, steps)
        Jm = run_J(p, u - eps * du, bx - eps * dbx, by - eps * dby, steps)
        fd = (Jp - Jm) / (2 * eps)
        ad = float((gu * du).sum() + (gbx * dbx).sum() + (gby * dby).sum())
        errs.append(abs(ad - fd) / max(abs(fd), 1e-300))
        print(f"  direction {k}: adjoint {ad:.12e} fd {fd:.12e} rel {errs[-1]:.2e}", flush=True)
    return np.array(errs)


@pytest.mark.parametrize("op,nc,eos", [(im.OP_MAG, 3, im.EOS_IDEAL), (im.OP_P_CARRIER, 1, im.EOS_IDEAL),
                                        (im.OP_P_JACOBIAN, 1, im.EOS_IDEAL), (im.OP_P_CARRIER, 1, im.EOS_RK),
                                        (im.OP_P_JACOBIAN, 1, im.EOS_RK)])
def test_transposed_operator_dot_product(op, nc, eos):
    """<y, A x> = <A^T y, x>; the Jacobian mode with p_freeze on both sides of p_floor (V_slope = Mbig
    below it), the Redlich-Kwong operators with their viscosity plane and V_slope."""
    print(f"[adjoint] dot-product test op={op} eos={eos}", flush=True)
    p, u, bx, by = vortex(32, 24)
    p.eos = eos
    p.p_floor = 0.99
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        s.step(2)
        N = p.nx * p.ny
        rng = np.random.default_rng(3)
        x = rng.standard_normal(nc * N)
        y = rng.standard_normal(nc * N)
        pf, mbig = None, 0.0
        if op == im.OP_P_JACOBIAN:
            pf = 1.0 + 0.02 * rng.standard_normal(N)
            mbig = 40.0
            assert (pf < p.p_floor).any() and (pf >= p.p_floor).any()
        Ax = s.apply_op(op, x, pf, mbig)
        ATy = s.apply_op_T(op, y, pf, mbig)
    finally:
        s.close()
    lhs, rhs = float(y @ Ax), float(ATy @ x)
    print(f"  <y,Ax> {lhs:.15e}  <A^T y,x> {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_adjoint_matches_central_differences_small():
    print("[adjoint] vortex 24x24, 4 steps, checkpoint every 2", flush=True)
    errs = fd_check(24, 4, 2, 4, seed=11)
    assert np.median(errs) <= 1e-5, errs
    assert errs.max() <= 1e-4, errs


def test_adjoint_one_step_uneven_grid():
    print("[adjoint] vortex 32x20, 1 step", flush=True)
    errs = fd_check(32, 1, 1, 3, seed=5, m=20)
    assert errs.max() <= 1e-5, errs


def test_adjoint_weno5():
    """WENO5 reconstruction (reconstruction.cpp:184-266): the flux element's stencil is six cells."""
    print("[adjoint] vortex 24x24 WENO5, 3 steps, checkpoint every 2", flush=True)

    def weno5(p):
        p.recon = im.RECON_WENO5
    errs = fd_check(24, 3, 2, 4, seed=21, setup=weno5)
    assert np.median(errs) <= 1e-5, errs
    assert errs.max() <= 1e-4, errs


def floored_cells(n, steps, setup):
    """Largest number of cells the nested Newton floors in one solve over the run, and min p."""
    p, u, bx, by = vortex(n)
    setup(p)
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        s.step(steps)
        st = s.stats()
    finally:
        s.close()
    return st["max_floored"], st["min_p"]


@pytest.mark.parametrize("mbig_fac", [10.0, 1e-3])
def test_adjoint_active_pressure_floor(mbig_fac):
    """p_floor above the vortex's lowest pressures: the nested Newton floors an annulus of cells
    (implicit_operators.cpp:625-738).  The transposed solves use the Newton Jacobian (V_slope = Mbig
    below the floor) and Mbig's cotangent reaches the floored cells' p; with a small Mbig_fac the Cor. 1
    slope M_cor(min b) decides Mbig, so the cotangent also reaches b at the cell of min b."""
    def floor(p):
        p.floor_on = 1
        p.p_floor = 0.99
        p.mbig_fac = mbig_fac
    nfl, pmin = floored_cells(24, 3, floor)
    print(f"[adjoint] vortex 24x24 with p_floor 0.99, Mbig_fac {mbig_fac}: up to {nfl:.0f} floored cells, min p {pmin:.6f}",
          flush=True)
    assert nfl > 0
    errs = fd_check(24, 3, 2, 4, seed=31, setup=floor)
    assert np.median(errs) <= 3e-5, errs
    assert errs.max() <= 3e-4, errs


def test_adjoint_redlich_kwong():
    """Redlich-Kwong EOS (config.hpp:130-173): temperature solves inside p(Qe), the sound speeds and the
    energy; the pressure solve is the nested Newton's on Vbase(p, rho) = rho e(rho, p), and the
    operator's viscosity scale is Vbase_slope(p(Qe), rho)."""
    print("[adjoint] vortex 24x24 Redlich-Kwong, 3 steps, checkpoint every 2", flush=True)

    def rk(p):
        p.eos = im.EOS_RK
    errs = fd_check(24, 3, 2, 4, seed=41, setup=rk)
    assert np.median(errs) <= 3e-5, errs
    assert errs.max() <= 3e-4, errs


def test_adjoint_redlich_kwong_active_floor():
    """Redlich-Kwong with the floor active and Mbig decided by Cor. 1, whose Vpe is max rho e(rho, p_floor)
    over the cells: Mbig's cotangent reaches rho at the cell that attains it.  The nested Newton stops at a
    relative residual of 1e-8 (implicit_operators.cpp:703-706); with the floored cells and the
    Redlich-Kwong energy its last iterate keeps a truncation error that differences of eps = 1e-6 amplify
    to ~1e-4 (measured: max 3.5e-4 at eps = 1e-6, 3.5e-5 at 1e-5, 7e-4 at 1e-4 where the differences cross
    the floor's kink), so this check differences with eps = 1e-5."""
    def rkf(p):
        p.eos = im.EOS_RK
        p.floor_on = 1
        p.p_floor = 0.99
        p.mbig_fac = 1e-3
    nfl, pmin = floored_cells(24, 3, rkf)
    print(f"[adjoint] vortex 24x24 Redlich-Kwong with p_floor 0.99: up to {nfl:.0f} floored cells, min p {pmin:.6f}",
          flush=True)
    assert nfl > 0
    errs = fd_check(24, 3, 2, 4, seed=43, eps=1e-5, setup=rkf)
    assert np.median(errs) <= 3e-5, errs
    assert errs.max() <= 2e-4, errs


def test_adjoint_last_step_clipped_to_time_stop():
    """dt = time_stop - t on the last step (main.cpp:341): its cotangent reaches every earlier dt
    through t, and the clipped step's own CFL dt gets none."""
    n, steps = 24, 4
    p, u, bx, by = vortex(n)
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        s.step(steps - 1)
        t3 = s.get_state()[4]["t"]
        s.step(1)
        dt4 = s.get_state()[4]["dt"]
    finally:
        s.close()
    ts = t3 + 0.6 * dt4
    print(f"[adjoint] vortex {n}x{n}, {steps} steps, time_stop {ts:.9e} clips the last step (t3 {t3:.9e}, dt4 {dt4:.3e})",
          flush=True)
    errs = fd_check(n, steps, 2, 4, seed=17, time_stop=ts)
    assert np.median(errs) <= 1e-5, errs
    assert errs.max() <= 1e-4, errs


def test_adjoint_resident_tapes_equal_recomputed_segments():
    """The forward can keep every step's tape in device memory (the default when it fits) or keep
    checkpoints only and re-run each segment in the backward pass (adjoint_recompute = 1): the taped
    steps are the same computations, so the gradients agree to the rounding of the transposed element
    functions' atomic accumulation (whose order is not fixed)."""
    n, steps, every = 32, 5, 2
    print(f"[adjoint] vortex {n}x{n}, {steps} steps: resident tapes against recomputed segments (every {every})", flush=True)
    grads = []
    for recompute in (0, 1):
        p, u, bx, by = vortex(n)
        p.adjoint_recompute = recompute
        s = im.Solver(p)
        try:
            s.set_state(u, bx, by)
            s.adjoint_forward(steps, every)
            J, g = kinetic(p, s.get_state()[0])
            au, abx, aby, info = s.adjoint_backward(g)
        finally:
            s.close()
        print(f"  adjoint_recompute={recompute}: J {J:.15e}, recomputed steps {info['recomputed_steps']}", flush=True)
        assert info["recomputed_steps"] == (steps if recompute else 0)
        grads.append((au, abx, aby))
    for a, b in zip(*grads):
        d = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"  relative difference {d:.2e}", flush=True)
        assert d <= 1e-12


def test_adjoint_unsupported_settings():
    p, u, bx, by = make_case("blast", 32, 32)
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        with pytest.raises(im.ImexError):
            s.adjoint_forward(2, 1)
    finally:
        s.close()
