"""Pins of the oracle's composed algebra — CPU only.

These parts of the reference are sums of many face and corner terms; a restatement can slip
a sign, a coefficient or a neighbour index in one of them and still pass identity, linearity
and constant-state checks.  Each is pinned here by what the mathematics fixes:

  part (reference)                                  pin
  Magnetic_Subsystem_AX face terms                  the 3x3 Fourier symbol on a uniform background
    (implicit_operators.cpp:266-306, 311-322)        (B0, rho), derived from the discrete face and
                                                    corner averages in flux form, for modes parallel,
                                                    perpendicular and oblique to B0; the implicit
                                                    Alfven eigenvalue 1 + h^2 B0^2/rho (4 sin^2(k/2)/dx^2)
                                                    and the Picard magnetosonic one with its 1/2
  MagSubsystem_UpdateB right-hand side              b = Qs_B + h curl(v' x B0), v' = (rho v* - h grad p)/rho,
    (imex_subsystems.cpp:594-602)                   exact for uniform B0 and rho, linear rho v* and
                                                    quadratic p (the solve then returns b itself)
  CollectMagFluxes (ct_flux.cpp:5-127)              the flux form of the magnetic solve:
                                                    Qs_By - (h/dx) dx Fx - B~y = -h (nu_y/dy^2) d2y B~y and
                                                    Qs_Bx - (h/dy) dy Fy - B~x = -h (nu_x/dx^2) d2x B~x on
                                                    whole runs; F = -E_z / +E_z on a uniform state
  ShockDetector (flux.cpp:5-76)                     eta = 0 on a uniform state; the closed-form chi~ of a
                                                    uniform compression; the 3x3 maximum of a single
                                                    compressive cell; c_f,min of a 3x3 neighbourhood
  PreSubsystem_Pressure, two passes                 the closed-form solution for one Fourier mode of
    (imex_subsystems.cpp:123-302)                   momentum on a uniform ideal-gas state: momentum from
                                                    the first pass's pressure, E from the second, whose
                                                    right-hand side keeps rho v* in its divergence term
  IMEX-RK2 stage algebra (main.cpp:273-289,         one whole step of a small velocity perturbation of a
    386-409, 539-555) with the pressure             uniform gas equals, per Fourier mode, the linear map
    subsystem's linear part                         composed from the Rusanov fluxes, the pressure solve
                                                    and the two stages (Ae(1,0) = 1/(2b), Ai = [[b,0],[1-b,b]])

Each pin fails for a plausible slip in the restatement (checked by planting them, one at a
time, in a copy of oracle/imex_oracle.cpp): the sign of Magnetic_Subsystem_AX's Bx_2 term, a
halved Fdiff_x in CollectMagFluxes, Ae(1,0) = 1/(2.1 b) in the tableau, a doubled h div u in
ShockDetector, the sign of the right-hand side's B_y v_x term or a halved h/dx in front of its
corner pressure difference, the B_z right-hand side's B_x v_z sign, and a second pressure pass
whose kinetic term reuses rho v*.
"""
import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import G


def _params(nx, ny, Lx, Ly, bc=im.BC_PERIODIC):
    p = im.default_params()
    p.nx, p.ny = nx, ny
    p.x0, p.x1, p.y0, p.y1 = 0.0, Lx, 0.0, Ly
    p.bc_x = p.bc_y = bc
    return p


def _uniform(p, pri):
    return orc.pri2con(p, np.tile(np.asarray(pri, dtype=float), (p.ny + 2 * G, p.nx + 2 * G, 1)))


def _stage(p, u_con, Qe=None, Qs=None):
    st = im.new_stage(p)
    st["u_con"][:] = u_con
    st["Qe"][:] = u_con if Qe is None else Qe
    st["Qs"][:] = u_con if Qs is None else Qs
    return st


# ---------------------------------------------------------------- magnetic operator
def _mag_symbol(k, l, B0, rho, h, dx, dy, nux, nuy):
    """3x3 matrix M with A (X e^{i(k i + l j)}) = (M X) e^{i(k i + l j)} for a uniform background.

    Flux form: the x and y components of the induction equation difference an electric field
    that lives on the y faces (B_x) and on the x faces (B_y), and B_z differences both.  Each
    face field is B0 times a face force built from the linearised magnetic stress, with the
    normal derivative taken between the two cells of the face and the tangential one between
    the face's two corners, a corner value being the mean of its four cells."""
    a, b, c = B0
    S, T = np.exp(1j * k), np.exp(1j * l)
    Dx = (S - 1) / dx                              # cells i, i+1 -> x face i+1/2
    Dy = (T - 1) / dy                              # cells j, j+1 -> y face j+1/2
    Ey = (1 + S) * (T - 1 / T) / (4 * dy)          # corners (i+1/2, j+-1/2) -> x face, y difference
    Ex = (1 + T) * (S - 1 / S) / (4 * dx)          # corners (i+-1/2, j+1/2) -> y face, x difference
    divx = (1 - 1 / S) / dx                        # x faces -> cell
    divy = (1 - 1 / T) / dy                        # y faces -> cell
    I3 = np.eye(3)
    P = 0.5 * np.array([a, b, c], dtype=complex)   # Picard magnetic pressure B0 . X / 2 as a row
    # face fields as rows acting on X = (X1, X2, X3)
    Fy_face = (a / rho) * (Ex * a * I3[1] - Dy * (P - b * I3[1])) + (b / rho) * (Ex * (P - a * I3[0]) - Dy * b * I3[0])
    Fx_face = (b / rho) * (-Dx * (P - a * I3[0]) + Ey * a * I3[1]) + (a / rho) * (-Dx * a * I3[1] + Ey * (P - b * I3[1]))
    Fzx = (c / rho) * (-Dx * (P - a * I3[0]) + Ey * b * I3[0]) + (a / rho) * (-Dx * a * I3[2] - Ey * b * I3[2])
    Fzy = (c / rho) * (Ex * a * I3[1] - Dy * (P - b * I3[1])) + (b / rho) * (-Ex * a * I3[2] - Dy * b * I3[2])
    M = np.eye(3, dtype=complex)
    M[0] += h * h * divy * Fy_face
    M[1] += h * h * divx * Fx_face
    M[2] += h * h * (divx * Fzx + divy * Fzy)
    glf = h * (nux * 4 * np.sin(k / 2) ** 2 / dx ** 2 + nuy * 4 * np.sin(l / 2) ** 2 / dy ** 2)
    return M + glf * np.eye(3)


def _mag_apply_mode(p, st, sc, k, l, X):
    I, J = np.meshgrid(np.arange(p.nx), np.arange(p.ny))
    ph = np.exp(1j * (k * I + l * J)).ravel()
    x = np.concatenate([(X[c] * ph).real for c in range(3)])
    return orc.apply_op(p, sc, st, im.OP_MAG, x), ph


@pytest.mark.parametrize("B0", [(0.5, -0.7, 0.3), (0.9, 0.0, 0.0), (0.0, 0.6, -0.4)])
def test_magnetic_operator_fourier_symbol(B0):
    p = _params(12, 10, 1.5, 0.7)
    rho = 1.3
    con = _uniform(p, [rho, 0.2, -0.1, 0.05, 0.9, *B0])
    st = _stage(p, con)
    dx, dy = 1.5 / p.nx, 0.7 / p.ny
    sc = dict(adt=0.04, B_max_x=1.7, B_max_y=0.6)
    rng = np.random.default_rng(11)
    N = p.nx * p.ny
    for kx, ky in [(1, 0), (0, 2), (3, 1), (5, 4), (2, -3), (6, 5)]:
        k, l = 2 * np.pi * kx / p.nx, 2 * np.pi * ky / p.ny
        X = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        Ax, ph = _mag_apply_mode(p, st, sc, k, l, X)
        M = _mag_symbol(k, l, B0, rho, sc["adt"], dx, dy, sc["B_max_x"] * dx, sc["B_max_y"] * dy)
        MX = M @ X
        want = np.concatenate([(MX[c] * ph).real for c in range(3)])
        assert np.allclose(Ax, want, rtol=0, atol=1e-13 * np.abs(want).max()), (kx, ky)
        Mh = _mag_symbol(k, l, B0, rho, sc["adt"], dx, dy, 0.0, 0.0)     # the face terms alone matter
        assert np.abs(Mh - np.eye(3)).max() > 1e-3, (kx, ky)


def test_magnetic_operator_alfven_and_magnetosonic_modes():
    # B0 = (B, 0, 0): a transverse (y or z) perturbation varying along B0 is the implicit Alfven
    # wave, eigenvalue 1 + h^2 B^2/rho * 4 sin^2(k/2)/dx^2; a B_x perturbation varying across B0
    # sees the Picard magnetic pressure B0 . X / 2, eigenvalue 1 + h^2 B^2/(2 rho) * 4 sin^2(l/2)/dy^2;
    # plus the GLF Laplacian h nu 4 sin^2(./2)/d^2 in every case
    p = _params(16, 12, 2.0, 1.2)
    rho, B = 0.8, 1.1
    st = _stage(p, _uniform(p, [rho, 0.0, 0.0, 0.0, 1.0, B, 0.0, 0.0]))
    dx, dy = 2.0 / p.nx, 1.2 / p.ny
    h, bmx, bmy = 0.03, 1.4, 0.9
    sc = dict(adt=h, B_max_x=bmx, B_max_y=bmy)
    N = p.nx * p.ny
    for kx in (1, 3, 7):
        k = 2 * np.pi * kx / p.nx
        s2 = 4 * np.sin(k / 2) ** 2 / dx ** 2
        for comp in (1, 2):
            X = np.zeros(3, complex)
            X[comp] = 1.0
            Ax, ph = _mag_apply_mode(p, st, sc, k, 0.0, X)
            lam = 1 + h * h * B * B / rho * s2 + h * bmx * dx * s2
            assert np.allclose(Ax[comp * N:(comp + 1) * N], lam * ph.real, rtol=0, atol=1e-13 * lam), (kx, comp)
            assert np.allclose(np.delete(Ax.reshape(3, N), comp, axis=0), 0.0, atol=1e-13)
    for ky in (1, 2, 5):
        l = 2 * np.pi * ky / p.ny
        s2 = 4 * np.sin(l / 2) ** 2 / dy ** 2
        Ax, ph = _mag_apply_mode(p, st, sc, 0.0, l, np.array([1.0, 0, 0], complex))
        lam = 1 + h * h * B * B / (2 * rho) * s2 + h * bmy * dy * s2
        assert np.allclose(Ax[:N], lam * ph.real, rtol=0, atol=1e-13 * lam), ky
        assert np.allclose(Ax[N:], 0.0, atol=1e-13)


# ---------------------------------------------------------------- magnetic right-hand side
def test_magnetic_rhs_closed_form_on_linear_flow():
    # uniform B0 (Qe) and rho (u), linear rho v* (u), quadratic p (Qe), ghosts included: the
    # face averages and the face / corner pressure differences are exact, the field
    # E = v' x B0 with v' = (rho v* - h grad p)/rho is linear, so b = Qs_B + h curl E is a
    # constant; the operator is the identity on constants, so the solve returns b
    p = _params(10, 8, 2.0, 1.0, bc=im.BC_TRANSMISSIVE)
    NX, NY = p.nx + 2 * G, p.ny + 2 * G
    dx, dy = 2.0 / p.nx, 1.0 / p.ny
    X = ((np.arange(NX) - G + 0.5) * dx)[None, :]
    Y = ((np.arange(NY) - G + 0.5) * dy)[:, None]
    rho, B0, Bs = 1.4, np.array([0.5, -0.4, 0.25]), np.array([0.3, 0.7, -0.2])
    m0 = np.array([0.2, -0.1, 0.3])
    Mx, My = np.array([0.35, -0.6, 0.45]), np.array([-0.25, 0.15, 0.5])
    p0, gx, gy, Hxx, Hxy, Hyy = 3.0, 0.7, -0.35, 0.4, -0.9, 0.6
    u = np.zeros((NY, NX, 8))
    u[..., 0] = rho
    for c in range(3):
        u[..., 1 + c] = m0[c] + Mx[c] * X + My[c] * Y
    u[..., 4] = 10.0
    qe_pri = np.zeros((NY, NX, 8))
    qe_pri[..., 0] = 1.0
    qe_pri[..., 4] = p0 + gx * X + gy * Y + 0.5 * Hxx * X * X + Hxy * X * Y + 0.5 * Hyy * Y * Y
    qe_pri[..., 5:8] = B0
    st = _stage(p, u, Qe=orc.pri2con(p, qe_pri), Qs=_uniform(p, [1.0, 0, 0, 0, 1.0, *Bs]))
    h = 0.05
    out, rc, stats = orc.sub(p, dict(adt=h, B_max_x=1.0, B_max_y=1.0), st, "mag_update_b")
    assert rc == 0 and stats["mag_solves"] == 1
    dvx = (Mx - h * np.array([Hxx, Hxy, 0.0])) / rho          # d v'/dx
    dvy = (My - h * np.array([Hxy, Hyy, 0.0])) / rho          # d v'/dy
    dE_dx, dE_dy = np.cross(dvx, B0), np.cross(dvy, B0)       # E = v' x B0 is linear
    curl = np.array([dE_dy[2], -dE_dx[2], dE_dx[1] - dE_dy[0]])
    b = Bs + h * curl
    assert np.allclose(out["u_con"][G:-G, G:-G, 5:8], b, rtol=0, atol=1e-13)
    assert np.abs(curl).min() > 0.01


# ---------------------------------------------------------------- CollectMagFluxes
@pytest.mark.parametrize("case,n,m,point", [("vortex", 24, 20, 1), ("ot", 24, 24, 11), ("blast", 24, 20, 11)])
def test_collect_mag_fluxes_flux_form_of_the_solve(case, n, m, point):
    # B~ solves A B~ = b; A - I and b - Qs_B are differences of face quantities in x (B_y) and y
    # (B_x) except the Laplacian's other direction, so CollectMagFluxes' face fluxes must give
    # Qs_By - h/dx (Fx_i - Fx_i-1) - B~y = -h nu_y/dy^2 (B~y_j+1 - 2 B~y_j + B~y_j-1) and the Bx analogue
    from helpers import make_case
    p, u, bx, by = make_case(case, n, m)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(2)
    st, sc, _ = o.step_capture(point)       # the stage after the explicit RusanovC update
    out, rc, _ = orc.sub(p, sc, st, "mag_update_b")
    assert rc == 0
    h = sc["adt"]
    dx, dy = (p.x1 - p.x0) / p.nx, (p.y1 - p.y0) / p.ny
    nux, nuy = sc["B_max_x"] * dx, sc["B_max_y"] * dy
    Bt = out["u_con"]
    Qs = st["Qs"]
    Fx, Fy = out["fx"], out["fy"]
    s = np.s_[G:-G, G:-G]
    sxm = np.s_[G:-G, G - 1:-G - 1]
    sym = np.s_[G - 1:-G - 1, G:-G]
    syp = np.s_[G + 1:-G + 1, G:-G]
    sxp = np.s_[G:-G, G + 1:-G + 1]
    res_y = Qs[s][..., 6] - h / dx * (Fx[s] - Fx[sxm]) - Bt[s][..., 6] + h * nuy / dy ** 2 * (Bt[syp][..., 6] - 2 * Bt[s][..., 6] + Bt[sym][..., 6])
    res_x = Qs[s][..., 5] - h / dy * (Fy[s] - Fy[sym]) - Bt[s][..., 5] + h * nux / dx ** 2 * (Bt[sxp][..., 5] - 2 * Bt[s][..., 5] + Bt[sxm][..., 5])
    scale = np.abs(Bt[s][..., 5:7]).max()
    if p.bc_x != im.BC_PERIODIC:
        # the flux of the low boundary face lives at a ghost index, which bound_cond_array then
        # overwrites with a copy of the first interior face (ct_flux.cpp:129-130)
        assert np.array_equal(Fx[G:-G, G - 1], Fx[G:-G, G])
        res_y = res_y[:, 1:]
    if p.bc_y != im.BC_PERIODIC:
        assert np.array_equal(Fy[G - 1, G:-G], Fy[G, G:-G])
        res_x = res_x[1:, :]
    assert np.abs(res_y).max() <= 1e-12 * scale and np.abs(res_x).max() <= 1e-12 * scale
    # the diffusion the fluxes cannot carry is a real part of the update
    assert np.abs(h * nuy / dy ** 2 * (Bt[syp][..., 6] - 2 * Bt[s][..., 6] + Bt[sym][..., 6])).max() > 1e-6 * scale


def test_collect_mag_fluxes_uniform_state_is_the_electric_field():
    # uniform state: every difference vanishes, Fx = v_x B_y - v_y B_x = -E_z and Fy = +E_z
    p = _params(8, 6, 1.0, 1.0)
    rho, v, B = 1.2, np.array([0.4, -0.3, 0.2]), np.array([0.7, 0.5, -0.6])
    con = _uniform(p, [rho, *v, 1.5, *B])
    st = _stage(p, con)
    out, rc, _ = orc.sub(p, dict(adt=0.02, B_max_x=1.0, B_max_y=1.0), st, "mag_update_b")
    assert rc == 0
    Ez = -(v[0] * B[1] - v[1] * B[0])
    assert np.allclose(out["fx"][G:-G, G:-G], -Ez, rtol=0, atol=1e-14)
    assert np.allclose(out["fy"][G:-G, G:-G], Ez, rtol=0, atol=1e-14)


# ---------------------------------------------------------------- ShockDetector
def test_shock_detector_closed_forms():
    p = _params(12, 10, 1.2, 1.0, bc=im.BC_TRANSMISSIVE)
    p.shock_kappa1 = 1e-3
    NX, NY = p.nx + 2 * G, p.ny + 2 * G
    dx, dy = 1.2 / p.nx, 1.0 / p.ny
    hmin = min(dx, dy)
    rho, pr, B = 1.1, 0.8, np.array([0.3, -0.2, 0.1])
    cf = np.sqrt(p.gamma * pr / rho + (B * B).sum() / rho)
    kap = p.shock_kappa1 * cf
    # uniform state: eta = 0
    assert np.array_equal(orc.shock(p, _uniform(p, [rho, 0.3, -0.2, 0.0, pr, *B]))[G:-G, G:-G], np.zeros((p.ny, p.nx)))
    X = ((np.arange(NX) - G + 0.5) * dx)[None, :]
    Y = ((np.arange(NY) - G + 0.5) * dy)[:, None]
    # uniform compression u = -al x, v = -be y: div u = -(al + be) exactly, chi~ = clamp(h (al+be)/(k1 cf) - 1)
    for target in (0.25, 0.6, 1.7):
        s = (1 + target) * kap / hmin
        al, be = 0.6 * s, 0.4 * s
        pri = np.zeros((NY, NX, 8))
        pri[..., 0], pri[..., 4], pri[..., 5:8] = rho, pr, B
        pri[..., 1] = -al * X + 0 * Y
        pri[..., 2] = -be * Y + 0 * X
        eta = orc.shock(p, orc.pri2con(p, pri))
        assert np.allclose(eta[G:-G, G:-G], min(target, 1.0), rtol=0, atol=1e-12), target
    # a single compressive cell (u = -c at i0+1, +c at i0-1): chi~ > 0 there only, eta its 3x3 maximum
    i0, j0 = G + 5, G + 4
    c = 1.5 * kap * dx / hmin
    pri = np.zeros((NY, NX, 8))
    pri[..., 0], pri[..., 4], pri[..., 5:8] = rho, pr, B
    pri[j0, i0 + 1, 1], pri[j0, i0 - 1, 1] = -c, c
    eta = orc.shock(p, orc.pri2con(p, pri))
    chi = min(1.0, max(0.0, -(hmin * (-2 * c / (2 * dx)) + kap) / kap))
    want = np.zeros((NY, NX))
    want[j0 - 1:j0 + 2, i0 - 1:i0 + 2] = chi
    assert 0 < chi < 1
    assert np.allclose(eta[G:-G, G:-G], want[G:-G, G:-G], rtol=0, atol=1e-13)
    # c_f,min over the 3x3 neighbourhood: a lower pressure two cells away raises chi~ of the cell
    pri[j0 + 1, i0 + 1, 4] = 0.6 * pr      # inside the 3x3 of (i0, j0), whose own div u stays 0
    eta2 = orc.shock(p, orc.pri2con(p, pri))
    cf_low = np.sqrt(p.gamma * 0.6 * pr / rho + (B * B).sum() / rho)
    k_low = p.shock_kappa1 * cf_low
    chi2 = min(1.0, max(0.0, -(hmin * (-2 * c / (2 * dx)) + k_low) / k_low))
    assert chi < chi2 < 1
    assert eta2[j0, i0] == pytest.approx(chi2, rel=1e-13, abs=1e-15)


# ---------------------------------------------------------------- PreSubsystem_Pressure
def test_pressure_subsystem_two_pass_fourier_mode():
    # uniform ideal-gas Qe and u except m_x = M sin(theta_i): the enthalpy ratio is the constant H,
    # eta = 0, and the pressure operator is a - h^2 H Laplacian (a = 1/(gamma-1)), with eigenvalue
    # lam_q = a + h^2 H 4 sin^2(q d/2)/dx^2 on cos(q theta).  Pass 1: b1 = E - B^2/2 - m*^2/(2 rho)
    # - h H (m*_i+1 - m*_i-1)/(2 dx) -> p1; momentum m = m* - h (p1_i+1 - p1_i-1)/(2 dx).  Pass 2:
    # the kinetic term from m, the divergence term still from m* -> p2; E = a p2 + m^2/(2 rho) + B^2/2
    p = _params(16, 6, 1.0, 1.0)
    g = p.gamma
    a = 1 / (g - 1)
    rho, pe, B = 1.1, 0.8, np.array([0.2, 0.3, -0.1])
    st = _stage(p, _uniform(p, [rho, 0, 0, 0, pe, *B]))
    M, E0 = 0.05, 2.0
    NX = p.nx + 2 * G
    th = 2 * np.pi * (np.arange(NX) - G) / p.nx
    st["u_con"][..., 1] = M * np.sin(th)[None, :]
    st["u_con"][..., 4] = E0
    h, dx = 0.02, 1.0 / p.nx
    d = 2 * np.pi / p.nx
    H = g * pe / ((g - 1) * rho)
    lam = lambda q: a + h * h * H * 4 * np.sin(q * d / 2) ** 2 / dx ** 2
    n = np.arange(p.nx)
    t = 2 * np.pi * n / p.nx
    B2 = (B * B).sum()
    # pass 1 in Fourier modes
    C0 = E0 - 0.5 * B2 - M * M / (4 * rho)
    al1 = -(h * H * M * np.sin(d) / dx) / lam(1)
    be1 = (M * M / (4 * rho)) / lam(2)
    p1 = C0 / lam(0) + al1 * np.cos(t) + be1 * np.cos(2 * t)
    A1 = M + (h / dx) * al1 * np.sin(d)
    A2 = (h / dx) * be1 * np.sin(2 * d)
    m_new = A1 * np.sin(t) + A2 * np.sin(2 * t)
    # pass 2: m_new^2 = A1^2 (1 - cos 2t)/2 + A1 A2 (cos t - cos 3t) + A2^2 (1 - cos 4t)/2
    c0 = E0 - 0.5 * B2 - (A1 * A1 + A2 * A2) / (4 * rho)
    coef = {1: -(h * H * M * np.sin(d) / dx) - A1 * A2 / (2 * rho), 2: A1 * A1 / (4 * rho), 3: A1 * A2 / (2 * rho),
            4: A2 * A2 / (4 * rho)}
    p2 = c0 / lam(0) + sum(cq / lam(q) * np.cos(q * t) for q, cq in coef.items())
    out, rc, stats = orc.sub(p, dict(adt=h, vx_max=0.3, vy_max=0.2), st, "pre_subsystem_pressure")
    assert rc == 0
    ui = out["u_con"][G:-G, G:-G]
    assert np.allclose(ui[..., 1], m_new[None, :], rtol=0, atol=1e-13)
    assert np.allclose(ui[..., 2], 0.0, atol=1e-14)
    E_want = a * p2 + m_new ** 2 / (2 * rho) + 0.5 * B2
    assert np.allclose(ui[..., 4], E_want[None, :], rtol=0, atol=1e-12 * E0)
    # the two passes differ: E from p1 would be off by far more than the tolerance
    assert np.abs(a * (p2 - p1)).max() > 1e-6


# ---------------------------------------------------------------- stage algebra
def test_one_step_linear_response_stage_algebra_and_pressure_solve():
    # A small velocity perturbation m = rho0 eps (Mx cos th, My sin(th + 0.4)) of a uniform gas at rest
    # (B = 0, NoRecon).  To first order in eps one step is a linear map of each Fourier mode of
    # (rho, m_x, m_y, E), built here from the step's parts:
    #   explicit Rusanov fluxes at Qe:  d rho = -div m + D rho, d m = D m, d E = D E, with the central
    #     divergence i sin(k)/dx and the GLF dissipation D = -(a/dx)(1 - cos k) - (a/dy)(1 - cos l);
    #   pressure subsystem:  b = E^ - h H div m^, (a_g + h^2 H L5) p = b, m = m^ - h grad p, E = a_g p
    #     (a_g = 1/(gamma-1), H the enthalpy ratio, L5 the 5-point Laplacian symbol; the second pass
    #     changes only the kinetic term, which is second order);
    #   IMEX-RK2 stages:  stage 1 from Qe = Qs = u^n with h = b dt, Hs = (u1 - u^n)/(b dt); stage 2 from
    #     Qe = u^n + dt Hs/(2b), Qs = u^n + (1 - b) dt Hs with h = b dt; b = 1 - 1/sqrt(2).
    # The CFL maxima are even in eps, so (u(+eps) - u(-eps))/2 removes every even-order term.
    n, m = 16, 12
    p = _params(n, m, 1.0, 0.75)
    p.recon = im.RECON_NONE
    NX, NY = n + 2 * G, m + 2 * G
    dx, dy = 1.0 / n, 0.75 / m
    rho0, p0, Mx, My = 1.2, 0.9, 0.7, -0.4
    g = p.gamma
    ag = 1 / (g - 1)
    I = (np.arange(NX) - G)[None, :]
    J = (np.arange(NY) - G)[:, None]
    for kx, ky in [(2, 1), (0, 3), (5, -2)]:
        k, l = 2 * np.pi * kx / n, 2 * np.pi * ky / m
        th = k * I + l * J

        def run(eps):
            u = np.zeros((NY, NX, 8))
            u[..., 0], u[..., 4] = rho0, p0
            u[..., 1] = eps * Mx * np.cos(th)
            u[..., 2] = eps * My * np.sin(th + 0.4)
            o = orc.Oracle(p)
            o.set_state(u, np.zeros((NY, NX)), np.zeros((NY, NX)))
            assert o.step(1) == 0
            uc, _, _, _, sc = o.get_state()
            return uc, sc

        eps = 1e-3
        (u_p, sc), (u_m, sc_m) = run(eps), run(-eps)
        assert sc["dt"] == sc_m["dt"] and sc["a_max"] == sc_m["a_max"]
        d = (u_p - u_m) / (2 * eps)
        b = 1 - 1 / np.sqrt(2)
        dt, am = sc["dt"], sc["a_max"]
        H = g * p0 / ((g - 1) * rho0)
        Cx, Cy = 1j * np.sin(k) / dx, 1j * np.sin(l) / dy
        D = -(am / dx) * (1 - np.cos(k)) - (am / dy) * (1 - np.cos(l))
        L5 = 4 * np.sin(k / 2) ** 2 / dx ** 2 + 4 * np.sin(l / 2) ** 2 / dy ** 2

        def stage(Qe, Qs, h):
            uh = Qs + h * np.array([-(Cx * Qe[1] + Cy * Qe[2]) + D * Qe[0], D * Qe[1], D * Qe[2], D * Qe[3]])
            pp = (uh[3] - h * H * (Cx * uh[1] + Cy * uh[2])) / (ag + h * h * H * L5)
            return np.array([uh[0], uh[1] - h * Cx * pp, uh[2] - h * Cy * pp, ag * pp])

        U0 = rho0 * np.array([0, Mx, -1j * My * np.exp(0.4j), 0], complex)
        u1 = stage(U0, U0, b * dt)
        Hs = (u1 - U0) / (b * dt)
        u2 = stage(U0 + dt / (2 * b) * Hs, U0 + (1 - b) * dt * Hs, b * dt)
        ph = np.exp(1j * th)
        for c, slot in enumerate((0, 1, 2, 4)):
            want = (u2[c] * ph).real[G:-G, G:-G]
            got = d[G:-G, G:-G, slot]
            assert np.allclose(got, want, rtol=0, atol=1e-6 * rho0 * max(abs(Mx), abs(My))), (kx, ky, slot)
        assert np.abs(d[G:-G, G:-G, 4]).max() > 1e-3 or kx == 0
