"""Output files and per-step diagnostics (src/output.cpp:127-326, src/main.cpp:562-720,
src/constrained_transport.cpp:165-222) — CPU only.

The library's writers are host code over host Grids; here they are compared byte for byte with the
oracle's restatement on the same arrays, and the formats are pinned against the reference's stream
conventions (default float format, precision 6; VTK fields width 24 with 16 significant digits,
looping over padded indices 0..n-1).  The oracle's diagnostics are pinned by brute force in numpy and,
for Diagnose_CT_EMF, by the closed form on a uniform state (corner EMF parts 2E, 0, -E, total E, no
face injection).
"""
import filecmp

import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import make_case, make_case_a, G


def evolved(case="ot", n=16, m=12, steps=2, **over):
    p, u, bx, by = make_case(case, n, m, **over)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(steps)
    uc, up, bx1, by1, sc = o.get_state()
    return p, uc, up, bx1, by1, sc


@pytest.mark.parametrize("case", ["ot", "blast", "vortex"])
def test_vtk_writer_bytes_match_oracle(tmp_path, case):
    p, uc, up, bx, by, sc = evolved(case)
    a, b = tmp_path / "lib.vtk", tmp_path / "orc.vtk"
    im.write_vtk_host(p, uc, up, sc["t"], 7, a)
    orc.write_vtk(p, uc, up, sc["t"], 7, b)
    assert filecmp.cmp(a, b, shallow=False)


def test_vtk_format(tmp_path):
    p, uc, up, bx, by, sc = evolved("ot", 10, 6)
    f = tmp_path / "x.vtk"
    im.write_vtk_host(p, uc, up, sc["t"], 42, f)
    lines = f.read_text().split("\n")
    assert lines[:6] == ["# vtk DataFile Version 3.0", "vtk output", "ASCII", "DATASET RECTILINEAR_GRID",
                         "FIELD FieldData 2", "CYCLE 1 1 int"]
    assert lines[6] == "42" and lines[7] == "TIME 1 1 double" and lines[8] == "%g" % sc["t"]
    assert lines[9] == "DIMENSIONS 11 7 1" and lines[10] == "X_COORDINATES 11 double"
    dx = 2 * np.pi / 10
    assert lines[11].split() == ["%g" % ((i + 0.5) * dx) for i in range(10)]      # n centres under a count of n + 1
    k = lines.index("FIELD FieldData 13")
    names = ["rho", "mom_X", "mom_Y", "mom_Z", "U", "B_X", "B_Y", "B_Z", "velX", "velY", "velZ", "p", "B"]
    pos = k + 1
    for var, name in enumerate(names):
        assert lines[pos] == f"{name} 1 60 double"
        rows = [np.array(lines[pos + 1 + j].split(), dtype=float) for j in range(6)]
        assert all(len(lines[pos + 1 + j]) == 10 * 25 for j in range(6))         # setw(24) and a space per value
        got = np.array(rows)
        # padded indices 0..n-1, as the reference's loops run
        if var < 8:
            want = uc[:6, :10, var]
        elif var < 12:
            want = up[:6, :10, var - 7]
        else:
            want = np.sqrt((uc[:6, :10, 5:8] ** 2).sum(-1))
        assert np.allclose(got, want, rtol=1e-15, atol=0), name
        assert lines[pos + 7] == ""
        pos += 8
    assert lines[pos] == "POINT_DATA 77"


@pytest.mark.parametrize("eos", [im.EOS_IDEAL, im.EOS_RK])
def test_field_files_bytes_match_oracle(tmp_path, eos):
    p, uc, up, bx, by, sc = evolved("vortex", 12, 10, eos=eos)
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    im.write_fields_host(p, uc, up, tmp_path / "a")
    orc.write_fields(p, uc, up, tmp_path / "b")
    names = ["RK_rho.dat", "RK_vx.dat", "RK_vy.dat", "RK_E.dat", "RK_p.dat", "RK_Bx.dat", "RK_By.dat", "RK_Bz.dat",
             "RK_v_mag.dat"] + (["RK_T.dat"] if eos == im.EOS_RK else [])
    assert sorted(x.name for x in (tmp_path / "a").iterdir()) == sorted(names)
    for nm in names:
        assert filecmp.cmp(tmp_path / "a" / nm, tmp_path / "b" / nm, shallow=False), nm
    rows = [l.split() for l in (tmp_path / "a" / "RK_p.dat").read_text().splitlines()]
    dx, dy = 10.0 / 12, 10.0 / 10
    # i outer, j inner; "x y value" with %g formatting
    assert rows[0] == ["%g" % (-5 + 0.5 * dx), "%g" % (-5 + 0.5 * dy), "%g" % up[G, G, 4]]
    assert rows[1] == ["%g" % (-5 + 0.5 * dx), "%g" % (-5 + 1.5 * dy), "%g" % up[G + 1, G, 4]]
    assert len(rows) == 120


def test_history_and_output_path(tmp_path):
    f = tmp_path / "RK_div_error.dat"
    steps, vals = [1, 2, 10], [1.5e-15, 2.25e-13, 0.125]
    im.write_history(f, steps, vals)
    assert f.read_text() == "".join("%d %g\n" % (s, v) for s, v in zip(steps, vals))
    p, _ = im.case_of("ot64")
    assert im.output_path(p, "out", "Orszag_Tang") == \
        "out/Ideal Gas/PreservingON/Orszag_Tang/PressureFormula/StaggerCT/definition/64x64_IMEXdt1_RK2_Order4_CFL080"
    for f_, mf, eos, floor in [(im.FORMULA_ENERGY, im.MAGFIELD_POTENTIAL_A, im.EOS_RK, 0),
                               (im.FORMULA_PRESSURE, im.MAGFIELD_NODIVFREE, im.EOS_IDEAL, 1)]:
        q, _ = im.case_of("vortex40")
        q.formula, q.magfield, q.eos, q.floor_on, q.cfl = f_, mf, eos, floor, 0.45
        assert im.output_path(q, "r", "MHD_vortex") == orc.output_path(q, "r", "MHD_vortex")
    q.formula, q.magfield = im.FORMULA_ENERGY, im.MAGFIELD_POTENTIAL_A
    assert im.output_path(q, "r", "MHD_vortex").endswith("/MHD_vortex/EnergyFormula/Potential//40x40_IMEXdt1_RK2_Order2_CFL045")


@pytest.mark.parametrize("case,mf", [("ot", im.MAGFIELD_STAGGER), ("blast", im.MAGFIELD_STAGGER),
                                     ("vortex", im.MAGFIELD_POTENTIAL_A)])
def test_oracle_diagnostics_brute_force(case, mf):
    p, u, bx, by, A = make_case_a(case, 16, 12, magfield=mf)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.set_potential(A)
    o.step(2)
    uc, up, bx1, by1, sc = o.get_state()
    d = orc.diagnose(p, sc, uc, bx1, by1)
    dx, dy = (p.x1 - p.x0) / p.nx, (p.y1 - p.y0) / p.ny
    q = uc[G:-G, G:-G]
    k = 0.5 * (q[..., 1:4] ** 2).sum(-1) / q[..., 0]
    mag = 0.5 * (q[..., 5:8] ** 2).sum(-1)
    e = q[..., 4] - k - mag
    assert d["mass"] == pytest.approx(q[..., 0].sum() * dx * dy, rel=1e-13)
    assert d["intE"] == pytest.approx(e.sum() * dx * dy, rel=1e-12)
    assert d["magE_max"] == mag.max() and d["emin"] == e.min()
    im0, jm0 = np.unravel_index(np.argmin(e.T), e.T.shape)               # first minimum scanning i outer, j inner
    assert d["emin_x"] == pytest.approx(p.x0 + (im0 + 0.5) * dx, rel=1e-14)
    assert d["emin_y"] == pytest.approx(p.y0 + (jm0 + 0.5) * dy, rel=1e-14)
    cdiv = (uc[G:-G, G + 1:-G + 1, 5] - uc[G:-G, G - 1:-G - 1, 5]) / (2 * dx) + \
           (uc[G + 1:-G + 1, G:-G, 6] - uc[G - 1:-G - 1, G:-G, 6]) / (2 * dy)
    assert d["B_inf"] == pytest.approx(np.abs(cdiv).max(), rel=1e-14)
    want_div = orc.div_b(p, bx1, by1) if mf == im.MAGFIELD_STAGGER else orc.div_b_cell(p, uc)
    assert d["div_error"] == want_div
    assert d["dt_comp"] == pytest.approx(sc["ch"] / sc["a_max"], rel=1e-14)
    assert (d["timestep"], d["t"], d["dt"]) == (sc["timestep"], sc["t"], sc["dt"])


def test_oracle_ct_emf_diagnostic_uniform_state():
    p = im.default_params()
    p.nx, p.ny = 12, 10
    p.x0, p.x1, p.y0, p.y1 = 0.0, 1.2, 0.0, 0.5
    NX, NY = p.nx + 2 * G, p.ny + 2 * G
    pri = np.array([1.3, 0.4, -0.25, 0.1, 0.9, 0.6, 0.35, -0.2])
    uc = orc.pri2con(p, np.tile(pri, (NY, NX, 1)))
    E0 = pri[2] * pri[5] - pri[1] * pri[6]
    out = orc.ct_emf_diag(p, uc)
    area = 1.2 * 0.5
    want = np.array([[2 * abs(E0) * np.sqrt(area), 2 * abs(E0), 0, 0],
                     [0, 0, 0, 0],
                     [abs(E0) * np.sqrt(area), abs(E0), 0, 0],
                     [abs(E0) * np.sqrt(area), abs(E0), 0, 0]])
    assert np.allclose(out, want, rtol=1e-13, atol=1e-12)
