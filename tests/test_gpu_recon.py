"""Reconstructions ENO3 (reconstruction.cpp:104-183) and NoRecon (:40-46), and ReconState PrimState
(:11-24, 272-281), on the GPU through the C ABI against the oracle.

Stage-level: the three explicit flux updates (RusanovC_Update, Pressure_flux_Update, Energy_flux_Update)
and the plain Staggered_CT, whose face EMFs come from the reconstruction, on the oracle's working set
captured right before them.  Whole runs: conserved fields after N steps (relative L2 per field <= 1e-9,
div B <= 1e-12) for every pair below, with both formulas and the RK EOS; the mixed form within 1e-6
(FP32 ENO3 of the conservative state; PrimState keeps its fluxes in FP64); Diagnose_CT_EMF.
"""
import numpy as np
import pytest

import imex_amd as im
import oracle as orc
from helpers import rel_l2, field_errors, make_case

pytestmark = pytest.mark.gpu

RECONS = [(im.RECON_ENO3, im.RECONSTATE_CONS), (im.RECON_NONE, im.RECONSTATE_CONS),
          (im.RECON_MINMOD, im.RECONSTATE_PRIM), (im.RECON_VANLEER, im.RECONSTATE_PRIM),
          (im.RECON_ENO3, im.RECONSTATE_PRIM), (im.RECON_WENO5, im.RECONSTATE_PRIM)]
RID = ["eno3", "none", "minmod-prim", "vanleer-prim", "eno3-prim", "weno5-prim"]
CASES = [("vortex", 40, 24, None), ("ot", 32, 32, None), ("blast", 24, 40, 10.0)]
# sub-step name, capture point, formula, CT mode, compared slots, tolerance
SUBS = {"explicit_conv": (0, im.FORMULA_PRESSURE, im.CT_CONSISTENT, [0, 1, 2, 3], 1e-14),
        "explicit_press": (2, im.FORMULA_PRESSURE, im.CT_CONSISTENT, [1, 2, 3, 4], 1e-14),
        "explicit_energy": (2, im.FORMULA_ENERGY, im.CT_CONSISTENT, [1, 2, 3, 4], 1e-14),
        "ct_update": (4, im.FORMULA_PRESSURE, im.CT_PLAIN, [5, 6], 1e-13)}
_snap = {}


def snapshot(case, n, m, amp, recon, state, formula, ct, point):
    key = (case, n, m, amp, recon, state, formula, ct, point)
    if key not in _snap:
        p, u, bx, by = make_case(case, n, m, amp=amp, recon=recon, recon_state=state, formula=formula, ct_mode=ct)
        o = orc.Oracle(p)
        o.set_state(u, bx, by)
        o.step(1)
        st, sc, _ = o.step_capture(point)
        _snap[key] = (p, st, sc)
    return _snap[key]


@pytest.mark.parametrize("recon,state", RECONS, ids=RID)
@pytest.mark.parametrize("case,n,m,amp", CASES)
@pytest.mark.parametrize("name", list(SUBS))
def test_recon_stage_parity(recon, state, case, n, m, amp, name):
    point, formula, ct, slots, tol = SUBS[name]
    print(f"[recon] recon={recon} state={state} {case} {n}x{m} {name}", flush=True)
    p, st, sc = snapshot(case, n, m, amp, recon, state, formula, ct, point)
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
    if name == "ct_update":
        for f in ("bx", "by"):
            e = rel_l2(out[f], ref[f])
            assert e <= 1e-13, f"{f}: {e:.3e}"


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
        ct = s.diagnose_ct_emf()
    finally:
        s.close()
    return ref, got, stats, ct


@pytest.mark.parametrize("recon,state", RECONS, ids=RID)
@pytest.mark.parametrize("case,n,m,steps,amp,over", [
    ("vortex", 64, 64, 8, None, {}), ("ot", 64, 48, 5, None, {}), ("blast", 48, 48, 4, 10.0, {}),
    ("ot", 48, 48, 4, None, dict(formula=im.FORMULA_ENERGY)), ("vortex", 32, 32, 3, None, dict(eos=im.EOS_RK)),
    ("vortex", 40, 24, 3, None, dict(ct_mode=im.CT_PLAIN, schedule=im.SCHED_HOST))])
def test_recon_full_run_parity(recon, state, case, n, m, steps, amp, over):
    print(f"[recon run] recon={recon} state={state} {case} {n}x{m} steps={steps} {over}", flush=True)
    p, u, bx, by = make_case(case, n, m, amp=amp, recon=recon, recon_state=state, **over)
    (uc_r, _, bx_r, by_r, sc_r), (uc, _, bxg, byg, sc), stats, ct = run_pair(p, u, bx, by, steps)
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-9, errs
    assert abs(sc["t"] - sc_r["t"]) <= 1e-12 * abs(sc_r["t"])
    assert stats["div_b"] <= 1e-12, stats["div_b"]
    ct_ref = orc.ct_emf_diag(p, uc_r)
    assert np.allclose(ct, ct_ref, rtol=1e-8, atol=1e-12 * np.abs(ct_ref).max())


@pytest.mark.parametrize("recon,state,flux", [(im.RECON_ENO3, im.RECONSTATE_CONS, 1), (im.RECON_ENO3, im.RECONSTATE_CONS, 0),
                                              (im.RECON_WENO5, im.RECONSTATE_PRIM, 1)])
def test_recon_mixed_form_within_1e6(recon, state, flux):
    """ENO3 with the opt-in FP32 fluxes chooses its stencil in FP64 and evaluates it in FP32."""
    print(f"[recon mixed] recon={recon} state={state} mixed_flux={flux} ot 64x64 steps=6", flush=True)
    p, u, bx, by = make_case("ot", 64, 64, recon=recon, recon_state=state)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    o.step(6)
    uc_r = o.get_state()[0]
    p.precision = im.PREC_MIXED
    p.mixed_flux = flux
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        s.step(6)
        uc = s.get_state()[0]
    finally:
        s.close()
    errs = field_errors(uc, uc_r)
    print("  rel L2 per field:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-6, errs


def test_adjoint_rejects_other_reconstructions():
    for recon, state in [(im.RECON_ENO3, im.RECONSTATE_CONS), (im.RECON_MINMOD, im.RECONSTATE_PRIM)]:
        p, u, bx, by = make_case("vortex", 24, recon=recon, recon_state=state)
        s = im.Solver(p)
        try:
            s.set_state(u, bx, by)
            with pytest.raises(im.ImexError):
                s.adjoint_forward(1, 1)
        finally:
            s.close()
