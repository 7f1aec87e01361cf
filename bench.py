"""Throughput of the IMEX-MHD step on one MI355X.

Workload: Orszag-Tang on periodic [0,2pi]^2 at 1024x1024, gamma 5/3, WENO5,
StaggerCT (EMF-consistent; --magfield: UnStaggerCT, PotentialA or NoDivFree),
PressureFormula (--formula energy: EnergyFormula),
floor_on with p_floor 1e-8, IMEX-RK2, CFL 0.8 — synthetic initial data
generated on the host.

A "step" is one full IMEX-RK2 step (two stages: explicit fluxes, magnetic GMRES,
EMF collection, two nested-Newton pressure solves, CT) with all fields resident in
HBM.  Prints one JSON line: Mcell-steps/s, the live per-kernel HIP-event roofline
of the dominant kernel class and the FP64 CPU oracle (faithful mode, 1 core) timed
on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


# IMEX_RECON_* values of the C ABI
RECON_NAMES = {"none": 0, "minmod": 1, "vanleer": 2, "eno3": 3, "weno5": 4}
RECON_LABEL = {0: "NoRecon", 1: "Minmod", 2: "VanLeer", 3: "ENO3", 4: "WENO5"}
WORKLOAD_NAME = {"ot": "orszag_tang", "vortex": "mhd_vortex", "blast": "magnetic_blast", "jet": "mhd_jet"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def uses_potential(p):
    import imex_amd as im
    return p.magfield in (im.MAGFIELD_UNSTAGGER, im.MAGFIELD_POTENTIAL_A)


def apply_over(p, formula=None, magfield=None, over=None):
    if formula is not None:
        p.formula = formula
    if magfield is not None:
        p.magfield = magfield
    for k, v in (over or {}).items():
        setattr(p, k, v)


def run_gpu(name, steps, warmup, precision, schedule, tols=None, formula=None, magfield=None, over=None):
    import imex_amd as im
    p, kw = im.case_of(name)
    p.precision = precision
    p.schedule = schedule
    apply_over(p, formula, magfield, over)
    if tols:
        p.mixed_inner_tol, p.mixed_outer_tol = tols
    u, bx, by = im.init_case(p, kw["case"], kw.get("amp", 1.0), kw.get("u0", 0.0), kw.get("v0", 0.0))
    s = im.Solver(p)
    s.set_state(u, bx, by)
    if uses_potential(p):
        s.set_potential(im.init_potential(p, kw["case"], kw.get("amp", 1.0)))
    tune = s.autotune(im.TUNE_EXPLICIT | im.TUNE_KRYLOV)       # launch shapes for this grid
    log(f"[bench] {name} autotune: {tune}")
    log(f"[bench] {name} warmup {warmup} steps")
    if warmup:
        s.step(warmup)
    state = s.get_state()
    A_state = s.get_potential() if uses_potential(p) else None
    s.reset_stats()
    s.sync()
    log(f"[bench] timing {steps} steps")
    t0 = time.perf_counter()
    s.step_async(steps)
    s.sync()
    t1 = time.perf_counter()
    stats = s.stats()
    final = s.get_state()
    # Per-kernel roofline: the same steps again, from the same post-warmup state, with every
    # launch bracketed by HIP events on the solver's stream.  Event timing needs launches queued
    # one by one, so this pass runs the budgeted schedule without the step graph (its first step,
    # before any iteration budget is known, is host scheduled); its wall time is reported apart.
    log(f"[bench] per-kernel event timing over the same {steps} steps")
    uc, up, bxs, bys, sc = state
    s.set_state(up, bxs, bys, sc["t"], sc["timestep"], sc["prev_dt"])
    if A_state is not None:
        s.set_potential(A_state)
    s.profile_enable(True)
    s.sync()
    t2 = time.perf_counter()
    s.step_async(steps)
    s.sync()
    t3 = time.perf_counter()
    prof = s.profile_read()
    s.profile_enable(False)
    s.close()
    return dict(p=p, elapsed=t1 - t0, prof_elapsed=t3 - t2, prof=prof, stats=stats, state=state, final=final, tune=tune,
                potential=A_state)


def run_pair(name, steps, warmup, schedule, formula=None, magfield=None, forms=("fp64", "mixed"), over=None):
    """FP64 and mixed forms of one workload (no per-kernel pass): throughput and the mixed form's error."""
    import numpy as np
    import imex_amd as im
    out = {}
    finals = {}
    for f, prec in (("fp64", im.PREC_FP64), ("mixed", im.PREC_MIXED)):
        if f not in forms:
            continue
        p, kw = im.case_of(name)
        p.precision = prec
        p.schedule = schedule
        apply_over(p, formula, magfield, over)
        u, bx, by = im.init_case(p, kw["case"], kw.get("amp", 1.0), kw.get("u0", 0.0), kw.get("v0", 0.0))
        s = im.Solver(p)
        try:
            s.set_state(u, bx, by)
            if uses_potential(p):
                s.set_potential(im.init_potential(p, kw["case"], kw.get("amp", 1.0)))
            s.autotune(im.TUNE_EXPLICIT | im.TUNE_KRYLOV)
            log(f"[bench] {name} {f}: warmup {warmup}, timing {steps} steps")
            if warmup:
                s.step(warmup)
            s.reset_stats()
            s.sync()
            t0 = time.perf_counter()
            s.step_async(steps)
            s.sync()
            t1 = time.perf_counter()
            st = s.stats()
            finals[f] = s.get_state()[0]
        finally:
            s.close()
        out[f] = {"value": round(p.nx * p.ny * steps / (t1 - t0) / 1e6, 3), "ms_per_step": round((t1 - t0) / steps * 1e3, 3),
                  "gmres": {k: st[k] for k in ("mag_gmres_iters", "mag_solves", "p_gmres_iters", "p_solves", "newton_iters",
                                               "mixed_refinements")},
                  "min_p": st["min_p"], "div_b": st["div_b"]}
        log(f"[bench] {name} {f}: {out[f]['value']} Mcell-steps/s")
    if len(finals) < 2:
        return out
    a, b = finals["mixed"], finals["fp64"]
    g = im.GC
    errs = [float(np.linalg.norm(a[g:-g, g:-g, k] - b[g:-g, g:-g, k]) / max(np.linalg.norm(b[g:-g, g:-g, k]), 1e-300))
            for k in range(b.shape[-1])]
    out["mixed_rel_l2_vs_fp64_max"] = float(f"{max(errs):.3e}")
    out["steps_compared"] = warmup + steps
    out["speedup_mixed_over_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 3)
    return out


def run_diag_cost(name, steps, warmup, schedule):
    """Cost of the per-step diagnostics (device ring), one Diagnose_CT_EMF and the output files at this size."""
    import tempfile
    import imex_amd as im
    p, kw = im.case_of(name)
    p.schedule = schedule
    u, bx, by = im.init_case(p, kw["case"], kw.get("amp", 1.0), kw.get("u0", 0.0), kw.get("v0", 0.0))
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        s.step(warmup)
        uc, up, bxs, bys, sc = s.get_state()
        times = {}
        for label, cap in (("off", 0), ("on", 64)):     # the same steps from the same state, without and with
            s.set_state(up, bxs, bys, sc["t"], sc["timestep"], sc["prev_dt"])
            s.diag_enable(cap)
            s.step(warmup)                               # budgets relearned after set_state
            s.sync()
            t0 = time.perf_counter()
            s.step_async(steps)
            s.sync()
            times[label] = (time.perf_counter() - t0) / steps * 1e3
        recs = s.diag_read()
        t0 = time.perf_counter()
        s.diagnose_ct_emf()
        t_ct = (time.perf_counter() - t0) * 1e3
        with tempfile.TemporaryDirectory() as tdir:
            t0 = time.perf_counter()
            s.write_vtk(os.path.join(tdir, "plt.vtk"), warmup + 2 * steps)
            t_vtk = (time.perf_counter() - t0) * 1e3
            t0 = time.perf_counter()
            s.write_fields(tdir)
            t_dat = (time.perf_counter() - t0) * 1e3
    finally:
        s.close()
    return {"ms_per_step_without": round(times["off"], 3), "ms_per_step_with_recording": round(times["on"], 3),
            "records": len(recs), "last_div_error": recs[-1]["div_error"] if recs else None,
            "ct_emf_diagnose_ms": round(t_ct, 2), "write_vtk_ms": round(t_vtk, 1), "write_fields_ms": round(t_dat, 1)}


BATCH_SIZES = (64, 96, 128, 192, 256, 384, 512)


def run_batch(n_inst, steps, warmup, eos, seed, schedule):
    """Independent MHD vortices of uneven size in one batch context; aggregate throughput."""
    import numpy as np
    import imex_amd as im
    rng = np.random.default_rng(seed)
    sizes = [(int(rng.choice(BATCH_SIZES)), int(rng.choice(BATCH_SIZES))) for _ in range(n_inst)]
    amps = rng.uniform(0.5, 2.0, size=(n_inst, 3))
    p = im.vortex_params(64, eos=eos)
    p.schedule = schedule
    cells = sum(a * b for a, b in sizes)
    log(f"[bench] batch: {n_inst} vortices, eos {'RK' if eos == im.EOS_RK else 'ideal'}, {cells / 1e6:.2f} Mcells")
    t_setup = time.perf_counter()
    b = im.Batch(p, sizes)
    try:
        for i, (nx, ny) in enumerate(sizes):
            q = b.params_of(i)
            u, bx, by = im.init_case(q, im.CASE_VORTEX, amps[i, 0], amps[i, 1], amps[i, 2], 0)
            b.set_instance(i, u, bx, by)
        tune = b.autotune(im.TUNE_EXPLICIT | im.TUNE_KRYLOV)
        log(f"[bench] batch autotune: {tune}")
        if warmup:
            b.step(warmup)
        b.reset_stats()
        b.sync()
        t_setup = time.perf_counter() - t_setup
        t0 = time.perf_counter()
        b.step_async(steps)
        b.sync()
        t1 = time.perf_counter()
        st = b.stats()
        dev_gb = b.L.imex_device_bytes(b.h) / 1e9
    finally:
        b.close()
    value = cells * steps / (t1 - t0) / 1e6
    log(f"[bench] batch eos {eos}: {value:.1f} Mcell-steps/s aggregate; {st}")
    return {"value": round(value, 3), "ms_per_step": round((t1 - t0) / steps * 1e3, 3), "instances": n_inst,
            "cells": cells, "seed": seed, "eos": "redlich_kwong" if eos == im.EOS_RK else "ideal",
            "autotuned": {"explicit_tile": list(tune["explicit_tile"]), "krylov_blocks": tune["krylov_blocks"],
                          "explicit_us": [round(x, 1) for x in tune["explicit_us"]],
                          "krylov_us": [round(x, 1) for x in tune["krylov_us"]]},
            "budget_steps": st["budget_steps"], "budget_reruns": st["budget_reruns"], "graph_builds": st["graph_builds"],
            "max_div_b": st["div_b"], "device_gb": round(dev_gb, 2), "setup_s": round(t_setup, 1)}


def run_adjoint(n, steps, every, ndir, seed, eps=1e-6, noise=1e-3):
    """Discrete adjoint of `steps` FP64 steps of the n^2 MHD vortex, J = total kinetic energy at the
    final step, against central differences along `ndir` seeded directions."""
    import numpy as np
    import imex_amd as im
    G = im.GC
    p, kw = im.case_of(f"vortex{n}")
    u, bx, by = im.init_case(p, kw["case"], kw["amp"], kw["u0"], kw["v0"])
    r = np.random.default_rng(seed + 1000)
    u[G:-G, G:-G, :] *= 1 + noise * r.standard_normal(u[G:-G, G:-G, :].shape)
    bx[G:-G, G:-G] *= 1 + noise * r.standard_normal(bx[G:-G, G:-G].shape)
    by[G:-G, G:-G] *= 1 + noise * r.standard_normal(by[G:-G, G:-G].shape)
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
        J_of(u, bx, by)                          # warm-up (graph capture)
        s.set_state(u, bx, by)
        s.sync()
        t0 = time.perf_counter()
        s.step_async(steps)
        s.sync()
        t_fwd = time.perf_counter() - t0
        s.set_state(u, bx, by)
        s.sync()
        t0 = time.perf_counter()
        s.adjoint_forward(steps, every)
        s.sync()
        t1 = time.perf_counter()
        J, g = kinetic(s.get_state()[0])
        t2 = time.perf_counter()
        au, abx, aby, info = s.adjoint_backward(g)
        t3 = time.perf_counter()
        log(f"[bench] adjoint {n}^2, {steps} steps: forward {t_fwd * 1e3:.1f} ms, checkpointed forward "
            f"{(t1 - t0) * 1e3:.1f} ms, backward {(t3 - t2) * 1e3:.1f} ms; {info}")
        rng = np.random.default_rng(seed)
        ad, fd = [], []
        for k in range(ndir):
            du = np.zeros_like(u); dbx = np.zeros_like(bx); dby = np.zeros_like(by)
            du[G:-G, G:-G, :] = rng.standard_normal(du[G:-G, G:-G, :].shape) * 1e-2
            dbx[G:-G, G:-G] = rng.standard_normal(dbx[G:-G, G:-G].shape) * 1e-2
            dby[G:-G, G:-G] = rng.standard_normal(dby[G:-G, G:-G].shape) * 1e-2
            f = (J_of(u + eps * du, bx + eps * dbx, by + eps * dby) - J_of(u - eps * du, bx - eps * dbx, by - eps * dby)) / (2 * eps)
            a = float((au * du).sum() + (abx * dbx).sum() + (aby * dby).sum())
            ad.append(a); fd.append(f)
            log(f"[bench] adjoint direction {k}: {a:.12e} fd {f:.12e} rel {abs(a - f) / abs(f):.2e}")
    finally:
        s.close()
    ad, fd = np.array(ad), np.array(fd)
    rel = np.abs(ad - fd) / np.abs(fd)
    return {"workload": f"MHD vortex {n}x{n} (seeded 1e-3 relative perturbation), FP64, {steps} steps, "
                        f"J = total kinetic energy at the final step, checkpoint every {every} steps "
                        f"(tapes of all steps kept in HBM: {info['recomputed_steps'] == 0})",
            "directions": ndir, "fd_eps": eps, "direction_seed": seed,
            "rel_error_median": float(f"{np.median(rel):.3e}"), "rel_error_max": float(f"{rel.max():.3e}"),
            "rel_error_all_directions": float(f"{np.linalg.norm(ad - fd) / np.linalg.norm(fd):.3e}"),
            "forward_ms": round(t_fwd * 1e3, 2), "checkpointed_forward_ms": round((t1 - t0) * 1e3, 2),
            "backward_ms": round((t3 - t2) * 1e3, 2),
            "adjoint_over_forward": round(((t1 - t0) + (t3 - t2)) / t_fwd, 2),
            "transposed_solves": info["transposed_solves"], "transposed_gmres_iters": info["mag_iters"] + info["p_iters"],
            "worst_solve_residual": info["worst_solve_residual"]}


def cpu_baseline(name, state, budget_steps=1, formula=None, magfield=None, potential=None, over=None, threads=1,
                 faithful=1):
    """FP64 oracle from the GPU's post-warmup state: faithful mode (per-apply recomputation like the reference)
    on one core, or the hoisted mode on `threads` cores."""
    import imex_amd as im
    import oracle as orc
    p, kw = im.case_of(name)
    apply_over(p, formula, magfield, over)
    uc, up, bx, by, sc = state
    used = orc.set_threads(threads)
    try:
        o = orc.Oracle(p, faithful)
        o.set_state(up, bx, by, sc["t"], sc["timestep"], sc["prev_dt"])
        if potential is not None:
            o.set_potential(potential)
        t0 = time.perf_counter()
        o.step(budget_steps)
        t1 = time.perf_counter()
    finally:
        orc.set_threads(os.cpu_count() or 1)
    return p.nx * p.ny * budget_steps / (t1 - t0) / 1e6, t1 - t0, used


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_accuracy(name, steps, formula=None, magfield=None, over=None):
    """Relative L2 per conserved field of the FP64 and mixed GPU forms against the oracle (all host cores,
    hoisted mode) after `steps` steps from the initial data."""
    import numpy as np
    import imex_amd as im
    import oracle as orc
    p, kw = im.case_of(name)
    apply_over(p, formula, magfield, over)
    u, bx, by = im.init_case(p, kw["case"], kw.get("amp", 1.0), kw.get("u0", 0.0), kw.get("v0", 0.0))
    orc.set_threads(os.cpu_count() or 1)
    o = orc.Oracle(p)
    o.set_state(u, bx, by)
    t0 = time.perf_counter()
    o.step(steps)
    t_orc = time.perf_counter() - t0
    ref = o.get_state()[0]
    g = im.GC
    out = {"steps": steps, "oracle_s": round(t_orc, 1)}
    for f, prec in (("fp64", im.PREC_FP64), ("mixed", im.PREC_MIXED)):
        q, _ = im.case_of(name)
        apply_over(q, formula, magfield, over)
        q.precision = prec
        s = im.Solver(q)
        try:
            s.set_state(u, bx, by)
            s.step(steps)
            uc = s.get_state()[0]
        finally:
            s.close()
        errs = [float(np.linalg.norm(uc[g:-g, g:-g, k] - ref[g:-g, g:-g, k]) / max(np.linalg.norm(ref[g:-g, g:-g, k]), 1e-300))
                for k in range(8)]
        out[f"{f}_rel_l2_vs_oracle_max"] = float(f"{max(errs):.3e}")
        out[f"{f}_rel_l2_vs_oracle_per_field"] = [float(f"{e:.3e}") for e in errs]
    return out


def copy_calibration(mib):
    """One read stream and one write stream of known size: the scale for rocprof's FETCH_SIZE/WRITE_SIZE."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": n * 8, "launches": 3}), flush=True)


def stream_ceiling():
    """Practical streaming rates on this GPU, the yardstick beside the 8 TB/s peak: torch elementwise kernels
    with one or two input streams and one output stream (negation, addition) over vectors of 1 Mi doubles
    (8 MB each, the size of one field at 1024^2; such working sets can stay in the 256 MB last-level cache
    across launches) and of 64 Mi doubles (512 MB, beyond it).  Each rate is bytes moved over the event-timed
    average of 50 back-to-back launches after 5 untimed ones."""
    import torch
    out = {}
    for n, tag in ((1 << 20, "8MB_vectors"), (64 << 20, "512MB_vectors")):
        a = torch.rand(n, dtype=torch.float64, device="cuda")
        b = torch.rand(n, dtype=torch.float64, device="cuda")
        c = torch.empty_like(a)
        cases = {"copy_1r1w": (lambda: torch.neg(a, out=c), 2), "add_2r1w": (lambda: torch.add(a, b, out=c), 3)}
        for name, (fn, streams) in cases.items():
            for _ in range(5):
                fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(50):
                fn()
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / 50
            out[f"{tag}_{name}"] = {"us": round(us, 2), "GB/s": round(streams * n * 8 / us / 1e3, 1)}
        del a, b, c
    print(json.dumps({"stream_ceiling": out}), flush=True)


KERNEL_OF_CLASS = {   # (class, form) or (variant, class, form) -> kernel symbol in the committed rocprof summary
    ("potential_a", "mag_arnoldi_apply", "fp64"): "k_arnoldi_apply<3, double, double>",
    ("potential_a", "mag_arnoldi_apply", "mixed"): "k_arnoldi_apply<3, float, float>",
    ("energy", "explicit_energy", "mixed"): "k_explicit<{tx}, {ty}, double, false, false>",   # the energy form keeps FP64 fluxes
    ("mag_arnoldi_apply", "fp64"): "k_arnoldi_apply<0, double, double>",
    ("mag_arnoldi_apply", "mixed"): "k_arnoldi_apply<0, float, float>",
    ("p_arnoldi_apply", "fp64"): "k_arnoldi_apply<1, double, double>",
    ("p_arnoldi_apply", "mixed"): "k_arnoldi_apply<1, double, double>",
    ("e_arnoldi_apply", "fp64"): "k_arnoldi_apply<2, double, double>",
    ("e_arnoldi_apply", "mixed"): "k_arnoldi_apply<2, double, double>",
    ("explicit_energy", "fp64"): "k_explicit<{tx}, {ty}, double, false, false>",
    ("explicit_energy", "mixed"): "k_explicit<{tx}, {ty}, float, false, false>",
}
TRAFFIC_SUMMARY = {"pressure": os.path.join(ROOT, "profiles", "imex_step_ot1024_kernels.md"),
                   "energy": os.path.join(ROOT, "profiles", "imex_step_ot1024_energy_kernels.md"),
                   "potential_a": os.path.join(ROOT, "profiles", "imex_step_ot1024_potential_kernels.md")}


def kernel_of(cls, form, variant, tile=(32, 16)):
    """Kernel symbol behind a profiler class; explicit kernels are named by their (autotuned) tile."""
    k = KERNEL_OF_CLASS.get((variant, cls, form), KERNEL_OF_CLASS.get((cls, form)))
    return k.format(tx=tile[0], ty=tile[1]) if k else None


def measured_traffic(cls, form, variant="pressure", tile=(32, 16)):
    """HBM bytes per active dispatch of the kernel behind `cls`, from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of the run variant (calibrated against a plain copy); None if it is not listed there."""
    import re
    name = kernel_of(cls, form, variant, tile)
    path = TRAFFIC_SUMMARY.get(variant)
    if not name or not path or not os.path.exists(path):
        return None
    in_traffic = False
    for line in open(path):
        if line.startswith("## HBM traffic"):
            in_traffic = True
        m = re.match(r"\| `(.+?)` \| (\d+) \| ([\d.]+) \| ([\d.]+) \|", line)
        if in_traffic and m and m.group(1) == name:
            return round((float(m.group(3)) + float(m.group(4))) * (1 << 20))
    return None


def summarize(r, steps, label, variant="pressure"):
    """Throughput, per-class table (stderr) and the roofline of the dominant kernel class."""
    p = r["p"]
    value = p.nx * p.ny * steps / r["elapsed"] / 1e6
    prof = [k for k in r["prof"] if k["launches"] > 0]
    tot_ms = sum(k["total_ms"] for k in prof)
    work = [k for k in prof if k["name"] not in ("early_exit", "ghost_fill", "misc", "krylov_finish")]
    dom = max(work, key=lambda k: k["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    log(f"[bench] {label}: {value:.1f} Mcell-steps/s")
    for k in sorted(prof, key=lambda k: -k["total_ms"]):
        gbs = k["bytes"] / (k["total_ms"] * 1e-3) / 1e9 if k["total_ms"] > 0 else 0.0
        log(f"  {k['name']:22s} n={k['launches']:6d} total {k['total_ms']:9.3f} ms  avg {k['total_ms'] / k['launches'] * 1e3:8.1f} us"
            f"  {gbs:8.1f} GB/s")
    log(f"[bench] {label} profiled pass: kernel time {tot_ms:.1f} ms of {r['prof_elapsed'] * 1e3:.1f} ms wall; "
        f"timed pass {r['elapsed'] * 1e3:.1f} ms; stats {r['stats']}")
    st = r["stats"]
    tile = tuple(r["tune"]["explicit_tile"])
    traffic = measured_traffic(dom["name"], label, variant, tile)
    return {
        "value": round(value, 3),
        "ms_per_step": round(r["elapsed"] / steps * 1e3, 3),
        "roofline": {"bound": "hbm", "kernel": dom["name"], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": ("rocprofv3 FETCH_SIZE + WRITE_SIZE per active dispatch of "
                                        f"{kernel_of(dom['name'], label, variant, tile)}, calibrated, "
                                        f"{os.path.relpath(TRAFFIC_SUMMARY[variant], ROOT)}") if traffic else None,
                     "avg_launch_us": round(avg_ms * 1e3, 2), "algorithmic_bytes_per_launch": round(bytes_per_launch)},
        "gmres": {k: st[k] for k in ("mag_gmres_iters", "mag_solves", "p_gmres_iters", "p_solves", "newton_iters",
                                     "mixed_refinements")},
        "autotuned": {"explicit_tile": list(r["tune"]["explicit_tile"]), "krylov_blocks": r["tune"]["krylov_blocks"],
                      "pressure_krylov_blocks": r["tune"]["krylov_p_blocks"]},
        "schedule": {"budget_steps": st["budget_steps"], "budget_reruns": st["budget_reruns"],
                     "graph_builds": st["graph_builds"],
                     "profiled_pass_ms_per_step": round(r["prof_elapsed"] / steps * 1e3, 3)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ot1024")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--schedule", default=None, choices=["auto", "graph", "budget", "host"],
                    help="step schedule; by default graph for the Orszag-Tang 1024^2 runs (fastest there: "
                         "steady iteration counts) and the library's automatic choice for the others")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--oracle-steps", type=int, default=3,
                    help="steps of the headline workload from its initial data for the forms' error against the oracle")
    ap.add_argument("--mixed-tols", type=float, nargs=2, default=None, metavar=("INNER", "OUTER"),
                    help="override the mixed form's inner / outer refinement tolerances")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the secondary workloads (blast 1024^2 FP64 vs mixed; the 256-vortex batches; the 256^2 adjoint)")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the 256^2 adjoint check and print it")
    ap.add_argument("--adjoint-dirs", type=int, default=32)
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the 256-vortex batches and print their results (for profiling)")
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--blast-steps", type=int, default=20)
    ap.add_argument("--formula", default="pressure", choices=["pressure", "energy"],
                    help="implicit subsystem of the headline run (PressureFormula or EnergyFormula)")
    ap.add_argument("--magfield", default="stagger", choices=["stagger", "unstagger", "potential_a", "nodivfree"],
                    help="magnetic treatment of the headline run (StaggerCT, UnStaggerCT, PotentialA, NoDivFree)")
    ap.add_argument("--magfield-steps", type=int, default=20,
                    help="timed FP64 steps of each other magnetic treatment among the secondary workloads")
    ap.add_argument("--energy-steps", type=int, default=40,
                    help="timed steps of the energy-formula FP64 / mixed pair among the secondary workloads")
    ap.add_argument("--recon", default=None, choices=list(RECON_NAMES),
                    help="reconstruction of the headline workload (default: the workload's own, WENO5)")
    ap.add_argument("--recon-state", default="cons", choices=["cons", "prim"],
                    help="reconstruct the conservative (default) or the primitive state")
    ap.add_argument("--recon-steps", type=int, default=20,
                    help="timed FP64 steps of each other reconstruction among the secondary workloads")
    ap.add_argument("--jet-steps", type=int, default=20,
                    help="timed steps of the 400x600 MHD jet FP64 / mixed pair among the secondary workloads")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="print the practical streaming rates of plain copy / add kernels (8 MB and 512 MB vectors) and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain elementwise copy of MIB MiB (reference point for FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.stream_ceiling:
        stream_ceiling()
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU domain decomposition is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import imex_amd as im
    SCHEDS = {"auto": im.SCHED_AUTO, "graph": im.SCHED_GRAPH, "budget": im.SCHED_BUDGET, "host": im.SCHED_HOST}
    sched_name = args.schedule or ("graph" if args.workload.startswith("ot1024") else "auto")
    sched = SCHEDS[sched_name]
    sched_other = SCHEDS[args.schedule or "auto"]      # the blast, jet and batch workloads
    if args.adjoint_only:
        print(json.dumps({"adjoint": run_adjoint(256, 20, 5, args.adjoint_dirs, 1)}), flush=True)
        return
    if args.batch_only:
        res = {e: run_batch(args.batch_instances, args.batch_steps, 1, eos, 2024, sched_other)
               for e, eos in (("ideal", im.EOS_IDEAL), ("redlich_kwong", im.EOS_RK))}
        print(json.dumps({"batch": res}), flush=True)
        return
    formula = {"pressure": im.FORMULA_PRESSURE, "energy": im.FORMULA_ENERGY}[args.formula]
    MAGFIELDS = {"stagger": im.MAGFIELD_STAGGER, "unstagger": im.MAGFIELD_UNSTAGGER,
                 "potential_a": im.MAGFIELD_POTENTIAL_A, "nodivfree": im.MAGFIELD_NODIVFREE}
    magfield = MAGFIELDS[args.magfield]
    over = {"recon_state": im.RECONSTATE_PRIM if args.recon_state == "prim" else im.RECONSTATE_CONS}
    if args.recon is not None:
        over["recon"] = RECON_NAMES[args.recon]
    forms = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    runs = {}
    for f in forms:
        prec = im.PREC_FP64 if f == "fp64" else im.PREC_MIXED
        runs[f] = run_gpu(args.workload, args.steps, args.warmup, prec, sched, args.mixed_tols if f == "mixed" else None,
                          formula, magfield, over)
    variant = args.magfield if args.magfield != "stagger" else args.formula
    if not args.workload.startswith("ot1024"):
        variant = "other"                 # the committed traffic summaries are per-dispatch figures of the OT 1024^2 runs
    res = {f: summarize(runs[f], args.steps, f, variant) for f in forms}
    head = "mixed" if "mixed" in res else "fp64"
    p = runs[head]["p"]
    out = {
        "metric": "Mcell-steps/s at 1024^2 WENO5 Orszag-Tang, mixed vs FP64 GPU; % HBM peak",
        "value": res[head]["value"],
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32-magnetic-operator/f64",
        "data": "synthetic (Orszag-Tang initial data generated on the host)",
        "config": {"workload": f"{WORKLOAD_NAME.get(args.workload.rstrip('0123456789'), args.workload)}_{p.nx}x{p.ny}_"
                               f"{RECON_LABEL[p.recon].lower()}_staggerct_rk2", "nx": p.nx, "ny": p.ny,
                   "recon": RECON_LABEL[p.recon], "recon_state": args.recon_state, "p_floor": p.p_floor,
                   "gamma": p.gamma, "form": head,
                   "formula": args.formula, "magfield": args.magfield, "schedule": sched_name},
        "roofline": res[head]["roofline"],
        "gmres": res[head]["gmres"],
        "schedule": res[head]["schedule"],
        "autotuned": res[head]["autotuned"],
    }
    if head == "mixed":
        out["config"]["mixed_inner_tol"] = runs["mixed"]["p"].mixed_inner_tol
        out["config"]["mixed_outer_tol"] = runs["mixed"]["p"].mixed_outer_tol
        out["config"]["mixed_flux"] = runs["mixed"]["p"].mixed_flux
        out["config"]["mixed_pressure"] = runs["mixed"]["p"].mixed_pressure
    if "fp64" in res and head != "fp64":
        out["fp64"] = res["fp64"]
        a, b = runs["mixed"]["final"][0], runs["fp64"]["final"][0]
        g = im.GC
        errs = [float(np.linalg.norm(a[g:-g, g:-g, k] - b[g:-g, g:-g, k]) / max(np.linalg.norm(b[g:-g, g:-g, k]), 1e-300))
                for k in range(b.shape[-1])]
        out["mixed_accuracy"] = {"rel_l2_vs_fp64_max": max(errs), "rel_l2_vs_fp64_per_field": [float(f"{e:.3e}") for e in errs],
                                 "steps_compared": args.warmup + args.steps}
        out["speedup_mixed_over_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
    if not args.no_extra:
        if args.formula == "pressure" and args.recon_state == "cons":
            out["mixed_variants"] = {"workload": "the headline workload with the opt-in FP32 parts of the mixed form",
                                     "steps": args.steps, "warmup": args.warmup}
            for vname, vover in (("fp32_fluxes", {"mixed_flux": 1}), ("fp32_pressure_solves", {"mixed_pressure": 1})):
                out["mixed_variants"][vname] = run_pair(args.workload, args.steps, args.warmup, sched, formula, magfield,
                                                        forms=("mixed",), over=dict(over, **vover))["mixed"]
        out["blast"] = dict(workload="magnetic blast 1024x1024 VanLeer, transmissive, nested-Newton p_floor 1e-10",
                            steps=args.blast_steps, schedule=args.schedule or "auto",
                            **run_pair("blast1024", args.blast_steps, 0, sched_other))
        out["batch"] = {"workload": f"{args.batch_instances} MHD vortices, sizes from {list(BATCH_SIZES)}^2, "
                                    "amplitude and background velocity U[0.5,2], Minmod, StaggerCT, FP64",
                        "steps": args.batch_steps, "warmup": 1,
                        "ideal": run_batch(args.batch_instances, args.batch_steps, 1, im.EOS_IDEAL, 2024, sched_other),
                        "redlich_kwong": run_batch(args.batch_instances, args.batch_steps, 1, im.EOS_RK, 2024, sched_other)}
        if args.magfield == "stagger":
            out["magfield"] = {"workload": "Orszag-Tang 1024x1024 WENO5, FP64, with the other magnetic treatments "
                                           "(UnStaggerCT and PotentialA evolve the vector potential A_z)",
                               "steps": args.magfield_steps, "warmup": 3}
            for name_mf in ("unstagger", "potential_a", "nodivfree"):
                out["magfield"][name_mf] = run_pair(args.workload, args.magfield_steps, 3, sched, formula,
                                                    MAGFIELDS[name_mf], forms=("fp64",))["fp64"]
        if args.formula == "pressure":
            out["energy_formula"] = dict(workload="Orszag-Tang 1024x1024 WENO5 with EnergyFormula (Energy_flux_Update + "
                                                  "PreSubsystem_Energy in place of the pressure subsystem)",
                                         steps=args.energy_steps,
                                         warmup=3, **run_pair(args.workload, args.energy_steps, 3, sched, im.FORMULA_ENERGY))
        if args.recon is None and args.recon_state == "cons":
            out["reconstructions"] = {"workload": "Orszag-Tang 1024x1024, FP64, with the other reconstructions "
                                                  "(ENO3, NoRecon) and with PrimState (primitive reconstruction)",
                                      "steps": args.recon_steps, "warmup": 3}
            for rn, st in (("eno3", "cons"), ("none", "cons"), ("weno5", "prim"), ("minmod", "prim")):
                o = {"recon": RECON_NAMES[rn],
                     "recon_state": im.RECONSTATE_PRIM if st == "prim" else im.RECONSTATE_CONS}
                out["reconstructions"][f"{rn}_{st}"] = run_pair(args.workload, args.recon_steps, 3, sched, formula, magfield,
                                                                forms=("fp64",), over=o)["fp64"]
        out["jet"] = dict(workload="MHD jet 400x600 (testMHDJet's grid): jet inflow strip |x| < 0.05 on the bottom "
                                   "boundary (rho 1.4, v 1e4, B0 141.4, p 1), outflow elsewhere, Minmod, StaggerCT",
                          steps=args.jet_steps, warmup=3, schedule=args.schedule or "auto",
                          **run_pair("jet400", args.jet_steps, 3, sched_other))
        out["diagnostics"] = dict(workload="Orszag-Tang 1024x1024, FP64: per-step diagnostics recorded on the device, "
                                           "Diagnose_CT_EMF, VTK and RK_*.dat output", steps=20,
                                  **run_diag_cost(args.workload, 20, 3, sched))
        out["adjoint"] = run_adjoint(256, 20, 5, args.adjoint_dirs, 1)
    if not args.no_cpu:
        log("[bench] oracle accuracy of both forms")
        out["oracle_accuracy"] = dict(workload="the headline workload from its initial data, both GPU forms against "
                                               "the FP64 oracle", **oracle_accuracy(args.workload, args.oracle_steps,
                                                                                    formula, magfield, over))
        log("[bench] CPU oracle sample")
        v, secs, used = cpu_baseline(args.workload, runs[head]["state"], args.cpu_steps, formula, magfield,
                                     runs[head]["potential"], over)
        out["cpu_baseline"] = {"value": round(v, 5), "unit": "Mcell-steps/s", "cores": used, "kind": "port",
                               "sample": f"{args.cpu_steps} steps of {p.nx}x{p.ny} Orszag-Tang ({args.formula} formula, "
                                         f"{args.magfield}) from the post-warmup state, faithful mode (the reference's "
                                         f"per-apply recomputation), {secs:.1f} s on {used} core",
                               "cpu_model": cpu_model(), "nproc": os.cpu_count()}
        va, secs_a, used_a = cpu_baseline(args.workload, runs[head]["state"], 1, formula, magfield, runs[head]["potential"],
                                          over, threads=os.cpu_count() or 1, faithful=0)
        out["cpu_all_cores"] = {"value": round(va, 5), "unit": "Mcell-steps/s", "cores": used_a,
                                "sample": f"1 step, hoisted mode (invariants computed once per stage), OpenMP over the "
                                          f"per-cell loops, {secs_a:.1f} s"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
