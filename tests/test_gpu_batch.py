"""Batches of independent instances against single-instance runs.

A batch steps instances of uneven grid size (and their own initial data) in one set of
launches, on iteration budgets taken over all instances and replayed as a captured
graph.  Each instance converges and stops on its own inside every solve, so each must
end bit for bit where the same problem run alone ends (host-scheduled, the reference's
control flow), with the same iteration counts in total.  Single-instance runs are
themselves checked against the FP64 oracle by the parity tests.
"""
import numpy as np
import pytest

import imex_amd as im

pytestmark = pytest.mark.gpu

# (nx, ny, amplitude, background u0, v0) of each instance
INSTANCES = [(48, 40, 1.0, 1.0, 1.0), (64, 64, 1.7, 0.6, 1.3), (32, 32, 0.8, 1.9, 0.5), (40, 56, 1.2, 1.1, 1.6)]
COUNTS = ("mag_solves", "mag_gmres_iters", "p_solves", "p_gmres_iters", "newton_iters")


def base_params(eos, precision):
    p = im.vortex_params(64, eos=eos)
    p.precision = precision
    return p


def instance_data(p, nx, ny, amp, u0, v0):
    q = im.Params.from_buffer_copy(p)
    q.nx, q.ny = nx, ny
    u, bx, by = im.init_case(q, im.CASE_VORTEX, amp, u0, v0, 0)
    return q, u, bx, by


def run_single(q, u, bx, by, steps):
    q = im.Params.from_buffer_copy(q)
    q.schedule = im.SCHED_HOST
    s = im.Solver(q)
    try:
        s.set_state(u, bx, by)
        _, st = s.step(steps)
        uc, up, bxg, byg, sc = s.get_state()
    finally:
        s.close()
    return uc, bxg, byg, sc, st


@pytest.mark.parametrize("eos,precision,schedule", [
    (im.EOS_IDEAL, im.PREC_FP64, im.SCHED_GRAPH),
    (im.EOS_IDEAL, im.PREC_FP64, im.SCHED_BUDGET),
    (im.EOS_RK, im.PREC_FP64, im.SCHED_GRAPH),
    (im.EOS_IDEAL, im.PREC_MIXED, im.SCHED_GRAPH),
])
def test_batch_matches_single_runs_bitwise(eos, precision, schedule):
    steps = 6
    print(f"[batch] eos={eos} precision={precision} schedule={schedule} instances={len(INSTANCES)} steps={steps}", flush=True)
    p = base_params(eos, precision)
    p.schedule = schedule
    data = [instance_data(p, *inst) for inst in INSTANCES]
    b = im.Batch(p, [(nx, ny) for nx, ny, *_ in INSTANCES])
    try:
        assert len(b) == len(INSTANCES)
        for i, (q, u, bx, by) in enumerate(data):
            b.set_instance(i, u, bx, by)
        _, st = b.step(steps)
        outs = [b.get_instance(i) for i in range(len(INSTANCES))]
        divs = [b.div_b_of(i) for i in range(len(INSTANCES))]
    finally:
        b.close()
    print(f"  batch: budget steps {st['budget_steps']} reruns {st['budget_reruns']} graphs {st['graph_builds']}; "
          + " ".join(f"{k}={st[k]}" for k in COUNTS), flush=True)
    assert st["steps"] == steps
    assert st["budget_steps"] >= steps
    tot = dict.fromkeys(COUNTS, 0)
    for i, (q, u, bx, by) in enumerate(data):
        uc, bxg, byg, sc, sst = run_single(q, u, bx, by, steps)
        bu, _, bbx, bby, bsc = outs[i]
        assert np.array_equal(bu, uc), f"instance {i} state differs"
        assert np.array_equal(bbx, bxg) and np.array_equal(bby, byg), f"instance {i} faces differ"
        assert bsc["t"] == sc["t"] and bsc["dt"] == sc["dt"]
        assert divs[i] <= 1e-12
        for k in COUNTS:
            tot[k] += sst[k]
    for k in COUNTS:
        assert st[k] == tot[k], (k, st[k], tot[k])


def test_batch_with_restarting_instances_matches_single_runs():
    """Magnetic blasts of several strengths: the strong ones need GMRES restart cycles, which
    the batch runs inside its budgets (implicit_operators.cpp:899-910) while the weak ones
    have stopped; every instance still ends where its host-scheduled single run ends."""
    steps = 5
    inst = [(48, 40, 100.0), (40, 40, 10.0), (32, 48, 60.0)]
    print(f"[batch] magnetic blasts {inst} steps={steps}", flush=True)
    p = im.blast_params(48)
    p.schedule = im.SCHED_GRAPH
    data = []
    for nx, ny, amp in inst:
        q = im.Params.from_buffer_copy(p)
        q.nx, q.ny = nx, ny
        data.append((q,) + im.init_case(q, im.CASE_BLAST, amp, 0.0, 0.0, 0))
    b = im.Batch(p, [(nx, ny) for nx, ny, _ in inst])
    try:
        for i, (q, u, bx, by) in enumerate(data):
            b.set_instance(i, u, bx, by)
        _, st = b.step(steps)
        outs = [b.get_instance(i) for i in range(len(inst))]
    finally:
        b.close()
    print(f"  batch: budget steps {st['budget_steps']} reruns {st['budget_reruns']} graphs {st['graph_builds']}; "
          + " ".join(f"{k}={st[k]}" for k in COUNTS + ("mag_gmres_restarts", "p_gmres_restarts")), flush=True)
    assert st["budget_steps"] >= steps
    restarts = []
    tot = dict.fromkeys(COUNTS + ("mag_gmres_restarts", "p_gmres_restarts"), 0)
    for i, (q, u, bx, by) in enumerate(data):
        uc, bxg, byg, sc, sst = run_single(q, u, bx, by, steps)
        restarts.append(sst["mag_gmres_restarts"] + sst["p_gmres_restarts"])
        bu, _, bbx, bby, bsc = outs[i]
        assert np.array_equal(bu, uc), f"instance {i} state differs"
        assert np.array_equal(bbx, bxg) and np.array_equal(bby, byg), f"instance {i} faces differ"
        assert bsc["t"] == sc["t"] and bsc["dt"] == sc["dt"]
        for k in tot:
            tot[k] += sst[k]
    print(f"  restart cycles of the single runs: {restarts}", flush=True)
    assert max(restarts) > 0
    for k in tot:
        assert st[k] == tot[k], (k, st[k], tot[k])


def test_batch_of_one_equals_solver():
    print("[batch] one instance", flush=True)
    p = base_params(im.EOS_IDEAL, im.PREC_FP64)
    q, u, bx, by = instance_data(p, 40, 40, 1.0, 1.0, 1.0)
    b = im.Batch(q, [(40, 40)])
    try:
        b.set_instance(0, u, bx, by)
        b.step(5)
        bu = b.get_instance(0)[0]
    finally:
        b.close()
    assert np.array_equal(bu, run_single(q, u, bx, by, 5)[0])


def test_batch_set_state_after_steps():
    """Resetting one instance mid-run leaves the others on course."""
    print("[batch] set state mid-run", flush=True)
    p = base_params(im.EOS_IDEAL, im.PREC_FP64)
    data = [instance_data(p, *inst) for inst in INSTANCES[:2]]
    b = im.Batch(p, [(nx, ny) for nx, ny, *_ in INSTANCES[:2]])
    try:
        for i, (q, u, bx, by) in enumerate(data):
            b.set_instance(i, u, bx, by)
        b.step(3)
        q, u, bx, by = data[0]
        b.set_instance(0, u, bx, by)
        b.step(3)
        o0 = b.get_instance(0)
        o1 = b.get_instance(1)
    finally:
        b.close()
    ref0 = run_single(*data[0], 3)
    ref1 = run_single(*data[1], 6)
    np.testing.assert_allclose(o0[0], ref0[0], rtol=1e-12, atol=1e-13)
    assert np.array_equal(o1[0], ref1[0])
    assert o0[4]["timestep"] == 3 and o1[4]["timestep"] == 6


def test_autotune_keeps_results():
    """The explicit tile choice leaves results bit for bit; the Krylov grid cap only by rounding."""
    print("[batch] autotune", flush=True)
    p = base_params(im.EOS_IDEAL, im.PREC_FP64)
    data = [instance_data(p, *inst) for inst in INSTANCES]
    sizes = [(nx, ny) for nx, ny, *_ in INSTANCES]
    outs = {}
    for flags in (0, im.TUNE_EXPLICIT, im.TUNE_EXPLICIT | im.TUNE_KRYLOV):
        b = im.Batch(p, sizes)
        try:
            for i, (q, u, bx, by) in enumerate(data):
                b.set_instance(i, u, bx, by)
            if flags:
                info = b.autotune(flags)
                print(f"  flags {flags}: {info}", flush=True)
                assert info["explicit_tile"] in ((16, 16), (32, 8), (64, 4), (32, 16), (64, 8))
                assert info["krylov_blocks"] in (1024, 2048, 4096)
                if flags & im.TUNE_EXPLICIT:
                    assert all(t > 0 for t in info["explicit_us"])
            b.step(4)
            outs[flags] = [b.get_instance(i) for i in range(len(sizes))]
        finally:
            b.close()
    for i in range(len(sizes)):
        assert np.array_equal(outs[im.TUNE_EXPLICIT][i][0], outs[0][i][0])
        np.testing.assert_allclose(outs[im.TUNE_EXPLICIT | im.TUNE_KRYLOV][i][0], outs[0][i][0], rtol=1e-10, atol=1e-12)
