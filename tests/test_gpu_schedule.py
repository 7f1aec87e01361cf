"""The four step schedules give the same state bit for bit.

HOST polls GMRES/Newton convergence between launches (the reference's control
flow); BUDGET queues every solve with an iteration budget learned from earlier
steps, so no host synchronisation happens inside a step; GRAPH replays the
budgeted step as a captured hipGraph; AUTO times a few steps of GRAPH and HOST
and runs the faster.  A negative budget margin forces solves
to outrun their budgets, which exercises the discard-and-rerun path.
"""
import numpy as np
import pytest

import imex_amd as im
from helpers import make_case

pytestmark = pytest.mark.gpu

CASES = [("vortex", 64, 64, 12), ("ot", 64, 64, 8), ("blast", 48, 40, 6)]
COUNTS = ("steps", "mag_solves", "mag_gmres_iters", "p_solves", "p_gmres_iters", "newton_iters")


def run(case, n, m, steps, schedule, margin=1):
    p, u, bx, by = make_case(case, n, m)
    p.schedule = schedule
    p.budget_margin = margin
    s = im.Solver(p)
    try:
        s.set_state(u, bx, by)
        rc, stats = s.step(steps)
        uc, up, bxg, byg, sc = s.get_state()
    finally:
        s.close()
    return uc, bxg, byg, sc, stats


@pytest.mark.parametrize("case,n,m,steps", CASES)
def test_schedules_bitwise_identical(case, n, m, steps):
    print(f"[schedule] {case} {n}x{m} steps={steps}", flush=True)
    ref = run(case, n, m, steps, im.SCHED_HOST)
    assert ref[4]["budget_steps"] == 0
    print(f"  host: restarts mag {ref[4]['mag_gmres_restarts']} p {ref[4]['p_gmres_restarts']}", flush=True)
    if case == "blast":
        assert ref[4]["mag_gmres_restarts"] + ref[4]["p_gmres_restarts"] > 0
    for sched in (im.SCHED_BUDGET, im.SCHED_GRAPH, im.SCHED_AUTO):
        out = run(case, n, m, steps, sched)
        st = out[4]
        print(f"  sched {sched}: budget steps {st['budget_steps']} reruns {st['budget_reruns']} "
              f"graphs {st['graph_builds']}; {dict((k, st[k]) for k in COUNTS)}", flush=True)
        if sched == im.SCHED_AUTO:      # five host-polled steps, then a trial of five budgeted ones
            assert st["budget_steps"] >= min(5, steps - 5)
        else:
            assert st["budget_steps"] >= steps - 1   # the blast's restart cycles are budgeted too
        if sched != im.SCHED_BUDGET:
            assert st["graph_builds"] >= 1
        for a, b in zip(out[:3], ref[:3]):
            assert np.array_equal(a, b)
        assert out[3]["t"] == ref[3]["t"] and out[3]["dt"] == ref[3]["dt"]
        for k in COUNTS:
            assert st[k] == ref[4][k], (k, st[k], ref[4][k])
        assert st["min_p"] == ref[4]["min_p"]


@pytest.mark.parametrize("case,n,m,steps", CASES)
def test_budget_overrun_reruns_exactly(case, n, m, steps):
    print(f"[overrun] {case} {n}x{m} steps={steps}", flush=True)
    ref = run(case, n, m, steps, im.SCHED_HOST)
    for sched in (im.SCHED_BUDGET, im.SCHED_GRAPH):
        out = run(case, n, m, steps, sched, margin=-2)
        st = out[4]
        print(f"  sched {sched}: budget steps {st['budget_steps']} reruns {st['budget_reruns']}", flush=True)
        assert st["budget_reruns"] >= 1
        for a, b in zip(out[:3], ref[:3]):
            assert np.array_equal(a, b)
        for k in COUNTS:
            assert st[k] == ref[4][k], (k, st[k], ref[4][k])
