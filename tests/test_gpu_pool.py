"""Continuous batching: jobs of different Hilbert dimension share every launch.

A job's per-trajectory results must be bitwise identical alone and in the mix
(its trajectories' arithmetic does not depend on the other jobs), jobs queued
behind the HBM budget must be admitted as others retire, and a captured HIP
graph of the iterations must reproduce eager launches exactly.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


JOB_A = dict(N=30, num_trajectories=500, n_periods=3, seed=5, diss_gamma=0.2, steps_per_period=8)


def _alone(prec):
    from os_d_amd import qjx
    p = qjx.Pool(prec, max_trajectories=4096)
    j = p.submit(**JOB_A)
    p.drain()
    r = p.result(j, arrays=True)
    p.close()
    return r


@pytest.mark.parametrize("prec", [1, 0])
def test_job_alone_equals_job_in_mix(prec):
    from os_d_amd import qjx
    _say("case: pool alone vs mix prec=%d" % prec)
    ra = _alone(prec)
    assert ra["status"] == 2
    p = qjx.Pool(prec, max_trajectories=4096)
    jb = p.submit(N=50, num_trajectories=800, n_periods=2, seed=1, drv_ampl=2.5, steps_per_period=8)
    p.step()
    p.step()                     # B is running when A arrives
    ja = p.submit(**JOB_A)
    jc = p.submit(N=80, num_trajectories=300, n_periods=4, seed=9, prm_U=0.9, steps_per_period=6)
    p.drain()
    rb, rm, rc = p.result(jb), p.result(ja, arrays=True), p.result(jc)
    _say("mix: B %s, A %s, C %s" % (rb["status"], rm["status"], rc["status"]))
    assert rb["status"] == 2 and rm["status"] == 2 and rc["status"] == 2
    assert np.array_equal(rm["mean"], ra["mean"])
    assert np.array_equal(rm["energy"], ra["energy"])
    assert rm["matvecs"] == ra["matvecs"]
    p.close()


def test_admission_by_budget_queues_and_completes():
    from os_d_amd import qjx
    _say("case: pool admission budget")
    # each job needs ~8.5 MB (propagators + build workspace + ensemble): room for two
    p = qjx.Pool(1, hbm_budget_gib=0.02, max_trajectories=1 << 16)
    ids = [p.submit(N=100, num_trajectories=400, n_periods=1, seed=s, steps_per_period=8) for s in range(5)]
    peak_queue = 0
    for _ in range(100000):
        r, q = p.step()
        peak_queue = max(peak_queue, q)
        if r == 0 and q == 0:
            break
    res = [p.result(j) for j in ids]
    _say("peak queue %d; admits %s" % (peak_queue, [round(x["t_admit"], 3) for x in res]))
    assert peak_queue >= 1
    assert all(x["status"] == 2 for x in res)
    # jobs that waited were admitted after an earlier one finished
    assert max(x["t_admit"] for x in res) >= min(x["t_done"] for x in res)
    p.close()


def test_short_lane_cap(monkeypatch):
    """The short lane takes in at most QJX_POOL_SHORT_TRAJ trajectories; a short job
    above the cap runs when the lane is empty.  Every job completes."""
    from os_d_amd import qjx
    _say("case: short-lane trajectory cap")
    monkeypatch.setenv("QJX_POOL_SHORT_TRAJ", "1000")
    p = qjx.Pool(1, max_trajectories=1 << 16)
    sizes = [600, 600, 1500, 600, 600]
    ids = [p.submit(N=50, num_trajectories=m, n_periods=1, seed=s, steps_per_period=4) for s, m in enumerate(sizes)]
    peak = 0
    for _ in range(100000):
        r, q = p.step()
        u = p.usage()
        admitted = u["running"] + u["building"]
        peak = max(peak, u["trajectories"])
        # within the cap, or the large job alone
        assert u["trajectories"] <= 1000 or (u["trajectories"] == 1500 and admitted == 1), u
        if r == 0 and q == 0 and u["building"] == 0:
            break
    res = [p.result(j) for j in ids]
    _say("peak admitted trajectories %d" % peak)
    assert all(x["status"] == 2 for x in res)
    assert peak == 1500
    p.close()


def test_graph_replay_equals_eager(monkeypatch):
    from os_d_amd import qjx
    _say("case: graph vs eager")
    out = {}
    for g in ("1", "0"):
        monkeypatch.setenv("QJX_GRAPH", g)
        e = qjx.Engine(40, ex_deep=12, steps_per_period=8, precision=1)
        e.ensemble_init(700, seed=2, jump=1)
        e.trans(0)
        e.run_periods(3, 1)
        out[g] = (e.get_state(), [e.get_jumps(t)["times"] for t in range(700)], e.stats()["iterations"])
        e.close()
    (sa, ja, ia), (sb, jb, ib) = out["1"], out["0"]
    _say("iterations graph %d eager %d" % (ia, ib))
    assert np.array_equal(sa["psi"], sb["psi"])
    assert all(np.array_equal(x, y) for x, y in zip(ja, jb))
