"""The GPU path at the benchmark configurations' own workloads (dimer J = 1, E0 = 1,
omega = 1, start e_0, MCG59 streams leap-frog(seed + tr, 1e6) seeded 777).

c2  d = 501, FP64: per-trajectory jump times, etas, RNG states and states against the
    CPU oracle with the same propagator tree and streams, on the reference's QJ tree
    (S = 4, ex_deep 16; qj_proc.cpp:271-343, experiment.cpp:481-494), after checking
    that the horizon is well conditioned (one-ulp propagator changes leave the oracle's
    discrete path unchanged).
c3  d = 1001, 1e5 trajectories, S = 200, ex_deep 16: FP32 and 3xBF16 against the FP64
    run on the same streams (<n1-n2>/N and purity within 1e-4), size-free invariants,
    and independence from the tile width and from skipping the band's exact zeros.
c4  the 32-instance uneven batch (seeded manifest): every member equals its run alone.
c5  the serving mix as specified (Poisson 20 jobs/s, N in {100..2000}, 1e3-1e5
    trajectories, 50-500 periods): admitted footprint within the budget, no job lost
    or failed, and a job's results the same alone and inside the mix.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MCG59_A, MCG59_M = 302875106592253, 1 << 59
STRIDE = pow(MCG59_A, 1000000, MCG59_M)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _perturbed_copy(oracle_mod, m, M, tree, D):
    o = oracle_mod.Ensemble(m, M, seed=0, jump=2, build_propagators=False)
    rng = np.random.default_rng(0)
    eps = 2.0 ** -53
    for b in range(2):
        for lvl in range(D):
            U = tree[b][lvl]
            re = U.real * (1 + eps * rng.choice([-1.0, 1.0], U.shape))
            im = U.imag * (1 + eps * rng.choice([-1.0, 1.0], U.shape))
            o.set_propagator(b, lvl, re + 1j * im)
    return o


def test_c2_fp64_d501_jump_parity(oracle_mod):
    from os_d_amd import qjx
    N, S, D, M, steps = 500, 4, 16, 16, 4
    _say("case: c2 FP64 d=%d S=%d D=%d M=%d steps=%d" % (N + 1, S, D, M, steps))
    e = qjx.Engine(N, prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0, diss_gamma=0.1,
                   ex_deep=D, steps_per_period=S, precision=qjx.FP64)
    e.ensemble_init(M, seed=0, jump=2, log_capacity=4096)
    tree = [[e.get_propagator(b, lvl) for lvl in range(D)] for b in range(2)]
    m = oracle_mod.model(N, gamma=0.1, S=S, D=D)
    o = oracle_mod.Ensemble(m, M, seed=0, jump=2, build_propagators=False)
    for b in range(2):
        for lvl in range(D):
            o.set_propagator(b, lvl, tree[b][lvl])
    # conditioning of the horizon: the oracle against itself with one-ulp changes
    p = _perturbed_copy(oracle_mod, m, M, tree, D)
    for x in (o, p):
        x.trans(0)
        x.steps(steps, 1)
    so, sp = o.get_state(), p.get_state()
    for k in ("rng", "eta", "t", "njumps"):
        assert np.array_equal(so[k], sp[k]), ("ill-conditioned horizon", k)
    assert np.abs(so["psi"] - sp["psi"]).max() < 1e-12
    _say("oracle: %d jumps, %d matvecs" % (so["njumps"].sum(), o.num_matvecs()))
    e.trans(0)
    e.run_steps(steps, 1)
    se = e.get_state()
    for k in ("rng", "eta", "t", "njumps"):
        assert np.array_equal(so[k], se[k]), k
    for tr in range(M):
        jo, je = o.get_jumps(tr), e.get_jumps(tr)
        assert np.array_equal(jo["times"], je["times"]), tr
        assert np.array_equal(jo["etas"], je["etas"]), tr
        assert np.array_equal(jo["types"], je["types"]), tr
        np.testing.assert_allclose(je["norms"], jo["norms"], rtol=1e-12)
    np.testing.assert_allclose(se["psi"], so["psi"], rtol=0, atol=1e-10)
    assert e.stats()["matvecs"] == o.num_matvecs()
    # ensemble averages of the same trajectories agree to 1e-10 relative
    co, ce = o.chars(), e.chars()
    for k in ("mean", "energy", "m2"):
        assert abs(ce[k].mean() - co[k].mean()) <= 1e-10 * abs(co[k].mean()), k
    e.close()


C3 = dict(prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0, diss_gamma=0.1)


def _c3_run(qjx, prec, M, steps):
    e = qjx.Engine(1000, ex_deep=16, steps_per_period=200, precision=prec, **C3)
    e.ensemble_init(M, seed=0, jump=1, log_capacity=256)
    e.trans(0)
    rng0 = e.get_state(psi=False)["rng"].copy()
    e.run_steps(steps, 1)
    return e, rng0


def _obs(e, N=1000):
    c = e.chars()
    _, pur = e.density(rho=False)
    return float(((N - 2.0 * c["mean"]) / N).mean()), pur, c["norm"]


def test_c3_precision_forms_d1001_1e5(monkeypatch):
    from os_d_amd import qjx
    M, steps = 100000, 4
    res = {}
    for prec in (qjx.FP64, qjx.FP32, qjx.BF16X3):
        _say("case: c3 d=1001 M=%d steps=%d prec=%d" % (M, steps, prec))
        e, rng0 = _c3_run(qjx, prec, M, steps)
        s = e.get_state(psi=False)
        nj = s["njumps"].astype(np.int64)
        assert nj.sum() > M // 2
        # RNG draws: eta_0 (before rng0), then (ran, eta) per jump
        for tr in range(M):
            assert int(s["rng"][tr]) == int(rng0[tr]) * pow(STRIDE, 2 * int(nj[tr]), MCG59_M) % MCG59_M, tr
        for tr in range(0, M, 997):
            jt = e.get_jumps(tr)["times"]
            assert len(jt) == nj[tr] and (np.diff(jt) > 0).all(), tr
        assert np.allclose(s["t"], steps * 2 * np.pi / 200, rtol=1e-12)
        x, pur, norm = _obs(e)
        tol = 1e-12 if prec == qjx.FP64 else (1e-4 if prec == qjx.FP32 else 1e-3)
        assert ((norm > 0.0) & (norm <= 1.0 + tol)).all()
        res[prec] = (x, pur, s)
        _say("prec %d: <n1-n2>/N %.8f purity %.8f jumps %d" % (prec, x, pur, nj.sum()))
        if prec != qjx.FP64:
            # the same steps with another tile width and every K block multiplied
            e.close()
            monkeypatch.setenv("QJX_BN", "64")
            monkeypatch.setenv("QJX_DENSE", "1")
            e2, _ = _c3_run(qjx, prec, M, steps)
            s2 = e2.get_state(psi=False)
            for k in ("eta", "t", "njumps", "rng"):
                assert np.array_equal(s[k], s2[k]), (prec, k)
            x2, pur2, _ = _obs(e2)
            assert x2 == x and pur2 == pur
            e2.close()
            monkeypatch.delenv("QJX_BN")
            monkeypatch.delenv("QJX_DENSE")
        else:
            e.close()
    x64, p64, _ = res[qjx.FP64]
    for prec in (qjx.FP32, qjx.BF16X3):
        x, p, _ = res[prec]
        assert abs(x - x64) <= 1e-4, (prec, x, x64)
        assert abs(p - p64) <= 1e-4, (prec, p, p64)


def test_c4_uneven_batch_members_equal_alone():
    import bench
    from os_d_amd import qjx
    inst = bench.batched_instances()
    S, D, steps = 200, 16, 2
    _say("case: c4 %d instances, %d trajectories, dims %d..%d" % (
        len(inst), sum(p["M"] for p in inst), min(p["N"] for p in inst) + 1, max(p["N"] for p in inst) + 1))
    engines = [bench._instance_engine(qjx, p, S, D, qjx.FP32) for p in inst]
    b = qjx.Batch(engines)
    b.run_steps(steps, 0)
    got = []
    for e in engines:
        got.append((e.get_state(psi=False), e.chars(), e.stats()["matvecs"]))
    b.close()
    for e in engines:
        e.close()
    for i, p in enumerate(inst):
        e = bench._instance_engine(qjx, p, S, D, qjx.FP32)
        e.run_steps(steps, 0)
        s, c, mv = got[i]
        sa, ca = e.get_state(psi=False), e.chars()
        for k in ("eta", "t", "njumps", "rng"):
            assert np.array_equal(s[k], sa[k]), (i, k)
        for k in ("norm", "mean", "m2", "energy"):
            assert np.array_equal(c[k], ca[k]), (i, k)
        assert mv == e.stats()["matvecs"], i
        e.close()


def test_c5_specified_serving_mix_admission():
    import bench
    from os_d_amd import qjx
    seconds, budget_gib = 6.0, 24.0
    _say("case: c5 specified mix %.0f s, budget %.0f GiB" % (seconds, budget_gib))
    probe = dict(N=100, num_trajectories=1000, n_periods=1, seed=77, drv_ampl=1.2, prm_U=0.3, diss_gamma=0.05,
                 steps_per_period=4, ex_deep=16)
    alone = qjx.Pool(qjx.FP32, hbm_budget_gib=budget_gib, max_trajectories=2000000)
    ja = alone.submit(**probe)
    alone.drain()
    ra = alone.result(ja, arrays=True)
    alone.close()
    assert ra["status"] == 2

    pool = qjx.Pool(qjx.FP32, hbm_budget_gib=budget_gib, max_trajectories=2000000)
    jobs, peak = bench.serve_mix(qjx, pool, seconds, 20.0, 12345, "specified", probe=probe, probe_at=1.0,
                                 check=True)
    u = pool.usage()
    _say("c5: submitted %d, usage %s, peak used %.2f GiB" % (len(jobs), u, peak / 2 ** 30))
    assert len(jobs) > 50
    assert peak <= u["budget_bytes"]
    assert u["failed"] == 0
    # every submitted job (the probe included) is queued, building, running or done
    assert u["queued"] + u["building"] + u["running"] + u["done"] == len(jobs)
    jp = [j["id"] for j in jobs if j["probe"]][0]
    for _ in range(200000):
        if pool.result(jp)["status"] == 2:
            break
        pool.step()
    rm = pool.result(jp, arrays=True)
    assert rm["status"] == 2
    assert np.array_equal(rm["mean"], ra["mean"]) and np.array_equal(rm["energy"], ra["energy"])
    assert rm["matvecs"] == ra["matvecs"]
    pool.close()
