"""Runge-Kutta trajectories (prop_id 1) on the GPU against the CPU restatement.

The kernel and the oracle perform the same FP64 operations in the same order
(reference: rk_period core.cpp:439-450, rk_step_dimer core.cpp:11886-11948,
rk_int_dimer core.cpp:11592-11619, rk_right_part_dimer core.cpp:11033-11115,
rk_recovery rk_proc.cpp:198-306), so states, etas, times, RNG states and jump
records (times, norms, etas, channels) must be bit-identical over the horizon, for
every lane width E (d <= 128, 256, 512, 1024, 2048) and both drive types.  At the
benchmark's size the checks are size-independent: RNG draws = 1 + 2 x jumps per
trajectory, logged jumps = jump count, norms in (0, 1], strictly increasing jump
times.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = dict(prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_freq=1.0)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(orc, N, A, gamma, drv_type, rk_ns, M, seed=3, jump=2, phase=0.2):
    from os_d_amd import qjx
    m = orc.model(N, drv_ampl=A, gamma=gamma, D=1, **P)
    o = orc.Ensemble(m, M, seed=seed, jump=jump, build_propagators=False)
    o.rk_setup(rk_ns, drv_type, phase)
    e = qjx.Engine(N, drv_ampl=A, diss_gamma=gamma, prop="rk", rk_ns=rk_ns, drv_type=drv_type,
                   drv_phase=phase, **P)
    e.ensemble_init(M, seed=seed, jump=jump)
    return o, e


def _same(o, e, M, logs=True):
    so, sg = o.get_state(), e.get_state()
    for k in ("eta", "t", "njumps", "rng"):
        assert np.array_equal(so[k], sg[k]), k
    dpsi = np.abs(so["psi"] - sg["psi"]).max()
    _say("max |psi_gpu - psi_oracle| = %.3e, jumps %d" % (dpsi, int(sg["njumps"].sum())))
    assert np.array_equal(so["psi"], sg["psi"])
    if logs:
        for tr in range(M):
            jo, jg = o.get_jumps(tr), e.get_jumps(tr)
            for k in ("times", "types", "norms", "etas"):
                assert np.array_equal(jo[k], jg[k]), (tr, k)


@pytest.mark.parametrize("N,A,gamma,drv_type,rk_ns,M,tp,mode,amount", [
    (1, 1.5, 0.3, 1, 1000, 64, 1, "periods", 2),       # d = 2: E = 1, halo from zero lanes
    (20, 1.5, 0.3, 0, 1000, 96, 1, "periods", 2),      # E = 1, square drive
    (20, 2.0, 0.2, 1, 1000, 96, 1, "periods", 2),      # E = 1, sine drive
    (63, 1.5, 0.2, 1, 2000, 64, 0, "periods", 1),      # d = 64: E = 1, every lane used
    (64, 1.5, 0.2, 1, 2000, 63, 0, "periods", 1),      # d = 65: 32 lanes x 4, two per wavefront, odd M
    (100, 1.5, 0.1, 1, 4000, 64, 0, "periods", 1),     # 32 lanes x 4
    (127, 1.5, 0.1, 0, 4000, 33, 1, "periods", 1),     # d = 128: 32 lanes x 4, every lane used
    (128, 1.5, 0.1, 1, 4000, 32, 0, "steps", 2000),    # d = 129: 64 lanes x 4
    (200, 1.5, 0.1, 0, 8000, 32, 0, "steps", 3000),    # E = 4, across the half period
    (300, 1.5, 0.1, 1, 12000, 32, 0, "steps", 3000),   # E = 8
    (1000, 1.5, 0.1, 1, 40000, 16, 0, "steps", 1500),  # E = 16
    (2000, 1.5, 0.1, 1, 80000, 8, 0, "steps", 800),    # E = 32
])
def test_rk_bitwise_parity(oracle_mod, N, A, gamma, drv_type, rk_ns, M, tp, mode, amount):
    _say("case: rk parity N=%d drv_type=%d rk_ns=%d M=%d %s %d" % (N, drv_type, rk_ns, M, mode, amount))
    o, e = _pair(oracle_mod, N, A, gamma, drv_type, rk_ns, M)
    o.trans(tp)
    e.trans(tp)
    _same(o, e, M, logs=False)
    if mode == "periods":
        o.periods(amount, 1)
        e.run_periods(amount, 1)
    else:
        o.steps(amount, 1)
        e.run_steps(amount, 1)
    _same(o, e, M)
    s = e.stats()
    assert s["jumps"] == o.num_jumps()
    assert s["matvecs"] == o.num_matvecs()
    e.close()


def test_rk_split_runs_equal_one_run(oracle_mod):
    """Launch boundaries (cached norms, period numbering) do not change results."""
    _say("case: rk split runs")
    o, e = _pair(oracle_mod, 60, 1.5, 0.2, 1, 2000, 48)
    o.trans(0)
    e.trans(0)
    o.steps(2600, 1)
    for n in (700, 1, 1299, 600):
        e.run_steps(n, 1)
    _same(o, e, 48)
    e.close()


def test_rk_chars_and_density_on_rk_context(oracle_mod):
    _say("case: rk chars")
    o, e = _pair(oracle_mod, 50, 1.5, 0.2, 1, 2000, 64)
    o.trans(1)
    e.trans(1)
    o.periods(1, 1)
    e.run_periods(1, 1)
    co, cg = o.chars(), e.chars()
    for k in ("norm", "mean", "m2", "energy"):
        np.testing.assert_allclose(cg[k], co[k], rtol=1e-12, atol=1e-12)
    rho, pur = e.density()
    psi = o.get_state()["psi"]
    w = psi / np.linalg.norm(psi, axis=1, keepdims=True)
    ref = w.T @ w.conj() / len(w)
    assert np.abs(rho - ref).max() < 1e-13
    e.close()


def test_rk_norm_above_one_is_an_error_on_both_sides(oracle_mod):
    """Square drive, no dissipation: the step ending at T/2 evaluates its last stage
    with the other half period's E and lifts ||psi||^2 above 1; the reference throws
    (experiment.cpp:1247-1251), the oracle and the library report the error."""
    from os_d_amd import qjx
    _say("case: rk norm error")
    o, e = _pair(oracle_mod, 8, 1.5, 0.0, 0, 2000, 4)
    o.set_state(eta=np.full(4, 1e-3))
    e.set_state(eta=np.full(4, 1e-3))
    with pytest.raises(RuntimeError):
        o.steps(1200, 0)
    with pytest.raises(qjx.QJXError) as ei:
        e.run_steps(1200, 0)
    assert ei.value.code == -3
    e.close()


def test_rk_bench_size_invariants():
    """d = 1001, 2000 trajectories: invariants that need no CPU run."""
    from os_d_amd import qjx
    N, M, rk_ns, steps = 1000, 2000, 40000, 400
    _say("case: rk invariants N=%d M=%d steps=%d" % (N, M, steps))
    e = qjx.Engine(N, drv_ampl=1.5, diss_gamma=0.1, prop="rk", rk_ns=rk_ns, drv_type=1, **P)
    e.ensemble_init(M, seed=11, jump=2)
    e.trans(0)
    rng0 = e.get_state()["rng"].copy()          # after the eta_0 draw
    e.run_steps(steps, 1)
    s = e.get_state()
    nj = s["njumps"]
    assert nj.sum() > 0
    n = e.chars()["norm"]
    assert ((n > 0.0) & (n <= 1.0)).all()
    a, mod = 302875106592253, 1 << 59
    stride = pow(a, 1000000, mod)
    for tr in range(0, M, 97):
        assert int(s["rng"][tr]) == int(rng0[tr]) * pow(stride, 2 * int(nj[tr]), mod) % mod
        j = e.get_jumps(tr)
        assert len(j["times"]) == nj[tr]
        assert (np.diff(j["times"]) > 0).all() and (j["types"] == 0).all()
    assert np.allclose(s["t"], steps * (2 * np.pi / rk_ns), rtol=1e-9)
    st = e.stats()
    assert st["jumps"] == nj.sum()
    e.close()


def test_jump_interval_pdf_matches_histogram():
    """qjx_jump_interval_pdf bins the waiting times as times_between_jumps.m:41-131;
    the counts equal numpy's histogram over the same log-spaced borders and the pdf
    integrates to one."""
    from os_d_amd import qjx
    _say("case: jump interval pdf")
    e = qjx.Engine(30, drv_ampl=1.5, diss_gamma=0.3, prop="rk", rk_ns=2000, drv_type=1, **P)
    e.ensemble_init(64, seed=7, jump=1)
    e.trans(0)
    e.run_periods(2, 1)
    borders, pdf, cnt, outside = e.jump_interval_pdf(1e-10, 15, 10)
    diffs = np.concatenate([np.diff(e.get_jumps(tr)["times"]) for tr in range(64)])
    inside = diffs[(diffs >= borders[0]) & (diffs <= borders[-1])]
    ref, _ = np.histogram(inside, bins=borders)
    assert cnt.sum() > 100
    assert np.array_equal(cnt, ref)
    assert outside == len(diffs) - len(inside)
    assert abs(np.sum(pdf * np.diff(borders)) - 1.0) < 1e-12
    e.close()
