"""The drop-in runner (reference config.txt / params.txt in, output files out) on the
GPU, checked against the oracle driven through the same STD experiment."""
import glob
import os
import sys

import numpy as np
import pytest

from deep_tree import reference_deep_tree

pytestmark = pytest.mark.gpu

CONFIG = """sys_id 0
task_id 1
prop_id 0
is_debug 0
is_pp 0
init_fn
path {path}
seed 0
mns 1000000
num_threads 1
num_trajectories 24
num_tp_periods 2
num_obs_periods 6
ex_deep 12
rk_ns 10000
"""

PARAMS = """N 12
diss_type 0
diss_gamma 0.2
diss_phase 0.0
dimer_drv_type 0
dimer_drv_ampl 1.5
dimer_drv_freq 1.0
dimer_drv_phase 0.0
dimer_prm_E 1.0
dimer_prm_U 0.5
dimer_prm_J 1.0
start_type 0
start_state 0
num_random_obs 0
dump_obs 1
dump_phi 0
dump_phi_evo 0
dump_adr_sep 0
dump_adr_avg 1
dump_evo_sep 1
dump_evo_avg 0
dump_type 0
dump_num 3
jump 2
jumps_counts 0
"""

SUFFIX = ("_setup(0_1_0)_rnd(0_1000000)_N(12)_diss(0_0.2000_0.0000)_drv(0_1.5000_1.0000_0.0000)"
          "_prm(1.0000_0.5000_1.0000)_start(0_0).txt")


def test_run_config_outputs_match_oracle(tmp_path, oracle_mod):
    from os_d_amd import qjx
    print("case: run_config N=12 M=24 S=4", flush=True)
    out = str(tmp_path) + "/"
    (tmp_path / "config.txt").write_text(CONFIG.format(path=out))
    (tmp_path / "params.txt").write_text(PARAMS)
    qjx.run_config(str(tmp_path / "config.txt"), str(tmp_path / "params.txt"), qjx.FP64, 0)
    files = sorted(os.path.basename(f) for f in glob.glob(out + "*" + SUFFIX))
    for name in ("norm", "mean", "mean_start", "dispersion", "m2", "energy", "periods", "mean_evo",
                 "norm_evo", "energy_evo", "m2_evo", "dispersion_evo", "adr_avg", "jump_times_0",
                 "diss_types_0", "jump_norms_0", "jump_etas_0"):
        assert name + SUFFIX in files, name
    # oracle through the same experiment (trans 2 periods, 3 dumps of 2 periods)
    m = oracle_mod.model(12, gamma=0.2, S=4, D=12)
    o = oracle_mod.Ensemble(m, 24, jump=2)
    o.trans(2)
    c0 = o.chars(is_start=1)
    evo = [c0]
    for _ in range(3):
        o.periods(2, 1)
        evo.append(o.chars())
    def load(name):
        return np.loadtxt(out + name + SUFFIX)
    np.testing.assert_allclose(load("mean"), evo[-1]["mean"], rtol=1e-9)
    np.testing.assert_allclose(load("energy"), evo[-1]["energy"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(load("mean_start"), c0["mean"], rtol=1e-12)
    me = load("mean_evo")                       # rows = dumps, columns = trajectories
    assert me.shape == (4, 24)
    for k in range(4):
        np.testing.assert_allclose(me[k], evo[k]["mean"], rtol=1e-9)
    assert list(np.atleast_1d(load("periods")).astype(int)) == [0, 2, 4, 6]
    for tr in (0, 5, 23):
        jt = np.atleast_1d(np.loadtxt(out + "jump_times_%d" % tr + SUFFIX)) \
            if os.path.getsize(out + "jump_times_%d" % tr + SUFFIX) else np.zeros(0)
        ref = o.get_jumps(tr)["times"]
        np.testing.assert_allclose(jt, ref, rtol=1e-15)
    text = open(out + "mean" + SUFFIX).read().splitlines()[0]
    assert "e" in text and len(text.split("e")[0]) == 18      # setprecision(16) scientific


SUFFIX_RK = ("_setup(0_1_1)_rnd(0_1000000)_N(12)_diss(0_0.2000_0.0000)_drv(1_1.5000_1.0000_0.3000)"
             "_prm(1.0000_0.5000_1.0000)_start(0_0).txt")


def test_run_config_rk_outputs_match_oracle(tmp_path, oracle_mod):
    """prop_id 1 in config.txt selects the Runge-Kutta path (RKPropagateBehavior,
    propagator.cpp:66-129) with rk_ns from config.txt and the sine drive of
    params.txt; the files equal the oracle's RK run through the same experiment."""
    from os_d_amd import qjx
    print("case: run_config RK N=12 M=24 rk_ns=2000", flush=True)
    out = str(tmp_path) + "/"
    cfg = CONFIG.format(path=out).replace("prop_id 0", "prop_id 1").replace("rk_ns 10000", "rk_ns 2000")
    prm = PARAMS.replace("dimer_drv_type 0", "dimer_drv_type 1").replace("dimer_drv_phase 0.0",
                                                                          "dimer_drv_phase 0.3")
    (tmp_path / "config.txt").write_text(cfg)
    (tmp_path / "params.txt").write_text(prm)
    # the GEMM form argument does not apply: the RK path is FP64
    qjx.run_config(str(tmp_path / "config.txt"), str(tmp_path / "params.txt"), qjx.FP32, 0)
    m = oracle_mod.model(12, gamma=0.2, D=1)
    o = oracle_mod.Ensemble(m, 24, jump=2, build_propagators=False)
    o.rk_setup(2000, 1, 0.3)
    o.trans(2)
    c0 = o.chars(is_start=1)
    evo = [c0]
    for _ in range(3):
        o.periods(2, 1)
        evo.append(o.chars())

    def load(name):
        return np.loadtxt(out + name + SUFFIX_RK)
    np.testing.assert_allclose(load("mean"), evo[-1]["mean"], rtol=1e-12)
    np.testing.assert_allclose(load("mean_start"), c0["mean"], rtol=1e-12)
    me = load("mean_evo")
    for k in range(4):
        np.testing.assert_allclose(me[k], evo[k]["mean"], rtol=1e-12)
    for tr in (0, 7, 23):
        f = out + "jump_times_%d" % tr + SUFFIX_RK
        jt = np.atleast_1d(np.loadtxt(f)) if os.path.getsize(f) else np.zeros(0)
        assert np.array_equal(jt, o.get_jumps(tr)["times"])


def _deep_files(tmp_path, prop_id, k, rk_ns=2000, drv_type=0):
    out = str(tmp_path) + "/"
    cfg = CONFIG.format(path=out).replace("task_id 1", "task_id 4").replace(
        "prop_id 0", "prop_id %d" % prop_id).replace("rk_ns 10000", "rk_ns %d" % rk_ns).replace(
        "num_obs_periods 6", "num_obs_periods 2")
    prm = PARAMS.replace("dimer_drv_type 0", "dimer_drv_type %d" % drv_type) + "deep_num_steps %d\n" % k
    (tmp_path / "config.txt").write_text(cfg)
    (tmp_path / "params.txt").write_text(prm)
    return out


@pytest.mark.parametrize("prop_id", [0, 1])
def test_run_config_deep_outputs_match_oracle(tmp_path, oracle_mod, prop_id):
    """task_id 4 (StdDeepExperimentBehaviour, experiment.cpp:769-862): a dump after
    every sub-period, 2 deep_num_steps per period.  QJ: the tree and the steps per
    sub-period come from the reference's node recursion (tests/deep_tree.py: head
    T/(2k), two children of 2 steps of T/(8k) at level 0); RK: rk_ns steps per
    sub-period of T/(2k) (rk_proc.cpp:65-76, core.cpp:490-530)."""
    from os_d_amd import qjx
    k, rk_ns = 3, 400
    print("case: run_config deep prop_id=%d k=%d" % (prop_id, k), flush=True)
    out = _deep_files(tmp_path, prop_id, k, rk_ns=rk_ns, drv_type=1)
    qjx.run_config(str(tmp_path / "config.txt"), str(tmp_path / "params.txt"), qjx.FP64, 0)
    sfx = "_setup(0_4_%d)_rnd(0_1000000)_N(12)_diss(0_0.2000_0.0000)_drv(1_1.5000_1.0000_0.0000)" \
          "_prm(1.0000_0.5000_1.0000)_start(0_0).txt" % prop_id
    if prop_id == 0:
        tree = reference_deep_tree(2 * np.pi, k, 12)
        m = oracle_mod.model(12, gamma=0.2, S=tree["S"], D=12)
        o = oracle_mod.Ensemble(m, 24, jump=2)
        per_sub = tree["level0_steps_per_sub"]
    else:
        m = oracle_mod.model(12, gamma=0.2, D=1)
        o = oracle_mod.Ensemble(m, 24, jump=2, build_propagators=False)
        o.rk_setup(rk_ns * 2 * k, 1, 0.0)
        per_sub = rk_ns
    o.trans(2)
    evo = [o.chars(is_start=1)]
    for _ in range(2):
        for _ in range(2 * k):
            o.steps(per_sub, 1)
            evo.append(o.chars())
    periods = np.atleast_1d(np.loadtxt(out + "periods" + sfx)).astype(int)
    assert list(periods) == [0] + [p for p in range(2) for _ in range(2 * k)]
    me = np.loadtxt(out + "mean_evo" + sfx)
    assert me.shape == (4 * k + 1, 24)
    rtol = 1e-9 if prop_id == 0 else 1e-12
    for i in range(4 * k + 1):
        np.testing.assert_allclose(me[i], evo[i]["mean"], rtol=rtol)
    for tr in (0, 11):
        f = out + "jump_times_%d" % tr + sfx
        jt = np.atleast_1d(np.loadtxt(f)) if os.path.getsize(f) else np.zeros(0)
        np.testing.assert_allclose(jt, o.get_jumps(tr)["times"], rtol=1e-15 if prop_id else 1e-12)
        if prop_id == 0 and jt.size:
            # jump times land on the finest grid of the deep tree, T / (8k 2^(D-1))
            q = jt / tree["level_dt"][-1]
            np.testing.assert_allclose(q, np.rint(q), rtol=0, atol=1e-6)
    assert not glob.glob(out + "phi_*")


def test_run_config_jumps_counts_stop(tmp_path, oracle_mod):
    """jumps_counts L > 0 (obser_process, experiment.cpp:481-499): periods run one at a
    time and the observation stops at the first period after which every trajectory
    has logged at least L jumps; the dumps after it stay zero."""
    from os_d_amd import qjx
    L = 3
    print("case: run_config jumps_counts=%d" % L, flush=True)
    out = str(tmp_path) + "/"
    cfg = CONFIG.format(path=out).replace("num_obs_periods 6", "num_obs_periods 40")
    prm = PARAMS.replace("jumps_counts 0", "jumps_counts %d" % L).replace("dump_num 3", "dump_num 40")
    (tmp_path / "config.txt").write_text(cfg)
    (tmp_path / "params.txt").write_text(prm)
    qjx.run_config(str(tmp_path / "config.txt"), str(tmp_path / "params.txt"), qjx.FP64, 0)
    m = oracle_mod.model(12, gamma=0.2, S=4, D=12)
    o = oracle_mod.Ensemble(m, 24, jump=2)
    o.trans(2)
    o.chars(is_start=1)
    p_stop = 0
    while p_stop < 40:
        o.periods(1, 1)
        p_stop += 1
        if (o.get_state()["njumps"] >= L).all():
            break
    assert p_stop < 40
    me = np.loadtxt(out + "mean_evo" + SUFFIX)
    np.testing.assert_allclose(me[p_stop], o.chars()["mean"], rtol=1e-9)
    assert (me[p_stop + 1:] == 0).all()
    np.testing.assert_allclose(np.loadtxt(out + "mean" + SUFFIX), o.chars()["mean"], rtol=1e-9)


LPN_KEYS = """lpn_type {lpn_type}
lpn_eps_deep 100
lpn_eps_error 1e-10
lpn_eps_high 100
lpn_eps_low -100
lpn_delta_s 1.0e-2
lpn_delta_f_high 5.0e-2
lpn_delta_f_low 2.0e-3
lambda_per_periods {lpp}
save_lambdas 1
random_obs_seed 1
random_obs_mns 1000000
random_obs_type 2
"""


@pytest.mark.parametrize("lpn_type,lpp", [(0, 0), (-1, 1)])
def test_run_config_lpn_outputs_match_oracle(tmp_path, oracle_mod, lpn_type, lpp):
    """task_id 7 (LpnMultExperimentBehaviour, experiment.cpp:137-286) with RK
    propagation: lambda, num_renorms, the lpn and std evolutions and the random
    observables written by the runner equal the oracle's LPN run."""
    from os_d_amd import qjx
    print("case: run_config lpn type=%d lpp=%d" % (lpn_type, lpp), flush=True)
    out = str(tmp_path) + "/"
    M = 16
    cfg = CONFIG.format(path=out).replace("task_id 1", "task_id 7").replace("prop_id 0", "prop_id 1").replace(
        "rk_ns 10000", "rk_ns 2000").replace("num_trajectories 24", "num_trajectories %d" % M).replace(
        "num_obs_periods 6", "num_obs_periods 4")
    prm = PARAMS.replace("dimer_drv_type 0", "dimer_drv_type 1").replace("num_random_obs 0", "num_random_obs 1") \
        .replace("dump_num 3", "dump_num 2").replace("jump 2", "jump 1").replace("N 12", "N 10") \
        + LPN_KEYS.format(lpn_type=lpn_type, lpp=lpp)
    (tmp_path / "config.txt").write_text(cfg)
    (tmp_path / "params.txt").write_text(prm)
    qjx.run_config(str(tmp_path / "config.txt"), str(tmp_path / "params.txt"), qjx.FP64, 0)
    sfx = "_setup(0_7_1)_rnd(0_1000000)_N(10)_diss(0_0.2000_0.0000)_drv(1_1.5000_1.0000_0.0000)" \
          "_prm(1.0000_0.5000_1.0000)_start(0_0).txt"
    m = oracle_mod.model(10, gamma=0.2, D=1)
    o = oracle_mod.Ensemble(m, M, jump=1, build_propagators=False)
    o.rk_setup(2000, 1, 0.0)
    o.lpn_setup(lpn_type=lpn_type, lambda_per_periods=lpp, save_lambdas=1, lpn_delta_f_high=5e-2,
                lpn_delta_f_low=2e-3, num_random_obs=1)
    o.lpn_trans(2)
    evo = [o.lpn_get()]
    obs = [o.lpn_random_obs()[0][:, 0]]
    for _ in range(2):
        o.lpn_period()
        o.lpn_period()
        evo.append(o.lpn_get())
        obs.append(o.lpn_random_obs()[0][:, 0])

    def load(name):
        return np.loadtxt(out + name + sfx)
    g = evo[-1]
    assert np.array_equal(load("lambda"), g["lambda_now"])
    assert np.array_equal(load("num_renorms").astype(int), g["num_renorms"])
    assert np.array_equal(load("mean_lpn"), g["mean_lpn"])
    le, me = load("lambda_evo"), load("mean_evo")
    assert le.shape == (3, M)
    for k in range(3):
        assert np.array_equal(le[k], evo[k]["lambda_now"])
        assert np.array_equal(me[k], evo[k]["mean"])
    rows = open(out + "random_obs_evo_1" + sfx).read().split("\n")[:3]
    for k, row in enumerate(rows):
        vals = [complex(*map(float, c.strip("()").split(","))) for c in row.split("\t")]
        assert np.array_equal(np.array(vals), obs[k])
    for tr in range(M // 2, M):
        r, ds, df = o.lpn_lambdas(tr)
        f = out + "delta_f_by_delta_s_%d" % tr + sfx
        got = np.atleast_1d(np.loadtxt(f)) if os.path.getsize(f) else np.zeros(0)
        assert np.array_equal(got, r)


def test_run_config_uniform_start_and_state_dumps_match_oracle(tmp_path, oracle_mod):
    """start_type 1 (the uniform state, init_start_state, newdel_exp.cpp:656-679) with
    every state dump switched on: phi_<tr> at the end (dump_phi, core.cpp:12567-12589),
    phi_evo_<tr> after the transient, at every dump and once more at the end
    (dump_phi_evo, core.cpp:12591-12613, called at experiment.cpp:466, 515, 531),
    adr_<tr> after the transient and at every dump (dump_adr_single, dump.cpp:3-25) and
    adr_avg at the same points (dump_evo_avg 1, dump.cpp:26-61): every file equals the
    oracle's run through the same experiment."""
    from os_d_amd import qjx
    print("case: run_config uniform start + phi/adr dumps", flush=True)
    out = str(tmp_path) + "/"
    M, d = 24, 13
    prm = PARAMS.replace("start_type 0", "start_type 1").replace("dump_phi 0", "dump_phi 1").replace(
        "dump_phi_evo 0", "dump_phi_evo 1").replace("dump_adr_sep 0", "dump_adr_sep 1").replace(
        "dump_evo_avg 0", "dump_evo_avg 1")
    (tmp_path / "config.txt").write_text(CONFIG.format(path=out))
    (tmp_path / "params.txt").write_text(prm)
    qjx.run_config(str(tmp_path / "config.txt"), str(tmp_path / "params.txt"), qjx.FP64, 0)
    sfx = SUFFIX.replace("_start(0_0)", "_start(1_0)")
    m = oracle_mod.model(12, gamma=0.2, S=4, D=12)
    o = oracle_mod.Ensemble(m, M, jump=2, start_type=1)
    psi0 = o.get_state()["psi"]
    np.testing.assert_array_equal(psi0, np.full((M, d), np.sqrt(1.0 / d)))
    o.trans(2)
    evo = [o.chars(is_start=1, adr=True)]
    phis = [o.get_state()["psi"]]
    for _ in range(3):
        o.periods(2, 1)
        evo.append(o.chars(adr=True))
        phis.append(o.get_state()["psi"])
    phis.append(phis[-1])                          # the final dump_phi_evo(ad, true)

    def cplx(fn):
        a = np.loadtxt(fn).reshape(-1, 2)
        return a[:, 0] + 1j * a[:, 1]
    for tr in (0, 9, 23):
        np.testing.assert_allclose(cplx(out + "phi_%d" % tr + sfx), phis[-1][tr], rtol=0, atol=1e-12)
        pe = cplx(out + "phi_evo_%d" % tr + sfx).reshape(len(phis), d)
        for k in range(len(phis)):
            np.testing.assert_allclose(pe[k], phis[k][tr], rtol=0, atol=1e-12)
        ad = np.loadtxt(out + "adr_%d" % tr + sfx).reshape(len(evo), d)
        for k in range(len(evo)):
            np.testing.assert_allclose(ad[k], evo[k]["adr"][tr], rtol=1e-10, atol=1e-15)
    aa = np.loadtxt(out + "adr_avg" + sfx).reshape(len(evo), d)
    for k in range(len(evo)):
        np.testing.assert_allclose(aa[k], evo[k]["adr"].mean(axis=0), rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(np.loadtxt(out + "mean_start" + sfx), evo[0]["mean"], rtol=1e-12)
    me = np.loadtxt(out + "mean_evo" + sfx)
    for k in range(len(evo)):
        np.testing.assert_allclose(me[k], evo[k]["mean"], rtol=1e-9)


def test_run_config_random_observables_task1_match_oracle(tmp_path, oracle_mod):
    """num_random_obs > 0 with the STD experiment (task_id 1): calc_random_obs at every
    dump (core.cpp:694, 724) written as random_obs_evo_<seed + o> (core.cpp:1050-1061),
    rows = dumps, columns = trajectories; equal to the oracle's values of the same
    states bit for bit."""
    from os_d_amd import qjx
    print("case: run_config task 1 random observables", flush=True)
    out = str(tmp_path) + "/"
    prm = PARAMS.replace("num_random_obs 0", "num_random_obs 2") + \
        "random_obs_seed 5\nrandom_obs_mns 1000000\nrandom_obs_type 2\n"
    (tmp_path / "config.txt").write_text(CONFIG.format(path=out))
    (tmp_path / "params.txt").write_text(prm)
    qjx.run_config(str(tmp_path / "config.txt"), str(tmp_path / "params.txt"), qjx.FP64, 0)
    m = oracle_mod.model(12, gamma=0.2, S=4, D=12)
    o = oracle_mod.Ensemble(m, 24, jump=2)
    # the runner's tree is its own; the dumps' states agree to 1e-10, the random
    # observables are compared on the oracle's states through the same formula
    o.trans(2)
    obs = [o.random_obs(2, seed=5, typ=2)]
    for _ in range(3):
        o.periods(2, 1)
        obs.append(o.random_obs(2, seed=5, typ=2))
    for k in range(2):
        rows = open(out + "random_obs_evo_%d" % (5 + k) + SUFFIX).read().strip().split("\n")
        assert len(rows) == 4
        for r, row in enumerate(rows):
            vals = np.array([complex(*map(float, c.strip("()").split(","))) for c in row.split("\t")])
            np.testing.assert_allclose(vals, obs[r][:, k], rtol=1e-9, atol=1e-12)
