"""The LPN task (task_id 7) on the GPU against the CPU restatement.

Propagation and the per-pair bookkeeping both run on the device, the bookkeeping in
the restatement's arithmetic (reference: LpnMultExperimentBehaviour, experiment.cpp:
137-286; var_trajectory_lpn, lambda_lpn(_now), experiment.cpp:2278-2393, 3065-3134).
With the RK path the propagation is bit-identical, so every LPN quantity must be
too; with the FP64 quantum-jump path the GEMM's summation order differs from the
oracle's and the results agree to rounding.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# the reference's own LPN parameter set (source/cpp/QJX/QJX/params.txt): E = 0,
# U = 0.3, J = 1, sine drive A = 3.4, omega = 1, gamma = 0.1
P = dict(prm_E=0.0, prm_U=0.3, prm_J=1.0, drv_ampl=3.4, drv_freq=1.0)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(orc, prop, M, lpn, N=10, rk_ns=2000):
    from os_d_amd import qjx
    m = orc.model(N, gamma=0.1, S=4, D=12, **P)
    o = orc.Ensemble(m, M, seed=1, jump=1, build_propagators=prop == "qj")
    if prop == "rk":
        o.rk_setup(rk_ns, 1, 0.0)
        e = qjx.Engine(N, diss_gamma=0.1, prop="rk", rk_ns=rk_ns, drv_type=1, **P)
    else:
        e = qjx.Engine(N, diss_gamma=0.1, ex_deep=12, steps_per_period=4, precision=qjx.FP64, **P)
    e.ensemble_init(M, seed=1, jump=1)
    if prop == "qj":          # one propagator tree on both sides: only summation orders differ
        for b in range(2):
            for lvl in range(12):
                e.set_propagator(b, lvl, o.get_propagator(b, lvl))
    o.lpn_setup(**lpn)
    L = qjx.Lpn(e, **lpn)
    return o, e, L


@pytest.mark.parametrize("lpn", [
    dict(lpn_type=-1, lambda_per_periods=1, save_lambdas=1),
    dict(lpn_type=0, num_random_obs=2, lambda_per_periods=0, lpn_delta_f_high=5e-2, lpn_delta_f_low=2e-3,
         save_lambdas=1),
    dict(lpn_type=-2, lambda_per_periods=0, lpn_delta_f_high=5e-2, lpn_delta_f_low=2e-3),
])
def test_lpn_rk_bitwise(oracle_mod, lpn):
    _say("case: lpn rk %s" % lpn)
    M = 16
    o, e, L = _pair(oracle_mod, "rk", M, lpn)
    o.lpn_trans(1)
    L.trans(1)
    for _ in range(4):
        o.lpn_period()
    L.run_periods(4)
    go, gg = o.lpn_get(), L.get()
    for k in ("lambda_now", "lambda", "num_renorms", "mean_lpn", "energy_lpn", "delta_s"):
        assert np.array_equal(go[k], gg[k]), k
    so, sg = o.get_state(), e.get_state()
    assert np.array_equal(so["psi"], sg["psi"]) and np.array_equal(so["rng"], sg["rng"])
    if lpn.get("num_random_obs", 0):
        ao, bo = o.lpn_random_obs()
        ag, bg = L.random_obs()
        assert np.array_equal(ao, ag) and np.array_equal(bo, bg)
        assert np.array_equal(L.random_obs_matrix(1), oracle_mod.random_obs_matrix(11, 1, 1000000, 2, 1))
    if lpn.get("save_lambdas", 0):
        for tr in range(M // 2, M):
            for a, b in zip(o.lpn_lambdas(tr), L.lambdas(tr)):
                assert np.array_equal(a, b)
    _say("lambda_now (partners): %s" % np.array2string(gg["lambda_now"][M // 2:], precision=4))
    L.close()
    e.close()


def test_lpn_qj_fp64_to_rounding(oracle_mod):
    _say("case: lpn qj fp64")
    M = 16
    lpn = dict(lpn_type=-1, lambda_per_periods=1)
    o, e, L = _pair(oracle_mod, "qj", M, lpn)
    o.lpn_trans(2)
    L.trans(2)
    for _ in range(3):
        o.lpn_period()
    L.run_periods(3)
    go, gg = o.lpn_get(), L.get()
    assert np.array_equal(go["num_renorms"], gg["num_renorms"])
    np.testing.assert_allclose(gg["lambda"], go["lambda"], rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(gg["mean_lpn"], go["mean_lpn"], rtol=1e-9)
    L.close()
    e.close()


@pytest.mark.parametrize("N,M,num,typ", [(10, 40, 2, 2), (63, 130, 1, 1), (200, 70, 1, 0)])
def test_random_obs_equal_oracle_bitwise(oracle_mod, N, M, num, typ):
    """qjx_random_obs (the STD path's calc_random_obs) on the device equals the
    oracle's get_random_obs of the same states bit for bit: the matrices R = X X^dag
    and the products <phi|R|phi> / ||phi||^2 run in the reference's order."""
    from os_d_amd import qjx
    print("case: random obs N=%d M=%d num=%d type=%d" % (N, M, num, typ), flush=True)
    m = oracle_mod.model(N, gamma=0.2, D=1)
    o = oracle_mod.Ensemble(m, M, jump=0, build_propagators=False)
    rng = np.random.default_rng(N)
    psi = rng.normal(size=(M, N + 1)) + 1j * rng.normal(size=(M, N + 1))
    psi *= rng.uniform(0.3, 1.0, size=(M, 1)) / np.linalg.norm(psi, axis=1, keepdims=True)
    o.set_state(psi=psi)
    e = qjx.Engine(N, ex_deep=2, steps_per_period=4, precision=qjx.FP64, diss_gamma=0.2)
    e.ensemble_init(M, jump=0)
    e.set_state(psi=psi)
    want = o.random_obs(num, seed=3, typ=typ)
    got = e.random_obs(num, seed=3, typ=typ)
    assert np.array_equal(got, want)
    # the matrices agree with the LPN object's (the same init_random_obs)
    for k in range(num):
        R = oracle_mod.random_obs_matrix(N + 1, 3, 1000000, typ, k)
        v = psi[5] / np.linalg.norm(psi[5])
        np.testing.assert_allclose(got[5, k], np.vdot(v, R @ v), rtol=1e-12)
    e.close()
