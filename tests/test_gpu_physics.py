"""GPU ensembles against the Lindblad rho-space oracle (statistical parity).

For every precision form, ensemble means of <k> (= <n2>), energy and the
ensemble density matrix's purity agree with the Lindblad solution within 3 sigma
(sigma = sample std / sqrt(M)); the FP32 / BF16x3 forms additionally state their
observable error against an FP64 run with the same streams.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_gpu_ensemble_matches_lindblad(oracle_mod, prec):
    from os_d_amd import qjx
    N, M, P, S = 20, 20000, 3, 8
    _say("case: gpu vs lindblad prec=%d N=%d M=%d" % (prec, N, M))
    m = oracle_mod.model(N, S=S, D=16)
    e = qjx.Engine(N, ex_deep=16, steps_per_period=S, precision=prec)
    e.ensemble_init(M, seed=11, jump=0)
    e.trans(0)
    rho = np.zeros((N + 1, N + 1), complex)
    rho[0, 0] = 1.0
    for p in range(P):
        e.run_periods(1, 1)
        rho = oracle_mod.lindblad_evolve(m, rho, 1, steps_per_period=2000)
        c = e.chars()
        lind = oracle_mod.lindblad_observables(m, rho)
        for k in ("mean", "energy"):
            v, sig = c[k].mean(), c[k].std() / np.sqrt(M)
            _say("period %d %s: qj %.6f lindblad %.6f (%.2f sigma)" % (p + 1, k, v, lind[k], (v - lind[k]) / sig))
            assert abs(v - lind[k]) < 3.0 * sig
    # ensemble density matrix on the GPU: rho_qj = mean_tr psi psi^dag / <psi|psi>
    rho_qj, pur_qj = e.density()
    err = np.abs(rho_qj - rho).max()
    _say("purity: qj %.5f lindblad %.5f; max |rho_qj - rho_lindblad| %.2e" % (pur_qj, lind["purity"], err))
    assert abs(pur_qj - lind["purity"]) < 0.01
    assert err < 0.02
    e.close()


def test_c1_ensemble_matches_cqdiss_lindblad_at_n100():
    """The plain CPU configuration's physics (N = 100, d = 101, J = 1, U = 0.5, E0 = 1,
    A = 1.5, gamma = 0.1, S = 200, ex_deep 16, complex128) against its oracle, the
    Lindblad RK4 with 1e4 steps per period and CQdiss's per-stage square drive: the
    FP64 ensemble's <k> and energy agree with the GPU Lindblad solution within 3 sigma
    at every period of the horizon."""
    from os_d_amd import qjx
    N, M, P = 100, 10000, 4
    prm = dict(prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0)
    _say("case: c1 physics N=%d M=%d P=%d" % (N, M, P))
    lb = qjx.Lindblad(N, steps_per_period=10000, diss_gamma=0.1, per_stage=1, **prm)
    e = qjx.Engine(N, ex_deep=16, steps_per_period=200, precision=qjx.FP64, diss_gamma=0.1, **prm)
    e.ensemble_init(M, seed=0, jump=0)
    e.trans(0)
    worst = 0.0
    for p in range(P):
        lb.run_periods(1)
        e.run_periods(1, 1)
        ob = lb.observables()
        c = e.chars()
        for k in ("mean", "energy"):
            v, sig = c[k].mean(), c[k].std() / np.sqrt(M)
            worst = max(worst, abs(v - ob[k]) / sig)
            _say("period %d %s: qj %.6f lindblad %.6f (%.2f sigma)" % (p + 1, k, v, ob[k], (v - ob[k]) / sig))
            assert abs(v - ob[k]) < 3.0 * sig
    _, pur = e.density(rho=False)
    _say("purity: qj %.5f lindblad %.5f; worst %.2f sigma" % (pur, ob["purity"], worst))
    assert abs(pur - ob["purity"]) < 0.02
    e.close()
    lb.close()
