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
