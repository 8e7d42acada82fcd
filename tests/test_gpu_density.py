"""The ensemble density matrix and purity computed on the GPU.

rho = (1/M) sum_tr psi psi^+ / ||psi||^2 over the current states must equal the
same sum taken in numpy over the exported states (FP64, to rounding), for every
precision form; rho is Hermitian with unit trace and Tr rho^2 is its purity.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_density_matches_exported_states(prec):
    from os_d_amd import qjx
    N, M = 90, 700
    _say("case: density prec=%d N=%d M=%d" % (prec, N, M))
    e = qjx.Engine(N, ex_deep=10, steps_per_period=8, precision=prec, diss_gamma=0.3)
    e.ensemble_init(M, seed=4, jump=0)
    e.trans(0)
    e.run_periods(2, 0)
    rho, pur = e.density()
    psi = e.get_state()["psi"]
    w = 1.0 / np.sum(np.abs(psi) ** 2, axis=1)
    ref = (psi.T * w) @ psi.conj() / M
    err = np.abs(rho - ref).max()
    _say("max |rho_gpu - rho_numpy| = %.2e, purity %.6f" % (err, pur))
    assert err < 1e-12
    assert np.abs(rho - rho.conj().T).max() < 1e-14
    assert abs(np.trace(rho).real - 1.0) < 1e-12
    assert abs(pur - np.real(np.trace(ref @ ref))) < 1e-12
    _, pur2 = e.density(rho=False)
    assert pur2 == pur
    e.close()


def test_density_of_identical_pure_states_is_pure():
    from os_d_amd import qjx
    _say("case: density of the start state")
    e = qjx.Engine(40, ex_deep=8, steps_per_period=4, precision=0)
    e.ensemble_init(333, seed=1, start_type=0, start_state=7, jump=0)
    rho, pur = e.density()
    expect = np.zeros_like(rho)
    expect[7, 7] = 1.0
    assert np.abs(rho - expect).max() < 1e-15
    assert abs(pur - 1.0) < 1e-14
    e.close()
