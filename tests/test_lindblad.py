"""The Lindblad rho-space oracle, and quantum-jump ensembles against it.

The QJ unravelling reproduces the Lindblad evolution on average:
E[|psi><psi| / <psi|psi>] = rho(t).  Ensemble means of the per-trajectory
observables (core.cpp:665-725) must therefore agree with Tr(rho O) within
their statistical error sigma = std / sqrt(M).
"""
import numpy as np
import pytest
import scipy.linalg as sl


def _liouvillian(H, L, g):
    d = H.shape[0]
    I = np.eye(d)
    # row-major vec: vec(A X B) = (A kron B^T) vec(X)
    LdL = L.conj().T @ L
    return (-1j * (np.kron(H, I) - np.kron(I, H.T))
            + g * (np.kron(L, L.conj()) - 0.5 * np.kron(LdL, I) - 0.5 * np.kron(I, LdL.T)))


@pytest.mark.parametrize("N", [1, 3, 7])
def test_rhs_matches_dense_liouvillian(oracle_mod, N):
    m = oracle_mod.model(N, prm_E=0.4, prm_U=0.7, prm_J=0.9, drv_ampl=1.3, gamma=0.37)
    mm = oracle_mod.model_matrices(m)
    rng = np.random.default_rng(N)
    X = rng.standard_normal((N + 1, N + 1)) + 1j * rng.standard_normal((N + 1, N + 1))
    rho = X @ X.conj().T
    rho /= np.trace(rho)
    E = 0.4 + 1.3
    Lv = _liouvillian(mm["H"] + E * mm["Hdrv"], mm["L"], 0.37 / N)
    ref = (Lv @ rho.reshape(-1)).reshape(N + 1, N + 1)
    got = oracle_mod.lindblad_rhs(m, E, rho)
    np.testing.assert_allclose(got, ref, atol=1e-11)


def test_rk4_period_matches_exact_exponential(oracle_mod):
    N = 3
    m = oracle_mod.model(N, gamma=0.5)
    mm = oracle_mod.model_matrices(m)
    T = 2 * np.pi
    g = 0.5 / N
    P0 = sl.expm(_liouvillian(mm["H"] + 2.5 * mm["Hdrv"], mm["L"], g) * T / 2)
    P1 = sl.expm(_liouvillian(mm["H"] - 0.5 * mm["Hdrv"], mm["L"], g) * T / 2)
    rho0 = np.zeros((N + 1, N + 1), complex)
    rho0[0, 0] = 1.0
    ref = (P1 @ (P0 @ rho0.reshape(-1))).reshape(N + 1, N + 1)
    got = oracle_mod.lindblad_evolve(m, rho0, 1, steps_per_period=4000)
    np.testing.assert_allclose(got, ref, atol=1e-10)
    assert abs(np.trace(got) - 1.0) < 1e-12
    np.testing.assert_allclose(got, got.conj().T, atol=1e-14)
    assert np.linalg.eigvalsh(got).min() > -1e-12


def test_dark_steady_state(oracle_mod):
    # U = E = A = 0: the dark state |n_- = 0> is the unique steady state: purity -> 1,
    # <n1 - n2> -> 0 (<k> -> N/2).
    N = 4
    m = oracle_mod.model(N, prm_E=0.0, prm_U=0.0, drv_ampl=0.0, gamma=2.0)
    rho = np.zeros((N + 1, N + 1), complex)
    rho[0, 0] = 1.0
    rho = oracle_mod.lindblad_evolve(m, rho, 40, steps_per_period=400)
    ob = oracle_mod.lindblad_observables(m, rho)
    assert ob["purity"] > 1 - 1e-6
    assert abs(ob["mean"] - N / 2) < 1e-6


def _qj_vs_lindblad(ens_means, ens_stds, M, lind, nsig):
    for k, v in ens_means.items():
        sig = ens_stds[k] / np.sqrt(M)
        assert abs(v - lind[k]) < nsig * sig + 1e-12, (k, v, lind[k], sig)


def test_oracle_qj_ensemble_matches_lindblad(oracle_mod):
    N, M, P = 10, 4000, 3
    m = oracle_mod.model(N, S=4, D=16)
    e = oracle_mod.Ensemble(m, M, jump=0)
    e.trans(0)
    rho = np.zeros((N + 1, N + 1), complex)
    rho[0, 0] = 1.0
    for p in range(P):
        e.periods(1, 1)
        rho = oracle_mod.lindblad_evolve(m, rho, 1, steps_per_period=2000)
        c = e.chars()
        lind = oracle_mod.lindblad_observables(m, rho)
        _qj_vs_lindblad({"mean": c["mean"].mean(), "energy": c["energy"].mean()},
                        {"mean": c["mean"].std(), "energy": c["energy"].std()}, M, lind, 3.0)
