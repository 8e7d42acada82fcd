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


def _dense_drive_solution(oracle_mod, m, N, E_of_t, rho0, t1, t_switch=None):
    """rho(t1) of d rho/dt = L(E(t)) rho from the dense Liouvillian by DOP853 (tight
    tolerances), restarted at the square drive's switch so that its jump is exact."""
    from scipy.integrate import solve_ivp
    mm = oracle_mod.model_matrices(m)
    g = m.gamma / N
    L0 = _liouvillian(mm["H"], mm["L"], g)
    Ld = -1j * (np.kron(mm["Hdrv"], np.eye(N + 1)) - np.kron(np.eye(N + 1), mm["Hdrv"].T))

    def f(t, y):
        return (L0 + E_of_t(t) * Ld) @ y
    y = rho0.reshape(-1).astype(complex)
    knots = [0.0] + ([t_switch] if t_switch else []) + [t1]
    for a, b in zip(knots[:-1], knots[1:]):
        y = solve_ivp(f, (a, b), y, method="DOP853", rtol=1e-12, atol=1e-14).y[:, -1]
    return y.reshape(N + 1, N + 1)


def test_cqdiss_sine_drive_converges_at_fourth_order(oracle_mod):
    """CQdiss's drive rule with the sine drive (calcODE.cpp:100-105): RK4 with the
    drive evaluated at each stage time converges to the exact solution at order 4
    (error ratio near 16 per halving), which fixes the stage times t, t + h/2, t + h
    and the phase."""
    N = 3
    m = oracle_mod.model(N, prm_E=0.3, drv_ampl=2.0, gamma=0.4)
    phase = 0.7
    rho0 = np.zeros((N + 1, N + 1), complex)
    rho0[0, 0] = 1.0
    T = 2 * np.pi
    ref = _dense_drive_solution(oracle_mod, m, N, lambda t: 0.3 + 2.0 * np.sin(t + phase), rho0, T)
    errs = []
    for S in (100, 200, 400):
        got = oracle_mod.lindblad_evolve_drive(m, rho0, 1, steps_per_period=S, drv_type=1, drv_phase=phase)
        errs.append(np.abs(got - ref).max())
    assert errs[-1] < 5e-7
    for a, b in zip(errs[:-1], errs[1:]):
        assert 12.0 < a / b < 20.0, errs


def test_cqdiss_square_drive_matches_a_dense_restatement(oracle_mod):
    """CQdiss's square drive (calcODE.cpp:106-116) takes the sign of sin(omega t) at
    every stage time t = p T + s h (+ h/2, + h).  Where rounding puts a stage time of
    a step next to a switch on the other side of it, that stage takes the other half
    period's value (here the last stage of each period, whose time rounds just above
    2 pi).  A second restatement written from calcODE.cpp -- RK4 on the dense
    Liouvillian with numpy's sine -- gives the same rho, and the rule matters: the
    result differs from holding the drive constant per step."""
    N = 3
    m = oracle_mod.model(N, prm_E=0.3, drv_ampl=2.0, gamma=0.4)
    mm = oracle_mod.model_matrices(m)
    g = m.gamma / N
    L0 = _liouvillian(mm["H"], mm["L"], g)
    Ld = -1j * (np.kron(mm["Hdrv"], np.eye(N + 1)) - np.kron(np.eye(N + 1), mm["Hdrv"].T))
    rho0 = np.zeros((N + 1, N + 1), complex)
    rho0[0, 0] = 1.0
    T = 2 * np.pi
    for S, P in ((400, 2), (800, 1)):
        h = T / S
        y = rho0.reshape(-1).copy()

        def E(t):
            return 0.3 + 2.0 * (-1.0 if np.sin(t) < 0.0 else 1.0)
        for p in range(P):
            for s in range(S):
                t = p * T + s * h
                k1 = (L0 + E(t) * Ld) @ y
                k2 = (L0 + E(t + h / 2.0) * Ld) @ (y + 0.5 * h * k1)
                k3 = (L0 + E(t + h / 2.0) * Ld) @ (y + 0.5 * h * k2)
                k4 = (L0 + E(t + h) * Ld) @ (y + h * k3)
                y = y + (h / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)
        got = oracle_mod.lindblad_evolve_drive(m, rho0, P, S, 0, 0.0, 1)
        np.testing.assert_allclose(got, y.reshape(N + 1, N + 1), rtol=0, atol=1e-12)
        per_step = oracle_mod.lindblad_evolve_drive(m, rho0, P, S, 0, 0.0, 0)
        assert np.abs(got - per_step).max() > 1e-4
    # time continues across calls: two periods equal one period then one from period 1
    one = oracle_mod.lindblad_evolve_drive(m, rho0, 2, 400, 0, 0.0, 1)
    two = oracle_mod.lindblad_evolve_drive(m, oracle_mod.lindblad_evolve_drive(m, rho0, 1, 400, 0, 0.0, 1), 1, 400,
                                           0, 0.0, 1, period0=1)
    assert np.array_equal(one, two)
