"""Pins of the oracle's model matrices and matrix exponential.

The dimer matrices are rebuilt here from the second-quantised operators (basis
|N-k, k>, k = bosons on site 2), so a wrong sign, factor or index in the oracle's
restatement of newdel.cpp:450-498, 1454-1498, 2285-2388 fails these checks.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "expm_dimer.npz")


def two_mode(N):
    d = N + 1
    b1d_b2 = np.zeros((d, d))
    for k in range(1, d):
        b1d_b2[k - 1, k] = np.sqrt((N - k + 1) * k)
    n2 = np.diag(np.arange(d, dtype=float))
    n1 = np.diag(N - np.arange(d, dtype=float))
    return b1d_b2, n1, n2


@pytest.mark.parametrize("N", [1, 2, 5, 30])
def test_matrices_from_second_quantisation(oracle_mod, N):
    J, U, E, A, g = 0.7, 0.45, 0.3, 1.2, 0.13
    m = oracle_mod.model(N, prm_E=E, prm_U=U, prm_J=J, drv_ampl=A, gamma=g)
    mm = oracle_mod.model_matrices(m)
    b1d_b2, n1, n2 = two_mode(N)
    I = np.eye(N + 1)
    H = -J * (b1d_b2 + b1d_b2.T) + 2.0 * (U / N) * (n1 @ (n1 - I) + n2 @ (n2 - I))
    np.testing.assert_allclose(mm["H"], H, atol=1e-12)
    np.testing.assert_allclose(mm["Hdrv"], n1 - n2, atol=0)
    L = -((n1 - n2) + b1d_b2.T - b1d_b2)          # -(b1^+ + b2^+)(b1 - b2)
    np.testing.assert_allclose(mm["L"], L, atol=1e-12)
    for b, Eb in enumerate((E + A, E - A)):
        Heff = H + Eb * (n1 - n2) - 0.5j * (g / N) * (L.T @ L)
        np.testing.assert_allclose(mm["Hqj%d" % b], -1j * Heff, atol=1e-12)


@pytest.mark.parametrize("N", [3, 8, 40])
def test_jump_operator_spectrum(oracle_mod, N):
    # L = 2 c+^dag c- in c+- = (b1 +- b2)/sqrt 2, so L^dag L = 4 n_-(n_+ + 1):
    # eigenvalues 4 m (N - m + 1), m = 0..N.
    mm = oracle_mod.model_matrices(oracle_mod.model(N))
    LdL = mm["L"].conj().T @ mm["L"]
    ev = np.sort(np.linalg.eigvalsh(LdL))
    m = np.arange(N + 1)
    np.testing.assert_allclose(ev, np.sort(4.0 * m * (N - m + 1)), atol=1e-9 * N * N)


def test_expm_against_scipy_fixtures(oracle_mod):
    fx = np.load(GOLD, allow_pickle=False)
    T = 2 * np.pi
    n = 0
    for key in fx.files:
        if not key.startswith("U_"):
            continue
        _, Ns, bs, dts = key.split("_")
        N, b, dt = int(Ns[1:]), int(bs[1:]), float(dts[2:])
        mm = oracle_mod.model_matrices(oracle_mod.model(N))
        U = oracle_mod.expm(mm["Hqj%d" % b] * dt)
        np.testing.assert_allclose(U, fx[key], atol=5e-14, err_msg=key)
        n += 1
    assert n == 32


def test_expm_two_by_two_closed_form(oracle_mod):
    # exp(a I + b . sigma) = e^a (cosh w I + sinh w / w  b . sigma), w^2 = b . b
    rng = np.random.default_rng(7)
    sx = np.array([[0, 1], [1, 0]], complex)
    sy = np.array([[0, -1j], [1j, 0]])
    sz = np.array([[1, 0], [0, -1]], complex)
    for scale in (1e-3, 0.5, 3.0, 40.0):
        a = complex(*rng.standard_normal(2)) * 0.1
        bv = (rng.standard_normal(3) + 1j * rng.standard_normal(3)) * scale
        B = bv[0] * sx + bv[1] * sy + bv[2] * sz
        w = np.sqrt(bv @ bv)
        ref = np.exp(a) * (np.cosh(w) * np.eye(2) + np.sinh(w) / w * B)
        got = oracle_mod.expm(a * np.eye(2) + B)
        np.testing.assert_allclose(got, ref, rtol=1e-12 * max(1.0, scale), atol=1e-13)


def test_expm_semigroup_and_unitarity(oracle_mod):
    N, D = 12, 6
    m = oracle_mod.model(N, gamma=0.0, S=4, D=D)
    e = oracle_mod.Ensemble(m, 1, threads=1)
    for b in range(2):
        for lvl in range(D):
            U = e.get_propagator(b, lvl)
            np.testing.assert_allclose(U.conj().T @ U, np.eye(N + 1), atol=1e-12)   # gamma = 0
            if lvl + 1 < D:
                Uh = e.get_propagator(b, lvl + 1)
                np.testing.assert_allclose(Uh @ Uh, U, atol=1e-12)
    m2 = oracle_mod.model(N, gamma=0.3, S=4, D=3)
    e2 = oracle_mod.Ensemble(m2, 1, threads=1)
    for lvl in range(3):
        s = np.linalg.svd(e2.get_propagator(0, lvl), compute_uv=False)
        assert s.max() <= 1.0 + 1e-12          # non-Hermitian evolution is contractive


def test_spin_rotation_limit(oracle_mod):
    # U = gamma = 0: the bosons are independent, each a two-level system with
    # h = E sz - J sx (basis site 1, site 2).  Starting from |N, 0> every boson
    # ends on site 2 with the Rabi probability (J/w)^2 sin^2(w t), w = sqrt(J^2 + E^2),
    # so <n2> = N (J/w)^2 sin^2(w t).
    N, J, E, t = 9, 0.8, 0.35, 0.77
    m = oracle_mod.model(N, prm_E=E, prm_U=0.0, prm_J=J, drv_ampl=0.0, gamma=0.0)
    Hq = oracle_mod.model_matrices(m)["Hqj0"]        # -i H
    psi = oracle_mod.expm(Hq * t)[:, 0]
    n2 = np.sum(np.arange(N + 1) * np.abs(psi) ** 2)
    w = np.sqrt(J * J + E * E)
    p_flip = (J / w) ** 2 * np.sin(w * t) ** 2
    assert abs(n2 - N * p_flip) < 1e-11
