"""Pins of the oracle's trajectory algorithm by brute force on tiny inputs.

Because ||psi(t)||^2 decays monotonically under exp(-i H_eff t), the reference's
bisection (qj_proc.cpp:271-343) lands on the FIRST point of the finest time grid
(spacing T / (S 2^(D-1))) where ||psi||^2 <= eta.  Here that point is found by
stepping a scipy propagator over every fine-grid point, the jump is applied as
L psi / ||L psi|| (experiment.cpp:1273-1382) and the RNG is consumed as eta_0,
then per jump `ran` and eta -- independently of the oracle's recursion.
"""
import numpy as np
import pytest
import scipy.linalg as sl


def _fine_grid_jumps(oracle_mod, m, tr, n_jumps, seed=0, mns=1000000):
    N, S, D = m.N, m.S, m.D
    T = 2 * np.pi / m.drv_freq
    mm = oracle_mod.model_matrices(m)
    nf = S * 2 ** (D - 1)                       # fine steps per period
    dtf = T / nf
    Uf = [sl.expm(mm["Hqj%d" % b] * dtf) for b in range(2)]
    draws = iter(oracle_mod.mcg59_uniform(seed + tr, mns, 4 * n_jumps + 16))

    def eta():
        e = next(draws)
        while e == 0.0:
            e = next(draws)
        return e

    psi = np.zeros(N + 1, complex)
    psi[0] = 1.0
    et = eta()
    k = 0
    out = []
    while len(out) < n_jumps:
        b = (k % nf) // (nf // 2)
        psi = Uf[b] @ psi
        k += 1
        n = np.vdot(psi, psi).real
        if n <= et:
            out.append((k * dtf, n, et))
            next(draws)                          # `ran` (one channel)
            v = mm["L"] @ (psi / np.sqrt(n))
            psi = v / np.linalg.norm(v)
            et = eta()
        if k > 200 * nf:
            break
    return out


@pytest.mark.parametrize("N,S,D", [(4, 4, 8), (6, 8, 6)])
def test_jump_times_are_first_fine_grid_crossings(oracle_mod, N, S, D):
    m = oracle_mod.model(N, gamma=0.6, S=S, D=D)
    M, P = 6, 6
    e = oracle_mod.Ensemble(m, M, jump=2, threads=1)
    e.trans(0)
    e.periods(P, 1)
    checked = 0
    for tr in range(M):
        j = e.get_jumps(tr)
        nj = min(len(j["times"]), 3)
        if nj == 0:
            continue
        ref = _fine_grid_jumps(oracle_mod, m, tr, nj)
        for i in range(nj):
            t, n, et = ref[i]
            assert abs(j["times"][i] - t) < 1e-9, (tr, i, j["times"][i], t)
            assert j["etas"][i] == et
            assert abs(j["norms"][i] - n) < 1e-10
            checked += 1
    assert checked >= 6


def test_no_jumps_without_dissipation(oracle_mod):
    # (gamma = 0 exactly makes ||psi||^2 exceed 1 by rounding, which the reference
    # treats as an error, experiment.cpp:1247-1251; a vanishing rate keeps the
    # evolution contractive.)
    m = oracle_mod.model(8, gamma=1e-9, S=4, D=6)
    e = oracle_mod.Ensemble(m, 16, jump=1, threads=2)
    e.trans(1)
    e.periods(3, 1)
    assert e.num_jumps() == 0
    c = e.chars()
    np.testing.assert_allclose(c["norm"], 1.0, atol=1e-6)


def test_dark_state_never_jumps(oracle_mod):
    # U = E = A = 0: H = -J(n_+ - n_-) keeps |n_- = 0> = (c_+^dag)^N |0> stationary
    # and L^dag L = 4 n_-(n_+ + 1) vanishes on it.
    N = 6
    m = oracle_mod.model(N, prm_E=0.0, prm_U=0.0, drv_ampl=0.0, gamma=0.5, S=4, D=6)
    mm = oracle_mod.model_matrices(m)
    w, V = np.linalg.eigh(mm["L"].conj().T @ mm["L"])
    dark = V[:, 0]
    assert abs(w[0]) < 1e-10
    M = 4
    e = oracle_mod.Ensemble(m, M, jump=1, threads=1)
    e.set_state(psi=np.tile(dark, (M, 1)))
    e.trans(0)
    e.periods(3, 1)
    assert e.num_jumps() == 0


def test_jump_counts_rise_with_rate(oracle_mod):
    # the mean number of jumps grows with gamma (rate gamma/N <L^dag L>)
    res = []
    for g in (0.05, 0.2, 0.8):
        e = oracle_mod.Ensemble(oracle_mod.model(6, gamma=g, S=4, D=8), 64, jump=1)
        e.trans(0)
        e.periods(4, 1)
        res.append(e.num_jumps())
    assert res[0] < res[1] < res[2]
