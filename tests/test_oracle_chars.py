"""Independent pin of the oracle's per-trajectory observables (qjo_chars, restating
calc_chars_std(_start), core.cpp:665-725, and experiment.cpp:1395-1469) and of the
ensemble average of abs_diag_rho that dump_adr_avg writes (dump.cpp:26-61).

The states are spin-coherent states, whose norm, <k>, m2, energy and |psi_k|^2 / ||psi||^2
have closed forms (tests/chars_closed_form.py); dispersion is (<k> - <k>_start)^2 between
two such states.  A dropped normalisation, a wrong centre in m2, the wrong start mean in
the dispersion or a sign in the energy's hopping, interaction or drive term fails here.
"""
import numpy as np
import pytest

from chars_closed_form import coherent_chars, coherent_ensemble


@pytest.mark.parametrize("N", [1, 7, 40])
def test_oracle_chars_closed_forms(oracle_mod, N):
    M = 12
    prm = dict(prm_E=0.7, prm_U=0.45, prm_J=1.3)
    m = oracle_mod.model(N, D=1, **prm)
    e = oracle_mod.Ensemble(m, M, jump=0, threads=1, build_propagators=False)
    p0, phi0, c0, psi0 = coherent_ensemble(N, M, seed=N)
    p1, phi1, c1, psi1 = coherent_ensemble(N, M, seed=100 + N)

    e.set_state(psi=psi0)
    s = e.chars(is_start=1, adr=True)
    for tr in range(M):
        want = coherent_chars(N, p0[tr], phi0[tr], c0[tr], **prm)
        assert s["norm"][tr] == pytest.approx(want["norm"], rel=1e-13)
        assert s["mean"][tr] == pytest.approx(want["mean"], rel=1e-12, abs=1e-13)
        assert s["m2"][tr] == pytest.approx(want["m2"], rel=1e-11, abs=1e-12)
        assert s["energy"][tr] == pytest.approx(want["energy"], rel=1e-11, abs=1e-11)
        np.testing.assert_allclose(s["adr"][tr], want["adr"], rtol=1e-12, atol=1e-15)
    assert np.all(s["dispersion"] == 0.0)

    e.set_state(psi=psi1)
    g = e.chars(adr=True)
    for tr in range(M):
        want = coherent_chars(N, p1[tr], phi1[tr], c1[tr], **prm)
        assert g["mean"][tr] == pytest.approx(want["mean"], rel=1e-12, abs=1e-13)
        assert g["m2"][tr] == pytest.approx(want["m2"], rel=1e-11, abs=1e-12)
        assert g["energy"][tr] == pytest.approx(want["energy"], rel=1e-11, abs=1e-11)
        # dispersion against the mean recorded by the start call, not the current one
        assert g["dispersion"][tr] == pytest.approx((N * p1[tr] - N * p0[tr]) ** 2, rel=1e-11, abs=1e-12)
    # dump_adr_avg: the trajectory average of abs_diag_rho is the mixture of the pmfs
    mix = np.mean([coherent_chars(N, p1[tr], phi1[tr], c1[tr], **prm)["adr"] for tr in range(M)], axis=0)
    np.testing.assert_allclose(g["adr"].mean(axis=0), mix, rtol=1e-12, atol=1e-15)
    assert g["adr"].sum(axis=1) == pytest.approx(np.ones(M), rel=1e-13)


def test_oracle_chars_follow_rabi_oscillation(oracle_mod):
    """At U = E = A = 0 and vanishing gamma the state from e_0 is the coherent state
    p = sin^2(J t), phi = -pi/2, so <k>(t) = N sin^2(J t), m2 = N sin^2 cos^2 and the
    energy stays 0 -- the observables of propagated states, not planted ones."""
    N, S = 10, 8
    m = oracle_mod.model(N, prm_E=0.0, prm_U=0.0, prm_J=1.0, drv_ampl=0.0, gamma=1e-12, S=S, D=2)
    e = oracle_mod.Ensemble(m, 3, jump=0, threads=1)
    e.trans(0)
    start = e.chars(is_start=1)
    assert np.allclose(start["mean"], 0.0)
    T = 2 * np.pi
    for step in range(1, S + 1):
        e.steps(1, 0)
        t = step * T / S
        c = e.chars()
        p = np.sin(t) ** 2
        np.testing.assert_allclose(c["mean"], N * p, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(c["m2"], N * p * (1 - p), rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(c["dispersion"], (N * p) ** 2, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(c["energy"], 0.0, atol=1e-9)
