"""Pins the Runge-Kutta restatement (prop_id 1) of the oracle.

Reference: rk_period (core.cpp:439-450), rk_step_dimer (core.cpp:11886-11948),
rk_int_dimer (core.cpp:11592-11619), rk_right_part_dimer (core.cpp:11033-11115),
arg_upd / rk_final / rk_recovery (rk_proc.cpp:141-306), all under
source/cpp/QJX/QJX/.  The reference ships no RK fixtures; each check below is
fixed by mathematics independent of the restatement's code:

* the band right-hand side equals the dense -i (H - i/2 gamma/N L^T L + E(t) H_drv) x
  built in numpy from the model matrices, for both drive types;
* the sine of the sinusoidal drive is within 2 ulp of the C library's;
* RK4 converges at order 4 to expm (constant drive) and to a DOP853 solution
  (sinusoidal drive, which also fixes frequency and phase);
* one integration over a step s != rk_step equals the closed-form polynomial of
  the reference's stage arguments, which use the global rk_step
  (rk_proc.cpp:150-173), and differs from an ordinary RK4 step of s;
* gamma = 0 gives no jumps; every trajectory consumes exactly 1 + 2 J draws of
  its RNG stream (eta_0, then ran and eta per jump);
* ensemble averages agree with the Lindblad oracle within 3 sigma, and the
  norm at a logged jump approaches its eta as the step shrinks.
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.linalg import expm

from oracle import oracle as O


def _hpart(m):
    mm = O.model_matrices(m)
    g = m.gamma / m.N
    return mm, mm["H"] - 0.5j * g * (mm["L"].conj().T @ mm["L"])


@pytest.mark.parametrize("drv_type,phase", [(0, 0.0), (1, 0.4), (7, 0.0)])
def test_rk_rhs_is_dense_heff(drv_type, phase):
    m = O.model(30, prm_E=1.0, prm_U=0.7, prm_J=1.0, drv_ampl=2.0, drv_freq=1.3, gamma=0.3, D=1)
    mm, Hp = _hpart(m)
    e = O.Ensemble(m, 1, build_propagators=False)
    e.rk_setup(200, drv_type, phase)
    T = 2 * math.pi / 1.3
    rng = np.random.default_rng(1)
    x = rng.normal(size=31) + 1j * rng.normal(size=31)
    for t in (0.0, 0.1, 2.0, T / 2 - 1e-9, T / 2 + 1e-9, 3.5, 4.7, 100.3):
        if drv_type == 1:
            E = 1.0 + 2.0 * math.sin(1.3 * t + phase)
            assert abs(e.rk_drive(t) - E) <= 4e-16 * 3
        else:   # any type but 1 is the square drive (core.cpp:11061-11088)
            E = 3.0 if math.fmod(t, T) < T / 2 else -1.0
            assert e.rk_drive(t) == E
        k = e.rk_rhs(t, x)
        kd = -1j * ((Hp + e.rk_drive(t) * mm["Hdrv"]) @ x)
        assert np.abs(k - kd).max() <= 1e-14 * np.abs(kd).max()


def test_rk_sine_within_two_ulp():
    xs = np.concatenate([np.random.default_rng(0).uniform(-1e4, 1e4, 50000), np.linspace(-10, 10, 4001),
                         np.arange(-40, 41) * (math.pi / 4)])
    worst = 0.0
    for x in xs:
        a, b = O.rk_sin(x), math.sin(x)
        if b != 0.0:
            worst = max(worst, abs(a - b) / math.ulp(abs(b)))
    assert worst <= 2.0
    assert O.rk_sin(0.0) == 0.0


def test_rk_norm_is_the_sum_of_squares():
    rng = np.random.default_rng(4)
    for d in (2, 21, 64, 65, 129, 1001, 2048):
        x = rng.normal(size=d) + 1j * rng.normal(size=d)
        ref = math.fsum(np.abs(x) ** 2)
        assert abs(O.rk_norm(x) - ref) <= 1e-15 * d * ref


def _final_state(m, ns, typ, phase):
    e = O.Ensemble(m, 1, build_propagators=False)
    e.rk_setup(ns, typ, phase)
    e.set_state(eta=np.array([1e-300]))        # no crossing: pure RK4 propagation
    e.periods(1, 0)
    s = e.get_state()
    return s["psi"][0], s["t"][0]


@pytest.mark.parametrize("drv_type", [0, 1])
def test_rk4_order_four(drv_type):
    N, A, phase = 12, (0.0 if drv_type == 0 else 1.5), 0.3
    m = O.model(N, prm_E=1.0, prm_U=0.7, prm_J=1.0, drv_ampl=A, drv_freq=1.0, gamma=0.05, D=1)
    mm, Hp = _hpart(m)
    T = 2 * math.pi
    psi0 = np.zeros(N + 1, complex)
    psi0[0] = 1.0
    if drv_type == 0:
        ref = expm(-1j * (Hp + 1.0 * mm["Hdrv"]) * T) @ psi0
    else:
        f = lambda t, y: -1j * ((Hp + (1.0 + A * math.sin(t + phase)) * mm["Hdrv"]) @ y)
        ref = solve_ivp(f, (0.0, T), psi0, method="DOP853", rtol=1e-13, atol=1e-15).y[:, -1]
    errs = []
    for ns in (800, 1600):
        psi, t = _final_state(m, ns, drv_type, phase)
        assert abs(t - T) < 1e-12
        errs.append(np.abs(psi - ref).max())
    assert 13.0 < errs[0] / errs[1] < 19.0
    assert errs[1] < 1e-4


def test_rk_substep_uses_global_rk_step():
    """A diagonal model (J = 0, gamma = 0, A = 0): each amplitude evolves by the
    scalar polynomial of the stage arguments phi + h/2 k1, phi + h/2 k2, phi + h k3
    with h = rk_step and the weights s/6 (1, 2, 2, 1) of the actual step s."""
    N, ns = 5, 40
    m = O.model(N, prm_E=1.3, prm_U=0.7, prm_J=0.0, drv_ampl=0.0, drv_freq=1.0, gamma=0.0, D=1)
    mm = O.model_matrices(m)
    a = -1j * (np.diag(mm["H"]) + 1.3 * np.diag(mm["Hdrv"]))
    h = 2 * math.pi / ns
    e = O.Ensemble(m, 1, start_type=1, build_propagators=False)
    e.rk_setup(ns, 0, 0.0)
    phi0 = e.get_state()["psi"][0]

    def poly(s, hh):
        k1 = a
        k2 = a * (1 + hh / 2 * a)
        k3 = a * (1 + hh / 2 * a * (1 + hh / 2 * a))
        k4 = a * (1 + hh * a * (1 + hh / 2 * a * (1 + hh / 2 * a)))
        return 1 + s / 6 * (k1 + 2 * k2 + 2 * k3 + k4)

    s = 0.37 * h
    e.rk_int(0, s)
    got = e.get_state()["psi"][0]
    want = poly(s, h) * phi0
    honest = poly(s, s) * phi0
    assert np.abs(got - want).max() < 1e-14
    assert np.abs(honest - want).max() > 1e-6
    assert abs(e.get_state()["t"][0] - s) < 1e-16
    # a full step is the ordinary RK4 polynomial
    e.set_state(psi=phi0[None, :], t=np.zeros(1))
    e.rk_int(0, h)
    z = h * a
    assert np.abs(e.get_state()["psi"][0] - (1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24) * phi0).max() < 1e-14


def test_rk_no_dissipation_no_jumps():
    m = O.model(8, prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0, gamma=0.0, D=1)
    e = O.Ensemble(m, 20, seed=2, jump=2, build_propagators=False)
    # RK4 itself loses norm, (h |H|)^4 per unit time: at 2000 steps per period
    # ||psi||^2 stays above 1 - 1e-4, so only an eta that close to 1 could cross.
    # (The square drive switches E inside the step that ends at T/2, and that
    # step can raise ||psi||^2 above 1 when nothing dissipates: the reference
    # then throws "norm more than 1", experiment.cpp:1247-1251.)
    e.rk_setup(2000, 1, 0.0)
    e.trans(1)
    e.periods(1, 1)
    eta_max = max(e.get_state()["eta"])
    n = e.chars()["norm"]
    assert (n <= 1.0).all() and (n > 1.0 - 1e-4).all()
    if eta_max < 1.0 - 1e-4:
        assert e.num_jumps() == 0
        assert (e.get_state()["njumps"] == 0).all()


def test_rk_jumps_rng_and_lindblad_averages():
    N, d, M = 10, 11, 4000
    m = O.model(N, prm_E=1.0, prm_U=0.7, prm_J=1.0, drv_ampl=2.0, drv_freq=1.0, gamma=0.3, D=1)
    rho0 = np.zeros((d, d), complex)
    rho0[0, 0] = 1.0
    ob = O.lindblad_observables(m, O.lindblad_evolve(m, rho0, 2, steps_per_period=4000))
    e = O.Ensemble(m, M, seed=5, jump=2, build_propagators=False)
    e.rk_setup(800, 0, 0.0)
    e.trans(0)
    e.periods(2, 1)
    c = e.chars()
    for k in ("mean", "energy"):
        v, sig = c[k].mean(), c[k].std() / math.sqrt(M)
        assert abs(v - ob[k]) < 3.0 * sig, (k, v, ob[k], sig)
    s = e.get_state()
    nj = s["njumps"]
    assert nj.sum() == e.num_jumps() > M
    a, mod = 302875106592253, 1 << 59
    stride = pow(a, 1000000, mod)
    for tr in range(0, M, 37):
        x0 = 777 * pow(a, 5 + tr, mod) % mod
        assert int(s["rng"][tr]) == x0 * pow(stride, 1 + 2 * int(nj[tr]), mod) % mod
    j = e.get_jumps(3)
    assert len(j["times"]) == nj[3] and (np.diff(j["times"]) > 0).all() and (j["types"] == 0).all()
    assert (j["times"] > 0).all() and (j["times"] < 2 * 2 * math.pi).all()
    assert (j["norms"] <= j["etas"] + 0.02).all()


def test_rk_jump_norm_approaches_eta():
    m = O.model(10, prm_E=1.0, prm_U=0.7, prm_J=1.0, drv_ampl=2.0, drv_freq=1.0, gamma=0.3, D=1)
    gaps = []
    for ns in (400, 1600):
        e = O.Ensemble(m, 200, seed=9, jump=2, build_propagators=False)
        e.rk_setup(ns, 0, 0.0)
        e.trans(0)
        e.periods(1, 1)
        g = [np.abs(j["norms"] - j["etas"]).max() for j in (e.get_jumps(tr) for tr in range(200)) if len(j["norms"])]
        gaps.append(max(g))
    assert gaps[1] < gaps[0] / 3.0
