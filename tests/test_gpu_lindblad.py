"""The GPU Lindblad integrator against the CPU restatement, and the trajectory
ensemble against the GPU Lindblad solution.

* rho after whole periods equals the oracle's rho-space RK4 bit for bit (same
  operations in the same order, no contraction on either side).
* Observables equal the oracle's to rounding; the trace and Hermiticity hold.
* The FP32 trajectory ensemble's <k> and energy agree with the GPU Lindblad
  solution within 3 sigma at a size the CPU solver would take minutes for.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PARAMS = dict(prm_E=1.0, prm_U=0.7, prm_J=1.0, drv_ampl=2.0, drv_freq=1.0)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# the step kernel's shapes (tile edge x RK4 steps per launch): d = 21 and 65 end in
# partial tiles (65: a one-row tile), d = 32 is whole tiles of both sizes
@pytest.mark.parametrize("shape", ["8x2", "8x1", "16x1"])
@pytest.mark.parametrize("N,steps,periods", [(20, 200, 2), (64, 2000, 1), (31, 400, 1)])
def test_gpu_lindblad_matches_oracle_bitwise(oracle_mod, N, steps, periods, shape, monkeypatch):
    from os_d_amd import qjx
    tile, fuse = shape.split("x")
    monkeypatch.setenv("QJX_LB_TILE", tile)
    monkeypatch.setenv("QJX_LB_FUSE", fuse)
    _say("case: lindblad N=%d steps=%d periods=%d shape=%s" % (N, steps, periods, shape))
    m = oracle_mod.model(N, gamma=0.3, **PARAMS)
    lb = qjx.Lindblad(N, steps_per_period=steps, diss_gamma=0.3, **PARAMS)
    rho0 = np.zeros((N + 1, N + 1), complex)
    rho0[0, 0] = 1.0
    assert np.array_equal(lb.get_rho(), rho0)
    lb.run_periods(periods)
    got = lb.get_rho()
    ref = oracle_mod.lindblad_evolve(m, rho0, periods, steps_per_period=steps)
    diff = np.abs(got - ref).max()
    _say("max |rho_gpu - rho_oracle| = %.3e" % diff)
    assert np.array_equal(got, ref)
    lb.close()


def test_gpu_lindblad_eager_matches_oracle_bitwise(oracle_mod, monkeypatch):
    """Eager launches (no graph): the step kernels alternate rho between two buffers
    launch by launch, and a sequence with an odd number of launches (the last 10 of
    202 steps, two per launch) copies it back; rho after the period equals the
    oracle's bit for bit."""
    from os_d_amd import qjx
    monkeypatch.setenv("QJX_GRAPH", "0")
    N, steps = 40, 202
    _say("case: lindblad eager N=%d steps=%d" % (N, steps))
    m = oracle_mod.model(N, gamma=0.3, **PARAMS)
    lb = qjx.Lindblad(N, steps_per_period=steps, diss_gamma=0.3, **PARAMS)
    rho0 = np.zeros((N + 1, N + 1), complex)
    rho0[0, 0] = 1.0
    lb.run_periods(1)
    ref = oracle_mod.lindblad_evolve(m, rho0, 1, steps_per_period=steps)
    assert np.array_equal(lb.get_rho(), ref)
    lb.close()


def test_gpu_lindblad_observables(oracle_mod):
    from os_d_amd import qjx
    N = 40
    _say("case: lindblad observables N=%d" % N)
    m = oracle_mod.model(N, gamma=0.2, **PARAMS)
    lb = qjx.Lindblad(N, steps_per_period=1000, diss_gamma=0.2, **PARAMS)   # RK4-stable step
    # a mixed, non-diagonal start: set_rho must be honoured
    rng = np.random.default_rng(3)
    X = rng.normal(size=(N + 1, N + 1)) + 1j * rng.normal(size=(N + 1, N + 1))
    rho0 = X @ X.conj().T
    rho0 /= np.trace(rho0).real
    lb.set_rho(rho0)
    assert np.array_equal(lb.get_rho(), rho0)
    lb.run_periods(1)
    rho = lb.get_rho()
    ob = lb.observables()
    ref = oracle_mod.lindblad_observables(m, rho)
    for k in ("mean", "m2", "energy", "purity"):
        _say("%s: gpu %.12f oracle %.12f" % (k, ob[k], ref[k]))
        assert abs(ob[k] - ref[k]) <= 1e-10 * max(1.0, abs(ref[k]))
    np.testing.assert_allclose(ob["adr"], ref["adr"], rtol=0, atol=1e-15)
    assert abs(np.trace(rho).real - 1.0) < 1e-10
    assert np.abs(rho - rho.conj().T).max() < 1e-12
    lb.close()


def test_fp32_ensemble_matches_gpu_lindblad():
    from os_d_amd import qjx
    N, M, P = 80, 20000, 2
    _say("case: fp32 ensemble vs gpu lindblad N=%d M=%d" % (N, M))
    lb = qjx.Lindblad(N, steps_per_period=2000, diss_gamma=0.1, **PARAMS)
    e = qjx.Engine(N, ex_deep=16, steps_per_period=40, precision=qjx.FP32, diss_gamma=0.1, **PARAMS)
    e.ensemble_init(M, seed=21, jump=0)
    e.trans(0)
    for p in range(P):
        lb.run_periods(1)
        e.run_periods(1, 1)
        ob = lb.observables()
        c = e.chars()
        for k in ("mean", "energy"):
            v, sig = c[k].mean(), c[k].std() / np.sqrt(M)
            _say("period %d %s: qj %.6f lindblad %.6f (%.2f sigma)" % (p + 1, k, v, ob[k], (v - ob[k]) / sig))
            assert abs(v - ob[k]) < 3.0 * sig
    _, pur = e.density(rho=False)
    _say("purity: qj %.5f lindblad %.5f" % (pur, ob["purity"]))
    assert abs(pur - ob["purity"]) < 0.02
    e.close()
    lb.close()


@pytest.mark.parametrize("drv_type,phase,steps,periods", [(0, 0.0, 400, 2), (1, 0.3, 200, 2), (0, 0.0, 2000, 1)])
def test_gpu_lindblad_cqdiss_drive_matches_oracle_bitwise(oracle_mod, drv_type, phase, steps, periods):
    """CQdiss's drive (calcODE.cpp:100-116) evaluated at every stage time on the GPU:
    rho equals the oracle's bit for bit, for the square drive (its rounding-dependent
    switch stages included) and the sine drive with a phase; running the periods in
    two calls continues the time and gives the same bits; the rho file is CQdiss's."""
    from os_d_amd import qjx
    N = 20
    _say("case: lindblad CQdiss drive type=%d phase=%.1f steps=%d" % (drv_type, phase, steps))
    m = oracle_mod.model(N, gamma=0.3, **PARAMS)
    rho0 = np.zeros((N + 1, N + 1), complex)
    rho0[0, 0] = 1.0
    ref = oracle_mod.lindblad_evolve_drive(m, rho0, periods, steps, drv_type, phase, 1)
    lb = qjx.Lindblad(N, steps_per_period=steps, diss_gamma=0.3, drv_type=drv_type, drv_phase=phase, per_stage=1,
                      **PARAMS)
    lb.run_periods(periods)
    got = lb.get_rho()
    _say("max |rho_gpu - rho_oracle| = %.3e" % np.abs(got - ref).max())
    assert np.array_equal(got, ref)
    lb.set_rho(rho0)                         # time restarts at 0
    for _ in range(periods):
        lb.run_periods(1)
    assert np.array_equal(lb.get_rho(), ref)
    lb.close()


def test_gpu_lindblad_rho_file(tmp_path):
    """qjx_lindblad_save_rho writes every entry as CQdiss writes its final rho
    (save_sparse_complex_mtx): "row col re im", 1-based, setprecision(16) scientific."""
    from os_d_amd import qjx
    N = 6
    lb = qjx.Lindblad(N, steps_per_period=200, diss_gamma=0.3, per_stage=1, **PARAMS)
    lb.run_periods(1)
    rho = lb.get_rho()
    fn = str(tmp_path / "rho.txt")
    lb.save_rho(fn)
    rows = np.loadtxt(fn)
    assert rows.shape == ((N + 1) ** 2, 4)
    assert np.array_equal(rows[:, 0].astype(int), np.repeat(np.arange(1, N + 2), N + 1))
    assert np.array_equal(rows[:, 1].astype(int), np.tile(np.arange(1, N + 2), N + 1))
    np.testing.assert_allclose(rows[:, 2] + 1j * rows[:, 3], rho.reshape(-1), rtol=1e-15, atol=1e-300)
    first = open(fn).readline().split()
    assert len(first[2].split("e")[0]) in (18, 19)            # 16 digits after the point
    lb.close()
