"""Batches and launch shapes: results do not depend on how the work is launched.

* A batch advances several ensembles of different Hilbert dimension in the same
  grouped GEMMs; each member's trajectories must end bitwise identical to a run
  of that ensemble alone (states, etas, times, jump counts, RNG states).
* The step GEMM's column tile width (32 to 256), persistent grid and row tiles
  per work unit are picked by the launch-shape tuner; every shape must give
  bitwise identical results.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MEMBERS = [
    dict(N=20, M=64, seed=3, drv_ampl=1.0, diss_gamma=0.3),
    dict(N=70, M=200, seed=4, drv_ampl=2.5, diss_gamma=0.1),
    dict(N=150, M=96, seed=5, drv_ampl=1.5, diss_gamma=0.05),
]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(prec, N, M, seed, drv_ampl, diss_gamma, S=8, D=8):
    from os_d_amd import qjx
    e = qjx.Engine(N, prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=drv_ampl, drv_freq=1.0,
                   diss_gamma=diss_gamma, ex_deep=D, steps_per_period=S, precision=prec)
    e.ensemble_init(M, seed=seed, mns=1000000, start_type=0, start_state=0, jump=0)
    e.trans(0)
    return e


def _same_state(a, b):
    for k in ("psi", "eta", "t", "njumps", "rng"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("prec", [1, 0, 2])
def test_batch_member_equals_alone(prec):
    from os_d_amd import qjx
    _say("case: batch vs alone prec=%d" % prec)
    alone = []
    for m in MEMBERS:
        e = _engine(prec, **m)
        e.run_steps(6, 0)
        alone.append((e.get_state(), e.stats()["matvecs"]))
        e.close()
    engines = [_engine(prec, **m) for m in MEMBERS]
    b = qjx.Batch(engines)
    b.run_steps(4, 0)
    b.run_steps(2, 0)
    st = b.stats()
    _say("batch: matvecs %d, iterations %d, bn %d" % (st["matvecs"], st["iterations"], st["gemm_bn"]))
    for e, (s_alone, mv) in zip(engines, alone):
        _same_state(e.get_state(), s_alone)
        assert e.stats()["matvecs"] == mv
    assert st["matvecs"] == sum(mv for _, mv in alone)
    b.close()
    for e in engines:
        e.close()


@pytest.mark.parametrize("prec", [1, 0, 2])
def test_tile_width_does_not_change_results(prec, monkeypatch):
    _say("case: tile widths %s prec=%d" % ("32/64/128" if prec == 0 else "64/128/256", prec))
    res = {}
    widths = ("64", "128", "256") if prec != 0 else ("32", "64", "128")
    for bn in widths:
        monkeypatch.setenv("QJX_BN", bn)
        e = _engine(prec, N=150, M=300, seed=11, drv_ampl=1.5, diss_gamma=0.1)
        e.run_steps(5, 0)
        s = e.stats()
        assert s["gemm_bn"] == int(bn)
        res[bn] = e.get_state()
        e.close()
    for bn in widths[1:]:
        _same_state(res[widths[0]], res[bn])


@pytest.mark.parametrize("prec", [1, 2])
def test_row_tiles_per_unit_do_not_change_results(prec, monkeypatch):
    # the complex-pair GEMM's work units (1, 2 or 3 row tiles, which the tuner picks
    # from) must not enter any sum
    _say("case: row tiles per unit 1, 2, 3 prec=%d" % prec)
    res = {}
    for rc in ("1", "2", "3"):
        monkeypatch.setenv("QJX_RC", rc)
        e = _engine(prec, N=300, M=500, seed=5, drv_ampl=1.5, diss_gamma=0.1)
        e.run_steps(5, 0)
        res[rc] = e.get_state()
        e.close()
    for rc in ("2", "3"):
        _same_state(res["1"], res[rc])


def test_autotuner_picks_a_shape_with_identical_results(monkeypatch):
    _say("case: autotuner")
    monkeypatch.setenv("QJX_BN", "128")
    e = _engine(1, N=200, M=2000, seed=2, drv_ampl=1.5, diss_gamma=0.2)
    e.run_steps(12, 0)
    ref = e.get_state()
    e.close()
    monkeypatch.delenv("QJX_BN")
    monkeypatch.delenv("QJX_TUNE", raising=False)
    # tune right after the first chunk (a run this short ends before the default 512)
    monkeypatch.setenv("QJX_TUNE_AFTER", "0")
    e = _engine(1, N=200, M=2000, seed=2, drv_ampl=1.5, diss_gamma=0.2)
    e.run_steps(12, 0)
    s = e.stats()
    _say("tuned: bn %d grid %d" % (s["gemm_bn"], s["gemm_grid"]))
    assert s["gemm_bn"] in (64, 128, 256)
    assert s["gemm_grid"] > 0 and s["gemm_grid"] % 8 == 0
    _same_state(e.get_state(), ref)
    e.close()


def test_batch_rejects_mixed_precision():
    from os_d_amd import qjx
    _say("case: batch precision check")
    a = _engine(1, **MEMBERS[0])
    b = _engine(0, **MEMBERS[1])
    with pytest.raises(qjx.QJXError):
        qjx.Batch([a, b])
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [1, 0, 2])
def test_band_skipping_does_not_change_results(prec, monkeypatch):
    """The step GEMM multiplies only the K blocks inside each propagator row tile's
    band (the rest are exact zeros): states must equal a dense run bitwise."""
    _say("case: band vs dense prec=%d" % prec)
    res, work = {}, {}
    for dense in ("1", "0"):
        monkeypatch.setenv("QJX_DENSE", dense)
        e = _engine(prec, N=150, M=300, seed=8, drv_ampl=1.5, diss_gamma=0.1, S=200, D=16)
        e.run_steps(6, 0)
        res[dense] = e.get_state()
        work[dense] = e.stats()["gemm_band_flops"]
        e.close()
    _say("GEMM flops dense %.3e band %.3e (ratio %.3f)" % (work["1"], work["0"], work["0"] / work["1"]))
    _same_state(res["0"], res["1"])
    assert work["0"] <= work["1"]
    if prec != 0:
        # FP32 storage underflows far from the diagonal: the band is narrow
        assert work["0"] < 0.8 * work["1"]


@pytest.mark.parametrize("N", [150, 700])
def test_bf16x3_kernel_variants_match(N, monkeypatch):
    """The 3xBF16 GEMM variants (4 or 8 waves, 128 or 256 columns) give identical
    results: each output's MFMA sequence and norm sums are the same."""
    _say("case: bf16x3 variants N=%d" % N)
    res = {}
    shapes = {"128x4": ("128", None), "128x8": ("128", "8"), "256x8": ("256", None)}
    for name, (bn, waves) in shapes.items():
        monkeypatch.setenv("QJX_BN", bn)
        if waves is None:
            monkeypatch.delenv("QJX_WAVES", raising=False)
        else:
            monkeypatch.setenv("QJX_WAVES", waves)
        e = _engine(2, N=N, M=300, seed=13, drv_ampl=1.5, diss_gamma=0.1, S=40)
        e.run_steps(5, 0)
        res[name] = e.get_state()
        e.close()
    for name in list(shapes)[1:]:
        _same_state(res["128x4"], res[name])
