"""Kernel-level GPU checks: matrix-core operand layouts, the two GEMM forms
against each other, and the FP32 form against the FP64 form step by step."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_mfma_layout_probe():
    from os_d_amd import qjx
    _say("case: mfma layout probe")
    rng = np.random.default_rng(1)
    A32 = rng.standard_normal((32, 2)).astype(np.float32)
    B32 = rng.standard_normal((2, 32)).astype(np.float32)
    A64 = rng.standard_normal((16, 4))
    B64 = rng.standard_normal((4, 16))
    # small integers are exact in bf16, so the bf16 product is exact in FP32
    Abf = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    Bbf = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    D32, D64, Dbf = qjx.mfma_probe(A32, B32, A64, B64, Abf, Bbf)
    np.testing.assert_allclose(D32, A32.astype(np.float64) @ B32.astype(np.float64), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(D64, A64 @ B64, rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(Dbf, Abf @ Bbf)
    # the FP32 step GEMM's 16x16x4 product; small integers keep every sum exact
    A16 = rng.integers(-8, 9, (16, 4)).astype(np.float32)
    B16 = rng.integers(-8, 9, (4, 16)).astype(np.float32)
    np.testing.assert_array_equal(qjx.mfma_probe_f32x16(A16, B16), A16 @ B16)


@pytest.mark.parametrize("prec", [0, 1])
def test_mfma_gemm_matches_fma_gemm(prec, monkeypatch):
    """Same ensemble, same RNG streams, the two GEMM kernels: identical jump
    sequences and states equal to the arithmetic's rounding."""
    from os_d_amd import qjx
    N, S, D, M = 60, 8, 12, 300
    _say("case: mfma vs fma prec=%d N=%d M=%d" % (prec, N, M))
    out = {}
    for kind in ("fma", "mfma"):
        monkeypatch.setenv("QJX_GEMM", kind)
        e = qjx.Engine(N, ex_deep=D, steps_per_period=S, precision=prec)
        e.ensemble_init(M, seed=3, jump=1)
        e.trans(0)
        e.run_periods(2, 1)
        st = e.get_state()
        out[kind] = (st, [e.get_jumps(tr)["times"] for tr in range(M)], e.stats()["matvecs"])
        e.close()
    (sa, ja, ma), (sb, jb, mb) = out["fma"], out["mfma"]
    same = sum(np.array_equal(x, y) for x, y in zip(ja, jb))
    _say("identical jump sequences: %d / %d; matvecs fma %d mfma %d" % (same, M, ma, mb))
    if prec == 0:
        assert same == M
        np.testing.assert_allclose(sb["psi"], sa["psi"], atol=1e-11)
    else:
        assert same >= int(0.9 * M)
