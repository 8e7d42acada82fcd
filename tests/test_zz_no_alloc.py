"""qjx_density_noalloc (fixed device scratch, no allocation) against qjx_density, through
the C ABI, on consecutive calls that share the scratch: a larger ensemble, then a
smaller one in another precision, then the largest dimension (fewer split-K partial
sums fit the scratch, so the sums are grouped differently), and the size limit."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(qjx, N, M, prec, steps):
    e = qjx.Engine(N, ex_deep=8, steps_per_period=8, precision=prec, diss_gamma=0.2)
    e.ensemble_init(M, seed=3, jump=0)
    e.trans(0)
    e.run_steps(steps, 0)
    return e


def test_density_noalloc_equals_density():
    from os_d_amd import qjx
    for N, M, prec, exact in ((300, 3000, qjx.FP64, True), (40, 500, qjx.FP32, True),
                              (1000, 2000, qjx.FP32, False)):
        _say("case: density_noalloc N=%d M=%d prec=%d" % (N, M, prec))
        e = _engine(qjx, N, M, prec, 3)
        r0, p0 = e.density()
        r1, p1 = e.density_noalloc()
        if exact:
            assert np.array_equal(r0, r1) and p0 == p1
        else:
            np.testing.assert_allclose(r1, r0, rtol=0, atol=1e-13)
            assert p1 == pytest.approx(p0, rel=1e-12)
        assert abs(np.trace(r1).real - 1.0) < 1e-6
        e.close()


def test_density_noalloc_rejects_sizes_beyond_its_scratch():
    from os_d_amd import qjx
    _say("case: density_noalloc limit d=1101")
    e = qjx.Engine(1100, ex_deep=2, steps_per_period=4, precision=qjx.FP32)
    e.ensemble_init(8, jump=0)
    with pytest.raises(qjx.QJXError) as ei:
        e.density_noalloc()
    assert ei.value.code == -1 and "1024" in str(ei.value)
    e.close()
