"""qjx_chars on the GPU against the closed forms of spin-coherent states (norm, <k>,
m2, dispersion, energy, abs_diag_rho and its ensemble average as dump_adr_avg writes it,
core.cpp:665-725, experiment.cpp:1395-1469, dump.cpp:26-61), for every precision form
and across the 64-row padding edges of the state layout."""
import numpy as np
import pytest

from chars_closed_form import coherent_chars, coherent_ensemble

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,N", [(0, 1), (0, 63), (0, 64), (0, 200), (1, 64), (2, 129)])
def test_gpu_chars_closed_forms(precision, N):
    from os_d_amd import qjx
    print("case: chars closed forms precision=%d N=%d" % (precision, N), flush=True)
    M = 40
    prm = dict(prm_E=0.7, prm_U=0.45, prm_J=1.3)
    e = qjx.Engine(N, ex_deep=2, steps_per_period=4, precision=precision, **prm)
    e.ensemble_init(M, jump=0)
    p0, phi0, c0, psi0 = coherent_ensemble(N, M, seed=N)
    p1, phi1, c1, psi1 = coherent_ensemble(N, M, seed=100 + N)
    # FP64 states are exact; the FP32 forms store the state in FP32 (2^-24 relative)
    rt = 1e-11 if precision == 0 else 2e-5
    e.set_state(psi=psi0)
    s = e.chars(is_start=1, adr=True)
    e.set_state(psi=psi1)
    g = e.chars(adr=True, adr_avg=True)
    for tr in range(M):
        w0 = coherent_chars(N, p0[tr], phi0[tr], c0[tr], **prm)
        w1 = coherent_chars(N, p1[tr], phi1[tr], c1[tr], **prm)
        assert s["norm"][tr] == pytest.approx(w0["norm"], rel=rt)
        assert s["mean"][tr] == pytest.approx(w0["mean"], rel=rt, abs=rt)
        assert s["dispersion"][tr] == 0.0
        assert g["mean"][tr] == pytest.approx(w1["mean"], rel=rt, abs=rt)
        assert g["m2"][tr] == pytest.approx(w1["m2"], rel=rt, abs=rt * N)
        assert g["energy"][tr] == pytest.approx(w1["energy"], rel=rt, abs=rt * N)
        assert g["dispersion"][tr] == pytest.approx((N * (p1[tr] - p0[tr])) ** 2, rel=10 * rt, abs=rt * N)
        np.testing.assert_allclose(g["adr"][tr], w1["adr"], rtol=rt, atol=rt * 1e-3)
    mix = np.mean([coherent_chars(N, p1[tr], phi1[tr], c1[tr], **prm)["adr"] for tr in range(M)], axis=0)
    np.testing.assert_allclose(g["adr_avg"], mix, rtol=rt, atol=rt * 1e-3)
    e.close()
