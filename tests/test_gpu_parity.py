"""GPU engine vs the FP64 CPU oracle (parity proper; needs an MI355X).

The oracle restates the reference's per-trajectory recursion
(source/cpp/QJX/QJX/qj_proc.cpp:271-343, experiment.cpp:1239-1382) with identical
MCG59 leap-frog streams.  With the same propagator tree on both sides the only
difference is the summation order of the matrix-vector products, so jump times,
channels, eta values and RNG states must agree exactly over the horizon and the
observables to 1e-10 relative.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(N, S, D, M, prec=0, seed=0, jump=2, **kw):
    from os_d_amd import qjx
    e = qjx.Engine(N, ex_deep=D, steps_per_period=S, precision=prec, **kw)
    e.ensemble_init(M, seed=seed, jump=jump)
    return e


def _share_tree(orc_ens, eng, D):
    for b in range(2):
        for lvl in range(D):
            eng.set_propagator(b, lvl, orc_ens.get_propagator(b, lvl))


@pytest.mark.parametrize("N,S", [(10, 4), (100, 200)])
def test_propagators_match_oracle(oracle_mod, N, S):
    _say("case: propagators N=%d S=%d" % (N, S))
    D = 16
    m = oracle_mod.model(N, S=S, D=D)
    o = oracle_mod.Ensemble(m, 1)
    e = _engine(N, S, D, 1)
    worst = 0.0
    for b in range(2):
        for lvl in range(D):
            Ug = e.get_propagator(b, lvl)
            Uo = o.get_propagator(b, lvl)
            worst = max(worst, np.abs(Ug - Uo).max())
    _say("max |U_gpu - U_oracle| = %.3e" % worst)
    assert worst < 1e-12


@pytest.mark.parametrize("N,S", [(100, 4), (700, 4), (300, 200)])
def test_squaring_chain_tree_equals_direct_expm(N, S, monkeypatch):
    """The tree is built deepest level first, each level whose scaling step needs
    s >= 1 squarings as the square of the level below it; that is the full scaling and
    squaring of every level, bit for bit."""
    from os_d_amd import qjx
    _say("case: squaring chain vs direct expm N=%d S=%d" % (N, S))
    D = 16
    trees = {}
    for direct in ("0", "1"):
        monkeypatch.setenv("QJX_EXPM_DIRECT", direct)
        e = qjx.Engine(N, ex_deep=D, steps_per_period=S, precision=qjx.FP64)
        trees[direct] = [e.get_propagator(b, lvl) for b in range(2) for lvl in range(D)]
        e.close()
    for a, b in zip(trees["0"], trees["1"]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("N,S,M,P", [(10, 4, 64, 6), (100, 200, 256, 2)])
def test_fp64_jump_parity_shared_tree(oracle_mod, N, S, M, P):
    _say("case: fp64 parity shared tree N=%d S=%d M=%d P=%d" % (N, S, M, P))
    D = 16
    m = oracle_mod.model(N, S=S, D=D)
    o = oracle_mod.Ensemble(m, M, jump=2)
    e = _engine(N, S, D, M, jump=2)
    _share_tree(o, e, D)
    o.trans(1)
    e.trans(1)
    co = o.chars(is_start=1)
    ce = e.chars(is_start=1)
    for p in range(P):
        o.periods(1, 1)
        e.run_periods(1, 1)
    so, se = o.get_state(), e.get_state()
    njo = sum(len(o.get_jumps(tr)["times"]) for tr in range(M))
    _say("oracle jumps logged: %d" % njo)
    assert njo > 0
    for tr in range(M):
        jo, je = o.get_jumps(tr), e.get_jumps(tr)
        assert np.array_equal(jo["times"], je["times"]), "trajectory %d jump times" % tr
        assert np.array_equal(jo["types"], je["types"])
        assert np.array_equal(jo["etas"], je["etas"])
        np.testing.assert_allclose(je["norms"], jo["norms"], rtol=1e-12)
    assert np.array_equal(so["rng"], se["rng"])
    assert np.array_equal(so["eta"], se["eta"])
    assert np.array_equal(so["t"], se["t"])
    assert np.array_equal(so["njumps"], se["njumps"])
    np.testing.assert_allclose(se["psi"], so["psi"], rtol=0, atol=1e-10)
    co, ce = o.chars(), e.chars()
    for k in ("norm", "mean", "m2", "energy", "dispersion"):
        np.testing.assert_allclose(ce[k], co[k], rtol=1e-10, atol=1e-12, err_msg=k)
    st = e.stats()
    _say("gpu matvecs=%d oracle matvecs=%d iterations=%d" % (st["matvecs"], o.num_matvecs(),
                                                             st["iterations"]))
    assert st["matvecs"] == o.num_matvecs()


def test_fp64_own_tree_horizon(oracle_mod):
    """Each side builds its own tree (GPU: Taylor scaling-squaring; oracle: Pade):
    trajectories still agree exactly except for near-ties of ||psi||^2 and eta."""
    N, S, M, P, D = 20, 8, 128, 4, 16
    _say("case: fp64 own trees N=%d S=%d M=%d P=%d" % (N, S, M, P))
    m = oracle_mod.model(N, S=S, D=D)
    o = oracle_mod.Ensemble(m, M, jump=1)
    e = _engine(N, S, D, M, jump=1)
    o.trans(0)
    e.trans(0)
    o.periods(P, 1)
    for p in range(P):
        e.run_periods(1, 1)
    same = sum(np.array_equal(o.get_jumps(tr)["times"], e.get_jumps(tr)["times"]) for tr in range(M))
    _say("trajectories with identical jump times: %d / %d" % (same, M))
    assert same >= int(0.97 * M)


def test_fp32_observables_close_to_fp64():
    """FP32 propagation with FP64 norm bookkeeping: ensemble means agree with the
    FP64 run within statistical error."""
    N, S, M, D = 30, 8, 512, 16
    _say("case: fp32 vs fp64 N=%d S=%d M=%d" % (N, S, M))
    res = {}
    for prec in (0, 1):
        e = _engine(N, S, D, M, prec=prec, jump=0)
        e.trans(0)
        e.chars(is_start=1)
        e.run_periods(4, 1)
        res[prec] = e.chars()
        e.close()
    for k in ("mean", "energy"):
        a, b = res[0][k], res[1][k]
        sig = np.sqrt(a.var() / M + b.var() / M) + 1e-12
        _say("%s: fp64 %.6f fp32 %.6f sigma %.3e" % (k, a.mean(), b.mean(), sig))
        assert abs(a.mean() - b.mean()) < 5 * sig
    np.testing.assert_allclose(res[1]["norm"], res[0]["norm"], rtol=0.5)
