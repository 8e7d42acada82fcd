"""STD-deep (task_id 4) tree: the generalised tree the runner uses for the quantum-jump
path must be the reference's deep tree (core.cpp:9928-9981, qj_proc.cpp:4-28, 345-360).

The expected level dt's and steps per sub-period come from tests/deep_tree.py, which
follows the reference's node recursion; the oracle built with S = 8 deep_num_steps must
hold exp(H_eff,b dt_i) at exactly those dt's and cover a sub-period with 4 level-0
steps.  A tree with twice the dt (S = 4 deep_num_steps) fails both checks.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from deep_tree import reference_deep_tree


@pytest.mark.parametrize("k", [1, 3, 5])
def test_reference_deep_tree_shape(k):
    T = 2 * np.pi
    t = reference_deep_tree(T, k, 6)
    # both children of a head are the same level chain
    assert t["children"][0] == t["children"][1]
    assert t["level0_steps_per_sub"] == 4
    assert t["S"] == 8 * k
    # level i is T / (8k 2^i), bit for bit what T / S halved i times gives
    dt = T / float(t["S"])
    for i, x in enumerate(t["level_dt"]):
        assert x == dt, (i, x, dt)
        dt *= 0.5
    # a sub-period (4 level-0 steps) spans T / (2k): 2k of them make one period
    assert t["level0_steps_per_sub"] * t["level_dt"][0] * t["sub_per_period"] == pytest.approx(T, rel=1e-15)


@pytest.mark.parametrize("k", [1, 3])
def test_oracle_deep_tree_levels_match_reference_recursion(oracle_mod, k):
    N, D = 6, 5
    T = 2 * np.pi
    ref = reference_deep_tree(T, k, D)
    m = oracle_mod.model(N, gamma=0.3, S=ref["S"], D=D)
    mm = oracle_mod.model_matrices(m)
    e = oracle_mod.Ensemble(m, 4, jump=1, threads=1)
    for b in range(2):
        for lvl in range(D):
            want = sl.expm(mm["Hqj%d" % b] * ref["level_dt"][lvl])
            np.testing.assert_allclose(e.get_propagator(b, lvl), want, rtol=0, atol=1e-12)
            # the S = 4k tree (twice the dt) is far from it
            assert np.abs(sl.expm(mm["Hqj%d" % b] * 2 * ref["level_dt"][lvl]) - want).max() > 1e-4
    # one sub-period = level0_steps_per_sub top-level steps of the oracle
    e.trans(0)
    e.steps(ref["level0_steps_per_sub"], 1)
    t = e.get_state()["t"]
    np.testing.assert_allclose(t, T / (2 * k), rtol=1e-14)
    e.steps(ref["level0_steps_per_sub"] * (ref["sub_per_period"] - 1), 1)
    np.testing.assert_allclose(e.get_state()["t"], T, rtol=1e-14)
