"""Host side of the learned root schedule (g2-aimd_amd/sched.py): core
numbers from kcore shells, the (order, split) rule, the rank correlation used
to score the cost model, and the list-schedule makespan model."""
import importlib

import numpy as np
import pytest
from scipy import stats


@pytest.fixture(scope="module")
def S():
    return importlib.import_module("g2-aimd_amd.sched")


def test_core_numbers(S):
    # shells[k]..shells[k+1] are the vertices of core number k (kcore recode order)
    core = S.core_numbers(np.array([0, 3, 5, 5, 9]), 9)
    assert core.tolist() == [0, 0, 0, 1, 1, 3, 3, 3, 3]


def test_learned_schedule_rule(S):
    rng = np.random.default_rng(4)
    n, waves = 500, 64
    pred = rng.normal(3.0, 1.0, n)
    pred[7] = 12.0                     # one root far above the per-wave share
    later = rng.integers(0, 40, n)
    later[7] = 30
    order, split = S.learned_schedule(pred, later, waves)
    assert sorted(order.tolist()) == list(range(n))
    cost = np.expm1(pred)
    assert np.all(np.diff(cost[order]) <= 0)          # longest predicted first
    assert order[0] == 7
    share = cost.sum() / waves
    assert split[7] == min(int(np.ceil(cost[7] / share)), 30)
    assert np.all(split >= 1) and np.all(split <= np.maximum(later, 1))
    # uniform predictions: nothing exceeds its share, nothing is split
    o2, s2 = S.learned_schedule(np.full(n, 2.0), later, waves)
    assert np.all(s2 == 1) and o2.tolist() == list(range(n))


def test_spearman_matches_scipy(S):
    rng = np.random.default_rng(5)
    a = rng.normal(size=300)
    b = a + rng.normal(scale=0.7, size=300)
    assert S.spearman(a, b) == pytest.approx(stats.spearmanr(a, b).statistic, abs=1e-12)
    assert S.spearman(a, np.exp(a)) == pytest.approx(1.0)
    assert S.spearman(a, -a) == pytest.approx(-1.0)


def test_simulated_makespan(S):
    cost = np.array([4.0, 3.0, 3.0, 2.0])
    order = np.arange(4)
    # longest first on 2 waves: 4 | 3, then 3 -> the 3-wave, 2 -> the 4-wave
    assert S.simulated_makespan(cost, order, None, 2) == 6.0
    # a split root is dealt evenly over its parts
    assert S.simulated_makespan(np.array([10.0]), np.arange(1), np.array([5]), 5) == 2.0
    rng = np.random.default_rng(6)
    c = rng.exponential(size=200)
    m = S.simulated_makespan(c, np.argsort(-c), None, 16)
    assert max(c.sum() / 16, c.max()) <= m <= c.sum() / 16 + c.max()
