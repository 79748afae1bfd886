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


def _peel_cores(g):
    """core numbers by repeated removal of a minimum-degree vertex (O(V^2))"""
    deg = g.degrees().astype(np.int64)
    alive = np.ones(g.nv, bool)
    core = np.zeros(g.nv, np.int64)
    k = 0
    for _ in range(g.nv):
        v = int(np.argmin(np.where(alive, deg, np.iinfo(np.int64).max)))
        k = max(k, int(deg[v]))
        core[v] = k
        alive[v] = False
        for u in g.cols[g.row_ptrs[v]:g.row_ptrs[v + 1]]:
            if alive[u]:
                deg[u] -= 1
    return core


def test_core_numbers_csr(g2, S):
    import graphs as G
    g = g2.gen_rmat(9, 8, 0.57, 0.19, 0.19, 7).preprocess()
    assert S.core_numbers_csr(g).tolist() == _peel_cores(g).tolist()
    # closed forms: K_n has core n-1 everywhere, a cycle 2, K_{3,5} 3
    assert set(S.core_numbers_csr(G.complete(g2, 9)).tolist()) == {8}
    assert set(S.core_numbers_csr(G.cycle(g2, 11)).tolist()) == {2}
    assert set(S.core_numbers_csr(G.complete_bipartite(g2, 3, 5)).tolist()) == {3}
    # on a kcore-recoded graph the shell index is the core number
    r, _, shells, dg = g.kcore_recode()
    assert S.core_numbers_csr(r).tolist() == S.core_numbers(shells, r.nv).tolist()
    assert int(S.core_numbers_csr(r).max()) == dg


class _FakeExec:
    def __init__(self, f):
        self.f = f

    def features(self):
        return self.f


def test_root_features_formula(S):
    """g2h_root_features against the formula in float64, rounded once"""
    rng = np.random.default_rng(9)
    n, nv = 300, 1000
    deg = rng.integers(0, 50, n)
    f = np.stack([deg, rng.integers(0, 900, n), rng.integers(0, 200, n), rng.integers(0, 50, n),
                  rng.integers(0, 50, n)], axis=1).astype(np.int64)
    src = rng.choice(nv, n, replace=False).astype(np.uint32)
    core = rng.integers(0, 20, nv).astype(np.uint32)
    x = S.root_features(_FakeExec(f), src, core, nv)
    d = f[:, 0].astype(np.float64)
    pairs = d * (d - 1)
    ref = np.stack([np.log1p(d), np.log1p(core[src].astype(np.float64)), np.log1p(f[:, 1]), np.log1p(f[:, 2]),
                    np.where(pairs > 0, 2 * f[:, 2] / np.maximum(pairs, 1), 0.0), src / nv,
                    f[:, 3] / np.maximum(d, 1), np.log1p(f[:, 4])], axis=1).astype(np.float32)
    assert np.array_equal(x, ref)


def test_standardizer_and_model_file(S, tmp_path):
    rng = np.random.default_rng(10)
    x = rng.normal(3.0, 2.0, (1000, 8)).astype(np.float32)
    x[:, 5] = 1.5  # a constant feature: scale floored at 1e-6
    n = S.Standardizer(x)
    assert np.allclose(n.mu, x.astype(np.float64).mean(0), rtol=1e-6, atol=1e-6)
    assert np.allclose(n.sd[[0, 1, 2, 3, 4, 6, 7]], x.astype(np.float64).std(0)[[0, 1, 2, 3, 4, 6, 7]], rtol=1e-5)
    assert n.sd[5] == np.float32(1e-6)
    z = n(x)
    assert np.allclose(z[:, 0].mean(), 0, atol=1e-4) and np.allclose(z[:, 0].std(), 1, atol=1e-4)
    params = rng.normal(size=4801).astype(np.float32)
    path = str(tmp_path / "m.model")
    S.save_model(path, n, params)
    n2, p2 = S.load_model(path, 4801)
    assert np.array_equal(p2, params) and np.array_equal(n2.mu, n.mu) and np.array_equal(n2.sd, n.sd)
    with pytest.raises(RuntimeError):
        S.load_model(path, 4800)
