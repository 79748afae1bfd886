"""Pins of the oracle's 2-hop ball estimate (orc_two_hop, a HyperLogLog of
|{v} u N(v) u N(N(v))| with 64 registers), the integers the device feature
kernel must reproduce.

The estimate is checked against the exact ball sizes (orc_two_hop_exact, a
brute-force marker walk, itself checked against scipy's (I + A + A^2) pattern)
through what the method fixes: near-exact linear counting on balls of a few
vertices, the known ball of every vertex of a star and a complete graph, and
the estimator's bias and spread (relative standard error 1.04 / sqrt(64) = 13%)
over the vertices of skewed graphs.
"""
import numpy as np
import pytest

import graphs as G


def star(g2, d):
    return G.from_pairs(g2, [(0, j) for j in range(1, d + 1)], d + 1)


def path(g2, n):
    return G.from_pairs(g2, [(i, i + 1) for i in range(n - 1)], n)


def ball_scipy(g):
    import scipy.sparse as sp
    n = g.nv
    src = np.repeat(np.arange(n), np.diff(g.row_ptrs).astype(np.int64))
    A = sp.csr_matrix((np.ones(g.ne, np.int64), (src, g.cols.astype(np.int64))), shape=(n, n))
    B = sp.identity(n, dtype=np.int64, format="csr") + A + A @ A
    return np.diff((B > 0).tocsr().indptr).astype(np.int64)


def test_exact_checker(g2, orc):
    for g in (G.gnp(g2, 200, 0.03, 1), g2.gen_rmat(9, 8, 0.57, 0.19, 0.19, 2).preprocess(), star(g2, 30)):
        assert np.array_equal(orc.two_hop_exact(g), ball_scipy(g))


def test_tiny_balls_near_exact(g2, orc):
    """A ball of k <= 5 vertices fills k registers unless two collide; then
    linear counting 64 ln(64 / (64 - k)) rounds to k.  An isolated vertex and
    the vertices of short paths and triangles are such balls."""
    for g in (G.from_pairs(g2, [], 5), path(g2, 2), path(g2, 3), path(g2, 5), G.cycle(g2, 5), G.complete(g2, 3)):
        est, ex = orc.two_hop(g), orc.two_hop_exact(g)
        assert np.all(np.abs(est - ex) <= 1), (est, ex)
        assert np.mean(est == ex) >= 0.8, (est, ex)


@pytest.mark.parametrize("d", [20, 200, 2000])
def test_star_and_complete(g2, orc, d):
    """every vertex of the star K_{1,d} and of K_{d+1} has the ball of all
    d + 1 vertices: one set, so every estimate is the same and within three
    standard errors"""
    for g in (star(g2, d), G.complete(g2, d + 1) if d <= 200 else star(g2, d)):
        est = orc.two_hop(g)
        assert np.all(est == est[0])
        assert abs(est[0] / (d + 1) - 1) < 3 * 1.04 / 8, (est[0], d + 1)


def test_bias_and_spread_on_skewed_graphs(g2, orc):
    """over many vertices the ratio estimate / exact is unbiased within a few
    percent and spread by about the estimator's standard error"""
    for g in (g2.gen_rmat(12, 16, 0.57, 0.19, 0.19, 5).preprocess(),
              g2.gen_chung_lu(20000, 2.3, 2.0, 500.0, 100000, 3).preprocess()):
        est, ex = orc.two_hop(g), orc.two_hop_exact(g)
        r = est / ex
        assert abs(np.mean(r) - 1) < 0.05, np.mean(r)
        assert 0.04 < np.std(r) < 0.2, np.std(r)
        big = ex > 1000  # the HyperLogLog range (above 5m/2 = 160)
        assert big.sum() > 500
        assert abs(np.mean(r[big]) - 1) < 0.1, np.mean(r[big])


def test_sources_subset(g2, orc):
    g = g2.gen_rmat(10, 16, 0.57, 0.19, 0.19, 7).preprocess()
    src = np.arange(0, g.nv, 3, dtype=np.uint32)
    assert np.array_equal(orc.two_hop(g, src), orc.two_hop(g)[src])
    assert np.array_equal(orc.two_hop_exact(g, src), orc.two_hop_exact(g)[src])
