"""Pins the matching oracle (orc_gm, a restatement of app_gmatch/gmatch.h)
against brute-force injective homomorphisms / |Aut|, closed forms and an
independent k-clique counter."""
import json
import math
import os

import numpy as np
import pytest

import graphs as G

HERE = os.path.dirname(os.path.abspath(__file__))
PRESETS = list(range(0, 19)) + [23, 24, 25, 26, 27]


def orc_plan(g2, orc, qadj):
    p = g2.plan_build(qadj)
    return orc.OrcPlan.from_buffer_copy(bytes(p))


def bfs_order(qadj):
    n = len(qadj)
    order, seen = [0], {0}
    while len(order) < n:
        for u in list(order):
            for v in range(n):
                if qadj[u] >> v & 1 and v not in seen:
                    seen.add(v)
                    order.append(v)
    pos = {v: i for i, v in enumerate(order)}
    out = [0] * n
    for v in range(n):
        for w in range(n):
            if qadj[v] >> w & 1:
                out[pos[v]] |= 1 << pos[w]
    return out


@pytest.mark.parametrize("q", PRESETS)
def test_presets_vs_bruteforce(g2, orc, q):
    g = G.gnp(g2, 26 if q < 19 else 30, 0.35, 100 + q)
    n, qadj = g2.query_preset(q)
    homs = orc.bruteforce_pattern(g, bfs_order(qadj))
    aut = orc.automorphism_count(qadj)
    assert homs % aut == 0
    assert orc.gm(g, orc_plan(g2, orc, qadj))["count"] == homs // aut


def test_random_patterns_vs_bruteforce(g2, orc):
    with open(os.path.join(HERE, "golden", "plans.json")) as f:
        fixtures = json.load(f)["random"]
    g = G.gnp(g2, 24, 0.4, 7)
    for item in fixtures[:24]:
        qadj = g2.query_remap(item["qadj"])
        homs = orc.bruteforce_pattern(g, bfs_order(qadj))
        aut = orc.automorphism_count(qadj)
        assert orc.gm(g, orc_plan(g2, orc, qadj))["count"] == homs // aut, item["seed"]


@pytest.mark.parametrize("n", [3, 6, 11])
def test_cliques_closed_form(g2, orc, n):
    g = G.complete(g2, n)
    for q, k in ((0, 3), (24, 4), (25, 5)):
        _, qadj = g2.query_preset(q)
        assert orc.gm(g, orc_plan(g2, orc, qadj))["count"] == math.comb(n, k)


def test_kclique_independent_method(g2, orc):
    g = g2.gen_rmat(11, 16, 0.57, 0.19, 0.19, 77).preprocess()
    for q, k in ((0, 3), (24, 4), (25, 5)):
        _, qadj = g2.query_preset(q)
        assert orc.gm(g, orc_plan(g2, orc, qadj))["count"] == orc.kclique_count(g, k)


def test_per_root_counts_sum(g2, orc):
    g = G.gnp(g2, 60, 0.2, 5)
    _, qadj = g2.query_preset(24)
    r = orc.gm(g, orc_plan(g2, orc, qadj), per_root=True)
    assert int(r["per_root_count"].sum()) == r["count"]


def test_byte_model_hand_worked(g2, orc):
    # Triangles (preset 0) on K_4.  Plan: position 1 extends from 0 with
    # M1 > M0; position 2 from 0 and 1 (row of 0 first on the degree tie) with
    # M2 > M0, M2 > M1; position 1 stores its list by EXPAND, 4(i+1) bytes per
    # entry.  Per extension of a match of length i through parent row pi:
    # 4i + 4d(pi) + c(1, d(b)) per candidate passing the conditions for each
    # other back neighbour b, + the stored list.  All degrees are 3, so
    # 4d = 12 and c(1,3) = 4 min(4, 3) = 12.
    #   root 0: 4 | i=1: 4+12, list {1,2,3}: 24 | i=2 via 1: 8+12+2*12,
    #           via 2: 8+12+12, via 3: 8+12                         = 140
    #   root 1: 4 | i=1: 4+12, list {2,3}: 16 | via 2: 8+12+12, via 3: 20 = 88
    #   root 2: 4 | i=1: 4+12, list {3}: 8 | via 3: 20                 = 48
    #   root 3: 4 | i=1: 4+12, empty list                              = 20
    # (Pricing the list as PREFIX, 4(i + |list|), would give 288.)
    _, qadj = g2.query_preset(0)
    p = orc_plan(g2, orc, qadj)
    assert list(p.strategy)[:3] == [0, 0, 0] and list(p.bn[2])[:2] == [0, 1]
    r = orc.gm(G.complete(g2, 4), p)
    assert r["count"] == 4 and r["bytes"] == 140 + 88 + 48 + 20


def test_ranked_kclique_split(g2, orc):
    """orc_kclique_count_ranked: every source set of a partition counts the
    cliques whose lowest (degree, id) vertex it holds, so the parts add up to
    the whole-graph count of the id-ordered counter and of orc_gm; all
    sources give the whole count; a hand-worked case fixes the rule."""
    g = g2.gen_rmat(10, 16, 0.57, 0.19, 0.19, 0x6732A1D0C0000002).preprocess()
    bins = np.random.default_rng(1).integers(0, 4, g.nv)
    for q, k in ((0, 3), (24, 4), (25, 5)):
        whole = orc.gm(g, orc.OrcPlan.from_buffer_copy(bytes(g2.plan_for_preset(q))))["count"]
        assert orc.kclique_count(g, k) == whole
        assert orc.kclique_count_ranked(g, k) == whole
        parts = [orc.kclique_count_ranked(g, k, sources=np.nonzero(bins == b)[0].astype(np.uint32))
                 for b in range(4)]
        assert sum(parts) == whole
    # triangle {0,1,2} plus the edge 2-3: degrees 2,2,3,1.  The lowest key of
    # the triangle is vertex 0 (degree 2, lower id than 1)
    tri = G.from_pairs(g2, [(0, 1), (0, 2), (1, 2), (2, 3)], 4)
    assert [orc.kclique_count_ranked(tri, 3, sources=np.array([v], np.uint32)) for v in range(4)] == [1, 0, 0, 0]
    # relabelled so that the lowest id is the degree-3 vertex: the rule follows degree first
    tri2 = G.from_pairs(g2, [(1, 2), (1, 0), (2, 0), (0, 3)], 4)
    assert [orc.kclique_count_ranked(tri2, 3, sources=np.array([v], np.uint32)) for v in range(4)] == [0, 1, 0, 0]


def test_local_bytes_hand_worked(g2, orc):
    """The local-path byte model on K_4 with the 4-clique plan: all degrees
    are 3, so ranks follow ids.  Root 0: U = {1,2,3}, m = 3 >= 3: 16 + 12 +
    member 1 (d = 2, a = 2: 16 + 4 min(2, 2*3)) + member 2 (d = 1, a = 1:
    16 + 4 min(1, 1*2)) = 72; roots 1, 2, 3 have m < 3: 16 each.  Total 120,
    all in the first pass.  Triangles (n = 3): root 1 has U = {2,3}, m = 2:
    16 + 8 + (16 + 4) = 44; root 0: 16 + 12 + 24 + 20 = 72; roots 2, 3: 16."""
    k4 = G.complete(g2, 4)
    assert orc.local_bytes(k4, 4).tolist() == [120, 0, 0, 0, 0]
    assert orc.local_bytes(k4, 3).tolist() == [72 + 44 + 32, 0, 0, 0, 0]
    assert orc.local_bytes(k4, 2).tolist() == [64, 0, 0, 0, 0]
    # a star: the hub has the highest degree, every leaf's U is {hub}
    star = G.from_pairs(g2, [(0, i) for i in range(1, 6)], 6)
    assert orc.local_bytes(star, 3).tolist() == [6 * 16, 0, 0, 0, 0]
    # K_{300,320}: each of the 320 degree-300 vertices ranks below the 300
    # degree-320 ones, so its U is that whole side (m = 300, rows of stride 5:
    # 1500 words, the second pass): 16 + 4*300 + 299 members with no later
    # neighbour (16 each) = 6000 B; the 300 others have m = 0: 16 B each
    big = G.complete_bipartite(g2, 300, 320)
    assert orc.local_bytes(big, 3).tolist() == [300 * 16, 320 * 6000, 0, 0, 0]
