"""Pins the MCE oracle (oracle/oracle.cpp: orc_mce, a restatement of
app_BK/BK.h) with checks that do not reuse its formula: brute-force subset
enumeration, closed forms, additivity and relabelling invariance."""
import math

import numpy as np
import pytest

import graphs as G


def hist_dict(h):
    return {int(i): int(h[i]) for i in np.nonzero(h)[0]}


@pytest.mark.parametrize("n,p,seed", [(8, 0.5, 1), (12, 0.4, 2), (14, 0.6, 3), (16, 0.3, 4),
                                      (18, 0.5, 5), (18, 0.8, 6), (10, 0.9, 7), (17, 0.15, 8)])
def test_bruteforce(g2, orc, n, p, seed):
    g = G.gnp(g2, n, p, seed)
    tot, hist = orc.bruteforce_mce(g)
    r = orc.mce(g, hist_len=64)
    assert r["total"] == tot
    assert hist_dict(r["hist"]) == hist_dict(hist)


@pytest.mark.parametrize("n", [2, 3, 5, 9, 13])
def test_complete_graph(g2, orc, n):
    r = orc.mce(G.complete(g2, n))
    assert r["total"] == 1 and hist_dict(r["hist"]) == {n: 1}


@pytest.mark.parametrize("a,b", [(1, 1), (2, 5), (7, 4)])
def test_complete_bipartite(g2, orc, a, b):
    r = orc.mce(G.complete_bipartite(g2, a, b))
    assert r["total"] == a * b and hist_dict(r["hist"]) == {2: a * b}


@pytest.mark.parametrize("n", [4, 5, 17])
def test_cycle(g2, orc, n):
    r = orc.mce(G.cycle(g2, n))
    assert r["total"] == n and hist_dict(r["hist"]) == {2: n}


@pytest.mark.parametrize("k", [2, 3, 4, 6])
def test_moon_moser(g2, orc, k):
    r = orc.mce(G.moon_moser(g2, k))
    assert r["total"] == 3 ** k and hist_dict(r["hist"]) == {k: 3 ** k}


def test_turan(g2, orc):
    sizes = [2, 3, 4, 1]
    r = orc.mce(G.turan(g2, sizes))
    assert r["total"] == math.prod(sizes) and hist_dict(r["hist"]) == {4: math.prod(sizes)}


def test_disjoint_union_and_isolated(g2, orc):
    parts = [G.complete(g2, 5), G.cycle(g2, 6), G.moon_moser(g2, 3), G.gnp(g2, 15, 0.4, 9)]
    parts.append(g2.Graph(np.zeros(4, np.uint64), np.zeros(0, np.uint32)))  # 3 isolated vertices
    u = G.disjoint_union(g2, parts)
    r = orc.mce(u)
    tot, hist = 0, np.zeros(256, np.uint64)
    for p in parts:
        rp = orc.mce(p)
        tot += rp["total"]
        hist += rp["hist"]
    assert r["total"] == tot and hist_dict(r["hist"]) == hist_dict(hist)


def test_relabel_invariance_and_root_partition(g2, orc):
    g = G.gnp(g2, 120, 0.12, 10)
    r = orc.mce(g, per_root=True)
    s = orc.mce(G.shuffled(g2, g, 3))
    assert r["total"] == s["total"] and hist_dict(r["hist"]) == hist_dict(s["hist"])
    assert int(r["per_root_nodes"].sum()) == r["nodes"]
    # any partition of the roots into bins adds up to the whole
    bins = np.arange(g.nv) % 3
    tot = sum(orc.mce(g, sources=np.nonzero(bins == b)[0])["total"] for b in range(3))
    assert tot == r["total"]


def test_per_root_nodes_tiny(g2, orc):
    # path a-b-c (ids 0-1-2): roots 0 and 1 own the two maximal edges.
    # root 0: P={1}, X={} -> node, pivot 1, Q={1} -> child [1 0 | ] clique: 2 nodes
    # root 1: P={2}, X={0}: pivot 2 (count 0 ties -> largest id), Q={2}: child {1,2} clique: 2 nodes
    # root 2: P={}, X={1}: 1 node
    g = G.from_pairs(g2, [(0, 1), (1, 2)], 3)
    r = orc.mce(g, per_root=True)
    assert r["total"] == 2
    assert list(r["per_root_nodes"]) == [2, 2, 1]


def test_per_root_nodes_tie_rule(g2, orc):
    # Pins the pivot rule of app_BK/BK.h:282-321: the pivot ranges over P and X,
    # and ties go to the LAST maximum in list order (ids ascending).
    #
    # Triangle 0-1-2, root 1: R={1}, P={2}, X={0}.  Scores |P & N(u)|: 0 -> 1,
    # 2 -> 0, so the pivot is the X member 0 and Q = P \ N(0) = {} : 1 node.
    # A pivot taken from P alone would be 2, with one more child: 2 nodes.
    tri = G.from_pairs(g2, [(0, 1), (0, 2), (1, 2)], 3)
    assert list(orc.mce(tri, per_root=True)["per_root_nodes"]) == [3, 1, 1]

    # Vertex 0 joined to 1..5, with 1-2, 1-3, 2-3, 2-4, 3-4, 4-5 among them.
    # Root 0: P={1..5}.  Scores 1:2, 2:3, 3:3, 4:3, 5:1 -> last maximum 4,
    # Q = P \ N(4) = {1,4}.
    #   q=1: P={2,3}; scores 1,1 -> pivot 3, Q={3}; child R+3, P={2};
    #        pivot 2, Q={2}; child {0,1,3,2} is a clique.            3 nodes
    #   q=4: P={2,3,5} (1 is not adjacent to 4); scores 1,1,0 -> pivot 3,
    #        Q={3,5}; q=3: P={2} -> pivot 2 -> clique {0,4,3,2}: 2 nodes;
    #        q=5: nothing adjacent -> clique {0,4,5}: 1 node.          4 nodes
    #   root total 1 + 3 + 4 = 8.  (Pivot 2, the first maximum, gives 7.)
    # Roots 1..5 each have 0 in X, adjacent to all of P.  Its score |P| is the
    # maximum, and any member tying with it is also adjacent to all of P, so
    # Q is empty: 1 node each.
    g = G.from_pairs(g2, [(0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (1, 2), (1, 3), (2, 3),
                          (2, 4), (3, 4), (4, 5)], 6)
    r = orc.mce(g, per_root=True)
    assert r["total"] == 3 and hist_dict(r["hist"]) == {3: 1, 4: 2}
    assert list(r["per_root_nodes"]) == [8, 1, 1, 1, 1, 1]


def test_byte_model_hand_worked(g2, orc):
    # Modelled bytes per node: 5s + sum_{u in P u X} c(P,u) + c(P,u*)
    #   + sum_{q in Q} [c(P u X, q) + 5|child_q|],  c(a,d) = 4 min(a+d, a(1+bitlen d)).
    # Path 0-1-2 (degrees 1, 2, 1):
    #   root 0: [0r 1p] 10; candidate 1: c(1,2)=12; pivot 1: 12; Q={1}:
    #           child [1r 0r]: c(1,2)+10 = 22; child node 10.          66
    #   root 1: [1r 0x 2p] 15; candidates 0, 2: c(1,1)=8 each; pivot 2: 8;
    #           Q={2}: child [2r 1r]: c(2,1)+10 = 12+10; child node 10.  71
    #   root 2: [2r 1x] 10; candidate 1: c(0,2)=0; pivot 1: 0.          10
    g = G.from_pairs(g2, [(0, 1), (1, 2)], 3)
    r = orc.mce(g, per_root=True)
    assert list(r["per_root_bytes"]) == [66, 71, 10] and r["bytes"] == 147
