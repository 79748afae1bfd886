"""Small deterministic graphs with known clique structure (test helpers)."""
import itertools
import random

import numpy as np


def from_pairs(g2, pairs, n):
    if not pairs:
        return g2.Graph(np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32))
    a = np.array([p[0] for p in pairs], np.uint32)
    b = np.array([p[1] for p in pairs], np.uint32)
    return g2.Graph.from_edges(a, b, n)


def gnp(g2, n, p, seed):
    rng = random.Random(seed)
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p]
    return from_pairs(g2, pairs, n)


def complete(g2, n):
    return from_pairs(g2, list(itertools.combinations(range(n), 2)), n)


def complete_bipartite(g2, a, b):
    return from_pairs(g2, [(i, a + j) for i in range(a) for j in range(b)], a + b)


def cycle(g2, n):
    return from_pairs(g2, [(i, (i + 1) % n) for i in range(n)], n)


def turan(g2, sizes):
    """complete multipartite graph: every maximal clique picks one vertex per part"""
    part = []
    for k, s in enumerate(sizes):
        part += [k] * s
    n = len(part)
    return from_pairs(g2, [(i, j) for i in range(n) for j in range(i + 1, n) if part[i] != part[j]], n)


def moon_moser(g2, k):
    """complement of k disjoint triangles: 3^k maximal cliques of size k"""
    return turan(g2, [3] * k)


def disjoint_union(g2, graphs):
    pairs, off = [], 0
    for g in graphs:
        rp, cols = g.row_ptrs, g.cols
        for u in range(g.nv):
            for e in range(int(rp[u]), int(rp[u + 1])):
                if cols[e] > u:
                    pairs.append((u + off, int(cols[e]) + off))
        off += g.nv
    return from_pairs(g2, pairs, off)


def shuffled(g2, g, seed):
    """same graph with vertex ids permuted"""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(g.nv).astype(np.uint32)
    rp, cols = g.row_ptrs, g.cols
    src = np.repeat(np.arange(g.nv, dtype=np.uint32), np.diff(rp).astype(np.int64))
    return g2.Graph.from_edges(perm[src], perm[cols], g.nv)
