"""Batched matching of many query patterns (persistent launches + autotuner).

The reference's `run` matches one query per process (app_gmatch/main.cu,
system/worker.h:28-280); a workload of hundreds of small patterns then pays a
launch, a host sync and a ragged tail per query.  Here all queries go into
one persistent launch (g2_gm_run_batch): waves fetch (query, root block)
tasks from one counter, queries ordered heaviest first by a static cost key,
so only the last, lightest tasks leave waves idle.  Each query's count is the
sum over the same root tasks as its solo run and is checked equal to it.
"""
from __future__ import annotations

import numpy as np


def _splitmix(x: int):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x, z ^ (z >> 31)


def random_queries(g2, nq: int, seed: int, nmin: int = 3, nmax: int = 7, dmin: float = 0.4,
                   dmax: float = 1.0):
    """nq seeded random connected patterns: size uniform in [nmin, nmax], edge
    density uniform in [dmin, dmax] (a random spanning tree plus random
    edges up to round(density * n(n-1)/2)), remapped as the reference's
    Graph::ReMapVertexId does before planning."""
    out = []
    s = seed
    for _ in range(nq):
        s, a = _splitmix(s)
        s, b = _splitmix(s)
        s, c = _splitmix(s)
        n = nmin + a % (nmax - nmin + 1)
        dens = dmin + (dmax - dmin) * (b >> 11) / float(1 << 53)
        out.append(g2.query_remap(g2.query_random(n, dens, c)))
    return out


def cost_key(qadj) -> tuple:
    """static heaviness of a pattern: more vertices first, then fewer edges
    (sparser patterns of one size have more matches)"""
    n = len(qadj)
    m = sum(bin(r).count("1") for r in qadj) // 2
    return (-n, m)


def heavy_first(queries) -> np.ndarray:
    return np.array(sorted(range(len(queries)), key=lambda i: cost_key(queries[i])), dtype=np.uint32)


def isomorphism_classes(g2, queries):
    """(representatives, class index of every query): isomorphic patterns
    have the same count (injective homomorphisms / |Aut|), so a batch counts
    each class once and hands the count to every member"""
    codes, reps, cls = {}, [], []
    for q in queries:
        c = g2.query_canonical(q)
        if c not in codes:
            codes[c] = len(reps)
            reps.append(q)
        cls.append(codes[c])
    return reps, np.array(cls, dtype=np.int64)
