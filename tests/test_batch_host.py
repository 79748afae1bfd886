"""Host side of batched matching: the isomorphism code used to count each
class of patterns once, checked against brute-force isomorphism tests."""
import importlib
import itertools
import random

import numpy as np
import pytest


@pytest.fixture(scope="module")
def B():
    return importlib.import_module("g2-aimd_amd.batch")


def relabel(qadj, perm):
    n = len(qadj)
    out = [0] * n
    for u in range(n):
        for w in range(n):
            if qadj[u] >> w & 1:
                out[perm[u]] |= 1 << perm[w]
    return out


def isomorphic(a, b):
    if len(a) != len(b):
        return False
    return any(relabel(a, p) == b for p in itertools.permutations(range(len(a))))


def test_canonical_invariant_under_relabelling(g2, B):
    rng = random.Random(3)
    for q in B.random_queries(g2, 60, 11, nmin=3, nmax=8):
        c = g2.query_canonical(q)
        for _ in range(5):
            p = list(range(len(q)))
            rng.shuffle(p)
            assert g2.query_canonical(relabel(q, p)) == c


def test_canonical_separates_non_isomorphic(g2, B):
    qs = B.random_queries(g2, 80, 12, nmin=3, nmax=5)
    codes = [g2.query_canonical(q) for q in qs]
    for i in range(len(qs)):
        for j in range(i + 1, len(qs)):
            assert (codes[i] == codes[j]) == isomorphic(qs[i], qs[j])


def test_isomorphism_classes(g2, B):
    qs = B.random_queries(g2, 200, 13)
    reps, cls = B.isomorphism_classes(g2, qs)
    assert len(reps) == len(set(g2.query_canonical(q) for q in qs))
    for i, q in enumerate(qs):
        assert g2.query_canonical(q) == g2.query_canonical(reps[cls[i]])
    assert np.all(cls < len(reps))
