"""Host-side checks of the oriented local-bitmap matching path.

Its totals rely on one property: for an eligible plan the order conditions
pick exactly one member of every automorphism class of an embedding for ANY
total order of the data vertices, so the whole-graph count does not depend on
the vertex labelling.  The oracle (id order, app_gmatch/gmatch.h semantics)
pins that directly: relabelling the graph must not change the count."""
import ctypes
import importlib

import numpy as np
import pytest

import graphs as G

CLIQUES = [0, 23, 24, 25, 26, 27]


@pytest.fixture(scope="module")
def dev():
    return importlib.import_module("g2-aimd_amd.device")


def test_eligible_presets(g2, dev):
    ok = [q for q in list(range(1, 19)) + CLIQUES if dev.plan_local_eligible(g2.plan_for_preset(q))]
    assert ok == CLIQUES


@pytest.mark.parametrize("q", CLIQUES)
def test_count_invariant_under_relabelling(g2, orc, dev, q):
    p = g2.plan_for_preset(q)
    po = orc.OrcPlan.from_buffer_copy(bytes(p))
    g = g2.gen_rmat(10, 16, 0.57, 0.19, 0.19, 5).preprocess()
    base = orc.gm(g, po)["count"]
    for seed in (1, 2):
        assert orc.gm(G.shuffled(g2, g, seed), po)["count"] == base


def test_incomplete_conditions_rejected(g2, dev):
    # triangle plan with "u2 > u1" weakened to NON_EQUAL: both orders of
    # (u1,u2) now pass, so the total would depend on the labelling
    p = g2.plan_for_preset(0)
    assert dev.plan_local_eligible(p)
    u2 = p.seq[2]
    for k in range(p.cond_count[u2]):
        if p.cond[u2][k][1] == p.seq[1]:
            p.cond[u2][k][0] = 2
    assert not dev.plan_local_eligible(p)


def plan_copy(p):
    return type(p).from_buffer_copy(bytes(p))


def test_plan_tables_bounds_checked(g2, dev):
    """Plan compilation (host code shared with g2_gm_set_plan) accepts every
    plan the plan builder makes and refuses tables that would index outside
    their G2_MAXQ-sized arrays: seq/rev not inverse permutations, more than
    G2_MAXQ entries in a row, or a query vertex >= n."""
    plans = [g2.plan_for_preset(q) for q in list(range(19)) + CLIQUES[1:]]
    plans += [g2.plan_build(g2.query_random(n, 0.5, seed)) for n in (4, 6, 8) for seed in range(8)]
    for p in plans:
        dev.plan_local_eligible(p)
    shared = [(p, i) for p in plans for i in range(2, p.n) if p.shared[i]]
    assert shared

    def refused(p):
        with pytest.raises(RuntimeError, match="invalid plan"):
            dev.plan_local_eligible(p)

    def fresh():
        return g2.plan_for_preset(24)

    p = fresh()
    p.seq[1], p.seq[2] = p.seq[2], p.seq[1]          # rev no longer its inverse
    refused(p)
    p = fresh()
    p.seq[3] = p.n                                   # not a query vertex
    refused(p)
    p = fresh()
    p.cond_count[p.seq[2]] = 9
    refused(p)
    p = fresh()
    u = p.seq[2]
    p.cond[u][0][1] = p.n
    refused(p)
    base, i = shared[0]
    for field in ("pre_bn_count", "after_bn_count"):
        p = plan_copy(base)
        getattr(p, field)[i] = 9
        refused(p)
    p = plan_copy(base)
    p.pre_bn[i][0] = p.n
    refused(p)
    p = plan_copy(base)
    p.after_cond_count[p.seq[i]] = 9
    refused(p)
