"""GPU subgraph matching (libg2hip.so g2_gm_run) against the oracle: u64
counts and per-root counts, bit-exact."""
import json
import math
import os

import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PRESETS = list(range(0, 19)) + [23, 24, 25, 26, 27]


def both(g2, orc, ex, g, qadj, sources=None, account=False):
    p = g2.plan_build(qadj)
    ex.transfer(g, sources)
    ex.set_plan(p)
    got = ex.run_gm(per_root=True)
    ref = orc.gm(g, orc.OrcPlan.from_buffer_copy(bytes(p)), sources=sources, per_root=True)
    assert got["count"] == ref["count"]
    assert np.array_equal(got["per_root_count"], ref["per_root_count"])
    if account:  # the accounting variant counts the same and models the same bytes
        acc = ex.run_gm(per_root=True, account=True)
        assert acc["count"] == ref["count"]
        assert acc["stats"]["alg_bytes"] == ref["bytes"]
    return got


@pytest.fixture(scope="module")
def rmat12(g2):
    return g2.gen_rmat(12, 16, 0.57, 0.19, 0.19, 0x6732A1D0C0000002).preprocess()


@pytest.fixture(scope="module")
def rmat9(g2):
    return g2.gen_rmat(9, 8, 0.57, 0.19, 0.19, 0x6732A1D0C0000002).preprocess()


@pytest.mark.parametrize("q", PRESETS)
def test_presets_rmat(g2, orc, gpu_exec, rmat9, q):
    print("case: preset", q, flush=True)
    _, qadj = g2.query_preset(q)
    both(g2, orc, gpu_exec, rmat9, qadj, account=True)


def test_random_patterns(g2, orc, gpu_exec):
    print("case: random patterns", flush=True)
    with open(os.path.join(HERE, "golden", "plans.json")) as f:
        fixtures = json.load(f)["random"]
    g = G.gnp(g2, 300, 0.05, 8)
    for item in fixtures:
        both(g2, orc, gpu_exec, g, g2.query_remap(item["qadj"]))


def test_clique_closed_forms(g2, orc, gpu_exec):
    print("case: clique closed forms", flush=True)
    g = G.complete(g2, 30)
    for q, k in ((0, 3), (24, 4), (25, 5), (26, 6)):
        _, qadj = g2.query_preset(q)
        assert both(g2, orc, gpu_exec, g, qadj)["count"] == math.comb(30, k)


def test_triangles_4cliques_rmat15(g2, orc, gpu_exec):
    """Canonical traversal (per-root counts and byte accounting) and the
    local-bitmap path (whole-graph totals) on the same graph."""
    print("case: rmat-15 q0 q24", flush=True)
    g = g2.gen_rmat(15, 16, 0.57, 0.19, 0.19, 0x6732A1D0C0000002).preprocess()
    for q in (0, 24):
        _, qadj = g2.query_preset(q)
        got = both(g2, orc, gpu_exec, g, qadj, account=True)
        fast = gpu_exec.run_gm()
        assert fast["stats"]["local_bitmap"] == 1
        assert fast["count"] == got["count"]


def test_windows_and_views(g2, orc, rmat12):
    """Level-1 windows split hub roots across waves; hop-2 views of 4 bins add
    up to the whole graph."""
    print("case: windows and views", flush=True)
    _, qadj = g2.query_preset(24)
    whole = orc.gm(rmat12, orc.OrcPlan.from_buffer_copy(bytes(g2.plan_build(qadj))))["count"]
    with g2.Executor(device=0, gm_chunk=16) as ex:
        both(g2, orc, ex, rmat12, qadj, account=True)
        tot = 0
        bins = np.arange(rmat12.nv) % 4
        for b in range(4):
            src = np.nonzero(bins == b)[0].astype(np.uint32)
            view = g2.view_materialize(rmat12, src, 2, 3)
            tot += both(g2, orc, ex, view, qadj, sources=src)["count"]
        assert tot == whole


CLIQUES = [0, 23, 24, 25, 26, 27]


def test_local_path_cliques(g2, orc, gpu_exec):
    """Whole-graph clique totals through the oriented local-bitmap kernel
    equal the oracle's (reference id order) and the canonical kernel's."""
    print("case: local path cliques", flush=True)
    # sizes keep every case under ~20M matches (R-MAT 13 already holds 3.4e9 7-cliques)
    graphs = [g2.gen_rmat(10, 16, 0.57, 0.19, 0.19, 0x6732A1D0C0000002).preprocess(),
              G.gnp(g2, 400, 0.1, 3), G.complete(g2, 40),
              g2.gen_chung_lu(20000, 2.3, 2.0, 800.0, 120000, 9).preprocess()]
    for g in graphs:
        gpu_exec.transfer(g)
        for q in CLIQUES:
            p = g2.plan_for_preset(q)
            gpu_exec.set_plan(p)
            fast = gpu_exec.run_gm()
            canon = gpu_exec.run_gm(canonical=True)
            ref = orc.gm(g, orc.OrcPlan.from_buffer_copy(bytes(p)))["count"]
            assert fast["count"] == ref == canon["count"], (q, fast["count"], ref, canon["count"])


def test_local_path_large_root_pass(g2, orc):
    """Roots whose local rows exceed a wave's LDS (|U| > 149 here) are
    deferred to the workgroup pass (rows in the workgroup's LDS); 3-, 4- and
    5-cliques, the last walking a per-lane level below the shared r2."""
    print("case: local path workgroup pass", flush=True)
    g = G.gnp(g2, 400, 0.5, 7)
    with g2.Executor(device=0) as ex:
        ex.transfer(g)
        for q in (0, 24, 25):
            p = g2.plan_for_preset(q)
            ex.set_plan(p)
            got = ex.run_gm()
            assert got["stats"]["big_roots"] > 0 and got["stats"]["launches"] == 2
            assert got["count"] == orc.gm(g, orc.OrcPlan.from_buffer_copy(bytes(p)))["count"]


def test_local_path_complete_graph(g2, gpu_exec):
    """K_170: every root's U is the rest of the order (|U| up to 169, so the
    wave and 256-thread workgroup passes both run), and the k-clique totals
    are the binomials C(170, k), up to 6-cliques."""
    print("case: local path complete graph", flush=True)
    gpu_exec.transfer(G.complete(g2, 170))
    for q, k in ((0, 3), (24, 4), (25, 5), (26, 6)):
        gpu_exec.set_plan(g2.plan_for_preset(q))
        assert gpu_exec.run_gm()["count"] == math.comb(170, k)


def test_local_path_global_arena_pass(g2, gpu_exec):
    """Turan graph T(1200, 40): degree 1170 everywhere, so in (degree, id)
    order the roots' |U| runs from 1170 down to 0 and every pass gets roots:
    wave LDS, the 256-, 512- and 1024-thread workgroups, and global arenas
    beyond |U| = 1039.  k-cliques pick k of the 40 parts and one of 30 vertices in
    each: C(40, k) * 30^k."""
    print("case: local path global-arena pass", flush=True)
    gpu_exec.transfer(G.turan(g2, [30] * 40))
    for q, k in ((0, 3), (24, 4)):
        gpu_exec.set_plan(g2.plan_for_preset(q))
        got = gpu_exec.run_gm()
        assert got["stats"]["launches"] == 5
        assert got["count"] == math.comb(40, k) * 30 ** k
