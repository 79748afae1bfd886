"""GPU local-bitmap matching (the kernels the matching headline times:
gm_local_kernel and gm_local_block_kernel), called through g2_gm_run without
per-root counts, against the oracle:

* the R-MAT 17 and R-MAT 22 totals of triangles and 4-cliques, from the
  oracle run once at full size (tests/golden/headline_counts.json, written by
  tests/golden/make_headline_fixtures.py);
* a graph built so that every pass (rows in a wave's LDS; in a 256-, 512-
  and 1024-thread workgroup's LDS; in per-wave global arenas) takes roots;
* view bins (view/view_bin.h:61-159): each hop-2 view counts the cliques
  whose lowest (degree, id) vertex is one of its sources, and the bins add
  up to the whole-graph total.
"""
import json
import os

import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
RMAT_SEED = 0x6732A1D0C0000002


def fixtures():
    with open(os.path.join(HERE, "golden", "headline_counts.json")) as f:
        return json.load(f)


def fingerprint(g):
    w = np.arange(1, g.ne + 1, dtype=np.uint64)
    return {"nv": int(g.nv), "ne": int(g.ne),
            "cols_checksum": int(np.sum(g.cols.astype(np.uint64) * w, dtype=np.uint64))}


def check_rmat(g2, scale, ex):
    fx = fixtures()[f"rmat-{scale}"]
    g = g2.gen_rmat(scale, 16, 0.57, 0.19, 0.19, RMAT_SEED).preprocess()
    fp = fingerprint(g)
    assert fp == {k: fx[k] for k in fp}, "the generator no longer makes the fixture's graph"
    ex.transfer(g)
    for q in (0, 24):
        ex.set_plan(g2.plan_for_preset(q))
        r = ex.run_gm()
        print(f"rmat-{scale} q={q}: {r['count']} (oracle {fx[f'q{q}']}), passes {r['stats']['local_pass_roots']}",
              flush=True)
        assert r["stats"]["local_bitmap"] == 1
        assert r["count"] == fx[f"q{q}"]


def test_local_rmat17_oracle_totals(g2, gpu_exec):
    print("case: local path, rmat-17 q0 q24 vs oracle fixture", flush=True)
    check_rmat(g2, 17, gpu_exec)


def test_local_rmat22_headline_totals(g2):
    """The benchmarked workload at full size: R-MAT 22 triangles and 4-cliques."""
    print("case: local path, rmat-22 q0 q24 vs oracle fixture", flush=True)
    with g2.Executor(device=0) as ex:
        check_rmat(g2, 22, ex)


def gadget_graph(g2, seed):
    """R-MAT 12 plus four gadgets, ids shuffled.  Gadget (a, p): an a-vertex
    side A holding G(a, p), and a + 40 vertices each adjacent to all of A.
    Those have degree a, below A's (a + 40 + deg in G(a,p)), so each is the
    lowest-ranked vertex of U = A, |U| = a: a = 300, 600, 900, 1200 put roots
    into the 256-, 512-, 1024-thread workgroup passes and the global-arena
    pass; R-MAT 12's roots fill the first pass.  4-cliques: (a + 40) x
    triangles(G(a,p)) + 4-cliques(G(a,p)) per gadget."""
    rng = np.random.default_rng(seed)
    base = g2.gen_rmat(12, 16, 0.57, 0.19, 0.19, RMAT_SEED).preprocess()
    rp, cols = base.row_ptrs, base.cols
    src = [np.repeat(np.arange(base.nv, dtype=np.int64), np.diff(rp).astype(np.int64))]
    dst = [cols.astype(np.int64)]
    off = base.nv
    for a, p in ((300, 0.05), (600, 0.03), (900, 0.02), (1200, 0.015)):
        iu, ju = np.triu_indices(a, 1)
        keep = rng.random(len(iu)) < p
        src.append(off + iu[keep])
        dst.append(off + ju[keep])
        nb = a + 40
        hub = np.repeat(np.arange(nb), a) + off + a
        side = np.tile(np.arange(a), nb) + off
        src.append(hub)
        dst.append(side)
        off += a + nb
    s = np.concatenate(src)
    d = np.concatenate(dst)
    perm = rng.permutation(off).astype(np.uint32)
    return g2.Graph.from_edges(perm[s], perm[d], off).preprocess()


def test_local_every_pass_takes_roots(g2, orc, gpu_exec):
    print("case: local path, all five passes", flush=True)
    g = gadget_graph(g2, 5)
    gpu_exec.transfer(g)
    for q in (0, 24):
        plan = g2.plan_for_preset(q)
        gpu_exec.set_plan(plan)
        r = gpu_exec.run_gm()
        ref = orc.gm(g, orc.OrcPlan.from_buffer_copy(bytes(plan)))["count"]
        print(f"q={q}: {r['count']} (oracle {ref}), pass roots {r['stats']['local_pass_roots']}", flush=True)
        assert r["stats"]["local_bitmap"] == 1
        assert r["count"] == ref
        if q == 24:
            assert all(n > 0 for n in r["stats"]["local_pass_roots"]), r["stats"]["local_pass_roots"]
            assert r["stats"]["launches"] == 5
        # the canonical traversal agrees
        assert gpu_exec.run_gm(canonical=True)["count"] == ref


def descending(g2, plan):
    """The same clique plan with its symmetry breaking reversed: position
    i >= 2 below position i-1 (LESS_THAN) instead of above it.  Still a
    complete symmetry breaking, so the count is the same, but the plan is no
    longer an ascending chain and takes the general walker (order ranges
    over symmetric rows) instead of the clique kernels (rows of later members
    only)."""
    d = type(plan).from_buffer_copy(bytes(plan))
    for i in range(2, d.n):
        u = d.seq[i]
        for t in range(d.cond_count[u]):
            if d.rev[d.cond[u][t][1]] != 0:
                d.cond[u][t][0] = 0
    return d


def test_clique_kernels_equal_general_walker(g2, orc, gpu_exec):
    """Clique kernels and the general walker agree on totals and on partial
    matches (nodes) in every pass, for 3- to 6-cliques."""
    print("case: clique kernels vs general walker", flush=True)
    g = gadget_graph(g2, 11)
    gpu_exec.transfer(g)
    for q in (0, 24, 25, 26):
        plan = g2.plan_for_preset(q)
        d = descending(g2, plan)
        assert g2.plan_local_eligible(d)
        gpu_exec.set_plan(plan)
        a = gpu_exec.run_gm()
        gpu_exec.set_plan(d)
        b = gpu_exec.run_gm()
        print(f"q={q}: clique {a['count']} nodes {a['stats']['nodes']}, general {b['count']} "
              f"nodes {b['stats']['nodes']}, pass roots {a['stats']['local_pass_roots']}", flush=True)
        assert a["stats"]["local_bitmap"] == b["stats"]["local_bitmap"] == 1
        assert a["count"] == b["count"]
        assert a["stats"]["nodes"] == b["stats"]["nodes"]
        assert a["stats"]["local_pass_roots"] == b["stats"]["local_pass_roots"]
        if q == 24:
            assert a["count"] == orc.gm(g, orc.OrcPlan.from_buffer_copy(bytes(plan)))["count"]


@pytest.mark.parametrize("q,k,root_degree", [(0, 3, False), (24, 4, False), (25, 5, False), (24, 4, True),
                                             (25, 5, True)])
def test_local_view_bins(g2, orc, gpu_exec, q, k, root_degree):
    """Four bins of R-MAT 14 as hop-2 views, each on the local path; with
    root_degree, sources below the plan's root degree are not extended (as
    the run command line materialises them), so some of them sit next to
    rowless vertices."""
    print(f"case: local path over view bins, q={q}, root degree {root_degree}", flush=True)
    g = g2.gen_rmat(14, 16, 0.57, 0.19, 0.19, RMAT_SEED).preprocess()
    plan = g2.plan_for_preset(q)
    whole = orc.gm(g, orc.OrcPlan.from_buffer_copy(bytes(plan)))["count"]
    bins = np.random.default_rng(3).integers(0, 4, g.nv).astype(np.uint32)
    total = 0
    for b in range(4):
        srcs = np.nonzero(bins == b)[0].astype(np.uint32)
        view = g2.view_materialize(g, srcs, plan.hop, plan.root_degree if root_degree else 0)
        gpu_exec.transfer(view, srcs)
        gpu_exec.set_plan(plan)
        r = gpu_exec.run_gm()
        assert r["stats"]["local_bitmap"] == 1
        ref = orc.kclique_count_ranked(view, k, sources=srcs)
        assert r["count"] == ref
        # the same sources over the whole graph count the same cliques
        assert ref == orc.kclique_count_ranked(g, k, sources=srcs)
        total += r["count"]
    assert total == whole


def test_hop1_view_takes_canonical_path(g2, orc, gpu_exec):
    """A view whose sources' neighbours lack their rows (hop 1) is not
    counted by the local path; the canonical traversal gives the oracle's
    count of that view."""
    print("case: hop-1 view, canonical path", flush=True)
    g = g2.gen_rmat(11, 16, 0.57, 0.19, 0.19, RMAT_SEED).preprocess()
    srcs = np.arange(0, g.nv, 3, dtype=np.uint32)
    view = g2.view_materialize(g, srcs, 1)
    plan = g2.plan_for_preset(24)
    gpu_exec.transfer(view, srcs)
    gpu_exec.set_plan(plan)
    r = gpu_exec.run_gm()
    assert r["stats"]["local_bitmap"] == 0
    assert r["count"] == orc.gm(view, orc.OrcPlan.from_buffer_copy(bytes(plan)), sources=srcs)["count"]


def test_upload_over_budget_keeps_state_consistent(g2):
    """A view refused for the budget (checked for every view buffer before any
    is replaced) leaves the previous view in place and usable."""
    print("case: upload budget", flush=True)
    small = G.complete(g2, 12)
    big = g2.gen_rmat(14, 16, 0.57, 0.19, 0.19, RMAT_SEED).preprocess()  # 2 MB of cols
    with g2.Executor(device=0, mem_budget=1 << 20) as ex:
        ex.transfer(small)
        ex.set_plan(g2.plan_for_preset(24))
        assert ex.run_gm()["count"] == 495
        with pytest.raises(RuntimeError, match=r"\(-3\)"):
            ex.transfer(big)
        assert ex.run_gm()["count"] == 495


@pytest.mark.parametrize("q,n", [(0, 3), (24, 4), (25, 5)])
def test_local_byte_model_equals_oracle(g2, orc, gpu_exec, q, n):
    """The accounting kernel's per-pass algorithmic bytes equal the oracle's
    restatement of the byte model, on R-MAT 12 and on the graph that fills
    every pass."""
    print(f"case: local byte model q={q}", flush=True)
    for g in (g2.gen_rmat(12, 16, 0.57, 0.19, 0.19, RMAT_SEED).preprocess(), gadget_graph(g2, 5)):
        gpu_exec.transfer(g)
        gpu_exec.set_plan(g2.plan_for_preset(q))
        r = gpu_exec.run_gm(local_bytes=True)
        ref = orc.local_bytes(g, n)
        assert r["stats"]["local_bitmap"] == 1
        assert r["stats"]["local_pass_bytes"] == ref.tolist()
        assert r["stats"]["alg_bytes"] == int(ref.sum())
