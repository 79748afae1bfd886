"""Batched matching (g2_gm_run_batch, g2_gm_autotune): every query's count in
one persistent launch equals its solo canonical run and the oracle; launch
shape, query order and node budget never change an admitted count."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SEED = 0x6732A1D0C0000004


@pytest.fixture(scope="module")
def B():
    return importlib.import_module("g2-aimd_amd.batch")


@pytest.fixture(scope="module")
def rmat9(g2):
    return g2.gen_rmat(9, 8, 0.57, 0.19, 0.19, 0x6732A1D0C0000002).preprocess()


def solo(ex, plans):
    out = []
    for p in plans:
        ex.set_plan(p)
        out.append(ex.run_gm(canonical=True)["count"])
    return np.array(out, np.uint64)


def test_batch_equals_solo_and_oracle(g2, orc, gpu_exec, B, rmat9):
    print("case: batch of 48 random patterns (3-5 vertices) on rmat-9", flush=True)
    qs = B.random_queries(g2, 48, SEED, nmax=5)
    plans = [g2.plan_build(q) for q in qs]
    gpu_exec.transfer(rmat9)
    r = gpu_exec.run_gm_batch(plans)
    assert not r["over_budget"].any()
    ref = np.array([orc.gm(rmat9, orc.OrcPlan.from_buffer_copy(bytes(p)))["count"] for p in plans], np.uint64)
    assert np.array_equal(r["counts"], ref)
    assert np.array_equal(solo(gpu_exec, plans), ref)
    assert r["stats"]["nodes"] == int(r["nodes"].sum())


@pytest.fixture(scope="module")
def gnp80(g2):
    import graphs as G
    return G.gnp(g2, 80, 0.12, 5)


@pytest.fixture(scope="module")
def solo_exec(g2):
    """solo runs with level-1 windows of 16, so hub roots spread over waves"""
    ex = g2.Executor(device=0, gm_chunk=16)
    yield ex
    ex.close()


@pytest.mark.parametrize("shape", [dict(window=0, roots_per_task=1), dict(window=8, roots_per_task=1),
                                   dict(window=32, roots_per_task=8), dict(window=0, roots_per_task=64,
                                                                           waves_per_cu=2)])
def test_batch_shapes_and_order(g2, orc, gpu_exec, B, gnp80, shape):
    """5-7 vertex patterns: every launch shape and query order gives the oracle's counts"""
    print("case: batch shape", shape, flush=True)
    qs = B.random_queries(g2, 64, SEED + 1, nmin=5, nmax=7)
    plans = [g2.plan_build(q) for q in qs]
    gpu_exec.transfer(gnp80)
    r = gpu_exec.run_gm_batch(plans, order=B.heavy_first(qs), **shape)
    ref = np.array([orc.gm(gnp80, orc.OrcPlan.from_buffer_copy(bytes(p)))["count"] for p in plans], np.uint64)
    assert np.array_equal(r["counts"], ref)


def test_batch_node_budget(g2, solo_exec, B, rmat9):
    """Queries over the node budget are flagged (their charged nodes exceed it);
    every unflagged count equals the solo run, and a larger budget keeps them."""
    print("case: batch node budget", flush=True)
    qs = B.random_queries(g2, 64, SEED + 2, nmin=4, nmax=7)
    plans = [g2.plan_build(q) for q in qs]
    solo_exec.transfer(rmat9)
    limit = 200_000
    r1 = solo_exec.run_gm_batch(plans, node_budget=limit, window=16)
    over = r1["over_budget"]
    assert over.any() and not over.all()
    assert (r1["nodes"][over] > limit).all() and (r1["nodes"][~over] <= limit).all()
    keep = [p for p, o in zip(plans, over) if not o]
    assert np.array_equal(r1["counts"][~over], solo(solo_exec, keep))
    r2 = solo_exec.run_gm_batch(plans, node_budget=4 * limit, window=32, roots_per_task=4)
    assert not (r2["over_budget"] & ~over).any()
    assert np.array_equal(r2["counts"][~over], r1["counts"][~over])


def test_batch_subset_sources_and_empty(g2, orc, gpu_exec, B, rmat9):
    print("case: batch over a source subset; empty batch", flush=True)
    qs = B.random_queries(g2, 16, SEED + 3, nmax=5)
    plans = [g2.plan_build(q) for q in qs]
    src = np.arange(0, rmat9.nv, 3, dtype=np.uint32)
    gpu_exec.transfer(rmat9, src)
    r = gpu_exec.run_gm_batch(plans, window=4)
    ref = [orc.gm(rmat9, orc.OrcPlan.from_buffer_copy(bytes(p)), sources=src)["count"] for p in plans]
    assert r["counts"].tolist() == ref
    e = gpu_exec.run_gm_batch([])
    assert len(e["counts"]) == 0


def test_autotune(g2, orc, gpu_exec, B, gnp80):
    print("case: autotune", flush=True)
    qs = B.random_queries(g2, 40, SEED + 4, nmax=7)
    plans = [g2.plan_build(q) for q in qs]
    gpu_exec.transfer(gnp80)
    t = gpu_exec.autotune(plans, sample_queries=16, seed=3, node_budget=1 << 40)
    assert t["window"] in (0, 64, 256, 1024, 4096)
    assert t["roots_per_task"] in (1, 2, 8, 32)
    assert t["waves_per_cu"] in (8, 16, 24, 32)
    assert t["sample_ms"] > 0
    shape = {k: t[k] for k in ("window", "roots_per_task", "waves_per_cu")}
    r = gpu_exec.run_gm_batch(plans, **shape)
    ref = np.array([orc.gm(gnp80, orc.OrcPlan.from_buffer_copy(bytes(p)))["count"] for p in plans], np.uint64)
    assert np.array_equal(r["counts"], ref)
