"""GPU maximal clique enumeration (libg2hip.so g2_mce_run) against the oracle:
totals, size histograms and per-root search-tree node counts, bit-exact."""
import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu

CFG1_SEED = 0x6732A1D0C0000001


def check(ex, orc, g, sources=None, **kw):
    ex.transfer(g, sources)
    got = ex.run_mce(per_root=True, **kw)
    ref = orc.mce(g, sources=sources, per_root=True)
    assert got["total"] == ref["total"]
    assert np.array_equal(got["hist"], ref["hist"])
    assert np.array_equal(got["per_root_nodes"], ref["per_root_nodes"])
    assert got["stats"]["nodes"] == ref["nodes"]
    return got, ref


def test_small_graphs(g2, orc, gpu_exec):
    print("case: small graphs", flush=True)
    # every case stays below ~2M search-tree nodes: one wave walks one root
    for seed, (n, p) in enumerate([(8, 0.5), (30, 0.3), (60, 0.5), (150, 0.1), (60, 0.8)]):
        check(gpu_exec, orc, G.gnp(g2, n, p, seed))
    for g in (G.complete(g2, 20), G.complete_bipartite(g2, 9, 13), G.cycle(g2, 40),
              G.moon_moser(g2, 7), G.turan(g2, [5, 1, 7, 3])):
        check(gpu_exec, orc, g)


def test_wide_neighbourhoods(g2, orc, gpu_exec):
    # p0 > 64 exercises multi-word P sets; dense graphs give deep trees
    print("case: wide neighbourhoods", flush=True)
    check(gpu_exec, orc, G.gnp(g2, 200, 0.5, 11))  # p0 up to ~100: two P words
    check(gpu_exec, orc, G.moon_moser(g2, 9))


def test_empty_and_isolated(g2, orc, gpu_exec):
    print("case: empty and isolated", flush=True)
    g = g2.Graph(np.zeros(6, np.uint64), np.zeros(0, np.uint32))
    got, _ = check(gpu_exec, orc, g)
    assert got["total"] == 0
    check(gpu_exec, orc, G.gnp(g2, 40, 0.1, 3), sources=np.array([], np.uint32))


@pytest.fixture(scope="module")
def cfg1(g2):
    g = g2.gen_chung_lu(100000, 2.3, 2.0, 2000.0, 1000000, CFG1_SEED).preprocess()
    g, order, shells, dg = g.kcore_recode()
    return g, shells


def test_chung_lu_100k(g2, orc, gpu_exec, cfg1):
    print("case: chung-lu 100k", flush=True)
    g, _ = cfg1
    got, ref = check(gpu_exec, orc, g, account=True)
    assert got["stats"]["alg_bytes"] == ref["bytes"]
    check(gpu_exec, orc, g)  # register-resident path for roots of degree <= 64


@pytest.mark.parametrize("small_max", [1, 8, 33, 64])
def test_register_path_bounds(g2, orc, small_max):
    """Roots of degree <= small_max run in registers, the rest in the arena:
    every split point gives the oracle's results (degrees up to ~100 here)."""
    print("case: register path bound", small_max, flush=True)
    with g2.Executor(device=0, mce_small_max=small_max) as ex:
        for seed, (n, p) in enumerate([(120, 0.52), (300, 0.08)]):
            g, _, _, _ = G.gnp(g2, n, p, 40 + seed).kcore_recode()
            check(ex, orc, g)
        check(ex, orc, G.complete(g2, 66))  # d = 65: the arena path; a 66-clique
        check(ex, orc, G.complete(g2, 64))  # d = 63 everywhere


def test_chung_lu_views(g2, orc, gpu_exec, cfg1):
    """4 kcore round-robin bins as hop-2 views: per-view results equal the
    oracle on the same view and add up to the whole graph's."""
    print("case: chung-lu views", flush=True)
    g, shells = cfg1
    bins = g2.kcore_bins(shells, 4, g.nv)
    total = 0
    hist = np.zeros(256, np.uint64)
    for b in range(4):
        src = np.nonzero(bins == b)[0].astype(np.uint32)
        view = g2.view_materialize(g, src, 2)
        got, _ = check(gpu_exec, orc, view, sources=src)
        total += got["total"]
        hist += got["hist"]
    whole = orc.mce(g)
    assert total == whole["total"] and np.array_equal(hist, whole["hist"])


def test_large_root_pass_and_schedule(g2, orc):
    """Tiny per-wave arenas force every root through the large-root pass; a
    reversed schedule changes the order, never the results."""
    print("case: large-root pass", flush=True)
    g = G.gnp(g2, 400, 0.08, 21)
    ref = orc.mce(g, per_root=True)
    # mce_small_max=1 keeps these degree-~32 roots off the register path
    with g2.Executor(device=0, wave_scratch_words=64, big_waves=64, mce_small_max=1) as ex:
        ex.transfer(g)
        got = ex.run_mce(per_root=True)
        assert got["stats"]["big_roots"] > 0
        assert got["total"] == ref["total"] and np.array_equal(got["hist"], ref["hist"])
        assert np.array_equal(got["per_root_nodes"], ref["per_root_nodes"])
        got = ex.run_mce(per_root=True, order=np.arange(g.nv, dtype=np.uint32)[::-1].copy())
        assert np.array_equal(got["per_root_nodes"], ref["per_root_nodes"])


def test_split_roots(g2, orc, gpu_exec, cfg1):
    """Dealing each root's first-level children over several waves (the
    heavy-root split of the learned scheduler) never changes a result."""
    print("case: split roots", flush=True)
    g, _ = cfg1
    ref = orc.mce(g, per_root=True, hist_len=256)
    gpu_exec.transfer(g)
    rng = np.random.default_rng(5)
    split = rng.integers(1, 9, size=g.nv).astype(np.uint32)
    order = rng.permutation(g.nv).astype(np.uint32)
    for kw in ({"split": split}, {"split": split, "order": order}, {"split": np.full(g.nv, 64, np.uint32)}):
        got = gpu_exec.run_mce(per_root=True, **kw)
        assert got["total"] == ref["total"]
        assert np.array_equal(got["hist"], ref["hist"])
        assert np.array_equal(got["per_root_nodes"], ref["per_root_nodes"])
    acc = gpu_exec.run_mce(account=True, split=split)
    assert acc["stats"]["alg_bytes"] == ref["bytes"]
    dense = G.gnp(g2, 200, 0.5, 11)
    ref = orc.mce(dense, per_root=True)
    gpu_exec.transfer(dense)
    got = gpu_exec.run_mce(per_root=True, split=np.full(dense.nv, 7, np.uint32))
    assert got["total"] == ref["total"]
    assert np.array_equal(got["per_root_nodes"], ref["per_root_nodes"])


def test_rmat_mce(g2, orc, gpu_exec):
    """R-MAT hubs: neighbourhoods of thousands of vertices (large-root pass)."""
    print("case: rmat-12 mce", flush=True)
    g = g2.gen_rmat(12, 16, 0.57, 0.19, 0.19, 0x6732A1D0C0000005).preprocess()
    g, _, _, _ = g.kcore_recode()
    check(gpu_exec, orc, g, account=True)


@pytest.mark.parametrize("every", [1, 7, 0xFFFFFFFF])
def test_work_donation(g2, orc, every):
    """Busy waves hand open frames to idle waves (every `every` nodes they look
    for idle waves; 0xFFFFFFFF = never): the tree, every count and per-root
    node total are unchanged.  R-MAT roots are skewed, so idle waves appear."""
    print("case: work donation every", every, flush=True)
    g = g2.gen_rmat(12, 16, 0.57, 0.19, 0.19, 0x6732A1D0C0000005).preprocess()
    g, _, _, _ = g.kcore_recode()
    ref = orc.mce(g, per_root=True)
    for small in (64, 1):  # register path and arena path
        with g2.Executor(device=0, mce_donate_every=every, mce_small_max=small) as ex:
            ex.transfer(g)
            got = ex.run_mce(per_root=True)
            assert got["total"] == ref["total"]
            assert np.array_equal(got["hist"], ref["hist"])
            assert np.array_equal(got["per_root_nodes"], ref["per_root_nodes"])
            if every == 0xFFFFFFFF:
                assert got["stats"]["donations"] == 0
            else:
                assert got["stats"]["donations"] > 0
            # split roots and donation together
            split = np.full(g.nv, 3, np.uint32)
            got = ex.run_mce(per_root=True, split=split)
            assert np.array_equal(got["per_root_nodes"], ref["per_root_nodes"])
            assert np.array_equal(got["hist"], ref["hist"])


@pytest.mark.parametrize("rpw", [2, 4])
def test_roots_per_wave(g2, orc, cfg1, rpw):
    """Batched small roots (lane groups of 32 or 16, roots of degree <= 32 or
    <= 16) give the oracle's results, alone and with order + split schedules."""
    print("case: roots per wave", rpw, flush=True)
    g, _ = cfg1
    ref = orc.mce(g, per_root=True)
    with g2.Executor(device=0, mce_roots_per_wave=rpw) as ex:
        for h in (g, G.gnp(g2, 90, 0.25, 3), G.complete(g2, 17), G.complete(g2, 33), G.cycle(g2, 20)):
            check(ex, orc, h)
        ex.transfer(g)
        rng = np.random.default_rng(rpw)
        order = rng.permutation(g.nv).astype(np.uint32)
        split = rng.integers(1, 4, size=g.nv).astype(np.uint32)
        got = ex.run_mce(per_root=True, order=order, split=split)
        assert got["total"] == ref["total"] and np.array_equal(got["hist"], ref["hist"])
        assert np.array_equal(got["per_root_nodes"], ref["per_root_nodes"])
        assert got["stats"]["launches"] >= 2



def test_malformed_view_refused(g2, orc, gpu_exec):
    """A column id >= nv or a decreasing row_ptrs is refused on the host,
    before anything on the device changes: the view uploaded before stays
    in place and still gives the oracle's counts."""
    print("case: malformed view refused", flush=True)
    g = G.gnp(g2, 50, 0.2, 4)
    gpu_exec.transfer(g)
    ref = orc.mce(g, per_root=True)
    cols = g.cols.copy()
    cols[-1] = g.nv
    with pytest.raises(RuntimeError, match="column id out of range"):
        gpu_exec.transfer(g2.Graph(g.row_ptrs, cols))
    rp = g.row_ptrs.copy()
    rp[10] = rp[11] + 1
    assert rp[10] <= g.ne
    with pytest.raises(RuntimeError, match="non-decreasing"):
        gpu_exec.transfer(g2.Graph(rp, g.cols))
    got = gpu_exec.run_mce(per_root=True)
    assert got["total"] == ref["total"] and np.array_equal(got["hist"], ref["hist"])
    assert np.array_equal(got["per_root_nodes"], ref["per_root_nodes"])


def test_schedule_checks_and_ranges(g2, orc, gpu_exec):
    """The device-built task list: a repeated or out-of-range source in the
    order is rejected; schedule ranges cover exactly their positions."""
    print("case: schedule checks and ranges", flush=True)
    g = G.gnp(g2, 120, 0.2, 9)
    gpu_exec.transfer(g)
    bad = np.arange(g.nv, dtype=np.uint32)
    bad[5] = bad[6]
    with pytest.raises(RuntimeError, match="permutation"):
        gpu_exec.run_mce(order=bad)
    bad[5] = g.nv
    with pytest.raises(RuntimeError, match="permutation"):
        gpu_exec.run_mce(order=bad)
    ref = orc.mce(g, per_root=True)
    rng = np.random.default_rng(2)
    order = rng.permutation(g.nv).astype(np.uint32)
    split = rng.integers(1, 5, size=g.nv).astype(np.uint32)
    tot, hist, nodes = 0, np.zeros(256, np.uint64), np.zeros(g.nv, np.uint64)
    for b, e in ((0, 37), (37, 38), (38, g.nv)):
        r = gpu_exec.run_mce(per_root=True, order=order, split=split, begin=b, end=e)
        tot += r["total"]
        hist += r["hist"]
        nodes += r["per_root_nodes"]
    assert tot == ref["total"] and np.array_equal(hist, ref["hist"])
    assert np.array_equal(nodes, ref["per_root_nodes"])


def test_host_spill_tiny_budget(g2, orc, cfg1):
    """A budget that cannot hold one large-root arena: the large-root pass
    runs with its arenas in page-locked host memory (app_BK/BK.h:357-398's
    spill, rebuilt), and every count equals the oracle."""
    print("case: host spill", flush=True)
    g, _ = cfg1
    tune = dict(wave_scratch_words=256, big_waves=64)  # most roots above degree ~20 are deferred
    with g2.Executor(device=0, **tune) as ex:
        ex.transfer(g)
        free = ex.run_mce(per_root=True)  # the same buffers as the run below
    st = free["stats"]
    need, nbig = st["max_scratch_words"], st["big_roots"]
    assert nbig > 0 and st["host_spill_bytes"] == 0
    bwaves = min(64, nbig)
    bw = min(4, bwaves)
    big = -(-(bwaves // bw * bw) * need * 8 // 256) * 256
    budget = st["device_bytes"] - big + need * 4  # everything but half of one arena
    with g2.Executor(device=0, mem_budget=budget, **tune) as ex:
        ex.transfer(g)
        got = ex.run_mce(per_root=True)
    ref = orc.mce(g, per_root=True)
    print(f"budget {budget} B, host spill {got['stats']['host_spill_bytes']} B, {nbig} large roots", flush=True)
    assert got["stats"]["host_spill_bytes"] > 0
    assert got["stats"]["device_bytes"] <= budget
    assert got["total"] == ref["total"] == free["total"]
    assert np.array_equal(got["hist"], ref["hist"])
    assert np.array_equal(got["per_root_nodes"], ref["per_root_nodes"])
