"""Root features and the MFMA cost-model MLP against CPU references."""
import importlib

import numpy as np
import pytest

import graphs as G
import oracle
from oracle import mlp_ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    return importlib.import_module("g2-aimd_amd.device")


def test_features(g2, gpu_exec):
    print("case: features", flush=True)
    for g in (G.gnp(g2, 300, 0.05, 1), g2.gen_rmat(10, 16, 0.57, 0.19, 0.19, 3).preprocess()):
        gpu_exec.transfer(g)
        f = gpu_exec.features()
        n = g.nv
        A = np.zeros((n, n), np.int64)
        src = np.repeat(np.arange(n), np.diff(g.row_ptrs).astype(np.int64))
        A[src, g.cols] = 1
        deg = A.sum(1)
        A2 = A @ A
        tri = (A2 * A).sum(1) // 2
        two = oracle.two_hop(g)
        shared = ((A2 * A) > 0).sum(1)
        p0 = np.array([np.sum(g.cols[g.row_ptrs[v]:g.row_ptrs[v + 1]] > v) for v in range(n)])
        assert np.array_equal(f[:, 0], deg)
        assert np.array_equal(f[:, 1], two)
        assert np.array_equal(f[:, 2], tri)
        assert np.array_equal(f[:, 3], shared)
        assert np.array_equal(f[:, 4], p0)


def features_ref(g):
    """deg, 2-hop ball estimate (the oracle's HyperLogLog restatement),
    triangles, triangle-carrying neighbours, later neighbours of every vertex,
    the counts from scipy's sparse products."""
    import scipy.sparse as sp
    n = g.nv
    src = np.repeat(np.arange(n), np.diff(g.row_ptrs).astype(np.int64))
    A = sp.csr_matrix((np.ones(g.ne, np.int64), (src, g.cols.astype(np.int64))), shape=(n, n))
    deg = np.asarray(A.sum(1)).ravel()
    T = (A @ A).multiply(A)          # T[v, u] = common neighbours of the edge (v, u)
    tri = np.asarray(T.sum(1)).ravel() // 2
    shared = np.asarray((T > 0).sum(1)).ravel()
    two = oracle.two_hop(g)
    p0 = np.asarray((sp.triu(A, 1) > 0).sum(1)).ravel()
    return np.stack([deg, two, tri, shared, p0], axis=1)


def test_features_oriented_and_views(g2, gpu_exec):
    """Triangles through the oriented rows: R-MAT 13 whole graph; sources of
    hop-2 views (every neighbour has its row) get their whole-graph
    features; a hop-1 view (neighbours without rows) takes the per-neighbour
    count and still gives exact degrees and later-neighbour counts."""
    print("case: features, oriented rows and views", flush=True)
    g = g2.gen_rmat(13, 16, 0.57, 0.19, 0.19, 0x6732A1D0C0000004).preprocess()
    ref = features_ref(g)
    gpu_exec.transfer(g)
    assert np.array_equal(gpu_exec.features(), ref)
    bins = np.arange(g.nv) % 3
    for b in range(3):
        src = np.nonzero(bins == b)[0].astype(np.uint32)
        gpu_exec.transfer(g2.view_materialize(g, src, 2), sources=src)
        assert np.array_equal(gpu_exec.features(), ref[src])
    src = np.nonzero(bins == 0)[0].astype(np.uint32)
    gpu_exec.transfer(g2.view_materialize(g, src, 1), sources=src)
    f = gpu_exec.features()
    assert np.array_equal(f[:, [0, 4]], ref[src][:, [0, 4]])
    # a hop-1 view has no rows for its sources' neighbours: their closed rows
    # are themselves, so the ball is estimated over the closed neighbourhood
    v1 = g2.view_materialize(g, src, 1)
    assert np.array_equal(f[:, 1], oracle.two_hop(v1, src))


def test_mlp_forward_and_grad(dev):
    print("case: mlp forward/grad", flush=True)
    rng = np.random.default_rng(3)
    m = dev.Mlp(seed=7)
    p = m.params().astype(np.float64)
    for n in (16, 37, 1000):
        x = rng.normal(0, 1, (n, 8)).astype(np.float32)
        t = rng.normal(0, 1, n).astype(np.float32)
        y = m.infer(x)
        yr, _ = mlp_ref.forward(p, x, emulate=True)
        assert np.max(np.abs(y - yr)) <= 1e-3 * (1 + np.max(np.abs(yr)))
        loss, g = m.grad(x, t)
        lr, gr = mlp_ref.loss_grad(p, x, t, emulate=True)
        assert abs(loss - lr) <= 1e-3 * (1 + lr)
        # bf16 staging of the backward activations: relative error <= 1e-2
        assert np.linalg.norm(g - gr) <= 1e-2 * np.linalg.norm(gr), (n, np.linalg.norm(g - gr) / np.linalg.norm(gr))
    m.close()


def test_mlp_learns(dev):
    print("case: mlp training", flush=True)
    rng = np.random.default_rng(4)
    x = rng.normal(0, 1, (20000, 8)).astype(np.float32)
    t = (np.log1p(np.abs(x[:, 0] * 3)) + 0.5 * np.maximum(x[:, 1], 0) - 0.3 * x[:, 2] * x[:, 3]).astype(np.float32)
    m = dev.Mlp(seed=11)
    before = float(np.mean((m.infer(x) - t) ** 2))
    after = m.train(x, t, epochs=30, batch=512, lr=3e-3)
    pred = m.infer(x)
    assert after < 0.1 * before
    assert abs(float(np.mean((pred - t) ** 2)) - after) < 0.5 * after + 1e-2
    m.close()


def social(g2, S, n, seed):
    f = n / 65e6
    g = g2.gen_social(n, 2.1, 2.0, 5000.0, int(1.54e9 * f), int(2e6 * f), 10, 200, 3.0, 0.3, 0.9, 1e4, seed)
    g = g.preprocess()
    g, _, shells, _ = g.kcore_recode()
    return g, S.core_numbers(shells, g.nv)


def test_learned_schedule_gain(g2, dev):
    """The cost model, trained on the GPU by its own kernels from the per-root
    subtree sizes that plain runs log on two smaller social graphs, ranks the
    roots of a fresh graph (Spearman > 0.9).  Its longest-predicted-first
    order then shortens the makespan, in logged per-root nodes on the
    kernel's wave count, against the plain degeneracy order (whose heaviest
    roots come last): the gain the network brings over the method without
    it.  Counts stay bit-exact under the learned schedule."""
    print("case: learned schedule gain", flush=True)
    S = importlib.import_module("g2-aimd_amd.sched")
    base = 0x6732A1D0C0000300
    xs, ys = [], []
    for i in range(2):
        g, core = social(g2, S, 500_000, base + 1 + i)
        x, y, _ = S.logged_rows(dev, g, core)
        xs.append(x)
        ys.append(y)
    X, Y = np.concatenate(xs), np.concatenate(ys)
    norm = S.Standardizer(X)
    m = dev.Mlp(seed=5)
    before = float(np.mean((m.infer(norm(X)) - Y) ** 2))
    after = m.train(norm(X), Y, epochs=10, batch=1024, lr=3e-3)
    assert after < 0.2 * before
    g, core = social(g2, S, 1_000_000, base)
    with dev.Executor(device=0) as ex:
        ex.transfer(g)
        x = S.root_features(ex, np.arange(g.nv, dtype=np.uint32), core, g.nv)
        pred = m.infer(norm(x))
        plain = ex.run_mce(per_root=True)
        nodes = plain["per_root_nodes"].astype(np.float64)
        assert S.spearman(pred, np.log1p(nodes)) > 0.9
        waves = dev.device_info(0)["num_cus"] * 16
        order, split = S.learned_schedule(pred, ex.features()[:, 4], waves)
        lrn = ex.run_mce(per_root=True, order=order, split=split)
        assert lrn["total"] == plain["total"] and np.array_equal(lrn["hist"], plain["hist"])
        assert np.array_equal(lrn["per_root_nodes"], plain["per_root_nodes"])
    m.close()
    plain_span = S.simulated_makespan(nodes, np.arange(g.nv), None, waves)
    learned_span = S.simulated_makespan(nodes, order, None, waves)
    print("makespan plain", plain_span, "learned order", learned_span, "bound", nodes.sum() / waves, flush=True)
    assert learned_span < 0.8 * plain_span
    # measured: the same kernel and launch shape (one root per wave, no work
    # donation), plain order against the learned order and splits
    with dev.Executor(device=0, mce_donate_every=0xFFFFFFFF, mce_roots_per_wave=1) as ex:
        ex.transfer(g)
        t = {}
        for name, kw in (("plain", {}), ("learned", {"order": order, "split": split})):
            ex.run_mce(**kw)
            t[name] = float(np.median([ex.run_mce(**kw)["stats"]["kernel_ms"] for _ in range(3)]))
    print(f"measured kernel time: plain {t['plain']:.1f} ms, learned {t['learned']:.1f} ms "
          f"({t['plain'] / t['learned']:.2f}x)", flush=True)
    assert t["learned"] < 0.9 * t["plain"]


def test_learned_schedule_rmat_family(g2, orc, dev):
    """Learned root scheduling on the R-MAT family: the cost model trained on
    plain-run logs of 8 seeds (scale 12) schedules a fresh seed.  Total,
    histogram and every root's subtree node count equal the oracle's under
    the learned order and splits, with and without work donation; the
    prediction error, rank correlation, wave idle fractions and measured
    kernel times of both schedules are reported."""
    print("case: learned schedule, rmat family", flush=True)
    S = importlib.import_module("g2-aimd_amd.sched")
    xs, ys = [], []
    for i in range(8):
        g, core = S.prepare_graph(g2, 12, 0x6732A1D0C0002001 + i)
        x, y, _ = S.logged_rows(dev, g, core)
        xs.append(x)
        ys.append(y)
    X, Y = np.concatenate(xs), np.concatenate(ys)
    norm = S.Standardizer(X)
    m = dev.Mlp(seed=3)
    m.train(norm(X), Y, epochs=40, batch=256, lr=3e-3)
    g, core = S.prepare_graph(g2, 12, 0x6732A1D0C0001002)
    ref = orc.mce(g, per_root=True)
    truth = np.log1p(ref["per_root_nodes"].astype(np.float64))
    waves = dev.device_info(0)["num_cus"] * 16
    out = {}
    for dname, tune in (("donate", {"mce_donate_every": 64}), ("nodonate", {"mce_donate_every": 0xFFFFFFFF})):
        with dev.Executor(device=0, **tune) as ex:
            ex.transfer(g)
            x = S.root_features(ex, np.arange(g.nv, dtype=np.uint32), core, g.nv)
            pred = m.infer(norm(x))
            order, split = S.learned_schedule(pred, ex.features()[:, 4], waves)
            for sname, kw in (("plain", {}), ("learned", {"order": order, "split": split})):
                r = ex.run_mce(per_root=True, **kw)
                assert r["total"] == ref["total"]
                assert np.array_equal(r["hist"], ref["hist"])
                assert np.array_equal(r["per_root_nodes"], ref["per_root_nodes"])
                st = [ex.run_mce(**kw)["stats"] for _ in range(3)]
                out[f"{sname}_{dname}"] = (float(np.median([s["kernel_ms"] for s in st])),
                                           float(np.median([s["wave_idle_frac"] for s in st])))
    m.close()
    rmse = float(np.sqrt(np.mean((pred - truth) ** 2)))
    rho = S.spearman(pred, truth)
    print(f"rmse {rmse:.3f} spearman {rho:.3f} " +
          " ".join(f"{k}: {v[0]:.1f} ms idle {v[1]:.3f}" for k, v in out.items()), flush=True)
    assert rho > 0.8


def test_mlp_training_is_reproducible(dev):
    """Per-wave partial gradients reduced in a fixed order: the same seed and
    data give the same parameters bit for bit, and so does the gradient."""
    print("case: mlp reproducible", flush=True)
    rng = np.random.default_rng(12)
    x = rng.normal(0, 1, (30000, 8)).astype(np.float32)
    t = (np.tanh(x[:, 0]) + 0.3 * x[:, 1] * x[:, 2]).astype(np.float32)
    params, grads = [], []
    for _ in range(2):
        m = dev.Mlp(seed=21)
        grads.append(m.grad(x, t))
        m.train(x, t, epochs=3, batch=1024, lr=2e-3)
        params.append(m.params())
        m.close()
    assert np.array_equal(params[0], params[1])
    assert grads[0][0] == grads[1][0] and np.array_equal(grads[0][1], grads[1][1])


def _ulps(a, b):
    """distance in units in the last place between two float32 arrays"""
    ia = a.astype(np.float32).view(np.int32).astype(np.int64)
    ib = b.astype(np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def test_device_core_numbers_and_inputs(g2, dev):
    """Core numbers from the device's h-index iteration equal the host bucket
    peel; the 8 fp32 model inputs computed on the device are within 1 ulp of
    the same fp32 formula on the CPU (NumPy float32, one rounding per step)."""
    print("case: device core numbers and model inputs", flush=True)
    import graphs as G
    S = importlib.import_module("g2-aimd_amd.sched")
    cases = [S.prepare_graph(g2, 13, 0x6732A1D0C0002001)[0], social(g2, S, 200_000, 0x6732A1D0C0000310)[0],
             G.complete(g2, 12), G.cycle(g2, 9), G.complete_bipartite(g2, 3, 7)]
    for g in cases:
        with dev.Executor(device=0) as ex:
            ex.transfer(g)
            core = ex.core_numbers()
            assert np.array_equal(core, S.core_numbers_csr(g))
            x = ex.root_inputs()
            assert np.array_equal(x, ex.root_inputs(core))
            f = ex.features()
        f32 = np.float32
        d = f[:, 0].astype(f32)
        tri = f[:, 2].astype(f32)
        pairs = (f[:, 0] * (f[:, 0] - 1)).astype(f32)
        src = np.arange(g.nv)
        with np.errstate(divide="ignore", invalid="ignore"):
            clust = np.where(f[:, 0] > 1, (f32(2) * tri) / pairs, f32(0))
        ref = np.stack([np.log1p(d), np.log1p(core.astype(f32)), np.log1p(f[:, 1].astype(f32)), np.log1p(tri),
                        clust.astype(f32), src.astype(f32) / f32(max(g.nv, 1)),
                        f[:, 3].astype(f32) / np.maximum(d, f32(1)), np.log1p(f[:, 4].astype(f32))], axis=1)
        assert ref.dtype == np.float32
        assert int(_ulps(x, ref).max()) <= 1


def test_device_learned_schedule(g2, dev):
    """g2_learned_schedule keeps the whole schedule on the device: its
    predictions equal the cost model run on the host-standardised device
    inputs bit for bit, its order is the stable longest-predicted-first sort
    of them, and its parts are g2h_learned_schedule's (host C++) rule; the
    same holds for the sources of a hop-2 view."""
    print("case: learned schedule on the device", flush=True)
    S = importlib.import_module("g2-aimd_amd.sched")
    g, core = social(g2, S, 200_000, 0x6732A1D0C0000311)
    m = dev.Mlp(0, seed=5)
    waves = 256 * 16
    with dev.Executor(device=0) as ex:
        cases = [(g, None), (None, np.nonzero(np.arange(g.nv) % 3 == 1)[0].astype(np.uint32))]
        for view, src in cases:
            if src is None:
                ex.transfer(g)
            else:
                ex.transfer(g2.view_materialize(g, src, 2), sources=src)
            x = ex.root_inputs(core)
            norm = S.Standardizer(x)
            host_pred = m.infer(norm(x))
            order, split, pred = ex.learned_schedule(m, norm.mu, norm.sd, waves, core=core)
            assert np.array_equal(pred, host_pred)
            p = np.where(np.isfinite(pred), np.clip(pred, 0, 60), 0).astype(np.float32)
            assert np.array_equal(order, np.argsort(~p.view(np.uint32), kind="stable"))
            h_order, h_split = S.learned_schedule(pred, ex.features()[:, 4], waves)
            assert np.array_equal(order, h_order)
            assert np.array_equal(split, h_split)
            print(f"ns={len(order)}: {int((split > 1).sum())} split roots, max parts {int(split.max())}", flush=True)
    m.close()
