"""Root features and the MFMA cost-model MLP against CPU references."""
import importlib

import numpy as np
import pytest

import graphs as G
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
        two = A @ deg - deg
        shared = ((A2 * A) > 0).sum(1)
        p0 = np.array([np.sum(g.cols[g.row_ptrs[v]:g.row_ptrs[v + 1]] > v) for v in range(n)])
        assert np.array_equal(f[:, 0], deg)
        assert np.array_equal(f[:, 1], two)
        assert np.array_equal(f[:, 2], tri)
        assert np.array_equal(f[:, 3], shared)
        assert np.array_equal(f[:, 4], p0)


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
