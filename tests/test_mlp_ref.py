"""The CPU MLP reference (oracle/mlp_ref.py) is pinned by central finite
differences, so the device gradients are compared against a checked formula."""
import numpy as np

from oracle import mlp_ref


def test_fd_gradient():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mlp_ref.NPARAM)
    x = rng.normal(0, 1, (40, 8))
    t = rng.normal(0, 1, 40)
    loss, g = mlp_ref.loss_grad(p, x, t, emulate=False)
    idx = rng.choice(mlp_ref.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [mlp_ref.NPARAM - 1, 0, 512, 4736]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (mlp_ref.loss_grad(p + e, x, t, emulate=False)[0] - mlp_ref.loss_grad(p - e, x, t, emulate=False)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.0e-5, 65504.0], np.float32)
    # 1 + 2^-8 is exactly halfway between bf16 neighbours 1 and 1+2^-7: ties to even -> 1
    assert mlp_ref.bf16(v)[1] == 1.0
    assert mlp_ref.bf16(v)[2] == 1.0078125
    assert mlp_ref.bf16(v)[0] == 1.0
