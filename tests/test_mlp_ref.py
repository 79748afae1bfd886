"""The CPU MLP reference (oracle/mlp_ref.py) is pinned by central finite
differences, by the same network written with torch's autograd, and by torch's
bfloat16 conversion, so the device forward pass and gradients are compared
against checked formulas."""
import numpy as np

from oracle import mlp_ref


def test_fd_gradient():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mlp_ref.NPARAM)
    x = rng.normal(0, 1, (40, 8))
    t = rng.normal(0, 1, 40)
    loss, g = mlp_ref.loss_grad(p, x, t, emulate=False)
    idx = rng.choice(mlp_ref.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [mlp_ref.NPARAM - 1, 0, 512, 4672, 4736]])  # b3, W1, b1, b2, W3
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


def _torch_mlp(p, x):
    """The same network written with torch (float64, autograd), independent
    of mlp_ref's own forward and backward formulas."""
    import torch
    W1, b1, W2, b2, W3, b3 = (torch.tensor(np.array(a, dtype=np.float64), requires_grad=True)
                              for a in mlp_ref.unpack(p))
    xt = torch.tensor(x, dtype=torch.float64)
    y = torch.relu(torch.relu(xt @ W1 + b1) @ W2 + b2) @ W3 + b3
    return y, (W1, b1, W2, b2, W3, b3)


def test_forward_and_gradient_match_torch():
    import torch
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, mlp_ref.NPARAM)
    x = rng.normal(0, 1, (33, 8))
    t = rng.normal(0, 1, 33)
    y_ref, _ = mlp_ref.forward(p, x, emulate=False)
    y, params = _torch_mlp(p, x)
    assert np.allclose(y.detach().numpy(), y_ref, rtol=1e-12, atol=1e-12)
    loss = torch.mean((y - torch.tensor(t)) ** 2)
    loss.backward()
    g_t = np.concatenate([q.grad.numpy().ravel() for q in params])
    l_ref, g_ref = mlp_ref.loss_grad(p, x, t, emulate=False)
    assert abs(l_ref - loss.item()) <= 1e-12 * max(1.0, abs(l_ref))
    assert np.allclose(g_ref, g_t, rtol=1e-10, atol=1e-12)


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(2)
    v = np.concatenate([rng.normal(0, 10, 4000), rng.normal(0, 1e-3, 1000)]).astype(np.float32)
    # exact halfway cases between bf16 neighbours, both parities
    base = rng.normal(0, 1, 500).astype(np.float32).view(np.uint32) & np.uint32(0xFFFF0000)
    v = np.concatenate([v, (base | np.uint32(0x8000)).view(np.float32)])
    ref = torch.tensor(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp_ref.bf16(v), ref)
