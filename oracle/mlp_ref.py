"""CPU reference of the cost-model MLP (8 -> 64 -> 64 -> 1, ReLU).

TEST INFRASTRUCTURE ONLY.  float64 forward/backward; with emulate=True the
weights, the inputs and the layer-1 activations are rounded to bfloat16 (round
to nearest even) exactly where the device kernels round them, so the device's
forward pass and gradients can be compared within bf16 tolerance.  The
analytic gradient is pinned by central finite differences in the tests.
Parameter layout: W1[8][64], b1[64], W2[64][64], b2[64], W3[64], b3.
"""
import numpy as np

NPARAM = 8 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def bf16(x):
    x32 = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = x32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + 512].reshape(8, 64); o += 512
    b1 = p[o:o + 64]; o += 64
    W2 = p[o:o + 4096].reshape(64, 64); o += 4096
    b2 = p[o:o + 64]; o += 64
    W3 = p[o:o + 64]; o += 64
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, x, emulate=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    xx = r(x)
    z1 = xx @ r(W1) + b1
    h1 = r(np.maximum(z1, 0))
    z2 = h1 @ r(W2) + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ r(W3) + b3
    return y, (xx, z1, h1, z2, h2)


def loss_grad(p, x, t, emulate=True):
    """mean squared error and its gradient (straight-through for roundings)"""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    y, (xx, z1, h1, z2, h2) = forward(p, x, emulate)
    n = len(t)
    e = y - np.asarray(t, np.float64)
    loss = float(np.mean(e * e))
    dy = 2 * e / n
    gW3 = h2.T @ dy
    gb3 = dy.sum()
    dh2 = np.outer(dy, r(W3)) * (z2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ r(W2).T) * (z1 > 0)
    gW1 = xx.T @ dh1
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g
