"""CPU oracle of the learned-dispatch MLP -- TEST INFRASTRUCTURE (tests/ only).

numpy float64 forward pass, analytic backward pass and central finite differences of the
network in pairwise_aligner_amd/csrc/pa_learned.hip: 8 -> 64 (ReLU) -> 64 (ReLU) -> 2, weighted
mean-squared error.  Parameter layout [W1 (8x64), b1, W2 (64x64), b2, W3 (64x2), b3], W[in][out].
The analytic gradient is pinned against finite differences in the tests, then used to check the
GPU gradient.
"""
import numpy as np

F, H, O = 8, 64, 2


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    i = 0
    W1 = p[i:i + F * H].reshape(F, H); i += F * H
    b1 = p[i:i + H]; i += H
    W2 = p[i:i + H * H].reshape(H, H); i += H * H
    b2 = p[i:i + H]; i += H
    W3 = p[i:i + H * O].reshape(H, O); i += H * O
    b3 = p[i:i + O]; i += O
    assert i == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    return h2 @ W3 + b3, (x, z1, h1, z2, h2)


def loss(p, x, t, w=None):
    y, _ = forward(p, x)
    w = np.ones_like(y) if w is None else np.asarray(w, dtype=np.float64)
    return float(np.sum(w * (y - t) ** 2) / len(x))


def loss_grad(p, x, t, w=None):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (x, z1, h1, z2, h2) = forward(p, x)
    n = len(x)
    w = np.ones_like(y) if w is None else np.asarray(w, dtype=np.float64)
    dy = 2.0 * w * (y - t) / n
    gW3 = h2.T @ dy
    gb3 = dy.sum(0)
    dh2 = (dy @ W3.T) * (z2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (z1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return float(np.sum(w * (y - t) ** 2) / n), g


def finite_difference(p, x, t, idx, w=None, eps=1e-5):
    p = np.asarray(p, dtype=np.float64)
    g = np.zeros(len(idx))
    for k, i in enumerate(idx):
        a = p.copy(); a[i] += eps
        b = p.copy(); b[i] -= eps
        g[k] = (loss(a, x, t, w) - loss(b, x, t, w)) / (2 * eps)
    return g
