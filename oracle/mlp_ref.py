"""CPU oracle of the variant-selector MLP (float64 numpy).

TEST INFRASTRUCTURE ONLY: the product trains and evaluates the network with
its own HIP kernels (tempi_amd/csrc/mlp.hip); this restatement is what their
gradients and updates are checked against.  It restates, in matrix form, the
network and loss those kernels implement:

    h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), y = h2 W3 + b3
    loss = sum(m * (y - t)^2) / sum(m)

with the parameter vector laid out [W1 (in x h1), b1, W2 (h1 x h2), b2,
W3 (h2 x out), b3], row = input.  The analytic gradient is pinned by central
finite differences in tests/test_mlp.py.  The model has no counterpart in the
reference (which interpolates measured tables,
src/internal/measure_system.cpp:181-299).
"""
from __future__ import annotations

import numpy as np

IN, H1, H2 = 16, 64, 64


def shapes(out):
    return [(IN, H1), (H1,), (H1, H2), (H2,), (H2, out), (out,)]


def num_params(out):
    return sum(int(np.prod(s)) for s in shapes(out))


def unpack(params, out):
    p = np.asarray(params, dtype=np.float64)
    parts, o = [], 0
    for s in shapes(out):
        n = int(np.prod(s))
        parts.append(p[o:o + n].reshape(s))
        o += n
    return parts


def init_params(out, seed):
    """He-normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in shapes(out):
        if len(s) == 2:
            parts.append(rng.standard_normal(s) * np.sqrt(2.0 / s[0]))
        else:
            parts.append(np.zeros(s))
    return np.concatenate([q.reshape(-1) for q in parts]).astype(np.float32)


def forward(params, x, out):
    W1, b1, W2, b2, W3, b3 = unpack(params, out)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ W3 + b3
    return y, (x, h1, h2)


def loss_and_grad(params, x, t, m=None, out=None):
    out = out if out is not None else np.asarray(t).shape[1]
    y, (x, h1, h2) = forward(params, x, out)
    t = np.asarray(t, dtype=np.float64)
    m = np.ones_like(t) if m is None else np.asarray(m, dtype=np.float64)
    denom = m.sum()
    e = y - t
    loss = float((m * e * e).sum() / denom)
    dy = 2.0 * m * e / denom
    W1, b1, W2, b2, W3, b3 = unpack(params, out)
    gW3 = h2.T @ dy
    gb3 = dy.sum(0)
    d2 = (dy @ W3.T) * (h2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    gW1 = x.T @ d1
    gb1 = d1.sum(0)
    grad = np.concatenate([g.reshape(-1) for g in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, grad


def adam_step(params, m1, m2, grad, lr, t, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba, ICLR 2015), step number t >= 1."""
    m1 = b1 * m1 + (1 - b1) * grad
    m2 = b2 * m2 + (1 - b2) * grad * grad
    mh = m1 / (1 - b1 ** t)
    vh = m2 / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m1, m2
