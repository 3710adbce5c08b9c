"""Float64 restatement of the autotuner's cost-model network (the HIP kernels
in cuot_amd/csrc/mlp.hip): F -> H tanh -> H tanh -> 1, mean squared error,
Adam.  Parameters are one vector: W1[H][F], b1[H], W2[H][H], b2[H], w3[H], b3.

TEST INFRASTRUCTURE ONLY: the checker for the HIP forward/backward kernels
(imported by tests/), never by the cuot_amd product.  The analytic gradient
here is itself pinned by central finite differences in the tests.
"""
from __future__ import annotations

import numpy as np


def param_count(F: int, H: int) -> int:
    return H * F + H + H * H + H + H + 1


def unpack(p, F: int, H: int):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, F: int, H: int, x):
    W1, b1, W2, b2, w3, b3 = unpack(p, F, H)
    x = np.asarray(x, dtype=np.float64)
    a1 = np.tanh(x @ W1.T + b1)
    a2 = np.tanh(a1 @ W2.T + b2)
    return a2 @ w3 + b3, (a1, a2)


def loss_grad(p, F: int, H: int, x, t):
    """(mean squared error, gradient vector) in float64."""
    W1, b1, W2, b2, w3, b3 = unpack(p, F, H)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    N = x.shape[0]
    y, (a1, a2) = forward(p, F, H, x)
    e = y - t
    loss = float(np.mean(e * e))
    gy = 2.0 * e / N                       # dL/dy
    g2 = (gy[:, None] * w3[None, :]) * (1.0 - a2 * a2)   # dL/dz2
    g1 = (g2 @ W2) * (1.0 - a1 * a1)       # dL/dz1
    gW1 = g1.T @ x
    gb1 = g1.sum(0)
    gW2 = g2.T @ a1
    gb2 = g2.sum(0)
    gw3 = a2.T @ gy
    gb3 = gy.sum()
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def adam_step(p, g, m, v, step: int, lr: float, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    p = p - lr * (m / c1) / (np.sqrt(v / c2) + eps)
    return p, m, v
