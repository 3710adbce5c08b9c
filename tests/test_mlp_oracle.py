"""The float64 cost-model restatement (oracle/mlp_ref.py) is pinned by central
finite differences of its own loss, and its Adam step by the closed form of
the first step (a step of -lr * sign(g))."""
import numpy as np

from oracle import mlp_ref as mr


def _setup(F=6, H=8, N=40, seed=1):
    r = np.random.default_rng(seed)
    p = r.normal(0, 0.5, mr.param_count(F, H))
    x = r.normal(0, 1, (N, F))
    t = r.normal(0, 1, N)
    return F, H, p, x, t


def test_gradient_matches_finite_differences():
    F, H, p, x, t = _setup()
    _, g = mr.loss_grad(p, F, H, x, t)
    eps = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        q = p.copy()
        q[i] += eps
        lp, _ = mr.loss_grad(q, F, H, x, t)
        q[i] -= 2 * eps
        lm, _ = mr.loss_grad(q, F, H, x, t)
        fd[i] = (lp - lm) / (2 * eps)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))


def test_param_layout():
    F, H = 5, 7
    p = np.arange(mr.param_count(F, H), dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = mr.unpack(p, F, H)
    assert W1[1, 0] == F and b1[0] == H * F and W2[0, 0] == H * F + H
    assert b3 == p[-1] and w3[-1] == p[-2]


def test_adam_first_step_is_signed_lr():
    r = np.random.default_rng(0)
    p = r.normal(size=50)
    g = r.normal(size=50)
    q, m, v = mr.adam_step(p, g, np.zeros(50), np.zeros(50), 1, 0.01)
    assert np.allclose(q, p - 0.01 * np.sign(g), atol=1e-6)
