"""GPU parity of the cost-model kernels (cuot_mlp_*) with the float64
restatement: forward outputs and loss gradients agree to float32 accuracy
(relative 1e-4 on the gradient norm, 1e-5 on the outputs), one Adam step
matches, and training on a smooth target lowers the loss."""
import ctypes as C

import numpy as np
import pytest

from oracle import mlp_ref as mr

pytestmark = pytest.mark.gpu


def _t(a):
    import torch

    return torch.tensor(np.asarray(a, dtype=np.float32), device="cuda")


def _grad_gpu(ctx, p, F, H, x, t):
    import torch

    N = x.shape[0]
    P = mr.param_count(F, H)
    ws = torch.empty(ctx.lib.cuot_mlp_workspace_bytes(F, H, N), dtype=torch.uint8, device="cuda")
    g = torch.empty(P, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    pd, xd, td = _t(p), _t(x), _t(t)
    from cuot_amd import check
    from cuot_amd import _stream_ptr

    check(ctx.lib.cuot_mlp_grad(ctx.handle, pd.data_ptr(), F, H, xd.data_ptr(), td.data_ptr(), N, ws.data_ptr(),
                                ws.numel(), g.data_ptr(), loss.data_ptr(), _stream_ptr(None)), "cuot_mlp_grad")
    torch.cuda.synchronize()
    return float(loss.item()), g.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("F,H,N", [(6, 8, 40), (12, 32, 1000), (20, 64, 3000)])
def test_forward_and_gradient_match_float64(ctx, F, H, N):
    import torch

    from cuot_amd import _stream_ptr, check

    print(f"mlp parity F={F} H={H} N={N}", flush=True)
    r = np.random.default_rng(F * 100 + H)
    p = r.normal(0, 0.4, mr.param_count(F, H)).astype(np.float32).astype(np.float64)
    x = r.normal(0, 1, (N, F)).astype(np.float32).astype(np.float64)
    t = r.normal(0, 1, N).astype(np.float32).astype(np.float64)
    y = torch.empty(N, dtype=torch.float32, device="cuda")
    pd, xd = _t(p), _t(x)
    check(ctx.lib.cuot_mlp_forward(ctx.handle, pd.data_ptr(), F, H, xd.data_ptr(), N, y.data_ptr(),
                                   _stream_ptr(None)), "cuot_mlp_forward")
    torch.cuda.synchronize()
    y_ref, _ = mr.forward(p, F, H, x)
    assert np.max(np.abs(y.cpu().numpy() - y_ref)) <= 1e-5 * max(1.0, np.max(np.abs(y_ref)))
    loss, g = _grad_gpu(ctx, p, F, H, x, t)
    loss_ref, g_ref = mr.loss_grad(p, F, H, x, t)
    assert abs(loss - loss_ref) <= 1e-5 * max(1.0, loss_ref)
    assert np.linalg.norm(g - g_ref) <= 1e-4 * np.linalg.norm(g_ref)


def test_adam_step_matches(ctx):
    import torch

    from cuot_amd import _stream_ptr, check

    r = np.random.default_rng(5)
    P = 300
    p, g = r.normal(size=P), r.normal(size=P)
    m, v = r.normal(size=P) * 0.1, np.abs(r.normal(size=P)) * 0.1
    pd, gd, md, vd = _t(p), _t(g), _t(m), _t(v)
    check(ctx.lib.cuot_mlp_adam(ctx.handle, pd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(), P, 3,
                                C.c_float(0.01), _stream_ptr(None)), "cuot_mlp_adam")
    torch.cuda.synchronize()
    p2, m2, v2 = mr.adam_step(p.astype(np.float32).astype(np.float64), g.astype(np.float32),
                              m.astype(np.float32), v.astype(np.float32), 3, 0.01)
    assert np.allclose(pd.cpu().numpy(), p2, rtol=1e-5, atol=1e-6)
    assert np.allclose(md.cpu().numpy(), m2, rtol=1e-5, atol=1e-7)
    assert np.allclose(vd.cpu().numpy(), v2, rtol=1e-5, atol=1e-7)


def test_training_lowers_loss(ctx):
    from cuot_amd.tune import CostModel

    r = np.random.default_rng(9)
    x = r.uniform(-1, 1, (2000, 4))
    t = np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2] - 0.5 * x[:, 3] ** 2
    cm = CostModel(ctx, F=4, H=32, seed=1)
    l0 = cm.loss(x, t)
    cm.fit(x, t, epochs=300, lr=3e-3)
    l1 = cm.loss(x, t)
    print(f"loss {l0:.4f} -> {l1:.4f}", flush=True)
    assert l1 < 0.1 * l0
