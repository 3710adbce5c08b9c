"""GPU parity of the drop-in boundary: the emp-ot dev_layer replacements
(cuda_mpcot_sender / cuda_mpcot_recver / cuda_primal_lpn) and the SOT role
objects, called through the C ABI from the Python mirrors, against the oracle
and against the engine's whole-instance entry point."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import rand_blocks, rng, to_dev, to_np

pytestmark = pytest.mark.gpu


def _choices(alphas, h):
    return np.array([[1 - ((int(a) >> (h - 1 - d)) & 1) for d in range(h)] for a in alphas], dtype=np.uint8)


@pytest.mark.parametrize("t,h", [(3, 5), (5, 13), (2, 9)])
def test_ferret_dev_layer(ctx, t, h):
    import torch

    from cuot_amd import dev_layer as dl

    r = rng(100 + t * h)
    N = t << h
    roots = rand_blocks(r, t)
    delta = rand_blocks(r, 1)[0] | np.array([1, 0], dtype=np.uint64)
    expanded = torch.zeros((N, 2), dtype=torch.int64, device="cuda")
    expanded[:t] = to_dev(roots)
    buf = torch.zeros_like(expanded)
    sep = torch.zeros_like(expanded)
    lS, rS, ss = np.zeros((t * h, 2), np.uint64), np.zeros((t * h, 2), np.uint64), np.zeros((t, 2), np.uint64)
    dl.cuda_setdev(0)
    dl.cuda_mpcot_sender(ctx, [expanded], [buf], [sep], lS, rS, ss, [t], h, delta)
    ov, ols, ors, oss = po.ferret_mpcot_sender(roots, t, h, delta, N)
    assert np.array_equal(to_np(expanded), ov)
    assert np.array_equal(lS, ols) and np.array_equal(rS, ors) and np.array_equal(ss, oss)
    # receiver
    alphas = r.integers(0, 1 << h, size=t)
    ch = _choices(alphas, h).reshape(-1)
    cS = np.ascontiguousarray(np.where(ch[:, None] == 1, rS, lS))
    rexp = torch.zeros_like(expanded)
    dl.cuda_mpcot_recver(ctx, [rexp], [buf], [sep], cS, ss, [t], h, ch)
    ow = po.ferret_mpcot_recver(t, h, ch, cS, ss, N)
    assert np.array_equal(to_np(rexp), ow)


def test_primal_lpn(ctx):
    import torch

    from cuot_amd import dev_layer as dl

    r = rng(9)
    n, k = 50000, 2500
    seed = rand_blocks(r, 1)[0]
    noise, kk = rand_blocks(r, n), rand_blocks(r, k)
    key = np.zeros(44, dtype=np.uint32)
    key[:4] = seed.view(np.uint32)
    key[4:] = 0xDEADBEEF  # only the first 16 bytes (the seed) matter
    expanded = to_dev(noise)
    kk_d = torch.zeros((k, 2), dtype=torch.int64, device="cuda")
    nn = np.zeros((n, 2), dtype=np.uint64)
    dl.cuda_primal_lpn(ctx, 0, None, 10, [n], k, key, [expanded], nn, [kk_d], kk)
    ref = po.ferret_lpn(noise, kk, seed, nthreads=4)
    assert np.array_equal(nn, ref)
    assert np.array_equal(to_np(expanded), ref)


def test_sot_roles_match_engine(ctx):
    import torch

    import cuot_amd
    from cuot_amd.silent import SilentConfig, SOTRecver, SOTSender

    seed = np.array([11, 22], dtype=np.uint64)
    cfg = SilentConfig(logOT=15, nTree=32)
    snd, rcv = SOTSender(ctx, cfg, seed), SOTRecver(ctx, cfg, seed)
    snd.base_ot()
    rcv.base_ot()
    snd.seed_expand()
    rcv.get_punc_key(snd)
    rcv.seed_expand()
    snd.dual_lpn()
    rcv.dual_lpn()
    n = 1 << 14
    v = snd.fullVector.clone().view(torch.int64).reshape(n, 2)
    w = rcv.puncVector.clone().view(torch.int64).reshape(n, 2)
    b = rcv.choiceVector.clone()
    d = snd.delta.clone().view(torch.int64).reshape(1, 2)
    # engine whole-instance run with the same config
    ov, ow, ob, od = po.sot_run(n, 32, 10, np.array([3242342, 0], np.uint64), np.array([8993849, 0], np.uint64),
                                seed, nthreads=4)
    assert np.array_equal(to_np(v), ov)
    assert np.array_equal(to_np(w), ow)
    assert np.array_equal(b.cpu().numpy()[: n // 8], ob)
    assert np.array_equal(to_np(d)[0], od)
    snd.close()
    rcv.close()


def test_role_order_is_enforced(ctx):
    from cuot_amd import CuotError
    from cuot_amd.silent import SilentConfig, SOTRecver

    rcv = SOTRecver(ctx, SilentConfig(logOT=12, nTree=8), np.array([1, 2], np.uint64))
    with pytest.raises(CuotError):
        rcv.seed_expand()  # before base_ot / get_punc_key
    rcv.close()
