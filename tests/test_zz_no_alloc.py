"""GPU parity of cuot_ferret_run_noalloc (library-held workspace, no
allocation in the call) against cuot_ferret_run (caller workspace), through the
C ABI: two calls in a row on the shared scratch, the second on a smaller
instance, byte-for-byte; the smaller one also against the oracle; a config whose
workspace exceeds the scratch is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import to_np

pytestmark = pytest.mark.gpu

SEED = (0x0706050403020100, 0x0F0E0D0C0B0A0908)


def _run_both(ctx, n, k, t, h, seed):
    import torch

    import cuot_amd

    cfg = cuot_amd.FerretConfig(n=n, k=k, t=t, h=h, ntasks=1, seed=cuot_amd.to_block(seed))
    ws = torch.empty(ctx.ferret_workspace_bytes(cfg), dtype=torch.uint8, device="cuda:0")
    s0 = torch.empty((n, 2), dtype=torch.int64, device="cuda:0")
    r0 = torch.empty_like(s0)
    d0 = torch.empty((1, 2), dtype=torch.int64, device="cuda:0")
    ctx.ferret_run(cfg, ws, s0, r0, d0)
    s1 = torch.full_like(s0, -1)
    r1 = torch.full_like(r0, -1)
    d1 = torch.full_like(d0, -1)
    ctx.ferret_run_noalloc(cfg, s1, r1, d1)
    torch.cuda.synchronize()
    return (to_np(s0), to_np(r0), to_np(d0)), (to_np(s1), to_np(r1), to_np(d1))


def test_noalloc_matches_workspace_entry_two_calls(ctx):
    import cuot_amd

    # first call: the Ferret workload n = 10,168,320 (workspace 28.5 MiB of the 32 MiB scratch)
    big = dict(n=10168320, k=452160, t=1270, h=13)
    cfg = cuot_amd.FerretConfig(ntasks=1, seed=cuot_amd.to_block(SEED), **big)
    assert ctx.ferret_workspace_bytes(cfg) <= cuot_amd.FERRET_NOALLOC_WORKSPACE
    print("noalloc call 1: ferret n=10168320", flush=True)
    ref, got = _run_both(ctx, seed=SEED, **big)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    s, r, d = got
    lsb = (r[:, 0] & 1).astype(bool)
    assert np.array_equal(r, s ^ np.where(lsb[:, None], d[0], 0).astype(np.uint64))
    # second call: smaller instance on the same scratch (stale contents from call 1)
    small = dict(n=40000, k=2000, t=5, h=13)
    seed2 = (0x1111, 0x2222)
    print("noalloc call 2: ferret n=40000", flush=True)
    ref, got = _run_both(ctx, seed=seed2, **small)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    os_, or_, _ = po.ferret_run(small["n"], small["k"], small["t"], small["h"], np.array(seed2, dtype=np.uint64),
                                nthreads=4)
    assert np.array_equal(got[0], os_) and np.array_equal(got[1], or_)


def test_noalloc_refuses_oversize(ctx):
    import torch

    import cuot_amd
    from cuot_amd import CuotError

    # k = 2^21 bootstrap COTs need 4 * 2^21 * 16 B = 128 MiB of workspace
    cfg = cuot_amd.FerretConfig(n=1 << 23, k=1 << 21, t=1024, h=13, ntasks=1, seed=cuot_amd.to_block(SEED))
    assert ctx.ferret_workspace_bytes(cfg) > cuot_amd.FERRET_NOALLOC_WORKSPACE
    out = torch.empty((16, 2), dtype=torch.int64, device="cuda:0")
    with pytest.raises(CuotError, match="exceeds the library scratch"):
        ctx.ferret_run_noalloc(cfg, out, out)
    lib = cuot_amd._lib.load()
    assert lib.cuot_ferret_run_noalloc(ctx.handle, C.byref(cfg), C.c_void_p(out.data_ptr()),
                                       C.c_void_p(out.data_ptr()), None, None) == 1  # CUOT_ERR_ARG
