"""Failure handling forced from the host on the GPU path: malformed instances
are refused with CUOT_ERR_ARG before any launch, the serving loop rejects a
request larger than its whole HBM budget (CUOT_ERR_NOMEM on that request
only), defers a request while the arena is full and admits it once memory is
released, refuses to release a request that is still queued
(CUOT_ERR_BUSY), and refuses requests for resident clients it does not have.
Every kernel that runs here exits normally."""
import ctypes as C
import time

import numpy as np
import pytest

from tests.test_gpu_pipeline import SOT_K0, SOT_K1, SEED

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_NOMEM, ERR_BUSY = 1, 3, 5


def test_malformed_instances_refused(ctx):
    import torch

    import cuot_amd

    B = cuot_amd.to_block
    lib = ctx.lib
    buf = torch.empty(1 << 20, dtype=torch.uint8, device="cuda")
    p = buf.data_ptr()
    # SOT whose trees do not cover the 2n leaves (n > t * 2^h / 2)
    cfg = cuot_amd.SotConfig(n=1 << 12, t=15, h=9, reserved=0, k0=B(SOT_K0), k1=B(SOT_K1), seed=B(SEED))
    assert lib.cuot_sot_run(ctx.handle, C.byref(cfg), p, 1 << 20, p, p, p, None, None) == ERR_ARG
    assert b"cover" in lib.cuot_last_error()
    # Ferret with k >= n
    fcfg = cuot_amd.FerretConfig(n=4096, k=4096, t=1, h=12, ntasks=1, seed=B(SEED))
    assert lib.cuot_ferret_run(ctx.handle, C.byref(fcfg), p, 1 << 20, p, p, None, None) == ERR_ARG
    # 10-local LPN only: d != 10 is refused
    lp = cuot_amd.LpnParams(n=1024, k=100, d=9, ntasks=1, seed=B(SEED))
    assert lib.cuot_lpn(ctx.handle, C.byref(lp), p, p, p, None) == ERR_ARG
    # GGM depth beyond 16
    gp = cuot_amd.GgmParams(t=1, h=17, mode=0, n_out=1 << 17, k0=B(SOT_K0), k1=B(SOT_K1))
    assert lib.cuot_ggm_sender(ctx.handle, C.byref(gp), p, p, p, None) == ERR_ARG
    # bitsliced GGM forms with a work buffer one byte short of their query
    gp = cuot_amd.GgmParams(t=3, h=5, mode=0, n_out=3 << 5, k0=B(SOT_K0), k1=B(SOT_K1))
    need_s = lib.cuot_ggm_sender_bs_work_bytes(C.byref(gp))
    need_r = lib.cuot_ggm_recver_bs_work_bytes(C.byref(gp))
    assert 0 < need_s < need_r <= 1 << 20
    assert lib.cuot_ggm_sender_bs(ctx.handle, C.byref(gp), p, p, p, p, need_s - 1, None) == ERR_ARG
    assert b"work buffer" in lib.cuot_last_error()
    assert lib.cuot_ggm_recver_bs(ctx.handle, C.byref(gp), p, p, p, p, p, need_r - 1, None) == ERR_ARG
    assert b"work buffer" in lib.cuot_last_error()
    torch.cuda.synchronize()


def _inst(kind, n, seed):
    import cuot_amd
    from cuot_amd.workloads import make_instance

    wi = make_instance(kind, n, seed)
    return wi, cuot_amd.make_instance_struct(kind, n, wi.t, wi.h, wi.k, seed, k0=SOT_K0, k1=SOT_K1)


def test_server_budget_refusal_deferral_busy(ctx):
    import cuot_amd
    from cuot_amd.workloads import FERRET, SOT

    small_wi, small = _inst(FERRET, 50000, (5, 6))
    big_wi, big = _inst(SOT, 1 << 20, (7, 8))
    # a budget that holds one small Ferret request (outputs + workspace) but not two,
    # and far less than the 2^20 SOT request needs
    ws = ctx.ferret_workspace_bytes(cuot_amd.FerretConfig(n=small_wi.n, k=small_wi.k, t=small_wi.t, h=small_wi.h,
                                                          ntasks=1, seed=cuot_amd.to_block((5, 6))))
    a256 = lambda v: (v + 255) // 256 * 256  # noqa: E731
    outs = 2 * a256(16 * small_wi.n) + a256((small_wi.n + 31) // 32 * 4) + 256  # arena bytes of one result
    budget = max(a256(ws) + outs + 4096, 1 << 20)
    assert budget < a256(ws) + 2 * outs
    print(f"server budget {budget} bytes: refusal, deferral, busy", flush=True)
    srv = cuot_amd.Server(ctx, hbm_budget=budget, max_batch=1, streams=1)
    t_big = srv.submit(big)
    r = srv.poll(t_big, wait=True)
    assert r.done and r.status == ERR_NOMEM  # larger than the whole budget: rejected, not queued forever
    srv.release(t_big)
    t1 = srv.submit(small)
    r1 = srv.poll(t1, wait=True)
    assert r1.done and r1.status == 0
    t2 = srv.submit(small)  # its outputs and workspace do not fit beside t1's outputs
    time.sleep(0.05)
    r2 = srv.poll(t2, wait=False)
    assert not r2.done  # deferred while the arena is full
    with pytest.raises(cuot_amd.CuotError):
        srv.release(t2)  # still queued: CUOT_ERR_BUSY
    assert ctx.lib.cuot_server_release(srv.handle, t2) == ERR_BUSY
    srv.release(t1)  # frees t1's outputs: t2 is admitted
    r2 = srv.poll(t2, wait=True)
    assert r2.done and r2.status == 0
    from tests.test_gpu_serve import _served

    a = _served(r2, small_wi)
    srv.release(t2)
    nb, used = srv.stats()
    assert used == 0
    srv.close()
    # the deferred request's result equals a fresh run of the same request
    from tests.test_gpu_pipeline import run_ferret_gpu
    from tests.helpers import to_np

    s, rr, d = run_ferret_gpu(ctx, small_wi.n, small_wi.k, small_wi.t, small_wi.h,
                              np.array(small_wi.seed, dtype=np.uint64))
    assert np.array_equal(a[0], to_np(s)) and np.array_equal(a[1], to_np(rr)) and np.array_equal(a[3], to_np(d)[0])


def test_server_refuses_unknown_client(ctx):
    import cuot_amd
    from cuot_amd.workloads import FERRET, SOT, make_instance

    wi = make_instance(FERRET, 50000, (9, 10))
    srv = cuot_amd.Server(ctx, hbm_budget=1 << 26, max_batch=1, streams=1)
    try:
        for kind, client in ((FERRET, 1), (SOT, 1), (FERRET, -1)):
            inst = cuot_amd.make_instance_struct(kind, wi.n, wi.t, wi.h, wi.k, (9, 10), k0=SOT_K0, k1=SOT_K1,
                                                 client=client)
            t = C.c_int64()
            assert ctx.lib.cuot_server_submit(srv.handle, C.byref(inst), C.byref(t)) == ERR_ARG
        nb, used = srv.stats()
        assert used == 0
    finally:
        srv.close()
