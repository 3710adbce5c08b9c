"""GPU test of the serving workload itself: Poisson arrivals (200 req/s) of
mixed SOT / Ferret requests with sizes log-uniform in [2^14, 2^24] COTs
(cuot_amd.workloads.serving_arrivals, which here includes requests of more than
2^23 COTs), submitted in real time to a continuously batching server with a
24 GiB HBM budget.  Every served result satisfies its COT relation (SOT:
w = v ^ b * Delta; Ferret: r = s ^ lsb(r) * Delta), and a sample of them --
every request of at least 2^23 COTs and every fifth request -- is
byte-identical to the same request run alone."""
import time

import numpy as np
import pytest

from tests.test_gpu_pipeline import SOT_K0, SOT_K1, relation_holds_gpu, run_ferret_gpu, run_sot_gpu

pytestmark = pytest.mark.gpu

COUNT = 40


def _view(ptr, nblocks):
    import torch

    from cuot_amd.silent import device_bytes

    return device_bytes(ptr, nblocks * 16).view(torch.int64).view(nblocks, 2)


def _relation_ok(res, wi):
    import torch

    from cuot_amd.silent import device_bytes

    o0, o1, d = _view(res.out0, wi.n), _view(res.out1, wi.n), _view(res.delta, 1)
    if wi.kind == 0:
        bits = device_bytes(res.bits, (wi.n + 7) // 8)
        ok, weight = relation_holds_gpu(o0, o1, bits, d, wi.n)
        return ok and 0 < weight < wi.n
    lsb = (o1[:, 0] & 1).to(torch.bool)
    return bool(torch.equal(torch.where(lsb.unsqueeze(1), o0 ^ d, o0), o1)) and int(d[0, 0].item()) & 1 == 1


def _solo_equal(ctx, res, wi, ins):
    import torch

    from cuot_amd.silent import device_bytes

    seed = np.array(wi.seed, dtype=np.uint64)
    o0, o1, d = _view(res.out0, wi.n), _view(res.out1, wi.n), _view(res.delta, 1)
    if wi.kind == 0:
        v, w, b, dd = run_sot_gpu(ctx, wi.n, wi.t, wi.h, seed)
        nb = (wi.n + 7) // 8
        return (torch.equal(v, o0) and torch.equal(w, o1) and torch.equal(dd, d)
                and torch.equal(b[:nb], device_bytes(res.bits, nb)))
    s, r, dd = run_ferret_gpu(ctx, wi.n, wi.k, wi.t, wi.h, seed, ins.ntasks)
    return torch.equal(s, o0) and torch.equal(r, o1) and torch.equal(dd, d)


def test_served_poisson_mix_relation_and_solo(ctx):
    import torch

    import cuot_amd
    from cuot_amd.workloads import serving_arrivals

    arrivals = serving_arrivals(200.0, COUNT)
    big = [i for i, (_, wi) in enumerate(arrivals) if wi.n >= 1 << 23]
    assert big, "the drawn mix must hold a request of at least 2^23 COTs"
    reqs = [(t, wi, cuot_amd.make_instance_struct(wi.kind, wi.n, wi.t, wi.h, wi.k, wi.seed, k0=SOT_K0, k1=SOT_K1))
            for t, wi in arrivals]
    print(f"serving {COUNT} Poisson arrivals, largest n={max(wi.n for _, wi, _ in reqs)}, "
          f"{len(big)} requests >= 2^23", flush=True)
    srv = cuot_amd.Server(ctx, hbm_budget=24 << 30, max_batch=16, streams=2)
    t0 = time.perf_counter()
    tickets = []
    for t_arr, _, ins in reqs:
        dt = t0 + t_arr - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        tickets.append(srv.submit(ins))
    sample = set(big) | set(range(0, COUNT, 5))
    for i, ((_, wi, ins), tk) in enumerate(zip(reqs, tickets)):
        res = srv.poll(tk, wait=True)
        assert res.done and res.status == 0, f"request {i} failed with status {res.status}"
        assert _relation_ok(res, wi), f"request {i} (kind {wi.kind}, n={wi.n}) breaks the COT relation"
        if i in sample:
            print(f"solo check request {i} kind={wi.kind} n={wi.n}", flush=True)
            assert _solo_equal(ctx, res, wi, ins), f"request {i} (kind {wi.kind}, n={wi.n}) differs from solo"
        srv.release(tk)
    nb, used = srv.stats()
    assert used == 0 and nb < COUNT
    srv.close()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
