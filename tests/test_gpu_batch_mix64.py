"""GPU parity of the batched workload itself: the 64-instance mix of
cuot_amd.workloads.batch_mix (sizes log-uniform in [2^16, 2^22] COTs, SOT and
Ferret alternating, one seed per instance) run through cuot_batch_* as one
launch sequence, eagerly and as a replayed HIP graph with the two-stream
overlap.  Every member is byte-identical to its plain one-instance run, every
member satisfies its COT relation (SOT: w = v ^ b * Delta; Ferret: r = s ^
lsb(r) * Delta), and the members of at most 2^17 COTs equal the oracle."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import to_np
from tests.test_gpu_pipeline import SOT_K0, SOT_K1, relation_holds_gpu, run_ferret_gpu, run_sot_gpu

pytestmark = pytest.mark.gpu

ORACLE_MAX_N = 1 << 17


def _instances():
    from cuot_amd import make_instance_struct
    from cuot_amd.workloads import batch_mix

    mix = batch_mix(64)
    return [(wi, make_instance_struct(wi.kind, wi.n, wi.t, wi.h, wi.k, wi.seed, k0=SOT_K0, k1=SOT_K1)) for wi in mix]


def _relation_ok(bt, i, wi):
    import torch

    d = bt.delta[i:i + 1]
    if wi.kind == 0:
        ok, weight = relation_holds_gpu(bt.out0[i], bt.out1[i], bt.bits[i].view(torch.uint8), d, wi.n)
        return ok and 0 < weight < wi.n
    s, r = bt.out0[i], bt.out1[i]
    lsb = (r[:, 0] & 1).to(torch.bool)
    return bool(torch.equal(torch.where(lsb.unsqueeze(1), s ^ d, s), r)) and int(d[0, 0].item()) & 1 == 1


def _plain_equal(ctx, bt, i, wi, ins):
    import torch

    seed = np.array(wi.seed, dtype=np.uint64)
    if wi.kind == 0:
        v, w, b, d = run_sot_gpu(ctx, wi.n, wi.t, wi.h, seed)
        nb = (wi.n + 7) // 8
        return (torch.equal(v, bt.out0[i]) and torch.equal(w, bt.out1[i])
                and torch.equal(b[:nb], bt.bits[i].view(torch.uint8)[:nb]) and torch.equal(d[0], bt.delta[i]))
    s, r, d = run_ferret_gpu(ctx, wi.n, wi.k, wi.t, wi.h, seed, ins.ntasks)
    return torch.equal(s, bt.out0[i]) and torch.equal(r, bt.out1[i]) and torch.equal(d[0], bt.delta[i])


def _oracle_equal(bt, i, wi, ins):
    seed = np.array(wi.seed, dtype=np.uint64)
    if wi.kind == 0:
        ov, ow, ob, od = po.sot_run(wi.n, wi.t, wi.h, SOT_K0, SOT_K1, seed, nthreads=8)
        bits = bt.bits[i].cpu().numpy().view(np.uint8)[: (wi.n + 7) // 8]
        return (np.array_equal(to_np(bt.out0[i]), ov) and np.array_equal(to_np(bt.out1[i]), ow)
                and np.array_equal(bits, ob) and np.array_equal(to_np(bt.delta)[i], od))
    os_, or_, od = po.ferret_run(wi.n, wi.k, wi.t, wi.h, seed, ntasks=ins.ntasks, nthreads=8)
    return (np.array_equal(to_np(bt.out0[i]), os_) and np.array_equal(to_np(bt.out1[i]), or_)
            and np.array_equal(to_np(bt.delta)[i], od))


def test_batch_mix64_matches_plain_relation_oracle(ctx):
    import torch

    import cuot_amd

    mix = _instances()
    total = sum(wi.n for wi, _ in mix)
    assert len(mix) == 64 and min(wi.n for wi, _ in mix) >= 1 << 16 and max(wi.n for wi, _ in mix) <= 1 << 22
    print(f"64-instance batched mix, {total} COTs: graph replay with overlap", flush=True)
    bt = cuot_amd.Batch(ctx, [ins for _, ins in mix])
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        bt.run(graph=True, stream=s, overlap=True)
        bt.run(graph=True, stream=s, overlap=True)  # a replay of the captured graph
    s.synchronize()
    n_oracle = 0
    for i, (wi, ins) in enumerate(mix):
        print(f"member {i} kind={wi.kind} n={wi.n}", flush=True)
        assert _relation_ok(bt, i, wi), f"member {i} (kind {wi.kind}, n={wi.n}) breaks the COT relation"
        assert _plain_equal(ctx, bt, i, wi, ins), f"member {i} (kind {wi.kind}, n={wi.n}) differs from plain"
        if wi.n <= ORACLE_MAX_N:
            assert _oracle_equal(bt, i, wi, ins), f"member {i} (kind {wi.kind}, n={wi.n}) differs from the oracle"
            n_oracle += 1
    assert n_oracle >= 4
    bt.close()
    torch.cuda.empty_cache()
