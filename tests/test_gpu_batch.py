"""GPU parity of the batched engine (cuot_batch_*): every instance of a mixed
SOT / Ferret batch is byte-identical to the plain single-instance run of the
same instance (which the pipeline tests pin to the oracle), one small SOT and
one small Ferret instance of the batch are checked against the oracle
directly, and replays of the captured HIP graph repeat the eager result."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import to_np
from tests.test_gpu_pipeline import SOT_K0, SOT_K1, run_ferret_gpu, run_sot_gpu

pytestmark = pytest.mark.gpu


def _mix():
    from cuot_amd import make_instance_struct
    from cuot_amd.workloads import FERRET, SOT, make_instance

    rows = [(SOT, 1000), (FERRET, 20000), (SOT, 1 << 14), (FERRET, 50000), (SOT, 70000), (SOT, 3 << 14),
            (FERRET, 3000), (SOT, 1 << 14), (FERRET, 100000), (SOT, 5000)]
    out = []
    for i, (kind, n) in enumerate(rows):
        seed = (0x1111 * (i + 1), 0xABCDEF00 + i)
        wi = make_instance(kind, n, seed)
        out.append((wi, make_instance_struct(kind, n, wi.t, wi.h, wi.k, seed, ntasks=1 + (i % 3),
                                             k0=SOT_K0, k1=SOT_K1)))
    return out


def _plain(ctx, wi, ins):
    if wi.kind == 0:
        v, w, b, d = run_sot_gpu(ctx, wi.n, wi.t, wi.h, np.array(wi.seed, dtype=np.uint64))
        return to_np(v), to_np(w), b.cpu().numpy()[: (wi.n + 7) // 8], to_np(d)[0]
    s, r, d = run_ferret_gpu(ctx, wi.n, wi.k, wi.t, wi.h, np.array(wi.seed, dtype=np.uint64), ins.ntasks)
    return to_np(s), to_np(r), None, to_np(d)[0]


def _batch_out(bt, i, wi):
    o0, o1 = to_np(bt.out0[i]), to_np(bt.out1[i])
    bits = None
    if wi.kind == 0:
        bits = bt.bits[i].cpu().numpy().view(np.uint8)[: (wi.n + 7) // 8]
    return o0, o1, bits, to_np(bt.delta)[i]


def _same(a, b):
    return all((x is None and y is None) or np.array_equal(x, y) for x, y in zip(a, b))


def test_batch_matches_plain(ctx):
    import torch

    import cuot_amd

    mix = _mix()
    print(f"batch of {len(mix)} instances vs plain runs", flush=True)
    bt = cuot_amd.Batch(ctx, [ins for _, ins in mix])
    bt.run()
    torch.cuda.synchronize()
    for i, (wi, ins) in enumerate(mix):
        got = _batch_out(bt, i, wi)
        exp = _plain(ctx, wi, ins)
        assert _same(got, exp), f"instance {i} kind={wi.kind} n={wi.n} differs from the plain run"
    bt.close()


def test_batch_small_members_match_oracle(ctx):
    import torch

    import cuot_amd

    mix = _mix()
    bt = cuot_amd.Batch(ctx, [ins for _, ins in mix])
    bt.run()
    torch.cuda.synchronize()
    wi = mix[0][0]  # SOT n = 1000
    ov, ow, ob, od = po.sot_run(wi.n, wi.t, wi.h, SOT_K0, SOT_K1, np.array(wi.seed, dtype=np.uint64), nthreads=8)
    v, w, b, d = _batch_out(bt, 0, wi)
    assert np.array_equal(d, od) and np.array_equal(v, ov) and np.array_equal(w, ow) and np.array_equal(b, ob)
    wi, ins = mix[6]  # Ferret n = 3000
    os_, or_, od = po.ferret_run(wi.n, wi.k, wi.t, wi.h, np.array(wi.seed, dtype=np.uint64), ntasks=ins.ntasks,
                                 nthreads=8)
    s, r, _, d = _batch_out(bt, 6, wi)
    assert np.array_equal(d, od) and np.array_equal(s, os_) and np.array_equal(r, or_)
    bt.close()


def test_batch_graph_replays(ctx):
    import torch

    import cuot_amd

    mix = _mix()
    bt = cuot_amd.Batch(ctx, [ins for _, ins in mix])
    bt.run()
    torch.cuda.synchronize()
    ref = [_batch_out(bt, i, wi) for i, (wi, _) in enumerate(mix)]
    s = torch.cuda.Stream()
    for rep in range(3):
        for t in bt.out0 + bt.out1:
            t.fill_(-1)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            bt.run(graph=True, stream=s)
        s.synchronize()
        # eager work between replays
        junk = torch.randn(1 << 20, device="cuda").sum().item()
        assert junk == junk
        for i, (wi, _) in enumerate(mix):
            assert _same(_batch_out(bt, i, wi), ref[i]), f"replay {rep} instance {i} differs"
    bt.close()


def test_batch_overlap_matches(ctx):
    """CUOT_BATCH_OVERLAP (preparation, sparse encodes and LPN on the second
    stream): eager and graph replays byte-identical to the one-stream run."""
    import torch

    import cuot_amd

    mix = _mix()
    print("batch with the two-stream overlap", flush=True)
    bt = cuot_amd.Batch(ctx, [ins for _, ins in mix])
    bt.run()
    torch.cuda.synchronize()
    ref = [_batch_out(bt, i, wi) for i, (wi, _) in enumerate(mix)]
    s = torch.cuda.Stream()
    for rep, graph in enumerate((False, True, True, True)):
        for t in bt.out0 + bt.out1:
            t.fill_(-1)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            bt.run(graph=graph, stream=s, overlap=True)
        s.synchronize()
        junk = torch.randn(1 << 20, device="cuda").sum().item()
        assert junk == junk
        for i, (wi, _) in enumerate(mix):
            assert _same(_batch_out(bt, i, wi), ref[i]), f"run {rep} (graph={graph}) instance {i} differs"
    bt.close()
