"""GPU tests of the serving loop (cuot_server_*): every request's result in a
continuously batched mix is byte-identical to the same request run alone
(plain cuot_sot_run / cuot_ferret_run), the FIFO plain comparator gives the
same bytes, and a real-time Poisson replay completes every request."""
import time

import numpy as np
import pytest

from tests.helpers import to_np
from tests.test_gpu_pipeline import SOT_K0, SOT_K1, run_ferret_gpu, run_sot_gpu

pytestmark = pytest.mark.gpu


def _requests(count, lo_log=10, hi_log=16, salt=7):
    from cuot_amd import make_instance_struct
    from cuot_amd.workloads import FERRET, SOT, make_instance

    r = np.random.default_rng(salt)
    out = []
    for i in range(count):
        n = int(2 ** (lo_log + (hi_log - lo_log) * r.random()))
        kind = SOT if r.random() < 0.5 else FERRET
        seed = (int(r.integers(0, 2**62)), int(r.integers(0, 2**62)))
        wi = make_instance(kind, n, seed)
        out.append((wi, make_instance_struct(kind, n, wi.t, wi.h, wi.k, seed, k0=SOT_K0, k1=SOT_K1)))
    return out


def _dev_to_np(ptr, nblocks):
    from cuot_amd.silent import device_bytes

    return device_bytes(ptr, nblocks * 16).cpu().numpy().view(np.uint64).reshape(nblocks, 2)


def _bits_to_np(ptr, n):
    from cuot_amd.silent import device_bytes

    return device_bytes(ptr, (n + 7) // 8).cpu().numpy()


def _solo(ctx, wi, ins):
    if wi.kind == 0:
        v, w, b, d = run_sot_gpu(ctx, wi.n, wi.t, wi.h, np.array(wi.seed, dtype=np.uint64))
        return to_np(v), to_np(w), b.cpu().numpy()[: (wi.n + 7) // 8], to_np(d)[0]
    s, r, d = run_ferret_gpu(ctx, wi.n, wi.k, wi.t, wi.h, np.array(wi.seed, dtype=np.uint64), ins.ntasks)
    return to_np(s), to_np(r), None, to_np(d)[0]


def _served(res, wi):
    o0 = _dev_to_np(res.out0, wi.n)
    o1 = _dev_to_np(res.out1, wi.n)
    bits = _bits_to_np(res.bits, wi.n) if wi.kind == 0 else None
    d = _dev_to_np(res.delta, 1)[0]
    return o0, o1, bits, d


@pytest.mark.parametrize("plain", [False, True])
def test_served_equals_solo(ctx, plain):
    import cuot_amd

    reqs = _requests(12)
    print(f"serving {len(reqs)} requests plain={plain}", flush=True)
    srv = cuot_amd.Server(ctx, hbm_budget=1 << 30, max_batch=4, streams=2, plain=plain)
    tickets = []
    for i, (_, ins) in enumerate(reqs):
        tickets.append(srv.submit(ins))
        if i % 3 == 2:
            time.sleep(0.002)  # arrivals while earlier batches run
    results = [srv.poll(t, wait=True) for t in tickets]
    for (wi, ins), res, t in zip(reqs, results, tickets):
        assert res.done and res.status == 0
        got = _served(res, wi)
        exp = _solo(ctx, wi, ins)
        for g, e in zip(got, exp):
            assert (g is None and e is None) or np.array_equal(g, e), f"request n={wi.n} kind={wi.kind} differs"
        srv.release(t)
    nb, used = srv.stats()
    assert used == 0
    if not plain:
        assert nb < len(reqs)  # requests were batched
    srv.close()


def test_replay_completes(ctx):
    import cuot_amd

    reqs = _requests(40, lo_log=10, hi_log=15, salt=11)
    r = np.random.default_rng(3)
    arrivals = np.cumsum(r.exponential(1 / 400.0, size=len(reqs)))
    lat, wall, nb, fails = cuot_amd.serve_replay(ctx, [ins for _, ins in reqs], arrivals, hbm_budget=1 << 30,
                                                 max_batch=8, check_relation=True)
    print(f"replay: wall {wall:.3f} s, batches {nb}, p50 {np.percentile(lat, 50) * 1e3:.2f} ms, "
          f"p99 {np.percentile(lat, 99) * 1e3:.2f} ms", flush=True)
    assert np.all(np.isfinite(lat)) and np.all(lat > 0)
    assert wall >= arrivals[-1]
    assert fails == 0, f"{fails} indices of the served results break the COT relation"


@pytest.mark.parametrize("plain", [False, True])
def test_resident_ferret_clients_match_rcot(ctx, plain):
    """Ferret requests of resident clients are served from each client's
    iterating session (Delta and bootstrap stay in HBM between requests): the
    outputs of a client's requests, in arrival order, equal FerretCOT.rcot
    calls of the same sizes on a session with the client's seed, whatever SOT
    requests run in between; Delta is the session's and the COT relation holds."""
    import torch

    import cuot_amd
    from cuot_amd import FerretConfig, make_instance_struct
    from cuot_amd.ferret import FerretCOT
    from cuot_amd.workloads import FERRET, SOT, make_instance

    print(f"resident ferret clients plain={plain}", flush=True)
    P = dict(n=40000, k=2000, t=5, h=13)
    seed = (0x1234, 0x5678)
    fcfg = FerretConfig(ntasks=1, seed=cuot_amd.to_block(seed), **P)
    srv = cuot_amd.Server(ctx, hbm_budget=2 << 30, plain=plain, ferret_clients=2, ferret_session=fcfg)
    sot = make_instance(SOT, 5000, (9, 9))
    reqs = [(1, 1000), (2, 50000), (0, None), (1, 40000), (2, 10), (1, 37807), (0, None), (2, 3)]
    tickets = []
    for client, n in reqs:
        if client == 0:
            inst = make_instance_struct(SOT, sot.n, sot.t, sot.h, 0, (9, 9), k0=SOT_K0, k1=SOT_K1)
        else:
            inst = make_instance_struct(FERRET, n, 0, 0, 0, (0, 0), client=client)
        tickets.append(srv.submit(inst))
    res = [srv.poll(t, wait=True) for t in tickets]
    assert all(r.status == 0 for r in res)
    refs = {c: FerretCOT(ctx, seed=(seed[0] ^ c, seed[1]), **P) for c in (1, 2)}
    for (client, n), r in zip(reqs, res):
        if client == 0:
            continue
        s_ref, r_ref = refs[client].rcot(n)
        torch.cuda.synchronize()
        s = _dev_to_np(r.out0, n)
        rr = _dev_to_np(r.out1, n)
        d = _dev_to_np(r.delta, 1)[0]
        assert np.array_equal(s, to_np(s_ref)) and np.array_equal(rr, to_np(r_ref))
        assert np.array_equal(d, to_np(refs[client].Delta)[0])
        lsb = (rr[:, 0] & 1).astype(bool)
        assert np.array_equal(rr, s ^ np.where(lsb[:, None], d, 0).astype(np.uint64))
    for t in tickets:
        srv.release(t)
    for f in refs.values():
        f.close()
    srv.close()
