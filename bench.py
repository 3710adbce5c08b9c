"""Benchmark: COTs/s of the silent-OT engine on one MI355X.

Workload (headline): plain SOT, n = 2^24 random COTs of 128-bit blocks,
t = 1024 GGM trees of depth 15 (2n = 2^25 leaves), AES-128 two-key PRG,
scaler-2 quasi-cyclic code, both roles on the GPU, everything derived from one
128-bit seed.  One step = one whole instance: seed -> (v, w, b) with the inputs
already resident in HBM.  Ferret (n = 10,168,320, k = 452,160, t = 1,270 trees
of 8192 leaves, d = 10) is measured in the same run and reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  The cpu_baseline leg times the CPU oracle (the
project's C restatement) on the SOT n = 2^20 configuration on host threads.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SOT_K0 = (3242342, 0)  # SilentConfig leftKey (silent/main.cu:89)
SOT_K1 = (8993849, 0)
SEED = (0x0706050403020100, 0x0F0E0D0C0B0A0908)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-ferret", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--sot-n-log", type=int, default=24)
    ap.add_argument("--sequential", action="store_true",
                    help="time staged (non-overlapped) instances for the headline value")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched (64-instance) leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-sweep", type=str, default="50,100,200,400,800",
                    help="offered loads (req/s) of the serving sweep; empty to skip")
    ap.add_argument("--sweep-requests", type=int, default=300)
    ap.add_argument("--batch-first", action="store_true", help="run the batched leg before the serving leg")
    ap.add_argument("--qc-only", action="store_true",
                    help="time only the quasi-cyclic encode of one party at n = 2^sot-n-log (after a 256 MiB "
                         "device copy that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE) and print one JSON line")
    return ap.parse_args()


def qc_only(args):
    """The QC encoder alone: A^ of a random circulant, one untimed encode, then
    `steps` encodes of random 2n-block input timed with HIP events (median).
    A 256 MiB device-to-device copy first gives the hardware byte counters a
    kernel of known traffic in the same process."""
    import numpy as np
    import torch

    import cuot_amd

    ctx = cuot_amd.Context(0)
    n = 1 << args.sot_n_log
    dev = "cuda"
    g = torch.Generator().manual_seed(1)
    a_bits = torch.randint(0, 256, ((n + 7) // 8 + 16,), dtype=torch.uint8, generator=g).to(dev)
    ahat = torch.empty(ctx.qc_ahat_len(n), dtype=torch.int32, device=dev)
    scratch = torch.empty(ctx.qc_scratch_bytes(n), dtype=torch.uint8, device=dev)
    x = torch.randint(-2**62, 2**62, (2 * n, 2), dtype=torch.int64, generator=g).to(dev)
    y = torch.empty((n, 2), dtype=torch.int64, device=dev)
    ctx.qc_prepare(n, a_bits, ahat, scratch)
    src = torch.empty(1 << 28, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    for _ in range(max(1, args.warmup)):
        ctx.qc_encode(n, ahat, x, y, scratch)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, b in ev:
        a.record()
        ctx.qc_encode(n, ahat, x, y, scratch)
        b.record()
    torch.cuda.synchronize()
    ms = float(np.median([a.elapsed_time(b) for a, b in ev]))
    print(json.dumps({"workload": "qc encode, one party", "n": n, "steps": args.steps, "ms_per_encode": round(ms, 4),
                      "compulsory_GBps": round(48 * n / (ms / 1e3) / 1e9, 1)}), flush=True)
    ctx.close()
    return 0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(threads: int):
    """The oracle (C, OpenMP; AES-NI, CLMUL) on the host cores, following the
    reference's timing protocol (silent/main.cu:96-101: warm-up, then repeated
    runs): SOT at n = 2^20 (t = 256, h = 13, the CPU-path workload), median of
    5 runs after 1 warm-up with `threads` threads plus one single-thread run;
    the Ferret instance n = 10,168,320 (k = 452,160, t = 1,270) the same way."""
    import numpy as np

    from oracle import pyoracle as po

    k0 = np.array(SOT_K0, dtype=np.uint64)
    k1 = np.array(SOT_K1, dtype=np.uint64)
    seed = np.array(SEED, dtype=np.uint64)

    def med(fn, reps):
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
        times.sort()
        return times[len(times) // 2]

    n, t, h = 1 << 20, 256, 13
    po.sot_run(1 << 14, 16, 11, k0, k1, seed, nthreads=threads)
    sot_all = med(lambda: po.sot_run(n, t, h, k0, k1, seed, nthreads=threads), 5)
    sot_one = med(lambda: po.sot_run(n, t, h, k0, k1, seed, nthreads=1), 1)
    fn, fk, ft, fh = 10168320, 452160, 1270, 13
    po.ferret_run(20000, 1000, 3, 13, seed, nthreads=threads)
    fer_all = med(lambda: po.ferret_run(fn, fk, ft, fh, seed, nthreads=threads), 5)
    fer_one = med(lambda: po.ferret_run(fn, fk, ft, fh, seed, nthreads=1), 1)
    return {
        "value": round(n / sot_all, 1),
        "unit": "COTs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle SOT n=2^20 t=256 h=13 both roles + QC, median of 5 runs after 1 warm-up on {threads} "
                  f"threads ({sot_all:.2f} s/run); one thread {sot_one:.2f} s/run",
        "one_thread": round(n / sot_one, 1),
        "ferret": {"n": fn, "value": round(fn / fer_all, 1), "one_thread": round(fn / fer_one, 1),
                   "unit": "COTs/s", "s_per_run": round(fer_all, 3)},
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "affinity_cores": len(os.sched_getaffinity(0)),
    }


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = [p.strip() for p in line.strip().strip("|").split("|")]
        if len(parts) >= 4 and parts[0] == kernel:
            try:
                return float(parts[3])
            except ValueError:
                return None
    return None


def tune_for_mix(ctx, mix):
    """Execution knobs for the batched mix from the project's autotuner: QC
    encode and GGM times measured on a grid of sizes and knob settings, the
    learned cost model trained on them, and the knobs it predicts fastest for
    the SOT size class that holds the most COTs (knobs are process-wide).
    Returns (knobs, tuning seconds)."""
    from cuot_amd import tune
    from cuot_amd.workloads import SOT

    t0 = time.perf_counter()
    sots = [wi for wi in mix if wi.kind == SOT]
    work = {}
    for wi in sots:
        work[wi.n.bit_length()] = work.get(wi.n.bit_length(), 0) + wi.n
    lead = max(work, key=work.get)
    n_lead = max(wi.n for wi in sots if wi.n.bit_length() == lead)
    tu = tune.Autotuner(ctx, reps=2)
    tu.collect_qc([1 << 18, 1 << 20, 1 << 21, n_lead])
    tu.train("qc", epochs=800)
    lead_inst = max((wi for wi in sots if wi.n.bit_length() == lead), key=lambda w: w.n)
    shapes = sorted({(wi.t, wi.h) for wi in sots})[-3:]
    tu.collect_ggm(shapes)
    tu.train("ggm", epochs=800)
    knobs = dict(tu.choose_qc(n_lead))
    knobs.update(tu.choose_ggm(lead_inst.t, lead_inst.h))
    tune.set_knobs(**knobs)
    return knobs, time.perf_counter() - t0


def batch_leg(torch, cuot_amd, ctx, steps):
    """64 instances (log-uniform 2^16..2^22, SOT/Ferret alternating): plain
    one-instance-at-a-time runs vs the batched engine replayed as a graph, both
    with the knobs the autotuner picks for the mix."""
    from cuot_amd.workloads import SOT, batch_mix

    mix = batch_mix(64)
    knobs, tune_s = tune_for_mix(ctx, mix)
    knobs = dict(knobs, tuning_s=round(tune_s, 1))
    insts = [cuot_amd.make_instance_struct(wi.kind, wi.n, wi.t, wi.h, wi.k, wi.seed, k0=SOT_K0, k1=SOT_K1)
             for wi in mix]
    total = sum(wi.n for wi in mix)
    dev = "cuda"
    # plain: per-instance workspaces and outputs, one stream, instances in order
    plain = []
    for wi, ins in zip(mix, insts):
        if wi.kind == SOT:
            cfg = cuot_amd.SotConfig(n=wi.n, t=wi.t, h=wi.h, reserved=0, k0=ins.k0, k1=ins.k1, seed=ins.seed)
            ws = torch.empty(ctx.sot_workspace_bytes(cfg), dtype=torch.uint8, device=dev)
            b = torch.empty(((wi.n + 31) // 32) * 4, dtype=torch.uint8, device=dev)
        else:
            cfg = cuot_amd.FerretConfig(n=wi.n, k=wi.k, t=wi.t, h=wi.h, ntasks=ins.ntasks, seed=ins.seed)
            ws = torch.empty(ctx.ferret_workspace_bytes(cfg), dtype=torch.uint8, device=dev)
            b = None
        o0 = torch.empty((wi.n, 2), dtype=torch.int64, device=dev)
        o1 = torch.empty((wi.n, 2), dtype=torch.int64, device=dev)
        d = torch.empty((1, 2), dtype=torch.int64, device=dev)
        plain.append((wi, cfg, ws, o0, o1, b, d))
    stream = torch.cuda.Stream()

    def run_plain():
        for wi, cfg, ws, o0, o1, b, d in plain:
            if wi.kind == SOT:
                ctx.sot_run(cfg, ws, o0, o1, b, d, stream)
            else:
                ctx.ferret_run(cfg, ws, o0, o1, d, stream)

    bt = cuot_amd.Batch(ctx, insts)

    def run_batch():
        bt.run(graph=True, stream=stream, overlap=True)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    plain_ms = timed(run_plain)
    batch_ms = timed(run_batch)
    # the same launch sequence without graph capture: the launch gaps the graph removes
    eager_ms = timed(lambda: bt.run(graph=False, stream=stream, overlap=True))
    # where the plain form spends its time: per-instance events, by kind and size
    evs = []
    for item in plain:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        wi, cfg, ws, o0, o1, b, d = item
        e0.record(stream)
        if wi.kind == SOT:
            ctx.sot_run(cfg, ws, o0, o1, b, d, stream)
        else:
            ctx.ferret_run(cfg, ws, o0, o1, d, stream)
        e1.record(stream)
        evs.append((wi, e0, e1))
    torch.cuda.synchronize()
    by_size = {}
    for wi, e0, e1 in evs:
        key = f"{'sot' if wi.kind == SOT else 'ferret'}_log2n_{int(wi.n).bit_length() - 1}"
        ent = by_size.setdefault(key, {"instances": 0, "cots": 0, "ms": 0.0})
        ent["instances"] += 1
        ent["cots"] += wi.n
        ent["ms"] = round(ent["ms"] + e0.elapsed_time(e1), 4)
    # the batched form per size class: each class alone as one batch (graph replay)
    classes = {}
    for i, wi in enumerate(mix):
        key = f"{'sot' if wi.kind == SOT else 'ferret'}_log2n_{int(wi.n).bit_length() - 1}"
        classes.setdefault(key, []).append(i)
    batched_by_size = {}
    for key, idx in sorted(classes.items()):
        sub = cuot_amd.Batch(ctx, [insts[i] for i in idx])
        ms = timed(lambda: sub.run(graph=True, stream=stream, overlap=True))
        cots = sum(mix[i].n for i in idx)
        batched_by_size[key] = {"instances": len(idx), "cots": cots, "ms": round(ms, 4),
                                "value": round(cots / (ms / 1e3), 1),
                                "plain_value": round(cots / (by_size[key]["ms"] / 1e3), 1)}
        sub.close()
    exact = True
    for i, (wi, cfg, ws, o0, o1, b, d) in enumerate(plain):
        exact &= bool(torch.equal(o0, bt.out0[i])) and bool(torch.equal(o1, bt.out1[i]))
        exact &= bool(torch.equal(d[0], bt.delta[i]))
        if wi.kind == SOT:
            exact &= bool(torch.equal(b[: (wi.n + 7) // 8], bt.bits[i].view(torch.uint8)[: (wi.n + 7) // 8]))
    out = {
        "config": {"workload": "batched mix", "instances": len(mix), "sizes": "n = floor(2^(16 + 6u))",
                   "kinds": "SOT / Ferret alternating", "total_cots": total},
        "plain": {"value": round(total / (plain_ms / 1e3), 1), "unit": "COTs/s", "ms": round(plain_ms, 3)},
        "batched": {"value": round(total / (batch_ms / 1e3), 1), "unit": "COTs/s", "ms": round(batch_ms, 3),
                    "form": "one launch sequence per stage over all instances, HIP graph replay, "
                            "preparation, sparse encodes, LPN and half the QC chunks on a second stream"},
        "speedup": round(plain_ms / batch_ms, 3),
        "eager_ms": round(eager_ms, 3),
        "graph_saves_ms": round(eager_ms - batch_ms, 3),
        "plain_by_size": dict(sorted(by_size.items())),
        "batched_by_size": batched_by_size,
        "knobs": knobs,
        "byte_exact_vs_plain": exact,
    }
    bt.close()
    del plain
    from cuot_amd import tune

    tune.set_knobs()
    return out


def serve_leg(torch, cuot_amd, ctx, count, rate, budget_gb, sweep_rates=(), sweep_count=300, arena=None):
    """Poisson arrivals, n = floor(2^(14 + 10u)), 50/50 SOT/Ferret: continuous
    batching under an HBM budget vs the FIFO plain loop, real-time replay; then
    the same at each offered load of `sweep_rates` with `sweep_count` requests."""
    import numpy as np

    from cuot_amd.workloads import serving_arrivals

    if arena is None:
        # one arena of the HBM budget for the whole leg (every server of the
        # main run and the sweep uses it) instead of one allocation per server
        torch.cuda.empty_cache()
        arena = torch.empty(int(budget_gb * (1 << 30)), dtype=torch.uint8, device="cuda")
    res = serve_run(torch, cuot_amd, ctx, count, rate, budget_gb, arena)
    sweep = []
    for r in sweep_rates:
        row = serve_run(torch, cuot_amd, ctx, sweep_count, r, budget_gb, arena)
        sweep.append({"rate_hz": r, "requests": sweep_count,
                      "offered_cots_s": row["config"]["offered_cots_s"],
                      "served": {k: row["served"][k] for k in ("p50_ms", "p99_ms", "throughput_cots_s",
                                                               "relation_ok")},
                      "fifo_plain": {k: row["fifo_plain"][k] for k in ("p50_ms", "p99_ms", "throughput_cots_s",
                                                                       "relation_ok")}})
    if sweep:
        res["sweep"] = sweep
    del arena
    torch.cuda.empty_cache()
    return res


FERRET_CLIENTS = 8


def serve_run(torch, cuot_amd, ctx, count, rate, budget_gb, arena=None):
    import numpy as np

    from cuot_amd.workloads import serving_arrivals

    arr = serving_arrivals(rate, count)
    # Ferret requests come from FERRET_CLIENTS resident clients (round robin), each an
    # iterating session of the emp-ot ferret_b13 size whose Delta and bootstrap stay in HBM
    from cuot_amd.workloads import FERRET

    insts, nf = [], 0
    for _, wi in arr:
        if wi.kind == FERRET:
            nf += 1
            insts.append(cuot_amd.make_instance_struct(FERRET, wi.n, 0, 0, 0, (0, 0),
                                                       client=1 + nf % FERRET_CLIENTS))
        else:
            insts.append(cuot_amd.make_instance_struct(wi.kind, wi.n, wi.t, wi.h, wi.k, wi.seed, k0=SOT_K0,
                                                       k1=SOT_K1))
    fsess = cuot_amd.FerretConfig(n=10168320, k=452160, t=1270, h=13, ntasks=1, seed=cuot_amd.to_block(SEED))
    times = [a for a, _ in arr]
    total = sum(wi.n for _, wi in arr)
    budget = int(budget_gb * (1 << 30))
    span = max(times) - min(times) if len(times) > 1 else 1.0
    res = {"config": {"workload": "serving", "requests": count, "rate_hz": rate,
                      "offered_cots_s": round(total / span, 1) if span > 0 else None,
                      "sizes": "n = floor(2^(14 + 10u))", "kinds": "SOT / Ferret 50/50", "hbm_budget_bytes": budget,
                      "ferret": f"{FERRET_CLIENTS} resident clients (iterating sessions n=10,168,320 k=452,160 "
                                "t=1,270: Delta, bootstrap and the last extend's outputs in HBM between requests)",
                      "max_batch": 64, "max_batch_cots": 1 << 22, "streams": 2}}
    torch.cuda.empty_cache()  # the earlier legs' cached blocks go back to the device
    res["config"]["arena"] = "caller-owned, shared by the servers of the leg" if arena is not None else "server-owned"
    for name, plain in (("served", False), ("fifo_plain", True)):
        torch.cuda.synchronize()
        # batches are capped at 2^22 COTs so that a large request does not hold
        # small ones behind it for a whole 2^24-COT encode
        lat, wall, nb, fails = cuot_amd.serve_replay(ctx, insts, times, hbm_budget=budget, plain=plain,
                                                     max_batch=64, streams=2, max_batch_cots=1 << 22,
                                                     check_relation=True, arena=arena,
                                                     ferret_clients=FERRET_CLIENTS, ferret_session=fsess)
        res[name] = {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                     "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                     "mean_ms": round(float(lat.mean()) * 1e3, 3),
                     "throughput_cots_s": round(total / wall, 1), "wall_s": round(wall, 3), "batches": nb,
                     "relation_ok": fails == 0, "relation_checked_results": count}
    return res


def relation_ok(torch, v, w, b, d, n):
    sh = torch.arange(8, device=b.device, dtype=torch.uint8)
    bits = ((b[: (n + 7) // 8].unsqueeze(1) >> sh) & 1).reshape(-1)[:n].to(torch.bool)
    return bool(torch.equal(torch.where(bits.unsqueeze(1), v ^ d, v), w))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        return 2
    if args.qc_only:
        return qc_only(args)
    import torch

    import cuot_amd

    ctx = cuot_amd.Context(0)
    stream = torch.cuda.current_stream()
    B = cuot_amd.to_block


    # ---- SOT headline -----------------------------------------------------------
    n = 1 << args.sot_n_log
    h = 15 if args.sot_n_log >= 24 else max(1, round((args.sot_n_log + 1 + 5) / 2))
    t = (2 * n + (1 << h) - 1) >> h
    cfg = cuot_amd.SotConfig(n=n, t=t, h=h, reserved=0, k0=B(SOT_K0), k1=B(SOT_K1), seed=B(SEED))
    ws = torch.empty(ctx.sot_workspace_bytes(cfg), dtype=torch.uint8, device="cuda")
    v = torch.empty((n, 2), dtype=torch.int64, device="cuda")
    w = torch.empty((n, 2), dtype=torch.int64, device="cuda")
    b = torch.empty(((n + 31) // 32) * 4, dtype=torch.uint8, device="cuda")
    d = torch.empty((1, 2), dtype=torch.int64, device="cuda")
    stages = cuot_amd.SOT_STAGES

    def sot_staged(events):
        for i, (_, bit) in enumerate(stages):
            events[i].record(stream)
            ctx.sot_run_stages(cfg, ws, v, w, b, d, bit, stream)
        events[-1].record(stream)

    # timed region 1 (value): whole instances; independent stages of the two
    # roles overlap on two streams (staged and sequential with --sequential)
    def sot_whole():
        if args.sequential:
            ctx.sot_run_stages(cfg, ws, v, w, b, d, cuot_amd.SOT_ALL, stream)
        else:
            ctx.sot_run(cfg, ws, v, w, b, d, stream)

    for _ in range(args.warmup):
        sot_whole()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sot_whole()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps

    # timed region 2 (kernels): K staged, sequential instances with HIP events
    # between stages and around every kernel launch on its stream, so that each
    # kernel's duration is its own (no other work shares the GPU)
    lib = ctx.lib
    import ctypes as C

    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(len(stages) + 1)] for _ in range(args.steps)]
    cuot_amd.check(lib.cuot_timing_begin(ctx.handle), "cuot_timing_begin")
    for s_ in range(args.steps):
        sot_staged(evs[s_])
    recs = (cuot_amd._lib.LaunchTime * 8192)()
    nrec = C.c_int()
    cuot_amd.check(lib.cuot_timing_end(ctx.handle, recs, 8192, C.byref(nrec)), "cuot_timing_end")
    torch.cuda.synchronize()
    stage_ms = {name: round(sum(evs[s_][i].elapsed_time(evs[s_][i + 1]) for s_ in range(args.steps)) / args.steps, 4)
                for i, (name, _) in enumerate(stages)}
    kernels = {}
    for r in recs[: nrec.value]:
        k = kernels.setdefault(r.kernel.decode(), {"launches": 0, "ms": 0.0, "bytes": 0.0})
        k["launches"] += 1
        k["ms"] += r.ms
        k["bytes"] += r.bytes
    per_kernel = {name: {"launches_per_step": k["launches"] / args.steps,
                         "ms_per_step": round(k["ms"] / args.steps, 4),
                         "avg_launch_ms": round(k["ms"] / k["launches"], 4),
                         "algorithmic_GBps": round(k["bytes"] / (k["ms"] / 1e3) / 1e9, 1) if k["ms"] > 0 else None}
                  for name, k in kernels.items()}
    dom_name, dom = max(kernels.items(), key=lambda kv: kv[1]["ms"])
    sot_whole()
    torch.cuda.synchronize()
    ok = relation_ok(torch, v, w, b, d, n)

    qc_ms = (stage_ms["qc_sender"] + stage_ms["qc_recver"]) / 2
    qc_bytes = 48 * n  # compulsory per encode: read 2n blocks, write n blocks
    dom_achieved = dom["bytes"] / (dom["ms"] / 1e3) / 1e9
    traffic = measured_traffic(dom_name)
    ggm_ms = (stage_ms["ggm_sender"] + stage_ms["ggm_recver"]) / 2
    aes_per_party = 2 * t * ((1 << h) - 1)

    # ---- Ferret -----------------------------------------------------------------
    ferret = None
    if not args.no_ferret:
        fn, fk, ft, fh = 10168320, 452160, 1270, 13
        fcfg = cuot_amd.FerretConfig(n=fn, k=fk, t=ft, h=fh, ntasks=1, seed=B(SEED))
        fws = torch.empty(ctx.ferret_workspace_bytes(fcfg), dtype=torch.uint8, device="cuda")
        fs = torch.empty((fn, 2), dtype=torch.int64, device="cuda")
        fr = torch.empty((fn, 2), dtype=torch.int64, device="cuda")
        fd = torch.empty((1, 2), dtype=torch.int64, device="cuda")
        fst = cuot_amd.FERRET_STAGES
        for _ in range(args.warmup):
            ctx.ferret_run(fcfg, fws, fs, fr, fd, stream)
        fev = [[torch.cuda.Event(enable_timing=True) for _ in range(len(fst) + 1)] for _ in range(args.steps)]
        torch.cuda.synchronize()
        f0 = time.perf_counter()
        for s in range(args.steps):
            for i, (_, bit) in enumerate(fst):
                fev[s][i].record(stream)
                ctx.ferret_run_stages(fcfg, fws, fs, fr, fd, bit, stream)
            fev[s][-1].record(stream)
        torch.cuda.synchronize()
        f1 = time.perf_counter()
        fms = (f1 - f0) * 1e3 / args.steps
        lsb = (fr[:, 0] & 1).to(torch.bool)
        fok = bool(torch.equal(torch.where(lsb.unsqueeze(1), fs ^ fd, fs), fr))
        ferret = {
            "config": {"workload": "ferret", "n": fn, "k": fk, "t": ft, "leaves_per_tree": 1 << fh, "d": 10},
            "value": round(fn / (fms / 1e3), 1),
            "unit": "COTs/s",
            "ms_per_step": round(fms, 4),
            "stage_ms": {name: round(sum(fev[s][i].elapsed_time(fev[s][i + 1]) for s in range(args.steps))
                                     / args.steps, 4) for i, (name, _) in enumerate(fst)},
            "relation_ok": fok,
        }

    # iterating Ferret session (FerretCOT extend loop) at the C2 size
    session = None
    if not args.no_ferret:
        from cuot_amd.ferret import FerretCOT

        fc = FerretCOT(ctx, n=10168320, k=452160, t=1270, h=13, seed=SEED)
        fc.extend_f2k()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            fc.extend_f2k()
        e1.record(stream)
        torch.cuda.synchronize()
        sms = e0.elapsed_time(e1) / args.steps
        lsb = (fc.out_r[:, 0] & 1).to(torch.bool)
        sok = bool(torch.equal(torch.where(lsb.unsqueeze(1), fc.out_s ^ fc.Delta[0], fc.out_s), fc.out_r))
        session = {"config": {"workload": "ferret session (extend loop)", "n": 10168320, "k": 452160, "t": 1270,
                              "bootstrap_M": fc.M, "ot_limit": fc.ot_limit},
                   "ms_per_extend": round(sms, 4), "value": round(fc.ot_limit / (sms / 1e3), 1),
                   "unit": "usable COTs/s", "relation_ok": sok}
        # malicious mode: the consistency check after every MPCOT
        fc.set_malicious(True)
        fc.extend_f2k()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            fc.extend_f2k()
        e1.record(stream)
        torch.cuda.synchronize()
        mms = e0.elapsed_time(e1) / args.steps
        cok = fc.check_result()[0]
        fc.set_malicious(False)
        session["malicious"] = {"ms_per_extend": round(mms, 4), "value": round(fc.ot_limit / (mms / 1e3), 1),
                                "unit": "usable COTs/s", "check_ok": cok,
                                "check_ms": round(mms - sms, 4)}
        # online derandomisation of one extend's usable COTs (both parties)
        import ctypes as C

        num = fc.ot_limit
        g = torch.Generator().manual_seed(9)
        bits = torch.randint(0, 2, (num,), dtype=torch.uint8, generator=g).to(stream.device)
        m0 = torch.randint(-2**62, 2**62, (num, 2), dtype=torch.int64, generator=g).to(stream.device)
        m1 = torch.randint(-2**62, 2**62, (num, 2), dtype=torch.int64, generator=g).to(stream.device)
        bo = torch.empty(num, dtype=torch.uint8, device=stream.device)
        pad = torch.empty((2 * num, 2), dtype=torch.int64, device=stream.device)
        outr = torch.empty((num, 2), dtype=torch.int64, device=stream.device)
        lib = ctx.lib
        sblk = B((0x0123456789ABCDEF, 0x0FEDCBA987654321))
        ks, mr = fc.out_s[:num], fc.out_r[:num]
        dev_times = {}
        for name, call in (
                ("cot_online", lambda: lib.cuot_cot_online(ctx.handle, ks.data_ptr(), mr.data_ptr(), bits.data_ptr(),
                                                           fc.Delta.data_ptr(), bo.data_ptr(), num,
                                                           stream.cuda_stream)),
                ("cot_to_ot", lambda: lib.cuot_cot_to_ot(ctx.handle, ks.data_ptr(), mr.data_ptr(), fc.Delta.data_ptr(),
                                                         sblk, 0, bits.data_ptr(), m0.data_ptr(), m1.data_ptr(),
                                                         pad.data_ptr(), outr.data_ptr(), num, stream.cuda_stream))):
            cuot_amd.check(call(), name)
            torch.cuda.synchronize()
            e0.record(stream)
            for _ in range(args.steps):
                cuot_amd.check(call(), name)
            e1.record(stream)
            torch.cuda.synchronize()
            dev_times[name] = e0.elapsed_time(e1) / args.steps
            if name == "cot_online" and (1 + args.steps) % 2 == 0:
                cuot_amd.check(call(), name)  # each call toggles K by bo*Delta: end on the chosen COTs
        oot = bool(torch.equal(outr, torch.where(bits.bool().unsqueeze(1), m1, m0)))
        session["derandomisation"] = {
            "n": num, "cot_online_ms": round(dev_times["cot_online"], 4),
            "cot_to_ot_ms": round(dev_times["cot_to_ot"], 4),
            "ots_per_s": round(num / ((dev_times["cot_online"] + dev_times["cot_to_ot"]) / 1e3), 1),
            "form": "chosen-bit COT then chosen-message OT, both parties (3 AES + 1 key schedule per OT)",
            "ot_ok": oot}
        del bits, m0, m1, bo, pad, outr
        fc.close()
    # release the headline buffers before the other legs
    del ws, v, w, b
    torch.cuda.empty_cache()
    serving = None
    batched = None
    rates = [float(r) for r in args.serve_sweep.split(",") if r.strip()]
    if args.batch_first and not args.no_batch:
        batched = batch_leg(torch, cuot_amd, ctx, max(1, args.steps))
        torch.cuda.empty_cache()
    if not args.no_serve:
        serving = serve_leg(torch, cuot_amd, ctx, args.serve_requests, args.serve_rate, args.serve_budget_gb, rates,
                            args.sweep_requests)
        serving["leg_order"] = "batched leg first" if (args.batch_first and not args.no_batch) else "serving first"
    if not args.batch_first and not args.no_batch:
        batched = batch_leg(torch, cuot_amd, ctx, max(1, args.steps))
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        # every core this process may run on, within the CPU share the host grants one GPU (its
        # OMP_NUM_THREADS) when it sets one
        share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        cores = len(os.sched_getaffinity(0))
        threads = args.cpu_threads or (min(cores, share) if share > 0 else cores)
        cpu = cpu_baseline(threads)

    value = n / (ms_per_step / 1e3)
    out = {
        "metric": "COTs/sec at 2^24 SOT & Ferret, batched vs plain; p50/p99 request latency",
        "value": round(value, 1),
        "unit": "COTs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (all inputs derived from one 128-bit seed by AES-CTR)",
        "config": {"workload": "plain SOT", "n": n, "t": t, "h": h, "leaves": 2 * n, "qc_scaler": 2,
                   "form": ("plain single instance, both roles, stages in sequence" if args.sequential else
                         "plain single instance, both roles; independent stages overlap on two streams")},
        "relation_ok": ok,
        "stage_ms": stage_ms,
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(dom_achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom_achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "avg_launch_ms": round(dom["ms"] / dom["launches"], 4),
            "algorithmic_bytes_per_launch": round(dom["bytes"] / dom["launches"]),
        },
        "kernels": per_kernel,
        "qc_encode": {"ms_per_party": round(qc_ms, 4), "compulsory_bytes": qc_bytes,
                      "compulsory_GBps": round(qc_bytes / (qc_ms / 1e3) / 1e9, 1)},
        "ggm": {"aes_per_launch": aes_per_party, "ms_per_launch": round(ggm_ms, 4),
                "aes_per_s": round(aes_per_party / (ggm_ms / 1e3), 1)},
        "ferret": ferret,
        "ferret_session": session,
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
