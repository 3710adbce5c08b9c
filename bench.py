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
    ap.add_argument("--no-small-codes", action="store_true", help="skip the small-code QC comparison leg")
    ap.add_argument("--no-aes", action="store_true", help="skip the AES cipher comparison leg")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the two rocprofv3 --pmc passes that measure HBM traffic and VALU work per kernel")
    ap.add_argument("--counter-run", action="store_true",
                    help="(used under rocprofv3 by the counter passes) one QC encode at 2^24, one Ferret run, "
                         "AES-128 ECB of 2^24 blocks with both ciphers and one matrix-core QC encode; prints the "
                         "kernels' algorithmic bytes as one JSON line")
    return ap.parse_args()


# ---- hardware counters -------------------------------------------------------------
VALU_LANE_OPS_PER_S = 256 * 64 * 2.4e9  # one wave64 VALU instruction per CU per cycle (issue model)
COUNTER_PASSES = (
    ["FETCH_SIZE", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES",
     "SQ_WAVE_CYCLES"],
    ["WRITE_SIZE", "SQ_INSTS_VALU_MFMA_I8", "SQ_INSTS_VALU_MFMA_MOPS_I8", "SQ_VALU_MFMA_BUSY_CYCLES"],
)
CALIB_BYTES = 1 << 28  # the device copy that calibrates FETCH_SIZE / WRITE_SIZE


def progress(msg: str):
    """One line on stderr as each leg finishes (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_base(name: str) -> tuple[str, str]:
    """(base name, name with template arguments) of a rocprofv3 kernel name."""
    s = name.replace("(anonymous namespace)::", "").strip()
    if s.startswith("void "):
        s = s[5:]
    full = s.split("(")[0].strip()
    return full.split("<")[0], full


def counter_run(args):
    """The counter passes' workload (runs under rocprofv3): a 256 MiB device
    copy first (a kernel of known traffic), then the kernels the bench line
    reports counters for.  Prints {kernel: {launches, bytes}} from the
    library's launch recorder (algorithmic bytes)."""
    import ctypes as C

    import torch

    import cuot_amd
    from cuot_amd import tune

    ctx = cuot_amd.Context(0)
    lib = ctx.lib
    dev = "cuda"
    src = torch.empty(CALIB_BYTES, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    acc = {}

    def timed(fn):
        cuot_amd.check(lib.cuot_timing_begin(ctx.handle), "cuot_timing_begin")
        fn()
        recs = (cuot_amd._lib.LaunchTime * 4096)()
        nrec = C.c_int()
        cuot_amd.check(lib.cuot_timing_end(ctx.handle, recs, 4096, C.byref(nrec)), "cuot_timing_end")
        for r in recs[: nrec.value]:
            k = acc.setdefault(r.kernel.decode().split("<")[0], {"launches": 0, "bytes": 0.0})
            k["launches"] += 1
            k["bytes"] += r.bytes

    g = torch.Generator().manual_seed(1)
    n = 1 << 24
    a_bits = torch.randint(0, 256, ((n + 7) // 8 + 16,), dtype=torch.uint8, generator=g).to(dev)
    ahat = torch.empty(ctx.qc_ahat_len(n), dtype=torch.int32, device=dev)
    scratch = torch.empty(ctx.qc_scratch_bytes(n), dtype=torch.uint8, device=dev)
    x = torch.randint(-2**62, 2**62, (2 * n, 2), dtype=torch.int64, generator=g).to(dev)
    y = torch.empty((n, 2), dtype=torch.int64, device=dev)
    ctx.qc_prepare(n, a_bits, ahat, scratch)
    timed(lambda: ctx.qc_encode(n, ahat, x, y, scratch))
    del a_bits, ahat, scratch, x
    B = cuot_amd.to_block
    fn, fk, ft, fh = 10168320, 452160, 1270, 13
    fcfg = cuot_amd.FerretConfig(n=fn, k=fk, t=ft, h=fh, ntasks=1, seed=B(SEED))
    fws = torch.empty(ctx.ferret_workspace_bytes(fcfg), dtype=torch.uint8, device=dev)
    fs = torch.empty((fn, 2), dtype=torch.int64, device=dev)
    fr = torch.empty((fn, 2), dtype=torch.int64, device=dev)
    fd = torch.empty((1, 2), dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    timed(lambda: ctx.ferret_run(fcfg, fws, fs, fr, fd, stream))
    del fws, fs, fr
    key = (0x0123456789ABCDEF, 0x0FEDCBA987654321)
    for impl in (1, 2):
        tune.set_knobs(aes_impl=impl)
        timed(lambda: ctx.aes_ecb(key, y, y))
    tune.set_knobs()
    del y
    m = 32768
    a_bits = torch.randint(0, 256, ((m + 31) // 32 * 4,), dtype=torch.uint8, generator=g).to(dev)
    ahat = torch.empty(ctx.qc_ahat_len(m), dtype=torch.int32, device=dev)
    scratch = torch.empty(ctx.qc_scratch_bytes(m), dtype=torch.uint8, device=dev)
    x = torch.randint(-2**62, 2**62, (2 * m, 2), dtype=torch.int64, generator=g).to(dev)
    y = torch.empty((m, 2), dtype=torch.int64, device=dev)
    ctx.qc_prepare(m, a_bits, ahat, scratch)
    tune.set_knobs(qc_mfma_max=1 << 16)
    timed(lambda: ctx.qc_encode(m, ahat, x, y, scratch))
    tune.set_knobs()
    torch.cuda.synchronize()
    print(json.dumps({"counter_run": acc}), flush=True)
    ctx.close()
    return 0


def collect_counters():
    """Two rocprofv3 --pmc passes over the counter-run workload, started
    before this process touches the GPU.  Per kernel (averaged over its
    dispatches): HBM bytes (FETCH_SIZE, WRITE_SIZE in KiB, scaled by what the
    counters read for the 256 MiB copy), VALU and LDS instructions, LDS bank
    conflicts, the profiled duration; plus the algorithmic bytes per launch
    the library's recorder notes.  Returns (table, note)."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not on PATH"
    disp = {}  # dispatch id -> {"name", "ns", counters}
    algo = None
    for pi, ctrs in enumerate(COUNTER_PASSES):
        d = tempfile.mkdtemp(prefix="cuot_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "180", exe, "--pmc", *ctrs, "--output-format", "csv", "-d", d, "-o", "run",
               "--", sys.executable, os.path.join(ROOT, "bench.py"), "--counter-run"]
        env = dict(os.environ, TMPDIR="/tmp", PYTHONPATH=ROOT)
        progress(f"counter pass {pi + 1}/{len(COUNTER_PASSES)}: {' '.join(ctrs)}")
        r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True)
        for line in r.stdout.splitlines():
            if line.startswith('{"counter_run"'):
                algo = json.loads(line)["counter_run"]
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            return None, f"counter pass {pi + 1} wrote no counter file (exit {r.returncode})"
        for row in csv.DictReader(open(files[0])):
            key = (pi, row["Dispatch_Id"])
            e = disp.setdefault(key, {"name": row["Kernel_Name"],
                                      "ns": float(row["End_Timestamp"]) - float(row["Start_Timestamp"])})
            e[row["Counter_Name"]] = e.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
        shutil.rmtree(d, ignore_errors=True)

    def calib(counter, pi):
        vals = [e[counter] for (p, _), e in disp.items() if p == pi and "copyBuffer" in e["name"] and counter in e]
        return CALIB_BYTES / (max(vals) * 1024.0) if vals and max(vals) > 0 else None

    fscale, wscale = calib("FETCH_SIZE", 0), calib("WRITE_SIZE", 1)
    table = {}
    for (pi, _), e in disp.items():
        if "copyBuffer" in e["name"] or "rocclr" in e["name"]:
            continue
        base, full = kernel_base(e["name"])
        k = table.setdefault(full, {"base": base, "n0": 0, "n1": 0, "ns": 0.0, "ns1": 0.0, "fetch": 0.0, "write": 0.0,
                                    "valu": 0.0, "lds": 0.0, "conflict": 0.0, "mfma": 0.0, "mops": 0.0,
                                    "mbusy": 0.0})
        if pi == 0:
            k["n0"] += 1
            k["ns"] += e["ns"]
            k["fetch"] += e.get("FETCH_SIZE", 0.0)
            k["valu"] += e.get("SQ_INSTS_VALU", 0.0)
            k["lds"] += e.get("SQ_INSTS_LDS", 0.0)
            k["conflict"] += e.get("SQ_LDS_BANK_CONFLICT", 0.0)
        else:
            k["n1"] += 1
            k["ns1"] += e["ns"]
            k["write"] += e.get("WRITE_SIZE", 0.0)
            k["mfma"] += e.get("SQ_INSTS_VALU_MFMA_I8", 0.0)
            k["mops"] += e.get("SQ_INSTS_VALU_MFMA_MOPS_I8", 0.0)
            k["mbusy"] += e.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0)
    out = {}
    for full, k in table.items():
        if not k["n0"]:
            continue
        us = k["ns"] / k["n0"] / 1e3
        row = {"launches": k["n0"], "us_per_launch": round(us, 1),
               "valu_instr_per_launch": round(k["valu"] / k["n0"]),
               "valu_util": round(k["valu"] / k["n0"] * 64 / (us * 1e-6 * VALU_LANE_OPS_PER_S), 3) if us > 0 else None,
               "lds_instr_per_launch": round(k["lds"] / k["n0"]),
               "lds_bank_conflict_cycles_per_launch": round(k["conflict"] / k["n0"])}
        if k["n1"] and k["mfma"] > 0:
            ops = k["mops"] / k["n1"] * 512  # the counter counts int8 multiply-add ops / 512
            us1 = k["ns1"] / k["n1"] / 1e3
            row["mfma_i8_instr_per_launch"] = round(k["mfma"] / k["n1"])
            row["mfma_i8_ops_per_launch"] = round(ops)
            row["mfma_busy_cycles_per_launch"] = round(k["mbusy"] / k["n1"])
            row["mfma_i8_TOPs_profiled"] = round(ops / (us1 * 1e-6) / 1e12, 1) if us1 > 0 else None
        if fscale and wscale and k["n1"]:
            tb = k["fetch"] / k["n0"] * 1024 * fscale + k["write"] / k["n1"] * 1024 * wscale
            row["hbm_bytes_per_launch"] = round(tb)
            a = (algo or {}).get(k["base"])
            if a and a["launches"] and a["bytes"] > 0:
                row["algorithmic_bytes_per_launch"] = round(a["bytes"] / a["launches"])
                row["traffic_over_algorithmic"] = round(tb / (a["bytes"] / a["launches"]), 3)
        out[full] = row
    note = (f"rocprofv3 --pmc, two passes over bench.py --counter-run; FETCH_SIZE x {fscale:.3f}, "
            f"WRITE_SIZE x {wscale:.3f} (KiB -> bytes, from the 256 MiB copy); VALU util = VALU instructions x 64 "
            f"/ (profiled duration x {VALU_LANE_OPS_PER_S:.3g} lane-ops/s)") if fscale and wscale else \
        "copy kernel not found: no byte calibration"
    return out, note


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


QC_ENCODE_KERNELS = ("qc_pack_x", "transpose_fwd", "qc_conv_rows", "transpose_inv", "qc_fft_pass", "qc_unpack_x",
                     "qc_rows_fold", "qc_output_planes", "qc_output")


def qc_encode_bound(counters, qc_ms):
    """Per-kernel bounds of one QC encode at n = 2^24 (the counter run's
    encode): each kernel's time at full VALU issue and at the HBM peak, the
    larger of the two, summed over its launches; against the profiled and the
    live per-party times."""
    if not counters:
        return None
    bound = prof = 0.0
    parts = {}
    for name, v in counters.items():
        if kernel_base(name)[0] not in QC_ENCODE_KERNELS or "hbm_bytes_per_launch" not in v:
            continue
        valu_ms = v["valu_instr_per_launch"] * 64 / VALU_LANE_OPS_PER_S * 1e3
        hbm_ms = v["hbm_bytes_per_launch"] / (HBM_PEAK_GBS * 1e9) * 1e3
        b = max(valu_ms, hbm_ms) * v["launches"]
        bound += b
        prof += v["us_per_launch"] * v["launches"] / 1e3
        parts[name] = {"bound_ms": round(b, 4), "by": "valu" if valu_ms > hbm_ms else "hbm",
                       "profiled_ms": round(v["us_per_launch"] * v["launches"] / 1e3, 4)}
    return {"sum_of_kernel_bounds_ms": round(bound, 4), "profiled_ms": round(prof, 4),
            "live_ms_per_party": round(qc_ms, 4), "bound_over_live": round(bound / qc_ms, 4) if qc_ms else None,
            "kernels": parts}


def timed_kernels(cuot_amd, ctx, fn, reps=3):
    """Minimum over `reps` calls of fn of the summed per-launch times (HIP
    events around every launch), with the per-kernel split of that call."""
    import ctypes as C

    lib = ctx.lib
    best = None
    for _ in range(reps):
        cuot_amd.check(lib.cuot_timing_begin(ctx.handle), "cuot_timing_begin")
        fn()
        recs = (cuot_amd._lib.LaunchTime * 1024)()
        nrec = C.c_int()
        cuot_amd.check(lib.cuot_timing_end(ctx.handle, recs, 1024, C.byref(nrec)), "cuot_timing_end")
        rows = [(recs[i].kernel.decode(), recs[i].ms) for i in range(nrec.value)]
        tot = sum(ms for _, ms in rows)
        if best is None or tot < best[0]:
            best = (tot, rows)
    split = {}
    for k, ms in best[1]:
        split[k] = round(split.get(k, 0.0) + ms, 4)
    return best[0], split


MFMA_I8_PEAK_TOPS = 5000.0  # MI355X dense int8


def small_code_leg(torch, cuot_amd, ctx, counters):
    """QC encode of small codes: the additive FFT against the int8 matrix-core
    product on one prepared code per size (both exact; outputs compared), the
    matrix cores' int8 rate against the dense peak, and the crossover the
    library uses."""
    from cuot_amd import tune

    g = torch.Generator().manual_seed(5)
    rows = []
    for n in (4096, 16384, 32768, 49152, 65536):
        a_bits = torch.randint(0, 256, ((n + 31) // 32 * 4,), dtype=torch.uint8, generator=g).cuda()
        ahat = torch.empty(ctx.qc_ahat_len(n), dtype=torch.int32, device="cuda")
        scratch = torch.empty(ctx.qc_scratch_bytes(n), dtype=torch.uint8, device="cuda")
        x = torch.randint(-2**62, 2**62, (2 * n, 2), dtype=torch.int64, generator=g).cuda()
        y = torch.empty((n, 2), dtype=torch.int64, device="cuda")
        ctx.qc_prepare(n, a_bits, ahat, scratch)
        res, outs = {}, {}
        for name, lim in (("fft", -1), ("mfma", 1 << 16)):
            tune.set_knobs(qc_mfma_max=lim)
            ctx.qc_encode(n, ahat, x, y, scratch)
            torch.cuda.synchronize()
            outs[name] = y.clone()
            res[name] = timed_kernels(cuot_amd, ctx, lambda: ctx.qc_encode(n, ahat, x, y, scratch))
        tune.set_knobs()
        ops = 2.0 * 128 * n * n
        enc = res["mfma"][1].get("qc_mf_encode", res["mfma"][0])
        rows.append({"n": n, "fft_us": round(res["fft"][0] * 1e3, 1), "mfma_us": round(res["mfma"][0] * 1e3, 1),
                     "mfma_split_us": {k: round(v * 1e3, 1) for k, v in res["mfma"][1].items()},
                     "mfma_int8_TOPs": round(ops / (enc * 1e-3) / 1e12, 1),
                     "mfma_frac_of_dense_peak": round(ops / (enc * 1e-3) / 1e12 / MFMA_I8_PEAK_TOPS, 4),
                     "faster": "mfma" if res["mfma"][0] < res["fft"][0] else "fft",
                     "same_output": bool(torch.equal(outs["fft"], outs["mfma"]))})
        del a_bits, ahat, scratch, x, y
    mf = next((v for k, v in (counters or {}).items() if k.startswith("qc_mf_encode")), None)
    return {"workload": "QC encode of one party, small n, both encoders on one prepared code",
            "sizes": rows, "builtin_crossover_n": int(ctx.lib.cuot_qc_mfma_crossover()),
            "crossover_rule": "matrix cores for 17 <= n <= crossover (knob qc_mfma_max), additive FFT above",
            "roofline_mfma": {"bound": "mfma", "kernel": "qc_mf_encode", "unit": "TOP/s", "peak": MFMA_I8_PEAK_TOPS,
                              "achieved": rows[-1]["mfma_int8_TOPs"], "frac": rows[-1]["mfma_frac_of_dense_peak"],
                              "n": rows[-1]["n"]},
            "counters_n32768": mf,
            "mfma_counter_check": None if not mf or "mfma_i8_ops_per_launch" not in mf else {
                "n": 32768, "algorithmic_int8_ops": 2 * 128 * 32768 * 32768,
                "counted_int8_ops": mf["mfma_i8_ops_per_launch"],
                "counted_TOPs_at_profiled_time": mf.get("mfma_i8_TOPs_profiled"),
                "frac_of_dense_peak": round(mf["mfma_i8_TOPs_profiled"] / MFMA_I8_PEAK_TOPS, 4)
                if mf.get("mfma_i8_TOPs_profiled") else None}}


def aes_leg(torch, cuot_amd, ctx, counters):
    """AES-128 ECB of 2^24 blocks with the LDS T-table and the bitsliced
    cipher (same output): AES blocks/s and VALU lane-ops per block (counter
    pass), and which one the library uses."""
    from cuot_amd import tune

    n = 1 << 24
    g = torch.Generator().manual_seed(6)
    x = torch.randint(-2**62, 2**62, (n, 2), dtype=torch.int64, generator=g).cuda()
    y = torch.empty_like(x)
    key = (0x0123456789ABCDEF, 0x0FEDCBA987654321)
    out, outs = {}, {}
    ciphers = (("ttable", 1, "aes_stream_kernel<false>"), ("bitsliced", 2, "aes_bs_stream_kernel<false>"))
    for name, impl, kname in ciphers:
        tune.set_knobs(aes_impl=impl)
        ctx.aes_ecb(key, x, y)
        torch.cuda.synchronize()
        outs[name] = y.clone()
        ms, _ = timed_kernels(cuot_amd, ctx, lambda: ctx.aes_ecb(key, x, y))
        row = (counters or {}).get(kname)
        out[name] = {"ms": round(ms, 4), "aes_per_s": round(n / (ms * 1e-3), 1),
                     "valu_lane_ops_per_aes": round(row["valu_instr_per_launch"] * 64 / n, 1) if row else None,
                     "lds_lane_ops_per_aes": round(row["lds_instr_per_launch"] * 64 / n, 1) if row else None,
                     "valu_util": row.get("valu_util") if row else None}
    tune.set_knobs()
    out["same_output"] = bool(torch.equal(outs["ttable"], outs["bitsliced"]))
    out["faster"] = max(("ttable", "bitsliced"), key=lambda k: out[k]["aes_per_s"])
    out["library_rule"] = "ttable"
    out["blocks"] = n
    del x, y
    # the GGM sender of the SOT workload (t = 1024 trees of depth 15) with each cipher
    t, h = 1024, 15
    B = cuot_amd.to_block
    p = cuot_amd.GgmParams(t=t, h=h, mode=0, n_out=t << h, k0=B(SOT_K0), k1=B(SOT_K1))
    roots = torch.randint(-2**62, 2**62, (t, 2), dtype=torch.int64, generator=g).cuda()
    leaves = [torch.empty((t << h, 2), dtype=torch.int64, device="cuda") for _ in range(2)]
    sums = [torch.empty((t * h * 2, 2), dtype=torch.int64, device="cuda") for _ in range(2)]
    work = torch.empty(ctx.ggm_sender_bs_work_bytes(p), dtype=torch.uint8, device="cuda")
    calls = {"ttable": lambda: ctx.ggm_sender(p, roots, leaves[0], sums[0]),
             "bitsliced": lambda: ctx.ggm_sender_bs(p, roots, leaves[1], sums[1], work)}
    aes = 2 * t * ((1 << h) - 1)
    ggm = {"config": {"t": t, "h": h, "aes": aes}}
    for name, fn in calls.items():
        fn()
        torch.cuda.synchronize()
        ms, _ = timed_kernels(cuot_amd, ctx, fn)
        ggm[name] = {"ms": round(ms, 4), "aes_per_s": round(aes / (ms * 1e-3), 1)}
    ggm["same_output"] = bool(torch.equal(leaves[0], leaves[1]) and torch.equal(sums[0], sums[1]))
    ggm["form"] = ("ttable: one fused launch (breadth-first top in LDS, per-lane depth-first subtrees); "
                   "bitsliced: one launch per level through HBM, 32 nodes per lane")
    out["ggm_sender"] = ggm
    # the receiver of the same trees (one punctured path per tree) with each cipher
    choice = torch.randint(0, 2, (t * h,), dtype=torch.uint8, generator=g).cuda()
    csum = sums[0].view(t * h, 2, 2)[torch.arange(t * h, device="cuda"), choice.long()].contiguous()
    fin = torch.randint(-2**62, 2**62, (t, 2), dtype=torch.int64, generator=g).cuda()
    rwork = torch.empty(ctx.ggm_recver_bs_work_bytes(p), dtype=torch.uint8, device="cuda")
    rcalls = {"ttable": lambda: ctx.ggm_recver(p, choice, csum, fin, leaves[0]),
              "bitsliced": lambda: ctx.ggm_recver_bs(p, choice, csum, fin, leaves[1], rwork)}
    rec = {"config": ggm["config"]}
    for name, fn in rcalls.items():
        fn()
        torch.cuda.synchronize()
        ms, _ = timed_kernels(cuot_amd, ctx, fn)
        rec[name] = {"ms": round(ms, 4), "aes_per_s": round(aes / (ms * 1e-3), 1)}
    rec["same_output"] = bool(torch.equal(leaves[0], leaves[1]))
    out["ggm_recver"] = rec
    del leaves, sums, work, rwork, roots, choice, csum, fin
    return out


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
    if args.counter_run:
        return counter_run(args)
    counters, counters_note = (None, "skipped (--no-counters)")
    if not args.no_counters:
        try:
            counters, counters_note = collect_counters()
        except Exception as e:  # the counters are a report, never a reason to lose the bench line
            counters, counters_note = None, f"counter passes failed: {type(e).__name__}: {e}"
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
    progress(f"SOT leg done: {ms_per_step:.3f} ms per instance")

    qc_ms = (stage_ms["qc_sender"] + stage_ms["qc_recver"]) / 2
    qc_bytes = 48 * n  # compulsory per encode: read 2n blocks, write n blocks
    dom_achieved = dom["bytes"] / (dom["ms"] / 1e3) / 1e9
    # the dominant kernel's counters, averaged over its template variants by launches
    variants = [v for k, v in (counters or {}).items() if kernel_base(k)[0] == dom_name]
    nl = sum(v["launches"] for v in variants)
    dom_ctr = None
    if nl:
        valu = sum(v["valu_instr_per_launch"] * v["launches"] for v in variants)
        dom_ctr = {"valu_instr_per_launch": round(valu / nl)}
        if all("hbm_bytes_per_launch" in v for v in variants):
            dom_ctr["hbm_bytes_per_launch"] = round(sum(v["hbm_bytes_per_launch"] * v["launches"] for v in variants)
                                                    / nl)
    traffic = dom_ctr.get("hbm_bytes_per_launch") if dom_ctr else None
    dom_launch_ms = dom["ms"] / dom["launches"]
    dom_valu = (dom_ctr["valu_instr_per_launch"] * 64 / (dom_launch_ms * 1e-3 * VALU_LANE_OPS_PER_S)
                if dom_ctr and dom_ctr.get("valu_instr_per_launch") else None)
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
        # per-kernel launch times of one more staged run (events around each launch)
        cuot_amd.check(lib.cuot_timing_begin(ctx.handle), "cuot_timing_begin")
        for i, (_, bit) in enumerate(fst):
            ctx.ferret_run_stages(fcfg, fws, fs, fr, fd, bit, stream)
        frecs = (cuot_amd._lib.LaunchTime * 4096)()
        fnrec = C.c_int()
        cuot_amd.check(lib.cuot_timing_end(ctx.handle, frecs, 4096, C.byref(fnrec)), "cuot_timing_end")
        fk_ = {}
        for r in frecs[: fnrec.value]:
            k = fk_.setdefault(r.kernel.decode(), {"launches": 0, "ms": 0.0, "bytes": 0.0})
            k["launches"] += 1
            k["ms"] += r.ms
            k["bytes"] += r.bytes
        ferret_kernels = {name: {"launches": k["launches"], "ms": round(k["ms"], 4),
                                 "algorithmic_GBps": round(k["bytes"] / (k["ms"] / 1e3) / 1e9, 1) if k["ms"] else None,
                                 "algorithmic_bytes_per_launch": round(k["bytes"] / k["launches"])}
                          for name, k in fk_.items()}
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
            "kernels": ferret_kernels,
        }
        lk = ferret_kernels.get("lpn_kernel")
        if lk:
            lpn_row = next((v for k, v in (counters or {}).items() if k.startswith("lpn_kernel")), None)
            ferret["lpn"] = {
                "ms_per_launch": round(lk["ms"] / lk["launches"], 4),
                "algorithmic_GBps": lk["algorithmic_GBps"],
                "hbm_frac_algorithmic": round(lk["algorithmic_GBps"] / HBM_PEAK_GBS, 4),
                "algorithmic_bytes_per_launch": lk["algorithmic_bytes_per_launch"],
                "bytes_per_row_per_role": 16 * 12,
                "hbm_bytes_per_launch": lpn_row.get("hbm_bytes_per_launch") if lpn_row else None,
                "traffic_over_algorithmic": lpn_row.get("traffic_over_algorithmic") if lpn_row else None,
                "valu_util": lpn_row.get("valu_util") if lpn_row else None,
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
    small = aes = None
    if not args.no_small_codes:
        small = small_code_leg(torch, cuot_amd, ctx, counters)
        progress("small-code QC leg done")
    if not args.no_aes:
        aes = aes_leg(torch, cuot_amd, ctx, counters)
        progress("AES leg done")
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
        progress("serving leg done")
    if not args.batch_first and not args.no_batch:
        batched = batch_leg(torch, cuot_amd, ctx, max(1, args.steps))
        torch.cuda.empty_cache()
        progress("batched leg done")

    cpu = None
    if not args.no_cpu:
        # every core this process may run on, within the CPU share the host grants one GPU (its
        # OMP_NUM_THREADS) when it sets one
        share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        cores = len(os.sched_getaffinity(0))
        threads = args.cpu_threads or (min(cores, share) if share > 0 else cores)
        cpu = cpu_baseline(threads)
        progress("CPU baseline done")

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
            "bound": "valu" if dom_valu is not None and dom_valu > dom_achieved / HBM_PEAK_GBS else "hbm",
            "kernel": dom_name,
            "achieved": round(dom_achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom_achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "avg_launch_ms": round(dom_launch_ms, 4),
            "algorithmic_bytes_per_launch": round(dom["bytes"] / dom["launches"]),
            "valu_frac": round(dom_valu, 4) if dom_valu is not None else None,
            "valu_instr_per_launch": dom_ctr.get("valu_instr_per_launch") if dom_ctr else None,
            "valu_peak_lane_ops_per_s": VALU_LANE_OPS_PER_S,
        },
        "counters": {"note": counters_note, "kernels": counters},
        "kernels": per_kernel,
        "qc_encode": {"ms_per_party": round(qc_ms, 4), "compulsory_bytes": qc_bytes,
                      "compulsory_GBps": round(qc_bytes / (qc_ms / 1e3) / 1e9, 1),
                      "bounds": qc_encode_bound(counters, qc_ms)},
        "ggm": {"aes_per_launch": aes_per_party, "ms_per_launch": round(ggm_ms, 4),
                "aes_per_s": round(aes_per_party / (ggm_ms / 1e3), 1)},
        "qc_small_codes": small,
        "aes": aes,
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
