"""Time the quasi-cyclic encoder on its own (python -m cuot_amd.qcprobe).

Prepares A^ for a random circulant, runs one untimed encode, then REPS timed
encodes of random 2n-block input at n COTs and prints the median time.  Before
the encodes it copies 256 MiB device-to-device, a kernel of known traffic that
calibrates hardware byte counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)
collected over the same process.

    python -m cuot_amd.qcprobe [--n 16777216] [--reps 3] [--knobs radix,cols]
"""
from __future__ import annotations

import argparse

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--knobs", type=str, default="", help="qc_radix,qc_cols (default: built-in rule)")
    args = ap.parse_args()
    import torch

    import cuot_amd
    from cuot_amd.tune import set_knobs

    ctx = cuot_amd.Context(0)
    n = args.n
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
    if args.knobs:
        r, c = (int(v) for v in args.knobs.split(","))
        set_knobs(qc_radix=r, qc_cols=c)
    ctx.qc_encode(n, ahat, x, y, scratch)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.reps)]
    for a, b in ev:
        a.record()
        ctx.qc_encode(n, ahat, x, y, scratch)
        b.record()
    torch.cuda.synchronize()
    set_knobs()
    ms = float(np.median([a.elapsed_time(b) for a, b in ev]))
    print(f"qc_encode n={n}: {ms:.3f} ms per party (median of {args.reps})", flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
