"""Instance parameter rules for the workloads (host logic, shared by the
engine, the bench and the tests).

Plain SOT (n COTs, scaler-2 code): 2n leaves in t trees of depth h,
    h = round((log2(2n) + 5) / 2), t = ceil(2n / 2^h), last tree truncated
    (reproduces the reference-sized configurations: n = 2^20 -> h = 13, t = 256;
    n = 2^24 -> h = 15, t = 1024).
Ferret (n COTs): k = ceil(n / 22) bootstrap COTs, d = 10, trees of 2^13 leaves,
    t = ceil(n / 8192), truncated to n.
Batched mix (64 instances): n_i = floor(2^(16 + 6 u_i)), kinds alternate
    SOT / Ferret, per-instance 128-bit seeds; u_i and the seeds come from a
    PCG64 stream seeded with the 128-bit master seed.
Serving: Poisson arrivals at rate lambda, n = floor(2^(14 + 10 u)), 50/50 mix.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SOT, FERRET = 0, 1
SOT_K0 = (3242342, 0)  # SilentConfig leftKey (silent/main.cu:89)
SOT_K1 = (8993849, 0)
MASTER_SEED = (0x0706050403020100, 0x0F0E0D0C0B0A0908)


@dataclass(frozen=True)
class Instance:
    kind: int
    n: int
    t: int
    h: int
    k: int  # Ferret bootstrap size (0 for SOT)
    seed: tuple  # (lo, hi)

    @property
    def leaves(self) -> int:
        return 2 * self.n if self.kind == SOT else self.n


def sot_shape(n: int):
    h = int(round((math.log2(2 * n) + 5) / 2))
    h = max(1, min(16, h))
    t = -(-2 * n // (1 << h))
    return t, h


def ferret_shape(n: int):
    h = 13
    t = -(-n // (1 << h))
    k = -(-n // 22)
    return t, h, k


def make_instance(kind: int, n: int, seed) -> Instance:
    if kind == SOT:
        t, h = sot_shape(n)
        return Instance(SOT, n, t, h, 0, tuple(int(x) for x in seed))
    t, h, k = ferret_shape(n)
    return Instance(FERRET, n, t, h, k, tuple(int(x) for x in seed))


def _rng(master, salt: int):
    lo, hi = master
    return np.random.Generator(np.random.PCG64([int(lo), int(hi), salt]))


def batch_mix(count: int = 64, master=MASTER_SEED, lo_log: int = 16, hi_log: int = 22):
    """The batched workload: count instances, log-uniform sizes, alternating kinds."""
    r = _rng(master, 3)
    out = []
    for i in range(count):
        u = r.random()
        n = int(math.floor(2.0 ** (lo_log + (hi_log - lo_log) * u)))
        seed = (int(r.integers(0, 2**63)) | (int(r.integers(0, 2)) << 63),
                int(r.integers(0, 2**63)) | (int(r.integers(0, 2)) << 63))
        out.append(make_instance(SOT if i % 2 == 0 else FERRET, n, seed))
    return out


def serving_arrivals(rate_hz: float, count: int, master=MASTER_SEED, lo_log: int = 14, hi_log: int = 24):
    """(arrival_time_s, Instance) pairs of a Poisson process with 50/50 kinds."""
    r = _rng(master, 4)
    t = 0.0
    out = []
    for i in range(count):
        t += r.exponential(1.0 / rate_hz)
        u = r.random()
        n = int(math.floor(2.0 ** (lo_log + (hi_log - lo_log) * u)))
        kind = SOT if r.random() < 0.5 else FERRET
        seed = (int(r.integers(0, 2**63)), int(r.integers(0, 2**63)))
        out.append((t, make_instance(kind, n, seed)))
    return out
