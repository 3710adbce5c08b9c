"""Autotuner for the execution knobs (cuot_set_knobs), driven by a learned
cost model.

The model is a small tanh network (cuot_mlp_*: forward, backward and Adam are
the library's own HIP kernels) that maps (instance sizes, knobs) to the log of
the measured stage time.  The tuner measures a sample of (size, knob) points
with HIP events, trains the model on them, and afterwards picks knobs for a
new size by evaluating every valid knob setting through the model instead of
timing each one.  Knobs never change outputs, so the model only ever affects
speed.

Stages and knobs:
  * "qc"  -- quasi-cyclic encode of one party at n COTs: FFT layers per upper
             pass (qc_radix 4..7) and columns per tile (qc_cols 1, 2, 4, 8);
             features (log2 n, m, radix, log2 cols);
  * "ggm" -- GGM expansion of t trees of depth h: per-lane depth-first levels
             L (ggm_dfs, h-10 <= L <= min(h-8, 6)); features (h, log2 t, L).
"""
from __future__ import annotations

import ctypes as C
import itertools
import math
import time

import numpy as np

from . import _lib
from ._lib import KnobValues, check


def _dev():
    import torch

    return torch.device("cuda", torch.cuda.current_device())


class CostModel:
    """F -> H tanh -> H tanh -> 1 regression network trained on the GPU."""

    def __init__(self, ctx, F: int, H: int = 32, seed: int = 0):
        import torch

        self.ctx = ctx
        self.lib = ctx.lib
        self.F, self.H = F, H
        self.P = int(self.lib.cuot_mlp_param_count(F, H))
        if self.P <= 0:
            raise ValueError("unsupported network size")
        r = np.random.default_rng(seed)
        parts = [r.normal(0, 1 / math.sqrt(F), H * F), np.zeros(H), r.normal(0, 1 / math.sqrt(H), H * H),
                 np.zeros(H), r.normal(0, 1 / math.sqrt(H), H), np.zeros(1)]
        self.params = torch.tensor(np.concatenate(parts), dtype=torch.float32, device=_dev())
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.xm = self.xs = None
        self.tm, self.ts = 0.0, 1.0

    # ---- normalisation ---------------------------------------------------------
    def _ensure_norm(self, x, t=None):
        if self.xm is None:
            x = np.asarray(x, dtype=np.float64)
            self.xm = x.mean(0)
            self.xs = np.where(x.std(0) > 0, x.std(0), 1.0)
            if t is not None:
                t = np.asarray(t, dtype=np.float64)
                self.tm = float(t.mean())
                self.ts = float(t.std()) if t.std() > 0 else 1.0

    def _xdev(self, x):
        import torch

        xn = (np.asarray(x, dtype=np.float64) - self.xm) / self.xs
        return torch.tensor(xn, dtype=torch.float32, device=_dev()).contiguous()

    def _stream(self):
        import torch

        return torch.cuda.current_stream().cuda_stream

    # ---- training / inference ----------------------------------------------------
    def loss(self, x, t) -> float:
        """Mean squared error on the normalised target scale."""
        import torch

        self._ensure_norm(x, t)
        xd = self._xdev(x)
        td = torch.tensor((np.asarray(t, dtype=np.float64) - self.tm) / self.ts, dtype=torch.float32,
                          device=_dev())
        N = xd.shape[0]
        ws = torch.empty(int(self.lib.cuot_mlp_workspace_bytes(self.F, self.H, N)), dtype=torch.uint8, device=_dev())
        g = torch.empty(self.P, dtype=torch.float32, device=_dev())
        ls = torch.empty(1, dtype=torch.float32, device=_dev())
        check(self.lib.cuot_mlp_grad(self.ctx.handle, self.params.data_ptr(), self.F, self.H, xd.data_ptr(),
                                     td.data_ptr(), N, ws.data_ptr(), ws.numel(), g.data_ptr(), ls.data_ptr(),
                                     self._stream()), "cuot_mlp_grad")
        return float(ls.item())

    def fit(self, x, t, epochs: int = 500, lr: float = 3e-3):
        """Full-batch Adam on the GPU; returns the final loss."""
        import torch

        self._ensure_norm(x, t)
        xd = self._xdev(x)
        td = torch.tensor((np.asarray(t, dtype=np.float64) - self.tm) / self.ts, dtype=torch.float32,
                          device=_dev())
        N = xd.shape[0]
        ws = torch.empty(int(self.lib.cuot_mlp_workspace_bytes(self.F, self.H, N)), dtype=torch.uint8, device=_dev())
        g = torch.empty(self.P, dtype=torch.float32, device=_dev())
        ls = torch.empty(1, dtype=torch.float32, device=_dev())
        st = self._stream()
        for _ in range(epochs):
            self.step += 1
            check(self.lib.cuot_mlp_grad(self.ctx.handle, self.params.data_ptr(), self.F, self.H, xd.data_ptr(),
                                         td.data_ptr(), N, ws.data_ptr(), ws.numel(), g.data_ptr(), ls.data_ptr(),
                                         st), "cuot_mlp_grad")
            check(self.lib.cuot_mlp_adam(self.ctx.handle, self.params.data_ptr(), g.data_ptr(), self.m.data_ptr(),
                                         self.v.data_ptr(), self.P, self.step, C.c_float(lr), st), "cuot_mlp_adam")
        return float(ls.item())

    def predict(self, x) -> np.ndarray:
        import torch

        xd = self._xdev(x)
        N = xd.shape[0]
        y = torch.empty(N, dtype=torch.float32, device=_dev())
        check(self.lib.cuot_mlp_forward(self.ctx.handle, self.params.data_ptr(), self.F, self.H, xd.data_ptr(), N,
                                        y.data_ptr(), self._stream()), "cuot_mlp_forward")
        return y.cpu().numpy().astype(np.float64) * self.ts + self.tm


# ---- knobs -----------------------------------------------------------------------
def set_knobs(qc_radix=0, qc_cols=0, ggm_dfs=0, lpn_wgs=0, qc_mfma_max=0, aes_impl=0):
    k = KnobValues(qc_radix, qc_cols, ggm_dfs, lpn_wgs, qc_mfma_max, aes_impl)
    check(_lib.load().cuot_set_knobs(C.byref(k)), "cuot_set_knobs")


def get_knobs() -> dict:
    k = KnobValues()
    check(_lib.load().cuot_get_knobs(C.byref(k)), "cuot_get_knobs")
    return {"qc_radix": k.qc_radix, "qc_cols": k.qc_cols, "ggm_dfs": k.ggm_dfs, "lpn_wgs": k.lpn_wgs,
            "qc_mfma_max": k.qc_mfma_max, "aes_impl": k.aes_impl}


QC_SPACE = [dict(qc_radix=r, qc_cols=c) for r, c in itertools.product((4, 5, 6, 7), (1, 2, 4, 8))]
QC_DEFAULT = dict(qc_radix=7, qc_cols=2)


def ggm_space(h: int):
    if h <= 8:
        return [dict(ggm_dfs=0)]
    return [dict(ggm_dfs=L) for L in range(max(1, h - 10), min(h - 8, 6) + 1)]


def qc_log_len(n: int) -> int:
    """Transform length 2^m of the FFT encoder: the least m >= 1 with 2^m >= 2 ceil(n/16) - 1."""
    nc = (n + 15) // 16
    m = 1
    while (1 << m) < 2 * nc - 1:
        m += 1
    return m


def qc_features(n: int, knobs: dict):
    return [math.log2(n), qc_log_len(n), knobs["qc_radix"], math.log2(knobs["qc_cols"])]


def ggm_features(t: int, h: int, knobs: dict):
    return [h, math.log2(t), knobs["ggm_dfs"]]


class Autotuner:
    """Measures stage times under knob settings, trains the cost models, picks knobs."""

    def __init__(self, ctx, H: int = 32, seed: int = 0, reps: int = 3):
        self.ctx = ctx
        self.reps = reps
        self.models = {"qc": CostModel(ctx, 4, H, seed), "ggm": CostModel(ctx, 3, H, seed + 1)}
        self.data = {"qc": ([], []), "ggm": ([], [])}
        self._qc_cache = {}
        self._ggm_cache = {}

    # ---- measurement ---------------------------------------------------------------
    def _time(self, fn) -> float:
        import torch

        fn()  # warm-up
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(self.reps)]
        st = torch.cuda.current_stream()
        for a, b in evs:
            a.record(st)
            fn()
            b.record(st)
        torch.cuda.synchronize()
        return float(np.median([a.elapsed_time(b) for a, b in evs]))

    def _qc_buffers(self, n):
        import torch

        if n not in self._qc_cache:
            dev = _dev()
            g = torch.Generator(device="cpu").manual_seed(n)
            a_bits = torch.randint(0, 256, ((n + 7) // 8 + 16,), dtype=torch.uint8, generator=g).to(dev)
            ahat = torch.empty(self.ctx.qc_ahat_len(n), dtype=torch.int32, device=dev)
            scratch = torch.empty(self.ctx.qc_scratch_bytes(n), dtype=torch.uint8, device=dev)
            x = torch.randint(-2**62, 2**62, (2 * n, 2), dtype=torch.int64, generator=g).to(dev)
            y = torch.empty((n, 2), dtype=torch.int64, device=dev)
            self.ctx.qc_prepare(n, a_bits, ahat, scratch)
            self._qc_cache = {n: (ahat, x, y, scratch)}  # keep one size resident
        return self._qc_cache[n]

    def measure_qc(self, n: int, knobs: dict) -> float:
        ahat, x, y, scratch = self._qc_buffers(n)
        set_knobs(**knobs)
        try:
            return self._time(lambda: self.ctx.qc_encode(n, ahat, x, y, scratch))
        finally:
            set_knobs()

    def _ggm_buffers(self, t, h):
        import torch

        import cuot_amd

        key = (t, h)
        if key not in self._ggm_cache:
            dev = _dev()
            roots = torch.randint(-2**62, 2**62, (t, 2), dtype=torch.int64).to(dev)
            leaves = torch.empty((t << h, 2), dtype=torch.int64, device=dev)
            sums = torch.empty((t * h * 2, 2), dtype=torch.int64, device=dev)
            p = cuot_amd.GgmParams(t=t, h=h, mode=cuot_amd.MODE_SOT, n_out=t << h, k0=cuot_amd.to_block((3242342, 0)),
                                   k1=cuot_amd.to_block((8993849, 0)))
            self._ggm_cache = {key: (p, roots, leaves, sums)}
        return self._ggm_cache[key]

    def measure_ggm(self, t: int, h: int, knobs: dict) -> float:
        p, roots, leaves, sums = self._ggm_buffers(t, h)
        set_knobs(**knobs)
        try:
            return self._time(lambda: self.ctx.ggm_sender(p, roots, leaves, sums))
        finally:
            set_knobs()

    # ---- data, training, choice ----------------------------------------------------------
    def add(self, stage: str, feats, ms: float):
        self.data[stage][0].append(list(feats))
        self.data[stage][1].append(math.log(ms))

    def collect_qc(self, sizes, space=None):
        for n in sizes:
            for k in space or QC_SPACE:
                self.add("qc", qc_features(n, k), self.measure_qc(n, k))

    def collect_ggm(self, shapes):
        for t, h in shapes:
            for k in ggm_space(h):
                self.add("ggm", ggm_features(t, h, k), self.measure_ggm(t, h, k))

    def train(self, stage: str, epochs: int = 1500, lr: float = 3e-3) -> float:
        x, t = self.data[stage]
        return self.models[stage].fit(np.array(x), np.array(t), epochs=epochs, lr=lr)

    def choose_qc(self, n: int) -> dict:
        feats = np.array([qc_features(n, k) for k in QC_SPACE])
        return QC_SPACE[int(np.argmin(self.models["qc"].predict(feats)))]

    def choose_ggm(self, t: int, h: int) -> dict:
        space = ggm_space(h)
        feats = np.array([ggm_features(t, h, k) for k in space])
        return space[int(np.argmin(self.models["ggm"].predict(feats)))]

    def exhaustive_qc(self, n: int):
        """(best knobs, best ms, all ms, seconds spent measuring)"""
        t0 = time.perf_counter()
        ms = [self.measure_qc(n, k) for k in QC_SPACE]
        i = int(np.argmin(ms))
        return QC_SPACE[i], ms[i], ms, time.perf_counter() - t0
