"""cuot_amd -- MI355X (gfx950) silent-OT correlation engine.

Host-side mirror of the reference's interfaces for the silent-OT path:
``cuot_amd.dev_layer`` mirrors emp-ot Ferret's ``dev_layer.h`` calls and
``cuot_amd.silent`` mirrors the SOT sender/receiver roles (``silent/roles.h``).
All compute runs in the HIP library ``lib/libcuot_amd.so`` through its C ABI
(``include/cuot_amd.h``); torch is used only to own device memory and streams.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (Block, CuotError, FerretConfig, GgmParams, Instance, LpnParams, ServerConfig, ServerResult,
                   SotConfig, check, load)

MODE_SOT = 0
MODE_FERRET = 1

__all__ = ["Context", "Batch", "Server", "ServerConfig", "serve_replay", "Block", "GgmParams", "LpnParams", "SotConfig", "FerretConfig", "Instance",
           "CuotError", "MODE_SOT", "MODE_FERRET", "BATCH_GRAPH", "BATCH_OVERLAP", "to_block", "load"]


def to_block(b) -> Block:
    """numpy [lo, hi] (uint64) or bytes(16) -> Block."""
    if isinstance(b, (bytes, bytearray)):
        b = np.frombuffer(bytes(b), dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64).reshape(2)
    return Block(int(b[0]), int(b[1]))


def _ptr(t) -> int:
    if t is None:
        return None
    return t.data_ptr()


def _stream_ptr(stream):
    if stream is None:
        import torch

        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


class Context:
    """One engine context bound to one GPU (cuot_init / cuot_destroy)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.cuot_init(device, C.byref(h)), "cuot_init")
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            self.lib.cuot_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- stages (async on the torch current stream) -------------------------
    def aes_ecb(self, key, inp, out, stream=None):
        n = inp.numel() // 2
        check(self.lib.cuot_aes_ecb(self.handle, C.byref(to_block(key)), _ptr(inp), _ptr(out), n,
                                    _stream_ptr(stream)), "cuot_aes_ecb")

    def prg(self, seed, stream_id: int, start: int, out, stream=None):
        n = out.numel() // 2
        check(self.lib.cuot_prg(self.handle, C.byref(to_block(seed)), stream_id, start, n, _ptr(out),
                                _stream_ptr(stream)), "cuot_prg")

    def ggm_sender(self, params: GgmParams, roots, leaves, sums, stream=None):
        check(self.lib.cuot_ggm_sender(self.handle, C.byref(params), _ptr(roots), _ptr(leaves), _ptr(sums),
                                       _stream_ptr(stream)), "cuot_ggm_sender")

    def ggm_sender_bs_work_bytes(self, params: GgmParams) -> int:
        return int(self.lib.cuot_ggm_sender_bs_work_bytes(C.byref(params)))

    def ggm_sender_bs(self, params: GgmParams, roots, leaves, sums, work, stream=None):
        """The sender expansion on the bitsliced AES (breadth-first, one launch per level)."""
        check(self.lib.cuot_ggm_sender_bs(self.handle, C.byref(params), _ptr(roots), _ptr(leaves), _ptr(sums),
                                          _ptr(work), work.numel() * work.element_size(), _stream_ptr(stream)),
              "cuot_ggm_sender_bs")

    def ggm_recver_bs_work_bytes(self, params: GgmParams) -> int:
        return int(self.lib.cuot_ggm_recver_bs_work_bytes(C.byref(params)))

    def ggm_recver_bs(self, params: GgmParams, choice, csum, fin, leaves, work, stream=None):
        """The receiver expansion on the bitsliced AES (breadth-first, per-level sibling fix)."""
        check(self.lib.cuot_ggm_recver_bs(self.handle, C.byref(params), _ptr(choice), _ptr(csum), _ptr(fin),
                                          _ptr(leaves), _ptr(work), work.numel() * work.element_size(),
                                          _stream_ptr(stream)), "cuot_ggm_recver_bs")

    def ggm_recver(self, params: GgmParams, choice, csum, fin, leaves, stream=None):
        check(self.lib.cuot_ggm_recver(self.handle, C.byref(params), _ptr(choice), _ptr(csum), _ptr(fin),
                                       _ptr(leaves), _stream_ptr(stream)), "cuot_ggm_recver")

    def lpn(self, params: LpnParams, noise, kk, out, stream=None):
        check(self.lib.cuot_lpn(self.handle, C.byref(params), _ptr(noise), _ptr(kk), _ptr(out),
                                _stream_ptr(stream)), "cuot_lpn")

    def lpn_dual(self, params: LpnParams, noise_s, kk_s, out_s, noise_r, kk_r, out_r, stream=None):
        check(self.lib.cuot_lpn_dual(self.handle, C.byref(params), _ptr(noise_s), _ptr(kk_s), _ptr(out_s),
                                     _ptr(noise_r), _ptr(kk_r), _ptr(out_r), _stream_ptr(stream)),
              "cuot_lpn_dual")

    # ---- quasi-cyclic code -----------------------------------------------------
    def qc_ahat_len(self, n: int) -> int:
        return int(self.lib.cuot_qc_ahat_len(n))

    def qc_scratch_bytes(self, n: int) -> int:
        return int(self.lib.cuot_qc_scratch_bytes(n))

    def qc_prepare(self, n, a_bits, ahat, scratch, stream=None):
        check(self.lib.cuot_qc_prepare(self.handle, n, _ptr(a_bits), _ptr(ahat), _ptr(scratch),
                                       _stream_ptr(stream)), "cuot_qc_prepare")

    def qc_encode(self, n, ahat, x, y, scratch, stream=None):
        check(self.lib.cuot_qc_encode(self.handle, n, _ptr(ahat), _ptr(x), _ptr(y), _ptr(scratch),
                                      _stream_ptr(stream)), "cuot_qc_encode")

    def qc_sparse(self, n, a_bits, pos, out_bits, stream=None):
        w = pos.numel()
        check(self.lib.cuot_qc_sparse(self.handle, n, _ptr(a_bits), _ptr(pos) if w else None, w, _ptr(out_bits),
                                      _stream_ptr(stream)), "cuot_qc_sparse")

    def bit_transpose(self, x, rows, stream=None):
        check(self.lib.cuot_bit_transpose(self.handle, _ptr(x), x.shape[0], _ptr(rows), _stream_ptr(stream)),
              "cuot_bit_transpose")

    def bit_untranspose(self, rows, n_blocks, out, stream=None):
        check(self.lib.cuot_bit_untranspose(self.handle, _ptr(rows), n_blocks, _ptr(out), _stream_ptr(stream)),
              "cuot_bit_untranspose")

    # ---- whole instances --------------------------------------------------------
    def sot_workspace_bytes(self, cfg: SotConfig) -> int:
        return int(self.lib.cuot_sot_workspace_bytes(C.byref(cfg)))

    def sot_run(self, cfg: SotConfig, ws, v, w, b, delta=None, stream=None):
        check(self.lib.cuot_sot_run(self.handle, C.byref(cfg), _ptr(ws), ws.numel() * ws.element_size(), _ptr(v),
                                    _ptr(w), _ptr(b), _ptr(delta), _stream_ptr(stream)), "cuot_sot_run")

    def ferret_workspace_bytes(self, cfg: FerretConfig) -> int:
        return int(self.lib.cuot_ferret_workspace_bytes(C.byref(cfg)))

    def ferret_run(self, cfg: FerretConfig, ws, out_s, out_r, delta=None, stream=None):
        check(self.lib.cuot_ferret_run(self.handle, C.byref(cfg), _ptr(ws), ws.numel() * ws.element_size(),
                                       _ptr(out_s), _ptr(out_r), _ptr(delta), _stream_ptr(stream)),
              "cuot_ferret_run")

    def ferret_run_noalloc(self, cfg: FerretConfig, out_s, out_r, delta=None, stream=None):
        """cuot_ferret_run with the library-held workspace (no allocation in the call)."""
        check(self.lib.cuot_ferret_run_noalloc(self.handle, C.byref(cfg), _ptr(out_s), _ptr(out_r), _ptr(delta),
                                               _stream_ptr(stream)),
              "cuot_ferret_run_noalloc")

    def sot_run_stages(self, cfg: SotConfig, ws, v, w, b, delta, stages: int, stream=None):
        check(self.lib.cuot_sot_run_stages(self.handle, C.byref(cfg), _ptr(ws), ws.numel() * ws.element_size(),
                                           _ptr(v), _ptr(w), _ptr(b), _ptr(delta), stages, _stream_ptr(stream)),
              "cuot_sot_run_stages")

    def ferret_run_stages(self, cfg: FerretConfig, ws, out_s, out_r, delta, stages: int, stream=None):
        check(self.lib.cuot_ferret_run_stages(self.handle, C.byref(cfg), _ptr(ws), ws.numel() * ws.element_size(),
                                              _ptr(out_s), _ptr(out_r), _ptr(delta), stages, _stream_ptr(stream)),
              "cuot_ferret_run_stages")


BATCH_STAGES = [("setup", 1), ("qc_prepare", 2), ("ggm_sender", 4), ("ggm_recver", 8), ("qc", 16), ("sparse", 32),
                ("lpn", 64)]
BATCH_ALL = 127
BATCH_GRAPH = 1 << 16
BATCH_OVERLAP = 1 << 17


class Batch:
    """Many SOT / Ferret instances as one launch sequence (cuot_batch_*).

    ``instances`` holds ``Instance`` records; the workspace and the outputs are
    torch device tensors owned by this object: ``out0[i]`` (SOT v / Ferret
    sender), ``out1[i]`` (SOT w / Ferret receiver), ``bits[i]`` (SOT choice
    bits, int32 words) and ``delta`` (one block per instance).
    """

    def __init__(self, ctx: Context, instances, qc_chunk_bytes: int = 0):
        import torch

        self.ctx = ctx
        self.lib = ctx.lib
        self.instances = list(instances)
        arr = (Instance * len(self.instances))(*self.instances)
        h = C.c_void_p()
        check(self.lib.cuot_batch_new(ctx.handle, arr, len(self.instances), qc_chunk_bytes, C.byref(h)),
              "cuot_batch_new")
        self.handle = h
        dev = torch.device("cuda", ctx.device)
        self.ws_bytes = int(self.lib.cuot_batch_workspace_bytes(h))
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)
        self.out0, self.out1, self.bits = [], [], []
        for ins in self.instances:
            self.out0.append(torch.empty((ins.n, 2), dtype=torch.int64, device=dev))
            self.out1.append(torch.empty((ins.n, 2), dtype=torch.int64, device=dev))
            self.bits.append(torch.empty(((ins.n + 31) // 32,), dtype=torch.int32, device=dev)
                             if ins.kind == MODE_SOT else None)
        self.delta = torch.empty((len(self.instances), 2), dtype=torch.int64, device=dev)
        n = len(self.instances)
        o0 = (C.c_void_p * n)(*[t.data_ptr() for t in self.out0])
        o1 = (C.c_void_p * n)(*[t.data_ptr() for t in self.out1])
        bb = (C.c_void_p * n)(*[(t.data_ptr() if t is not None else None) for t in self.bits])
        check(self.lib.cuot_batch_bind(h, self.ws.data_ptr(), self.ws_bytes, o0, o1, bb, self.delta.data_ptr()),
              "cuot_batch_bind")

    def run(self, stages: int = BATCH_ALL, graph: bool = False, stream=None, overlap: bool = False):
        flags = stages | (BATCH_GRAPH if graph else 0) | (BATCH_OVERLAP if overlap else 0)
        check(self.lib.cuot_batch_run(self.handle, flags, _stream_ptr(stream)), "cuot_batch_run")

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cuot_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop (cuot_server_*): submit requests, poll results, release."""

    def __init__(self, ctx: Context, hbm_budget: int, max_batch: int = 0, streams: int = 0, qc_chunk_bytes: int = 0,
                 use_graph: bool = False, plain: bool = False, max_batch_cots: int = 0, ferret_clients: int = 0,
                 ferret_session: FerretConfig | None = None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.cfg = ServerConfig(hbm_budget=hbm_budget, max_batch=max_batch, streams=streams,
                                qc_chunk_bytes=qc_chunk_bytes, use_graph=int(use_graph), plain=int(plain),
                                max_batch_cots=max_batch_cots, ferret_clients=ferret_clients,
                                ferret_session=ferret_session or FerretConfig())
        h = C.c_void_p()
        check(self.lib.cuot_server_new(ctx.handle, C.byref(self.cfg), C.byref(h)), "cuot_server_new")
        self.handle = h

    def submit(self, inst: Instance) -> int:
        t = C.c_int64()
        check(self.lib.cuot_server_submit(self.handle, C.byref(inst), C.byref(t)), "cuot_server_submit")
        return t.value

    def poll(self, ticket: int, wait: bool = False) -> ServerResult:
        r = ServerResult()
        check(self.lib.cuot_server_poll(self.handle, ticket, int(wait), C.byref(r)), "cuot_server_poll")
        return r

    def release(self, ticket: int):
        check(self.lib.cuot_server_release(self.handle, ticket), "cuot_server_release")

    def stats(self):
        b = C.c_int64()
        u = C.c_size_t()
        check(self.lib.cuot_server_stats(self.handle, C.byref(b), C.byref(u)), "cuot_server_stats")
        return b.value, u.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cuot_server_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_replay(ctx: Context, instances, arrivals, hbm_budget: int, plain: bool = False, max_batch: int = 0,
                 streams: int = 0, qc_chunk_bytes: int = 0, use_graph: bool = False, max_batch_cots: int = 0,
                 check_relation: bool = False, arena=None, ferret_clients: int = 0,
                 ferret_session: FerretConfig | None = None):
    """Real-time replay of (arrival seconds, Instance) through a fresh server
    (arena: optional caller-owned device tensor of at least hbm_budget bytes).

    Returns (latencies in s as numpy array, wall seconds, batches launched,
    relation failures).  With check_relation every result is checked on the
    device against its COT relation before release (failures = the number of
    failing indices over all results); otherwise failures is None."""
    n = len(instances)
    cfg = ServerConfig(hbm_budget=hbm_budget, max_batch=max_batch, streams=streams, qc_chunk_bytes=qc_chunk_bytes,
                       use_graph=int(use_graph), plain=int(plain), max_batch_cots=max_batch_cots,
                       arena_d=arena.data_ptr() if arena is not None else None, ferret_clients=ferret_clients,
                       ferret_session=ferret_session or FerretConfig())
    if arena is not None and arena.numel() * arena.element_size() < hbm_budget:
        raise ValueError("arena smaller than hbm_budget")
    arr = (Instance * n)(*instances)
    t = (C.c_double * n)(*[float(a) for a in arrivals])
    lat = (C.c_double * n)()
    wall = C.c_double()
    nb = C.c_int64()
    fails = C.c_int64(-1)
    check(ctx.lib.cuot_serve_replay(ctx.handle, C.byref(cfg), arr, t, n, lat, C.byref(wall), C.byref(nb),
                                    C.byref(fails) if check_relation else None), "cuot_serve_replay")
    return np.array(lat[:], dtype=np.float64), wall.value, nb.value, (fails.value if check_relation else None)


def make_instance_struct(kind: int, n: int, t: int, h: int, k: int, seed, ntasks: int = 1, k0=(3242342, 0),
                         k1=(8993849, 0), client: int = 0) -> Instance:
    """Instance record (SOT keys default to the SilentConfig leftKey/rightKey;
    client > 0: a Ferret request served from that resident client's session)."""
    return Instance(kind=kind, h=h, n=n, t=t, k=k, ntasks=ntasks, client=client, k0=to_block(k0), k1=to_block(k1),
                    seed=to_block(seed))


SOT_ALL = 127
FERRET_ALL = 15
# bytes of the library-held workspace of Context.ferret_run_noalloc (CUOT_FERRET_NOALLOC_WORKSPACE)
FERRET_NOALLOC_WORKSPACE = 32 << 20
SOT_STAGES = [("setup", 1), ("qc_prepare", 2), ("ggm_sender", 4), ("ggm_recver", 8), ("qc_sender", 16),
              ("qc_recver", 32), ("sparse", 64)]
FERRET_STAGES = [("setup", 1), ("ggm_sender", 2), ("ggm_recver", 4), ("lpn", 8)]
