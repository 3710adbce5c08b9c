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
from ._lib import Block, CuotError, GgmParams, check, load

MODE_SOT = 0
MODE_FERRET = 1

__all__ = ["Context", "Block", "GgmParams", "CuotError", "MODE_SOT", "MODE_FERRET", "to_block", "load"]


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

    def ggm_recver(self, params: GgmParams, choice, csum, fin, leaves, stream=None):
        check(self.lib.cuot_ggm_recver(self.handle, C.byref(params), _ptr(choice), _ptr(csum), _ptr(fin),
                                       _ptr(leaves), _stream_ptr(stream)), "cuot_ggm_recver")
