"""Python mirror of the SOT roles (silent/roles.h:15-85): SilentConfig,
SOTSender and SOTRecver with base_ot / seed_expand / get_punc_key / dual_lpn.

Both roles run in-process on one GPU; base OTs are simulated from a shared
128-bit seed.  Outputs stay device-resident and are exposed as torch tensors
viewing the library-owned buffers.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import Block, SilentConfig as _SilentConfigC, check
from .logger import LPN, BaseOT, Log, Recver, SeedExp, Sender


@dataclass
class SilentConfig:
    """Fields of SilentConfig (silent/roles.h:15-25); logOT = log2(#leaves) = log2(2n)."""
    id: int = 0
    logOT: int = 21
    nTree: int = 256
    leftKey: tuple = (3242342, 0, 0, 0)
    rightKey: tuple = (8993849, 0, 0, 0)
    gpuPerParty: int = 1

    def to_c(self):
        c = _SilentConfigC()
        c.id, c.logOT, c.nTree, c.gpuPerParty = self.id, self.logOT, self.nTree, self.gpuPerParty
        for i in range(4):
            c.leftKey[i] = self.leftKey[i]
            c.rightKey[i] = self.rightKey[i]
        return c


def _torch_view(ptr: int, nbytes: int, device=0):
    """Wrap a library-owned device buffer as a torch uint8 tensor (no copy)."""
    import torch

    class _Holder:
        __cuda_array_interface__ = None

    st = torch.cuda.current_stream(device)
    iface = {"shape": (nbytes,), "typestr": "|u1", "data": (ptr, False), "version": 3, "stream": None}
    h = _Holder()
    h.__cuda_array_interface__ = iface
    return torch.as_tensor(h, device=f"cuda:{device}")


device_bytes = _torch_view


class _Role:
    _new = None
    _log_role = Sender

    def __init__(self, ctx, config: SilentConfig, seed):
        self.ctx = ctx
        self.config = config
        self.lib = _lib.load()
        h = C.c_void_p()
        sb = np.ascontiguousarray(seed, dtype=np.uint64).reshape(2)
        check(getattr(self.lib, self._new)(ctx.handle, C.byref(config.to_c()),
                                           C.byref(Block(int(sb[0]), int(sb[1]))), C.byref(h)), self._new)
        self.handle = h

    # each stage is one blocking call, logged as the reference's events
    # (gpu/logger.h:7-9) when a logger.Log is open for this role
    def base_ot(self):
        Log.start(self._log_role, BaseOT)
        check(self.lib.cuot_sot_base_ot(self.handle), "base_ot")
        Log.end(self._log_role, BaseOT)

    def seed_expand(self):
        Log.start(self._log_role, SeedExp)
        check(self.lib.cuot_sot_seed_expand(self.handle), "seed_expand")
        Log.mem(self._log_role, SeedExp)
        Log.end(self._log_role, SeedExp)

    def dual_lpn(self):
        Log.start(self._log_role, LPN)
        check(self.lib.cuot_sot_dual_lpn(self.handle), "dual_lpn")
        Log.mem(self._log_role, LPN)
        Log.end(self._log_role, LPN)

    def _outputs(self):
        v, b, d, n = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_int64()
        check(self.lib.cuot_sot_outputs(self.handle, C.byref(v), C.byref(b), C.byref(d), C.byref(n)), "outputs")
        return v.value, b.value, d.value, n.value

    def close(self):
        if self.handle:
            self.lib.cuot_sot_role_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SOTSender(_Role):
    """SOTSender (silent/roles.h:53-64): fullVector (n blocks) and delta after dual_lpn."""
    _new = "cuot_sot_sender_new"

    @property
    def fullVector(self):
        v, _, _, n = self._outputs()
        return _torch_view(v, n * 16).view(-1, 16)

    @property
    def delta(self):
        _, _, d, _ = self._outputs()
        return _torch_view(d, 16)


class SOTRecver(_Role):
    """SOTRecver (silent/roles.h:66-85): puncVector (n blocks), choiceVector (n bits)."""
    _new = "cuot_sot_recver_new"
    _log_role = Recver

    def get_punc_key(self, sender: SOTSender):
        check(self.lib.cuot_sot_get_punc_key(self.handle, sender.handle), "get_punc_key")

    @property
    def puncVector(self):
        v, _, _, n = self._outputs()
        return _torch_view(v, n * 16).view(-1, 16)

    @property
    def choiceVector(self):
        _, b, _, n = self._outputs()
        return _torch_view(b, ((n + 31) // 32) * 4)
