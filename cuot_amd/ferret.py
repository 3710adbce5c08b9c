"""Host-side mirror of emp-ot's FerretCOT extend loop for both parties on one
GPU (ferret/emp-ot/emp-ot/ferret/ferret_cot.hpp), over the iterating session
of the HIP library (cuot_ferret_session_*).

    fc = FerretCOT(ctx, n=..., k=..., t=..., h=13, seed=(lo, hi))
    s, r = fc.rcot(num)        # num COTs: r = s ^ lsb(r) * Delta

`rcot` follows FerretCOT::rcot (ferret_cot.hpp:171-209): it serves requests
from the current extend's ot_limit outputs and runs extend_f2k when they run
out; `silent_ot_left` is the number still unserved.  The outputs are torch
device tensors of shape (num, 2) int64 (one 16-byte block per row).
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import FerretConfig, check


class FerretCOT:
    def __init__(self, ctx, n: int, k: int, t: int, h: int, seed, ntasks: int = 1):
        import torch

        from . import to_block

        self.ctx = ctx
        self.lib = _lib.load()
        cfg = FerretConfig(n=n, k=k, t=t, h=h, ntasks=ntasks, seed=to_block(seed))
        hnd = C.c_void_p()
        check(self.lib.cuot_ferret_session_new(ctx.handle, C.byref(cfg), C.byref(hnd)), "cuot_ferret_session_new")
        self.handle = hnd
        M, lim, rnd = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.lib.cuot_ferret_session_info(hnd, C.byref(M), C.byref(lim), C.byref(rnd)),
              "cuot_ferret_session_info")
        self.n, self.M, self.ot_limit = n, M.value, lim.value
        dev = torch.device("cuda", ctx.device)
        self.out_s = torch.empty((n, 2), dtype=torch.int64, device=dev)
        self.out_r = torch.empty((n, 2), dtype=torch.int64, device=dev)
        self.Delta = torch.empty((1, 2), dtype=torch.int64, device=dev)
        check(self.lib.cuot_ferret_session_delta(hnd, self.Delta.data_ptr(), self._stream()),
              "cuot_ferret_session_delta")
        self.ot_used = self.ot_limit  # nothing generated yet (ferret_cot.hpp:77)

    def _stream(self):
        import torch

        return torch.cuda.current_stream().cuda_stream

    @property
    def rounds(self) -> int:
        rnd = C.c_int64()
        check(self.lib.cuot_ferret_session_info(self.handle, None, None, C.byref(rnd)), "cuot_ferret_session_info")
        return rnd.value

    def silent_ot_left(self) -> int:
        return self.ot_limit - self.ot_used

    def extend_f2k(self):
        """One extend: n fresh COTs per party, the last M kept as the bootstrap."""
        check(self.lib.cuot_ferret_session_extend(self.handle, self.out_s.data_ptr(), self.out_r.data_ptr(),
                                                  self._stream()), "cuot_ferret_session_extend")
        self.ot_used = 0

    def rcot(self, num: int):
        """num random COTs for both parties (sender blocks, receiver blocks)."""
        import torch

        parts_s, parts_r = [], []
        left = num
        while left > 0:
            if self.silent_ot_left() == 0:
                self.extend_f2k()
            take = min(left, self.silent_ot_left())
            a, b = self.ot_used, self.ot_used + take
            parts_s.append(self.out_s[a:b].clone())
            parts_r.append(self.out_r[a:b].clone())
            self.ot_used += take
            left -= take
        if not parts_s:
            return self.out_s[:0].clone(), self.out_r[:0].clone()
        return torch.cat(parts_s), torch.cat(parts_r)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cuot_ferret_session_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
