"""Host-side mirror of emp-ot's FerretCOT extend loop for both parties on one
GPU (ferret/emp-ot/emp-ot/ferret/ferret_cot.hpp), over the iterating session
of the HIP library (cuot_ferret_session_*).

    fc = FerretCOT(ctx, n=..., k=..., t=..., h=13, seed=(lo, hi))
    s, r = fc.rcot(num)        # num COTs: r = s ^ lsb(r) * Delta

`rcot` follows FerretCOT::rcot (ferret_cot.hpp:171-209): it serves requests
from the current extend's ot_limit outputs and runs extend_f2k when they run
out; `silent_ot_left` is the number still unserved.  The outputs are torch
device tensors of shape (num, 2) int64 (one 16-byte block per row).

Session state in emp-ot's byte layouts: `assemble_state` / `disassemble_state`
/ `state_size` (ferret_cot.hpp:309-347) and `write_pre_data128_to_file` /
`read_pre_data128_from_file` (ferret_cot.hpp:215-247), per party (ALICE = 1
sender, BOB = 2 receiver).  With a `logger.Log` open for a role, each extend
is timed as emp-ot logs it: SeedExp (pre-OT + MPCOT) and LPN
(ferret_cot.hpp:84-92), for both roles since both run here.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import FerretConfig, check
from .logger import LPN, Log, Recver, SeedExp, Sender

ALICE, BOB = 1, 2  # emp-ot party ids


class FerretCOT:
    def __init__(self, ctx, n: int, k: int, t: int, h: int, seed, ntasks: int = 1, n_pre: int = 0, _handle=None):
        import torch

        from . import to_block

        self.ctx = ctx
        self.lib = _lib.load()
        # n_pre: blocks of emp-ot's ot_pre_data in the state blob and pre-OT
        # files (PrimalLPNParameter::n_pre, constants.h:8-22); 0 = the bootstrap size M
        cfg = FerretConfig(n=n, k=k, t=t, h=h, ntasks=ntasks, seed=to_block(seed), n_pre=n_pre)
        self._cfg = cfg
        self._params = dict(n=n, k=k, t=t, h=h, seed=seed, ntasks=ntasks, n_pre=n_pre)
        if _handle is None:
            hnd = C.c_void_p()
            check(self.lib.cuot_ferret_session_new(ctx.handle, C.byref(cfg), C.byref(hnd)),
                  "cuot_ferret_session_new")
        else:
            hnd = _handle
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
        self.ot_used = self.ot_limit  # nothing generated yet (ferret_cot.hpp:76)

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
        args = (self.handle, self.out_s.data_ptr(), self.out_r.data_ptr())
        if not (Log.is_open(Sender) or Log.is_open(Recver)):
            check(self.lib.cuot_ferret_session_extend(*args, self._stream()), "cuot_ferret_session_extend")
        else:
            import torch

            for stage, event in ((1, SeedExp), (2, LPN)):
                for role in (Sender, Recver):
                    Log.start(role, event)
                check(self.lib.cuot_ferret_session_extend_stages(*args, stage, self._stream()),
                      "cuot_ferret_session_extend_stages")
                torch.cuda.synchronize()
                for role in (Sender, Recver):
                    Log.mem(role, event)
                    Log.end(role, event)
        self.ot_used = 0

    # ---- online derandomisation (both parties) --------------------------------
    def send_recv_cot(self, b):
        """FerretCOT::send_cot + recv_cot (ferret_cot.hpp:297-307): len(b) COTs
        with the receiver's chosen bits b (uint8 device tensor, one per COT):
        (sender K, receiver M = K ^ b*Delta, the receiver's message bo)."""
        import torch

        num = int(b.shape[0])
        ks, mr = self.rcot(num)
        bo = torch.empty(num, dtype=torch.uint8, device=ks.device)
        check(self.lib.cuot_cot_online(self.ctx.handle, ks.data_ptr(), mr.data_ptr(), b.data_ptr(),
                                       self.Delta.data_ptr(), bo.data_ptr(), num, self._stream()), "cuot_cot_online")
        return ks, mr, bo

    def send_recv_ot(self, m0, m1, r, s, key_start: int = 0):
        """COT::send / recv (cot.h:16-56): chosen-message OTs, the receiver
        learning m_r; s = the sender's MITCCRH block.  (sender pads, receiver
        outputs)."""
        import torch

        from . import to_block

        ks, mr, _ = self.send_recv_cot(r)
        num = int(r.shape[0])
        pad = torch.empty((2 * num, 2), dtype=torch.int64, device=ks.device)
        out = torch.empty((num, 2), dtype=torch.int64, device=ks.device)
        check(self.lib.cuot_cot_to_ot(self.ctx.handle, ks.data_ptr(), mr.data_ptr(), self.Delta.data_ptr(),
                                      to_block(s), key_start, r.data_ptr(), m0.data_ptr(), m1.data_ptr(),
                                      pad.data_ptr(), out.data_ptr(), num, self._stream()), "cuot_cot_to_ot")
        return pad, out

    def send_recv_rot(self, r, s, key_start: int = 0):
        """COT::send_rot / recv_rot (cot.h:58-91): random OTs with the
        receiver's choice r: (sender data0, data1, receiver data = data_r)."""
        import torch

        from . import to_block

        ks, mr, _ = self.send_recv_cot(r)
        num = int(r.shape[0])
        d0, d1, out = (torch.empty((num, 2), dtype=torch.int64, device=ks.device) for _ in range(3))
        check(self.lib.cuot_cot_to_rot(self.ctx.handle, ks.data_ptr(), mr.data_ptr(), self.Delta.data_ptr(),
                                       to_block(s), key_start, d0.data_ptr(), d1.data_ptr(), out.data_ptr(), num,
                                       self._stream()), "cuot_cot_to_rot")
        return d0, d1, out

    # ---- malicious mode ----------------------------------------------------
    def set_malicious(self, on: bool = True):
        """MpcotReg::set_malicious: check every MPCOT (both parties)."""
        check(self.lib.cuot_ferret_session_set_malicious(self.handle, 1 if on else 0),
              "cuot_ferret_session_set_malicious")

    def check_result(self):
        """The last consistency check: (ok, x' bytes, ALICE digest, BOB digest)."""
        ok = C.c_int()
        xp, da, db = (C.c_uint8 * 128)(), (C.c_uint8 * 32)(), (C.c_uint8 * 32)()
        check(self.lib.cuot_ferret_session_check(self.handle, C.byref(ok), xp, da, db), "cuot_ferret_session_check")
        return bool(ok.value), bytes(xp), bytes(da), bytes(db)

    # ---- state blob and pre-OT files (emp-ot layouts) ------------------------
    def state_size(self) -> int:
        return int(self.lib.cuot_ferret_state_size(C.byref(self._cfg)))

    def assemble_state(self, party: int) -> bytes:
        """One party's state blob: int64 party, n, t, k; Delta; bootstrap."""
        buf = C.create_string_buffer(self.state_size())
        check(self.lib.cuot_ferret_session_assemble_state(self.handle, party, buf, len(buf)),
              "cuot_ferret_session_assemble_state")
        return buf.raw

    @classmethod
    def disassemble_state(cls, ctx, alice: bytes, bob: bytes, n: int, k: int, t: int, h: int, seed,
                          round: int = 0, ntasks: int = 1, n_pre: int = 0):
        """A session resumed from both parties' state blobs at `round`."""
        from . import to_block

        cfg = FerretConfig(n=n, k=k, t=t, h=h, ntasks=ntasks, seed=to_block(seed), n_pre=n_pre)
        hnd = C.c_void_p()
        size = min(len(alice), len(bob))
        check(_lib.load().cuot_ferret_session_disassemble_state(ctx.handle, C.byref(cfg), round, alice, bob, size,
                                                                C.byref(hnd)),
              "cuot_ferret_session_disassemble_state")
        return cls(ctx, n, k, t, h, seed, ntasks, n_pre, _handle=hnd)

    def write_pre_data128_to_file(self, party: int, filename: str):
        check(self.lib.cuot_ferret_session_write_pre_file(self.handle, party, filename.encode()),
              "cuot_ferret_session_write_pre_file")

    @classmethod
    def read_pre_data128_from_file(cls, ctx, alice_file: str, bob_file: str, n: int, k: int, t: int, h: int, seed,
                                   round: int = 0, ntasks: int = 1, n_pre: int = 0):
        """A session from both parties' pre-OT files (which are then removed)."""
        from . import to_block

        cfg = FerretConfig(n=n, k=k, t=t, h=h, ntasks=ntasks, seed=to_block(seed), n_pre=n_pre)
        hnd = C.c_void_p()
        check(_lib.load().cuot_ferret_session_read_pre_files(ctx.handle, C.byref(cfg), round, alice_file.encode(),
                                                             bob_file.encode(), C.byref(hnd)),
              "cuot_ferret_session_read_pre_files")
        return cls(ctx, n, k, t, h, seed, ntasks, n_pre, _handle=hnd)

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
