"""ctypes bindings for the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the cuot_amd product.

Blocks are numpy uint64 arrays of shape (N, 2) holding [lo, hi] -- the same
memory layout as a 16-byte block in byte order (bytes 0..7 = lo).
Bit vectors are numpy uint8 arrays packed LSB-first (bit i = byte i//8, bit i%8).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

P = C.c_void_p
I64 = C.c_int64
INT = C.c_int


class ORB(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]


class SOTCFG(C.Structure):
    _fields_ = [("n", I64), ("t", I64), ("h", INT), ("k0", ORB), ("k1", ORB), ("seed", ORB)]


class FERRETCFG(C.Structure):
    _fields_ = [("n", I64), ("k", I64), ("t", I64), ("h", INT), ("ntasks", INT), ("seed", ORB)]


class FERRETSESSION(C.Structure):
    _fields_ = [("cfg", FERRETCFG), ("M", I64), ("ot_limit", I64), ("round", I64), ("delta", ORB),
                ("boot_s", P), ("boot_r", P), ("malicious", C.c_int), ("check_ok", C.c_int),
                ("xprime", C.c_uint8 * 128), ("dig_a", C.c_uint8 * 32), ("dig_b", C.c_uint8 * 32)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_aes_expand.argtypes = [P, P]
        L.or_aes_ecb.argtypes = [P, P, P, C.c_size_t, INT]
        L.or_have_aesni.restype = INT
        L.or_prg.argtypes = [P, C.c_uint64, C.c_uint64, C.c_size_t, P]
        L.or_ggm_sender_tree.argtypes = [P, P, ORB, INT, P, P, P]
        L.or_ferret_mpcot_sender.argtypes = [P, I64, INT, ORB, I64, P, P, P, P, INT]
        L.or_ferret_mpcot_recver.argtypes = [I64, INT, P, P, P, I64, P, INT]
        L.or_sot_sender_expand.argtypes = [P, P, P, I64, INT, ORB, P, P, I64, P, P, P, INT]
        L.or_sot_recver_expand.argtypes = [P, P, I64, INT, P, P, P, P, P, I64, P, INT]
        L.or_qc_encode_brute.argtypes = [P, P, I64, P]
        L.or_qc_encode_fast.argtypes = [P, P, I64, P, INT]
        L.or_qc_encode_bits_brute.argtypes = [P, P, I64, P]
        L.or_qc_sparse.argtypes = [P, I64, P, I64, P]
        L.or_bpmul.argtypes = [P, P, I64, P]
        L.or_bit_transpose.argtypes = [P, I64, P]
        L.or_bit_untranspose.argtypes = [P, I64, P]
        L.or_ferret_lpn.argtypes = [P, I64, P, I64, ORB, INT, INT, INT]
        L.or_ferret_lpn_indices.argtypes = [I64, I64, ORB, INT, INT, P]
        L.or_sot_run.argtypes = [C.POINTER(SOTCFG), INT, P, P, P, P, P, P]
        L.or_ferret_run.argtypes = [C.POINTER(FERRETCFG), INT, P, P, P]
        L.or_ferret_session_init.argtypes = [C.POINTER(FERRETSESSION), C.POINTER(FERRETCFG)]
        L.or_ferret_session_init.restype = INT
        L.or_ferret_session_extend.argtypes = [C.POINTER(FERRETSESSION), INT, P, P]
        L.or_ferret_session_free.argtypes = [C.POINTER(FERRETSESSION)]
        L.or_ccrh.argtypes = [P, P, C.c_size_t]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P)


def blocks(n: int) -> np.ndarray:
    return np.zeros((n, 2), dtype=np.uint64)


def as_orb(b) -> ORB:
    b = np.asarray(b, dtype=np.uint64).reshape(2)
    return ORB(int(b[0]), int(b[1]))


def block_from_bytes(bs: bytes) -> np.ndarray:
    return np.frombuffer(bytes(bs), dtype=np.uint64).reshape(-1, 2).copy()


def block_to_bytes(b: np.ndarray) -> bytes:
    return np.ascontiguousarray(b, dtype=np.uint64).tobytes()


# ---- AES --------------------------------------------------------------------
def aes_expand(key: np.ndarray) -> np.ndarray:
    rk = np.zeros(176, dtype=np.uint8)
    k = np.ascontiguousarray(key, dtype=np.uint64).reshape(1, 2)
    lib().or_aes_expand(_p(k), _p(rk))
    return rk


def aes_ecb(key: np.ndarray, x: np.ndarray, impl: int = 1) -> np.ndarray:
    rk = aes_expand(key)
    x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 2)
    out = blocks(x.shape[0])
    lib().or_aes_ecb(_p(rk), _p(x), _p(out), x.shape[0], impl)
    return out


def have_aesni() -> bool:
    return bool(lib().or_have_aesni())


def prg(seed: np.ndarray, stream: int, start: int, count: int) -> np.ndarray:
    out = blocks(count)
    s = np.ascontiguousarray(seed, dtype=np.uint64).reshape(1, 2)
    lib().or_prg(_p(s), stream, start, count, _p(out))
    return out


# ---- GGM ---------------------------------------------------------------------
def ggm_sender_tree(k0, k1, root, h):
    rk0, rk1 = aes_expand(k0), aes_expand(k1)
    leaves = blocks(1 << h)
    ls, rs = blocks(h), blocks(h)
    lib().or_ggm_sender_tree(_p(rk0), _p(rk1), as_orb(root), h, _p(leaves), _p(ls), _p(rs))
    return leaves, ls, rs


def ferret_mpcot_sender(roots, t, h, delta, n_out, nthreads=1):
    roots = np.ascontiguousarray(roots, dtype=np.uint64)
    leaves = blocks(n_out)
    ls, rs, ss = blocks(t * h), blocks(t * h), blocks(t)
    lib().or_ferret_mpcot_sender(_p(roots), t, h, as_orb(delta), n_out, _p(leaves), _p(ls), _p(rs),
                                 _p(ss), nthreads)
    return leaves, ls, rs, ss


def ferret_mpcot_recver(t, h, b, csum, secret_sum, n_out, nthreads=1):
    b = np.ascontiguousarray(b, dtype=np.uint8)
    csum = np.ascontiguousarray(csum, dtype=np.uint64)
    secret_sum = np.ascontiguousarray(secret_sum, dtype=np.uint64)
    leaves = blocks(n_out)
    lib().or_ferret_mpcot_recver(t, h, _p(b), _p(csum), _p(secret_sum), n_out, _p(leaves), nthreads)
    return leaves


def sot_sender_expand(k0, k1, roots, t, h, delta, m0b, m1b, n_leaves, nthreads=1):
    rk0, rk1 = aes_expand(k0), aes_expand(k1)
    roots = np.ascontiguousarray(roots, dtype=np.uint64)
    m0b = np.ascontiguousarray(m0b, dtype=np.uint64)
    m1b = np.ascontiguousarray(m1b, dtype=np.uint64)
    leaves = blocks(n_leaves)
    M0, M1 = blocks((h + 1) * t), blocks((h + 1) * t)
    lib().or_sot_sender_expand(_p(rk0), _p(rk1), _p(roots), t, h, as_orb(delta), _p(m0b), _p(m1b),
                               n_leaves, _p(leaves), _p(M0), _p(M1), nthreads)
    return leaves, M0, M1


def sot_recver_expand(k0, k1, t, h, c, M0, M1, m0b, m1b, n_leaves, nthreads=1):
    rk0, rk1 = aes_expand(k0), aes_expand(k1)
    args = [np.ascontiguousarray(v, dtype=np.uint64) for v in (M0, M1, m0b, m1b)]
    c = np.ascontiguousarray(c, dtype=np.uint8)
    leaves = blocks(n_leaves)
    lib().or_sot_recver_expand(_p(rk0), _p(rk1), t, h, _p(c), *[_p(v) for v in args], n_leaves,
                               _p(leaves), nthreads)
    return leaves


# ---- QC ----------------------------------------------------------------------
def qc_encode_brute(x, a_bits, n):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    a_bits = np.ascontiguousarray(a_bits, dtype=np.uint8)
    y = blocks(n)
    lib().or_qc_encode_brute(_p(x), _p(a_bits), n, _p(y))
    return y


def qc_encode_fast(x, a_bits, n, nthreads=1):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    a_bits = np.ascontiguousarray(a_bits, dtype=np.uint8)
    y = blocks(n)
    lib().or_qc_encode_fast(_p(x), _p(a_bits), n, _p(y), nthreads)
    return y


def qc_encode_bits_brute(e_bits, a_bits, n):
    out = np.zeros((n + 7) // 8, dtype=np.uint8)
    lib().or_qc_encode_bits_brute(_p(np.ascontiguousarray(e_bits, dtype=np.uint8)),
                                  _p(np.ascontiguousarray(a_bits, dtype=np.uint8)), n, _p(out))
    return out


def qc_sparse(pos, a_bits, n):
    pos = np.ascontiguousarray(pos, dtype=np.int64)
    out = np.zeros((n + 7) // 8, dtype=np.uint8)
    lib().or_qc_sparse(_p(pos), pos.shape[0], _p(np.ascontiguousarray(a_bits, dtype=np.uint8)), n, _p(out))
    return out


def bpmul(a_words, b_words):
    a = np.ascontiguousarray(a_words, dtype=np.uint64)
    b = np.ascontiguousarray(b_words, dtype=np.uint64)
    m = a.shape[0]
    out = np.zeros(2 * m, dtype=np.uint64)
    lib().or_bpmul(_p(a), _p(b), m, _p(out))
    return out


# ---- transpose ---------------------------------------------------------------
def bit_transpose(x):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    N = x.shape[0]
    rows = np.zeros((128, N // 8), dtype=np.uint8)
    lib().or_bit_transpose(_p(x), N, _p(rows))
    return rows


def bit_untranspose(rows, N):
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    out = blocks(N)
    lib().or_bit_untranspose(_p(rows), N, _p(out))
    return out


# ---- Ferret LPN --------------------------------------------------------------
def ferret_lpn(nn, kk, seed, d=10, ntasks=1, nthreads=1):
    nn = np.array(nn, dtype=np.uint64, copy=True, order="C")
    kk = np.ascontiguousarray(kk, dtype=np.uint64)
    lib().or_ferret_lpn(_p(nn), nn.shape[0], _p(kk), kk.shape[0], as_orb(seed), d, ntasks, nthreads)
    return nn


def ferret_lpn_indices(n, k, seed, d=10, ntasks=1):
    idx = np.zeros(n * d, dtype=np.uint32)
    lib().or_ferret_lpn_indices(n, k, as_orb(seed), d, ntasks, _p(idx))
    return idx.reshape(n, d)


# ---- whole instances -----------------------------------------------------------
def sot_run(n, t, h, k0, k1, seed, nthreads=1, want_leaves=False):
    cfg = SOTCFG(n, t, h, as_orb(k0), as_orb(k1), as_orb(seed))
    v, w, d = blocks(n), blocks(n), blocks(1)
    b = np.zeros((n + 7) // 8, dtype=np.uint8)
    ls = blocks(2 * n) if want_leaves else None
    lr = blocks(2 * n) if want_leaves else None
    lib().or_sot_run(C.byref(cfg), nthreads, _p(v), _p(w), _p(b), _p(d), _p(ls) if want_leaves else None,
                     _p(lr) if want_leaves else None)
    if want_leaves:
        return v, w, b, d[0], ls, lr
    return v, w, b, d[0]


def ferret_run(n, k, t, h, seed, ntasks=1, nthreads=1):
    cfg = FERRETCFG(n, k, t, h, ntasks, as_orb(seed))
    s, r, d = blocks(n), blocks(n), blocks(1)
    lib().or_ferret_run(C.byref(cfg), nthreads, _p(s), _p(r), _p(d))
    return s, r, d[0]


class FerretSession:
    """The oracle's iterating Ferret (or_ferret_session_*): extend() returns the
    n sender and receiver outputs of one round."""

    def __init__(self, n, k, t, h, seed, ntasks=1, nthreads=1):
        self.cfg = FERRETCFG(n, k, t, h, ntasks, as_orb(seed))
        self.s = FERRETSESSION()
        self.nthreads = nthreads
        if lib().or_ferret_session_init(C.byref(self.s), C.byref(self.cfg)) != 0:
            raise ValueError("n must exceed k + h*t + 128")
        self.n = n

    @property
    def M(self):
        return self.s.M

    @property
    def ot_limit(self):
        return self.s.ot_limit

    @property
    def delta(self):
        return np.array([self.s.delta.lo, self.s.delta.hi], dtype=np.uint64)

    def bootstrap(self):
        bs = np.ctypeslib.as_array(C.cast(self.s.boot_s, C.POINTER(C.c_uint64)), shape=(self.s.M, 2)).copy()
        br = np.ctypeslib.as_array(C.cast(self.s.boot_r, C.POINTER(C.c_uint64)), shape=(self.s.M, 2)).copy()
        return bs, br

    def set_malicious(self, on=True):
        self.s.malicious = 1 if on else 0

    def check_result(self):
        """The last extend's consistency check: (ok, x', digest ALICE, digest BOB)."""
        return (bool(self.s.check_ok), np.frombuffer(bytes(self.s.xprime), dtype=np.uint8).copy(),
                bytes(self.s.dig_a), bytes(self.s.dig_b))

    def extend(self):
        o_s, o_r = blocks(self.n), blocks(self.n)
        lib().or_ferret_session_extend(C.byref(self.s), self.nthreads, _p(o_s), _p(o_r))
        return o_s, o_r

    def close(self):
        if self.s.boot_s:
            lib().or_ferret_session_free(C.byref(self.s))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ccrh(x):
    x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 2)
    out = blocks(x.shape[0])
    lib().or_ccrh(_p(x), _p(out), x.shape[0])
    return out


# ---- emp-ot Ferret session byte layouts --------------------------------------
# (ferret/emp-ot/emp-ot/ferret/ferret_cot.hpp; party 1 = ALICE, 2 = BOB)
def ferret_state_blob(party, n, t, k, delta, boot):
    """FerretCOT::assemble_state (ferret_cot.hpp:309-322): int64 party at 0,
    n at 8, t at 16, k at 24, the 16-byte Delta at 32, ot_pre_data from 48."""
    head = np.array([party, n, t, k], dtype="<i8").tobytes()
    d = np.ascontiguousarray(delta, dtype="<u8").reshape(2).tobytes()
    return head + d + np.ascontiguousarray(boot, dtype="<u8").tobytes()


def ferret_state_parse(blob, n_pre):
    """FerretCOT::disassemble_state (ferret_cot.hpp:324-342): the header fields,
    Delta and n_pre blocks."""
    head = np.frombuffer(blob[:32], dtype="<i8")
    delta = np.frombuffer(blob[32:48], dtype="<u8")
    boot = np.frombuffer(blob[48:48 + 16 * n_pre], dtype="<u8").reshape(n_pre, 2)
    return [int(v) for v in head], delta, boot


def ferret_pre_file(party, n, t, k, delta, boot):
    """write_pre_data128_to_file (ferret_cot.hpp:215-227): int64 party; the
    16-byte delta for ALICE only; int64 n, t, k; n_pre blocks."""
    out = np.array([party], dtype="<i8").tobytes()
    if party == 1:
        out += np.ascontiguousarray(delta, dtype="<u8").reshape(2).tobytes()
    out += np.array([n, t, k], dtype="<i8").tobytes()
    return out + np.ascontiguousarray(boot, dtype="<u8").tobytes()


# ---- malicious consistency check (emp-tool arithmetic restated) ---------------
def sha256(msg: bytes) -> bytes:
    out = (C.c_uint8 * 32)()
    buf = C.create_string_buffer(bytes(msg), len(msg))
    lib().or_sha256(buf, C.c_size_t(len(msg)), out)
    return bytes(out)


def gf128_mul(a, b):
    L = lib()
    L.or_gf128_mul.restype = ORB
    L.or_gf128_mul.argtypes = [ORB, ORB]
    r = L.or_gf128_mul(as_orb(a), as_orb(b))
    return np.array([r.lo, r.hi], dtype=np.uint64)


def consist_trees(leaves_s, leaves_r, n, t, h, ssum, bits, nthreads=8):
    """Per tree (V, W, chi_alpha) of the f2k consistency check."""
    ls = np.ascontiguousarray(leaves_s, dtype=np.uint64).reshape(-1, 2)
    lr = np.ascontiguousarray(leaves_r, dtype=np.uint64).reshape(-1, 2)
    ss = np.ascontiguousarray(ssum, dtype=np.uint64).reshape(-1, 2)
    bb = np.ascontiguousarray(bits, dtype=np.uint8)
    V, W, ca = blocks(t), blocks(t), blocks(t)
    lib().or_ferret_consist_trees(_p(ls), _p(lr), C.c_int64(n), C.c_int64(t), C.c_int(h), _p(ss), _p(bb), _p(V),
                                  _p(W), _p(ca), C.c_int(nthreads))
    return V, W, ca


def consist_final(V, W, chi_alpha, pre_s, pre_r, delta):
    """consistency_check_f2k for both parties: (ok, x', digest ALICE, digest BOB)."""
    V, W, ca = (np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 2) for x in (V, W, chi_alpha))
    ps = np.ascontiguousarray(pre_s, dtype=np.uint64).reshape(-1, 2)
    pr = np.ascontiguousarray(pre_r, dtype=np.uint64).reshape(-1, 2)
    xp, da, db = (C.c_uint8 * 128)(), (C.c_uint8 * 32)(), (C.c_uint8 * 32)()
    ok = lib().or_ferret_consist_final(_p(V), _p(W), _p(ca), C.c_int64(V.shape[0]), _p(ps), _p(pr), as_orb(delta),
                                       xp, da, db)
    return bool(ok), np.frombuffer(bytes(xp), dtype=np.uint8).copy(), bytes(da), bytes(db)


# ---- online derandomisation ------------------------------------------------------
def cot_online(data_s, data_r, b, delta):
    """(data_s with bo*Delta applied, bo)"""
    ds = np.ascontiguousarray(data_s, dtype=np.uint64).reshape(-1, 2).copy()
    dr = np.ascontiguousarray(data_r, dtype=np.uint64).reshape(-1, 2)
    bb = np.ascontiguousarray(b, dtype=np.uint8)
    bo = np.zeros(ds.shape[0], dtype=np.uint8)
    lib().or_cot_online(_p(ds), _p(dr), _p(bb), as_orb(delta), _p(bo), C.c_int64(ds.shape[0]))
    return ds, bo


def cot_to_ot(data_s, data_r, delta, s, g0, r, m0, m1, nthreads=8):
    ds = np.ascontiguousarray(data_s, dtype=np.uint64).reshape(-1, 2)
    dr = np.ascontiguousarray(data_r, dtype=np.uint64).reshape(-1, 2)
    n = ds.shape[0]
    rr = np.ascontiguousarray(r, dtype=np.uint8)
    a0 = np.ascontiguousarray(m0, dtype=np.uint64).reshape(-1, 2)
    a1 = np.ascontiguousarray(m1, dtype=np.uint64).reshape(-1, 2)
    pad, out = blocks(2 * n), blocks(n)
    lib().or_cot_to_ot(_p(ds), _p(dr), as_orb(delta), as_orb(s), C.c_uint64(g0), _p(rr), _p(a0), _p(a1), _p(pad),
                       _p(out), C.c_int64(n), C.c_int(nthreads))
    return pad, out


def cot_to_rot(data_s, data_r, delta, s, g0, nthreads=8):
    ds = np.ascontiguousarray(data_s, dtype=np.uint64).reshape(-1, 2)
    dr = np.ascontiguousarray(data_r, dtype=np.uint64).reshape(-1, 2)
    n = ds.shape[0]
    d0, d1, out = blocks(n), blocks(n), blocks(n)
    lib().or_cot_to_rot(_p(ds), _p(dr), as_orb(delta), as_orb(s), C.c_uint64(g0), _p(d0), _p(d1), _p(out),
                        C.c_int64(n), C.c_int(nthreads))
    return d0, d1, out
