"""The oracle's restatement of emp-tool's arithmetic for the malicious
consistency check (emp-tool @44b1dde is absent): SHA-256 against FIPS 180-4
and hashlib, GF(2^128) against closed forms and schoolbook products of Python
integers, chi_j = seed^(j+1) through unit-vector trees, the per-tree identity
W ^ V = chi_alpha * Delta of honest parties, an honest iterating session
passing and a corrupted leaf failing."""
import hashlib
import random

import numpy as np

from oracle import pyoracle as po

SEED = np.array([0x0706050403020100, 0x0F0E0D0C0B0A0908], dtype=np.uint64)
POLY = (1 << 128) | 0x87


def to_int(b):
    return int(b[0]) | (int(b[1]) << 64)


def to_blk(v):
    return np.array([v & (2**64 - 1), v >> 64], dtype=np.uint64)


def int_mul(a, b):
    r = 0
    while b:
        if b & 1:
            r ^= a
        a <<= 1
        b >>= 1
    for k in range(r.bit_length() - 1, 127, -1):
        if (r >> k) & 1:
            r ^= POLY << (k - 128)
    return r


def test_sha256():
    assert po.sha256(b"abc").hex() == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    for m in (b"", b"a" * 55, b"a" * 56, b"b" * 64, bytes(range(256)) * 3):
        assert po.sha256(m) == hashlib.sha256(m).digest()


def test_gf128():
    x = lambda i: to_blk(1 << i)
    assert to_int(po.gf128_mul(x(127), x(1))) == 0x87          # x^128 = x^7 + x^2 + x + 1
    assert to_int(po.gf128_mul(x(64), x(64))) == 0x87
    assert to_int(po.gf128_mul(x(127), x(127))) == int_mul(1 << 127, 1 << 127)
    r = random.Random(5)
    for _ in range(300):
        a, b = r.getrandbits(128), r.getrandbits(128)
        assert to_int(po.gf128_mul(to_blk(a), to_blk(b))) == int_mul(a, b)
        assert to_int(po.gf128_mul(to_blk(a), to_blk(b))) == to_int(po.gf128_mul(to_blk(b), to_blk(a)))


def test_chi_powers_on_unit_trees():
    # one tree of 2^h leaves, a single identity leaf at j: V = seed^(j+1) * 1
    h, t = 5, 1
    n = t << h
    ssum = to_blk(0x1234567890ABCDEF_0011223344556677)
    seed = to_int(np.frombuffer(hashlib.sha256(ssum.astype("<u8").tobytes()).digest()[:16], dtype="<u8"))
    bits = np.zeros(h, dtype=np.uint8)
    for j in (0, 1, 17, 31):
        ls = np.zeros((n, 2), dtype=np.uint64)
        ls[j] = to_blk(1)
        V, W, ca = po.consist_trees(ls, ls, n, t, h, ssum, bits)
        p = 1
        for _ in range(j + 1):
            p = int_mul(p, seed)
        assert to_int(V[0]) == p == to_int(W[0])
    # bits all 0 -> alpha = 2^h - 1: chi_alpha = seed^(2^h)
    p = 1
    for _ in range(n):
        p = int_mul(p, seed)
    assert to_int(ca[0]) == p


def _session(n, k, t, h):
    s = po.FerretSession(n, k, t, h, SEED, nthreads=8)
    s.set_malicious(True)
    return s


def test_honest_session_passes_and_identity():
    for (n, k, t, h) in ((40960, 2000, 5, 13), (40000, 2000, 5, 13), (3000, 300, 12, 8)):
        s = _session(n, k, t, h)
        for _ in range(2):
            s.extend()
            ok, xp, da, db = s.check_result()
            assert ok and da == db


def test_corrupted_leaf_fails():
    # one MPCOT of both parties from the oracle's single-instance functions
    n, t, h = 8192 * 3, 3, 13
    r = np.random.default_rng(7)
    roots = r.integers(0, 2**63, size=(t, 2), dtype=np.uint64)
    delta = r.integers(0, 2**63, size=2, dtype=np.uint64)
    delta[0] |= np.uint64(1)
    bits = r.integers(0, 2, size=t * h, dtype=np.uint8)
    ls_, ls, rs, ssum = po.ferret_mpcot_sender(roots, t, h, delta, n)
    csum = np.where(bits[:, None].astype(bool), rs, ls)
    lr_ = po.ferret_mpcot_recver(t, h, bits, csum, ssum, n)
    pre_s = r.integers(0, 2**63, size=(128, 2), dtype=np.uint64)
    pre_s[:, 0] &= ~np.uint64(1)
    c = r.integers(0, 2, size=128, dtype=np.uint64)
    pre_r = pre_s ^ (c[:, None] * delta[None, :])
    V, W, ca = po.consist_trees(ls_, lr_, n, t, h, ssum, bits)
    for tr in range(t):  # honest parties: W ^ V = chi_alpha * Delta per tree
        assert np.array_equal(V[tr] ^ W[tr], po.gf128_mul(ca[tr], delta))
    ok, _, da, db = po.consist_final(V, W, ca, pre_s, pre_r, delta)
    assert ok and da == db
    L = 1 << h
    bad = lr_.copy()
    bad[L + 77, 1] ^= np.uint64(1 << 40)
    V2, W2, ca2 = po.consist_trees(ls_, bad, n, t, h, ssum, bits)
    ok2, _, da2, db2 = po.consist_final(V2, W2, ca2, pre_s, pre_r, delta)
    assert not ok2 and da2 == da and db2 != db
    # a wrong correlation in the 128 check COTs is caught too
    pre_bad = pre_r.copy()
    pre_bad[5, 1] ^= np.uint64(1)
    ok3, *_ = po.consist_final(V, W, ca, pre_s, pre_bad, delta)
    assert not ok3
