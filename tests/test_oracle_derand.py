"""The oracle's online derandomisation (ferret_cot.hpp:277-307, cot.h:16-91):
chosen-bit COTs satisfy M = K ^ b*Delta, chosen-message OTs give the receiver
m_r and random OTs data_r, on COTs of a known correlation; MITCCRH is AES
under s ^ makeBlock(g, 0) xor the input (emp-tool @44b1dde restated: the key
derivation is parity unpinned)."""
import numpy as np

from oracle import pyoracle as po


def _cots(n, seed=3):
    r = np.random.default_rng(seed)
    delta = r.integers(0, 2**63, size=2, dtype=np.uint64)
    delta[0] |= np.uint64(1)
    K = r.integers(0, 2**63, size=(n, 2), dtype=np.uint64)
    K[:, 0] &= ~np.uint64(1)
    c = r.integers(0, 2, size=n, dtype=np.uint64)
    M = K ^ (c[:, None] * delta[None, :])
    return r, delta, K, M


def test_online_correlation():
    r, delta, K, M = _cots(1000)
    b = r.integers(0, 2, size=1000, dtype=np.uint8)
    K2, bo = po.cot_online(K, M, b, delta)
    assert np.array_equal(bo, (b ^ (M[:, 0] & 1).astype(np.uint8)))
    assert np.array_equal(M, K2 ^ (b.astype(np.uint64)[:, None] * delta[None, :]))


def test_ot_and_rot():
    n = 777
    r, delta, K, M = _cots(n, 5)
    b = r.integers(0, 2, size=n, dtype=np.uint8)
    K2, _ = po.cot_online(K, M, b, delta)
    m0 = r.integers(0, 2**63, size=(n, 2), dtype=np.uint64)
    m1 = r.integers(0, 2**63, size=(n, 2), dtype=np.uint64)
    s = np.array([0x1111, 0x2222], dtype=np.uint64)
    pad, out = po.cot_to_ot(K2, M, delta, s, 16, b, m0, m1)
    assert np.array_equal(out, np.where(b[:, None].astype(bool), m1, m0))
    d0, d1, outr = po.cot_to_rot(K2, M, delta, s, 16)
    assert np.array_equal(outr, np.where(b[:, None].astype(bool), d1, d0))
    # H_i(x) = AES_{s ^ makeBlock(16 + i, 0)}(x) ^ x
    for i in (0, 5, 776):
        key = np.array([s[0], s[1] ^ np.uint64(16 + i)], dtype=np.uint64)
        assert np.array_equal(d0[i], po.aes_ecb(key, K2[i:i + 1])[0] ^ K2[i])
        assert np.array_equal(pad[2 * i + 1], po.aes_ecb(key, (K2[i] ^ delta)[None])[0] ^ K2[i] ^ delta ^ m1[i])
