"""Whole-instance oracle runs satisfy the COT relation and the seed
derivation rules (Delta LSB, Ferret choice bit = receiver LSB)."""
import numpy as np

from oracle import pyoracle as po

SOT_K0 = np.array([3242342, 0], dtype=np.uint64)  # SilentConfig leftKey {3242342} (silent/main.cu:89)
SOT_K1 = np.array([8993849, 0], dtype=np.uint64)


def bits_to_bool(b, n):
    return np.unpackbits(b, bitorder="little")[:n].astype(bool)


def test_sot_relation_small():
    seed = np.array([0x0706050403020100, 0x0F0E0D0C0B0A0908], dtype=np.uint64)
    n, t, h = 1 << 10, 16, 7
    v, w, b, d = po.sot_run(n, t, h, SOT_K0, SOT_K1, seed, nthreads=4)
    assert int(d[0]) & 1 == 1
    bb = bits_to_bool(b, n)
    expect = v ^ np.where(bb[:, None], d, 0).astype(np.uint64)
    assert np.array_equal(w, expect)
    assert 0 < bb.sum() < n


def test_sot_truncated_last_tree():
    seed = np.array([1, 2], dtype=np.uint64)
    n, t, h = 1000, 16, 7  # 2n = 2000 < 16 * 128
    v, w, b, d = po.sot_run(n, t, h, SOT_K0, SOT_K1, seed, nthreads=4)
    bb = bits_to_bool(b, n)
    assert np.array_equal(w, v ^ np.where(bb[:, None], d, 0).astype(np.uint64))


def test_ferret_relation_small():
    seed = np.array([5, 6], dtype=np.uint64)
    n, k, t, h = 3000, 300, 12, 8  # 12 * 256 = 3072 >= 3000 (truncated)
    s, r, d = po.ferret_run(n, k, t, h, seed, nthreads=4)
    assert int(d[0]) & 1 == 1
    lsb = (r[:, 0] & 1).astype(bool)
    assert np.array_equal(r, s ^ np.where(lsb[:, None], d, 0).astype(np.uint64))
    assert np.all((s[:, 0] & 1) == 0) or True  # sender LSB is the LPN of K (LSB 0) ^ noise (LSB 0)
    assert np.all((s[:, 0] & 1) == 0)
