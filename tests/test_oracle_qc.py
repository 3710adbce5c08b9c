"""Pin the oracle's quasi-cyclic encoder and bit transpose.

QC: the O(n^2) brute force and the CLMUL/Karatsuba product are two
independent algorithms; both must satisfy the closed-form special cases
a = 1 (y = x0 ^ x1), a = x^s (y_i = x0_i ^ x1_{(i-s) mod n}), linearity in x,
and agree with each other.  The bits-only and block versions agree on every
row.  Transpose: per-bit brute force and involution.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import rand_blocks, rng


def bits_from_int_positions(n, ones):
    a = np.zeros((n + 7) // 8, dtype=np.uint8)
    for j in ones:
        a[j >> 3] |= 1 << (j & 7)
    return a


def rand_bits(r, n):
    a = r.integers(0, 256, size=(n + 7) // 8, dtype=np.uint8)
    if n % 8:
        a[-1] &= (1 << (n % 8)) - 1
    return a


@pytest.mark.parametrize("n", [1, 2, 7, 64, 100, 129])
def test_special_cases(n):
    r = rng(n)
    x = rand_blocks(r, 2 * n)
    one = bits_from_int_positions(n, [0])
    assert np.array_equal(po.qc_encode_brute(x, one, n), x[:n] ^ x[n:])
    assert np.array_equal(po.qc_encode_fast(x, one, n), x[:n] ^ x[n:])
    s = (3 * n) // 4
    xs = bits_from_int_positions(n, [s])
    expect = x[:n] ^ x[n + (np.arange(n) - s) % n]
    assert np.array_equal(po.qc_encode_brute(x, xs, n), expect)
    assert np.array_equal(po.qc_encode_fast(x, xs, n), expect)


@pytest.mark.parametrize("n", [3, 64, 65, 500, 1024, 4099])
def test_brute_vs_karatsuba_and_linearity(n):
    r = rng(10 + n)
    a = rand_bits(r, n)
    x1, x2 = rand_blocks(r, 2 * n), rand_blocks(r, 2 * n)
    y1 = po.qc_encode_brute(x1, a, n)
    assert np.array_equal(y1, po.qc_encode_fast(x1, a, n, nthreads=4))
    y2 = po.qc_encode_fast(x2, a, n, nthreads=4)
    assert np.array_equal(po.qc_encode_fast(x1 ^ x2, a, n, nthreads=4), y1 ^ y2)


def test_bits_version_matches_block_rows():
    r = rng(5)
    n = 300
    a = rand_bits(r, n)
    x = rand_blocks(r, 2 * n)
    y = po.qc_encode_brute(x, a, n)
    for k in (0, 1, 63, 64, 127):
        word, bit = divmod(k, 64)
        e = ((x[:, word] >> np.uint64(bit)) & np.uint64(1)).astype(np.uint8)
        eb = np.packbits(e, bitorder="little")
        out = po.qc_encode_bits_brute(eb, a, n)
        got = np.unpackbits(out, bitorder="little")[:n]
        assert np.array_equal(got, ((y[:, word] >> np.uint64(bit)) & np.uint64(1)).astype(np.uint8))


def test_sparse_is_code_of_unit_vectors():
    r = rng(6)
    n = 257
    a = rand_bits(r, n)
    pos = np.array([3, 100, 256, 257, 300, 513, 2 * n + 5], dtype=np.int64)  # last one ignored
    e = np.zeros(2 * n, dtype=np.uint8)
    for p in pos:
        if p < 2 * n:
            e[p] ^= 1
    eb = np.packbits(e, bitorder="little")
    assert np.array_equal(po.qc_sparse(pos, a, n), po.qc_encode_bits_brute(eb, a, n))


def test_bpmul_small_closed_forms():
    # (x^63 + 1)^2 = x^126 + 1 over GF(2)
    a = np.array([(1 << 63) | 1], dtype=np.uint64)
    assert list(po.bpmul(a, a)) == [1, 1 << 62]
    r = rng(8)
    for m in (1, 3, 17, 40):
        a = r.integers(0, 2**63, size=m, dtype=np.uint64)
        b = r.integers(0, 2**63, size=m, dtype=np.uint64)
        ai = sum(int(v) << (64 * i) for i, v in enumerate(a))
        bi = sum(int(v) << (64 * i) for i, v in enumerate(b))
        p = 0
        while bi:
            if bi & 1:
                p ^= ai
            ai <<= 1
            bi >>= 1
        got = sum(int(v) << (64 * i) for i, v in enumerate(po.bpmul(a, b)))
        assert got == p


@pytest.mark.parametrize("N", [8, 64, 96])
def test_transpose_bits_and_involution(N):
    r = rng(N)
    x = rand_blocks(r, N)
    rows = po.bit_transpose(x)
    xb = np.unpackbits(x.view(np.uint8).reshape(N, 16), axis=1, bitorder="little")  # [m][k]
    rb = np.unpackbits(rows, axis=1, bitorder="little")  # [k][m]
    assert np.array_equal(rb, xb.T)
    assert np.array_equal(po.bit_untranspose(rows, N), x)
