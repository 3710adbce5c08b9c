"""Pin the oracle's GGM / puncturing restatements.

Checks are independent of the oracle's own loop structure: a recursive
brute-force GGM in Python (AES from the FIPS-pinned oracle AES), the level-sum
definition, and the correlation invariant w_i = v_i ^ [i = alpha] * Delta that
both receiver restatements must satisfy.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import rand_blocks, rng

FERRET_K0 = np.array([0, 0], dtype=np.uint64)
FERRET_K1 = np.array([1, 0], dtype=np.uint64)  # makeBlock(0, 1)
LSB_CLEAR = np.array([0xFFFFFFFFFFFFFFFE, 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)


def brute_tree(k0, k1, root, h):
    """Recursive definition: node -> (AES_k0(node)^node, AES_k1(node)^node)."""
    level = [np.asarray(root, dtype=np.uint64).reshape(2)]
    sums = []
    for _ in range(h):
        nxt = []
        for p in level:
            l = po.aes_ecb(k0, p.reshape(1, 2))[0] ^ p
            r = po.aes_ecb(k1, p.reshape(1, 2))[0] ^ p
            nxt += [l, r]
        level = nxt
        arr = np.array(level)
        sums.append((np.bitwise_xor.reduce(arr[0::2], axis=0), np.bitwise_xor.reduce(arr[1::2], axis=0)))
    return np.array(level), sums


@pytest.mark.parametrize("h", [1, 2, 5])
def test_tree_matches_recursive_definition(h):
    r = rng(10 + h)
    k0, k1, root = rand_blocks(r, 1), rand_blocks(r, 1), rand_blocks(r, 1)
    leaves, ls, rs = po.ggm_sender_tree(k0, k1, root, h)
    bl, bs = brute_tree(k0, k1, root, h)
    assert np.array_equal(leaves, bl)
    for d in range(h):
        assert np.array_equal(ls[d], bs[d][0]) and np.array_equal(rs[d], bs[d][1])


def _alpha_to_choice(alpha, h):
    # alpha bit (h-1-d) = 1 - c[d]  (spcot_recver.h:33-41, silent/recver.cu:78-87)
    return np.array([1 - ((alpha >> (h - 1 - d)) & 1) for d in range(h)], dtype=np.uint8)


@pytest.mark.parametrize("h,t,trunc", [(1, 3, 0), (4, 5, 0), (7, 3, 0), (6, 4, 37)])
def test_ferret_mpcot_correlation(h, t, trunc):
    r = rng(100 + h)
    roots = rand_blocks(r, t)
    delta = rand_blocks(r, 1)[0] | np.array([1, 0], dtype=np.uint64)
    n_out = t << h if trunc == 0 else (t << h) - trunc
    v, ls, rs, ss = po.ferret_mpcot_sender(roots, t, h, delta, n_out)
    alphas = r.integers(0, 1 << h, size=t)
    b = np.concatenate([_alpha_to_choice(int(a), h) for a in alphas])
    csum = np.where(b.reshape(t * h, 1) == 1, rs, ls)
    w = po.ferret_mpcot_recver(t, h, b, csum, ss, n_out)
    e = np.zeros(n_out, dtype=bool)
    for tr, a in enumerate(alphas):
        if tr * (1 << h) + a < n_out:
            e[tr * (1 << h) + a] = True
    expect = v ^ np.where(e[:, None], delta, 0).astype(np.uint64)
    assert np.array_equal(w, expect)
    # sender leaves have LSB 0; receiver LSB is the choice bit
    assert np.all((v[:, 0] & 1) == 0)
    assert np.array_equal((w[:, 0] & 1).astype(bool), e)
    # sums are over the unmasked tree: compare with the recursive definition on tree 0
    bl, bs = brute_tree(FERRET_K0, FERRET_K1, roots[0], h) if h <= 4 else (None, None)
    if bl is not None:
        assert np.array_equal(v[: 1 << h], bl & LSB_CLEAR)
        assert np.array_equal(ls[h - 1], bs[h - 1][0])
        expect_ss = np.bitwise_xor.reduce(bl & LSB_CLEAR, axis=0) ^ delta
        assert np.array_equal(ss[0], expect_ss)


@pytest.mark.parametrize("h,t", [(1, 2), (5, 3), (8, 2)])
def test_sot_expand_correlation(h, t):
    r = rng(200 + h)
    k0, k1 = rand_blocks(r, 1), rand_blocks(r, 1)
    roots = rand_blocks(r, t)
    delta = rand_blocks(r, 1)[0] | np.array([1, 0], dtype=np.uint64)
    m0b, m1b = rand_blocks(r, h * t), rand_blocks(r, h * t)
    n = t << h
    v, M0, M1 = po.sot_sender_expand(k0, k1, roots, t, h, delta, m0b, m1b, n)
    alphas = r.integers(0, 1 << h, size=t)
    c = np.zeros((h, t), dtype=np.uint8)
    for tr, a in enumerate(alphas):
        c[:, tr] = _alpha_to_choice(int(a), h)
    w = po.sot_recver_expand(k0, k1, t, h, c.reshape(-1), M0, M1, m0b, m1b, n)
    expect = v.copy()
    for tr, a in enumerate(alphas):
        expect[tr * (1 << h) + a] ^= delta
    assert np.array_equal(w, expect)
    # sender leaves equal the plain tree
    for tr in range(t):
        leaves, _, _ = po.ggm_sender_tree(k0, k1, roots[tr], h)
        assert np.array_equal(v[tr << h:(tr + 1) << h], leaves)


def test_ferret_receiver_wrong_sum_breaks_correlation():
    # a corrupted chosen sum must not yield a valid correlation (mistake planted
    # in the oracle inputs, on the CPU)
    r = rng(7)
    h, t = 4, 2
    roots = rand_blocks(r, t)
    delta = np.array([3, 5], dtype=np.uint64)
    v, ls, rs, ss = po.ferret_mpcot_sender(roots, t, h, delta, t << h)
    b = np.zeros(t * h, dtype=np.uint8)
    csum = ls.copy()
    csum[1] ^= np.array([2, 0], dtype=np.uint64)
    w = po.ferret_mpcot_recver(t, h, b, csum, ss, t << h)
    diff = np.any(w != v, axis=1).sum()
    assert diff > t  # more than the t punctured positions differ
