"""Pin the oracle's Ferret 10-local LPN (lpn_f2.h:46-79).

emp-tool is not available, so the index rule is checked against a second,
independent restatement written from lpn_f2.h in Python (AES from the
FIPS-pinned oracle AES), plus structural invariants of the task partition
and the linearity of the code in the noise and in kk.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import rand_blocks, rng


def py_indices(n, k, seed, ntasks=1, d=10):
    mask = 1
    while mask < k:
        mask = (mask << 1) | 1
    out = np.zeros((n, d), dtype=np.int64)
    width = n // ntasks
    for tk in range(ntasks):
        start = tk * width
        end = n if tk == ntasks - 1 else (tk + 1) * width
        j = start
        while j < end - 4:  # __compute4
            ctr = np.array([[m, j] for m in range(d)], dtype=np.uint64)
            words = po.aes_ecb(seed, ctr).view(np.uint32).reshape(-1)
            for q in range(4):
                for jj in range(d):
                    idx = int(words[q * d + jj]) & mask
                    out[j + q, jj] = idx - k if idx >= k else idx
            j += 4
        while j < end:  # __compute1
            nb = d // 4 + (d % 4 != 0)
            ctr = np.array([[m, j] for m in range(nb)], dtype=np.uint64)
            words = po.aes_ecb(seed, ctr).view(np.uint32).reshape(-1)
            for jj in range(d):
                out[j, jj] = int(words[jj]) % k
            j += 1
    return out


@pytest.mark.parametrize("n,k,ntasks", [(1, 7, 1), (5, 100, 1), (8, 37, 1), (33, 1000, 1), (50, 64, 3),
                                        (103, 513, 4)])
def test_indices_match_second_restatement(n, k, ntasks):
    seed = rand_blocks(rng(n + k), 1)
    a = po.ferret_lpn_indices(n, k, seed, ntasks=ntasks)
    b = py_indices(n, k, seed, ntasks=ntasks)
    assert np.array_equal(a.astype(np.int64), b)
    assert a.max() < k


def test_partition_changes_only_boundary_rows():
    # rows far from task boundaries use the same __compute4 groups when the
    # boundaries are 4-aligned relative to each other
    n, k = 64, 1000
    seed = rand_blocks(rng(5), 1)
    one = po.ferret_lpn_indices(n, k, seed, ntasks=1)
    two = po.ferret_lpn_indices(n, k, seed, ntasks=2)  # tasks [0,32) [32,64)
    # task 0 of the 2-task split: rows 0..27 are __compute4 with the same
    # counters as in the 1-task run
    assert np.array_equal(one[:28], two[:28])
    assert not np.array_equal(one[28:32], two[28:32])  # __compute1 tail of task 0


def test_linear_in_noise_and_kk():
    r = rng(9)
    n, k = 200, 300
    seed = rand_blocks(r, 1)
    kk1, kk2 = rand_blocks(r, k), rand_blocks(r, k)
    z1, z2 = rand_blocks(r, n), rand_blocks(r, n)
    f = lambda z, kk: po.ferret_lpn(z, kk, seed)
    assert np.array_equal(f(z1 ^ z2, kk1), f(z1, kk1) ^ z2)
    assert np.array_equal(f(z1, kk1 ^ kk2), f(z1, kk1) ^ f(np.zeros_like(z1), kk2))
    # and equals the explicit gather with the checked indices
    idx = py_indices(n, k, seed)
    expect = z1.copy()
    for i in range(n):
        for j in range(10):
            expect[i] ^= kk1[idx[i, j]]
    assert np.array_equal(f(z1, kk1), expect)
