"""GPU parity: exact QC encoder (additive FFT and, for small codes, the int8
matrix-core product), sparse encoder and the bit transpose against the CPU
oracle, bit-exact."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import dev_blocks, rand_blocks, rng, to_dev, to_np

pytestmark = pytest.mark.gpu


def rand_bits(r, n):
    a = r.integers(0, 256, size=(n + 7) // 8, dtype=np.uint8)
    if n % 8:
        a[-1] &= (1 << (n % 8)) - 1
    return a


def dev_bits(a, n):
    import torch

    words = (n + 31) // 32 + 2
    buf = np.zeros(words * 4, dtype=np.uint8)
    buf[: a.shape[0]] = a
    return torch.from_numpy(buf).cuda()


def qc_gpu(ctx, x, a, n):
    import torch

    ahat = torch.zeros(ctx.qc_ahat_len(n), dtype=torch.int32, device="cuda")
    scratch = torch.zeros(ctx.qc_scratch_bytes(n), dtype=torch.uint8, device="cuda")
    ctx.qc_prepare(n, dev_bits(a, n), ahat, scratch)
    y = dev_blocks(n)
    ctx.qc_encode(n, ahat, to_dev(x), y, scratch)
    return to_np(y)[:n]


@pytest.mark.parametrize("n", [1, 5, 16, 17, 100, 1000, 4096, 4099])
def test_qc_small_vs_brute(ctx, n):
    print(f"qc brute n={n}", flush=True)
    r = rng(n)
    a, x = rand_bits(r, n), rand_blocks(r, 2 * n)
    assert np.array_equal(qc_gpu(ctx, x, a, n), po.qc_encode_brute(x, a, n))


# 2^20 and 2^22 + 7 also cover the row conversion with several strings per
# workgroup at 2^17 .. 2^19-bit strings (folded and general output)
@pytest.mark.parametrize("n", [1 << 14, 65537, 1 << 18, 1 << 20, (1 << 22) + 7])
def test_qc_vs_karatsuba(ctx, n):
    print(f"qc karatsuba n={n}", flush=True)
    r = rng(n + 1)
    a, x = rand_bits(r, n), rand_blocks(r, 2 * n)
    assert np.array_equal(qc_gpu(ctx, x, a, n), po.qc_encode_fast(x, a, n, nthreads=8))


def test_qc_special_cases(ctx):
    n = 1 << 12
    r = rng(3)
    x = rand_blocks(r, 2 * n)
    one = np.zeros(n // 8, dtype=np.uint8)
    one[0] = 1
    assert np.array_equal(qc_gpu(ctx, x, one, n), x[:n] ^ x[n:])
    s = 1234
    xs = np.zeros(n // 8, dtype=np.uint8)
    xs[s >> 3] |= 1 << (s & 7)
    assert np.array_equal(qc_gpu(ctx, x, xs, n), x[:n] ^ x[n + (np.arange(n) - s) % n])


@pytest.fixture
def knob_mfma():
    """Sets the matrix-core crossover for one test and restores the rule after."""
    from cuot_amd import tune

    def set_(v):
        tune.set_knobs(qc_mfma_max=v)

    yield set_
    tune.set_knobs()


# small codes on the int8 matrix cores: one slice (n <= 256, direct write) and
# several (atomic XOR), n around the 32-row and 16-k tile edges
@pytest.mark.parametrize("n", [17, 31, 32, 33, 255, 256, 257, 1000, 4099, 10007])
def test_qc_mfma_vs_brute(ctx, knob_mfma, n):
    print(f"qc mfma brute n={n}", flush=True)
    knob_mfma(1 << 16)
    r = rng(n + 11)
    a, x = rand_bits(r, n), rand_blocks(r, 2 * n)
    assert np.array_equal(qc_gpu(ctx, x, a, n), po.qc_encode_brute(x, a, n))


# the two encoders on one prepared code (prepare keeps both forms), against
# the oracle's Karatsuba product, up to the largest matrix-core size
@pytest.mark.parametrize("n", [12345, 1 << 15, 40000, 1 << 16])
def test_qc_mfma_and_fft_agree(ctx, knob_mfma, n):
    import torch

    print(f"qc mfma/fft n={n}", flush=True)
    r = rng(n + 13)
    a, x = rand_bits(r, n), rand_blocks(r, 2 * n)
    want = po.qc_encode_fast(x, a, n, nthreads=8)
    ahat = torch.zeros(ctx.qc_ahat_len(n), dtype=torch.int32, device="cuda")
    scratch = torch.zeros(ctx.qc_scratch_bytes(n), dtype=torch.uint8, device="cuda")
    ctx.qc_prepare(n, dev_bits(a, n), ahat, scratch)
    xd = to_dev(x)
    for lim in (1 << 16, -1):
        knob_mfma(lim)
        y = dev_blocks(n)
        ctx.qc_encode(n, ahat, xd, y, scratch)
        assert np.array_equal(to_np(y)[:n], want), lim


# y over the first half of x (the C ABI allows the alias), both output modes
@pytest.mark.parametrize("n", [200, 5000])
def test_qc_mfma_in_place(ctx, knob_mfma, n):
    import torch

    knob_mfma(1 << 16)
    r = rng(n + 17)
    a, x = rand_bits(r, n), rand_blocks(r, 2 * n)
    ahat = torch.zeros(ctx.qc_ahat_len(n), dtype=torch.int32, device="cuda")
    scratch = torch.zeros(ctx.qc_scratch_bytes(n), dtype=torch.uint8, device="cuda")
    ctx.qc_prepare(n, dev_bits(a, n), ahat, scratch)
    xd = to_dev(x)
    ctx.qc_encode(n, ahat, xd, xd, scratch)
    assert np.array_equal(to_np(xd)[:n], po.qc_encode_brute(x, a, n))


def test_qc_mfma_special_cases(ctx, knob_mfma):
    knob_mfma(1 << 16)
    n = 3000
    r = rng(5)
    x = rand_blocks(r, 2 * n)
    one = np.zeros((n + 7) // 8, dtype=np.uint8)
    one[0] = 1
    assert np.array_equal(qc_gpu(ctx, x, one, n), x[:n] ^ x[n:])
    s = 2999
    xs = np.zeros((n + 7) // 8, dtype=np.uint8)
    xs[s >> 3] |= 1 << (s & 7)
    assert np.array_equal(qc_gpu(ctx, x, xs, n), x[:n] ^ x[n + (np.arange(n) - s) % n])
    ones = rand_bits(r, n)
    ones[:] = 0xFF
    ones[-1] &= (1 << (n % 8)) - 1 if n % 8 else 0xFF
    # all ones: y_i = x_i ^ XOR of the whole second half
    tot = np.bitwise_xor.reduce(x[n:], axis=0)
    assert np.array_equal(qc_gpu(ctx, x, ones, n), x[:n] ^ tot)


# 4096 and 2^16 take the word-aligned path (n a multiple of 32, >= 64 words)
@pytest.mark.parametrize("n", [1, 33, 1000, 4096, 100003, 1 << 16])
def test_sparse_vs_oracle(ctx, n):
    import torch

    r = rng(n + 7)
    a = rand_bits(r, n)
    w = 37
    pos = np.sort(r.choice(2 * n, size=min(w, 2 * n), replace=False)).astype(np.int64)
    out = torch.zeros(((n + 31) // 32) * 4, dtype=torch.uint8, device="cuda")
    ctx.qc_sparse(n, dev_bits(a, n), torch.from_numpy(pos).cuda(), out)
    got = out.cpu().numpy()[: (n + 7) // 8]
    assert np.array_equal(got, po.qc_sparse(pos, a, n))


@pytest.mark.parametrize("N", [32, 1024, 32 * 1000, 1 << 16])
def test_transpose(ctx, N):
    import torch

    r = rng(N)
    x = rand_blocks(r, N)
    rows = torch.zeros(128 * N // 8, dtype=torch.uint8, device="cuda")
    ctx.bit_transpose(to_dev(x), rows)
    assert np.array_equal(rows.cpu().numpy().reshape(128, N // 8), po.bit_transpose(x))
    back = dev_blocks(N)
    ctx.bit_untranspose(rows, N, back)
    assert np.array_equal(to_np(back), x)
