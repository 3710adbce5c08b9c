"""GPU parity: AES-128, the counter-mode PRG and the fused GGM expansion
(sender and receiver, SOT and Ferret) against the CPU oracle, bit-exact."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import dev_blocks, rand_blocks, rng, to_dev, to_np

pytestmark = pytest.mark.gpu

FERRET_K0 = np.array([0, 0], dtype=np.uint64)
FERRET_K1 = np.array([1, 0], dtype=np.uint64)


def test_aes_ecb_matches_oracle(ctx):
    r = rng(1)
    for n in (1, 255, 256, 1000, 70001):
        key = rand_blocks(r, 1)
        x = rand_blocks(r, n)
        xd, yd = to_dev(x), dev_blocks(n)
        ctx.aes_ecb(key[0], xd, yd)
        assert np.array_equal(to_np(yd), po.aes_ecb(key, x)), n


def test_aes_fips197_vector(ctx):
    key = np.frombuffer(bytes(range(16)), dtype=np.uint64)
    x = po.block_from_bytes(bytes.fromhex("00112233445566778899aabbccddeeff"))
    yd = dev_blocks(1)
    ctx.aes_ecb(key, to_dev(x), yd)
    assert po.block_to_bytes(to_np(yd)).hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"


@pytest.fixture
def aes_impl():
    """Selects the AES stream cipher variant for one test, restoring the rule after."""
    from cuot_amd import tune

    def set_(v):
        tune.set_knobs(aes_impl=v)

    yield set_
    tune.set_knobs()


# the bitsliced cipher (32 blocks per lane, 2048 per wave): ragged waves and
# workgroups, the FIPS-197 answer, and counter mode
@pytest.mark.parametrize("n", [1, 31, 2047, 2048, 2049, 8193, 70001])
def test_aes_bitsliced_matches_oracle(ctx, aes_impl, n):
    print(f"aes bitsliced n={n}", flush=True)
    aes_impl(2)
    r = rng(50 + n)
    key = rand_blocks(r, 1)
    x = rand_blocks(r, n)
    xd, yd = to_dev(x), dev_blocks(n)
    ctx.aes_ecb(key[0], xd, yd)
    assert np.array_equal(to_np(yd)[:n], po.aes_ecb(key, x))


def test_aes_bitsliced_fips197_and_prg(ctx, aes_impl):
    aes_impl(2)
    key = np.frombuffer(bytes(range(16)), dtype=np.uint64)
    x = po.block_from_bytes(bytes.fromhex("00112233445566778899aabbccddeeff"))
    yd = dev_blocks(1)
    ctx.aes_ecb(key, to_dev(x), yd)
    assert po.block_to_bytes(to_np(yd)).hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"
    r = rng(3)
    seed = rand_blocks(r, 1)
    yd = dev_blocks(5000)
    ctx.prg(seed[0], 3, 77, yd)
    assert np.array_equal(to_np(yd), po.prg(seed, 3, 77, 5000))


def test_prg_matches_oracle(ctx):
    r = rng(2)
    seed = rand_blocks(r, 1)
    yd = dev_blocks(5000)
    ctx.prg(seed[0], 3, 77, yd)
    assert np.array_equal(to_np(yd), po.prg(seed, 3, 77, 5000))


def _params(t, h, mode, n_out, k0, k1):
    import cuot_amd

    return cuot_amd.GgmParams(t=t, h=h, mode=mode, n_out=n_out, k0=cuot_amd.to_block(k0),
                              k1=cuot_amd.to_block(k1))


def _alpha_choice(alphas, h):
    a = np.asarray(alphas, dtype=np.int64).reshape(-1, 1)
    shifts = (h - 1 - np.arange(h, dtype=np.int64)).reshape(1, -1)
    return (1 - ((a >> shifts) & 1)).astype(np.uint8)


CASES = [(1, 3), (3, 2), (8, 3), (9, 2), (11, 3), (13, 2), (14, 2), (15, 2), (16, 1)]


@pytest.mark.parametrize("h,t", CASES)
@pytest.mark.parametrize("mode", [0, 1])
def test_ggm_sender_receiver(ctx, h, t, mode):
    import torch

    print(f"ggm case h={h} t={t} mode={mode}", flush=True)
    r = rng(1000 + 17 * h + t + mode)
    if mode == 0:
        k0, k1 = rand_blocks(r, 1)[0], rand_blocks(r, 1)[0]
    else:
        k0, k1 = FERRET_K0, FERRET_K1
    roots = rand_blocks(r, t)
    delta = rand_blocks(r, 1)[0] | np.array([1, 0], dtype=np.uint64)
    n_full = t << h
    n_out = n_full - (3 if h >= 4 else 0)
    p = _params(t, h, mode, n_out, k0, k1)
    leaves_d, sums_d = dev_blocks(n_out), dev_blocks(t * h * 2)
    ctx.ggm_sender(p, to_dev(roots), leaves_d, sums_d)
    v = to_np(leaves_d)[:n_out]
    sums = to_np(sums_d).reshape(t, h, 2, 2)
    # oracle
    if mode == 1:
        ov, ls, rs, ss = po.ferret_mpcot_sender(roots, t, h, delta, n_out)
    else:
        ov = np.zeros((n_out, 2), dtype=np.uint64)
        ls, rs = np.zeros((t * h, 2), np.uint64), np.zeros((t * h, 2), np.uint64)
        for tr in range(t):
            lv, l, rr = po.ggm_sender_tree(k0, k1, roots[tr], h)
            m = min(1 << h, n_out - (tr << h))
            ov[tr << h:(tr << h) + m] = lv[:m]
            ls[tr * h:(tr + 1) * h], rs[tr * h:(tr + 1) * h] = l, rr
    assert np.array_equal(v, ov)
    assert np.array_equal(sums[:, :, 0].reshape(-1, 2), ls)
    assert np.array_equal(sums[:, :, 1].reshape(-1, 2), rs)

    # receiver from its own view
    alphas = r.integers(0, 1 << h, size=t)
    c = _alpha_choice(alphas, h)
    csum = np.where(c.reshape(-1, 1) == 1, rs, ls)
    if mode == 1:
        fin = ss if mode == 1 else None
    else:
        last_l, last_r = ls.reshape(t, h, 2)[:, h - 1], rs.reshape(t, h, 2)[:, h - 1]
        fin = np.where(c[:, h - 1:h] == 1, last_l, last_r) ^ delta
    wd = dev_blocks(n_out)
    ctx.ggm_recver(p, torch.from_numpy(c.reshape(-1).copy()).cuda(), to_dev(csum), to_dev(fin), wd)
    w = to_np(wd)[:n_out]
    expect = ov.copy()
    for tr, a in enumerate(alphas):
        g = (tr << h) + int(a)
        if g < n_out:
            expect[g] ^= delta
    assert np.array_equal(w, expect)
    if mode == 1:
        ow = po.ferret_mpcot_recver(t, h, c.reshape(-1), csum, ss, n_out)
        assert np.array_equal(w, ow)


# the sender and receiver on the bitsliced cipher (breadth-first, one launch
# per level) equal the LDS T-table ones: leaves, truncation, Ferret LSB rule,
# the level sums of every tree and the receiver's punctured-path rules, for
# levels with fewer and more than 32 nodes per tree and for trees sharing and
# spanning waves, and more trees than one grid of the per-tree kernels covers
@pytest.mark.parametrize("h,t", [(1, 3), (3, 2), (5, 7), (6, 64), (9, 5), (11, 3), (13, 2), (15, 2),
                                 (1, (1 << 20) + 3)])
@pytest.mark.parametrize("mode", [0, 1])
def test_ggm_sender_bitsliced_matches_ttable(ctx, h, t, mode):
    import torch

    print(f"ggm bitsliced h={h} t={t} mode={mode}", flush=True)
    r = rng(3000 + 17 * h + t + mode)
    k0, k1 = (rand_blocks(r, 1)[0], rand_blocks(r, 1)[0]) if mode == 0 else (FERRET_K0, FERRET_K1)
    roots = rand_blocks(r, t)
    n_out = (t << h) - (5 if h >= 4 else 0)
    p = _params(t, h, mode, n_out, k0, k1)
    a_leaves, a_sums = dev_blocks(n_out), dev_blocks(t * h * 2)
    ctx.ggm_sender(p, to_dev(roots), a_leaves, a_sums)
    b_leaves, b_sums = dev_blocks(n_out), dev_blocks(t * h * 2)
    work = torch.empty(ctx.ggm_sender_bs_work_bytes(p), dtype=torch.uint8, device="cuda")
    ctx.ggm_sender_bs(p, to_dev(roots), b_leaves, b_sums, work)
    assert np.array_equal(to_np(b_leaves)[:n_out], to_np(a_leaves)[:n_out])
    assert np.array_equal(to_np(b_sums), to_np(a_sums))
    if mode == 0 and t <= 3:
        lv, l, rr = po.ggm_sender_tree(k0, k1, roots[0], h)
        assert np.array_equal(to_np(b_leaves)[: 1 << h], lv)
    # receiver from its own view, both ciphers
    sums = to_np(a_sums).reshape(t, h, 2, 2)
    ls, rs = sums[:, :, 0].reshape(-1, 2), sums[:, :, 1].reshape(-1, 2)
    alphas = r.integers(0, 1 << h, size=t)
    c = _alpha_choice(alphas, h)
    csum = np.where(c.reshape(-1, 1) == 1, rs, ls)
    delta = rand_blocks(r, 1)[0] | np.array([1, 0], dtype=np.uint64)
    if mode == 1:
        fin = rand_blocks(r, t)  # the secret sums: any blocks exercise the final rule
    else:
        last_l, last_r = ls.reshape(t, h, 2)[:, h - 1], rs.reshape(t, h, 2)[:, h - 1]
        fin = np.where(c[:, h - 1:h] == 1, last_l, last_r) ^ delta
    cd = torch.from_numpy(c.reshape(-1).copy()).cuda()
    wa, wb = dev_blocks(n_out), dev_blocks(n_out)
    ctx.ggm_recver(p, cd, to_dev(csum), to_dev(fin), wa)
    rwork = torch.empty(ctx.ggm_recver_bs_work_bytes(p), dtype=torch.uint8, device="cuda")
    ctx.ggm_recver_bs(p, cd, to_dev(csum), to_dev(fin), wb, rwork)
    assert np.array_equal(to_np(wb)[:n_out], to_np(wa)[:n_out])
