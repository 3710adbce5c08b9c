"""GPU parity: fused Ferret LPN against the oracle, bit-exact."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import dev_blocks, rand_blocks, rng, to_dev, to_np

pytestmark = pytest.mark.gpu


def _p(n, k, seed, ntasks=1):
    import cuot_amd

    return cuot_amd.LpnParams(n=n, k=k, d=10, ntasks=ntasks, seed=cuot_amd.to_block(seed))


@pytest.mark.parametrize("n,k,ntasks", [(1, 5, 1), (7, 11, 1), (1000, 452, 1), (12345, 999, 1), (4099, 3000, 3),
                                        (100000, 4523, 1), (100003, 4523, 7)])
def test_lpn_single_and_dual(ctx, n, k, ntasks):
    print(f"lpn n={n} k={k} ntasks={ntasks}", flush=True)
    r = rng(n + k)
    seed = rand_blocks(r, 1)[0]
    noise, kk = rand_blocks(r, n), rand_blocks(r, k)
    ref = po.ferret_lpn(noise, kk, seed, ntasks=ntasks, nthreads=4)
    out = dev_blocks(n)
    ctx.lpn(_p(n, k, seed, ntasks), to_dev(noise), to_dev(kk), out)
    assert np.array_equal(to_np(out)[:n], ref)
    # in place, dual role
    noise2, kk2 = rand_blocks(r, n), rand_blocks(r, k)
    ref2 = po.ferret_lpn(noise2, kk2, seed, ntasks=ntasks, nthreads=4)
    a, b = to_dev(noise), to_dev(noise2)
    ctx.lpn_dual(_p(n, k, seed, ntasks), a, to_dev(kk), a, b, to_dev(kk2), b)
    assert np.array_equal(to_np(a)[:n], ref)
    assert np.array_equal(to_np(b)[:n], ref2)
