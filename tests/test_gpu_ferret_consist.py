"""GPU parity of the malicious-mode consistency check of the Ferret session
(cuot_ferret_session_set_malicious / _check): outputs unchanged by the check,
x' and both digests byte-identical to the oracle's session over two extends
(full and truncated trees), and a corrupted receiver leaf failing the check
run between the MPCOT and the LPN."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import to_np

pytestmark = pytest.mark.gpu

SEED = (0x0706050403020100, 0x0F0E0D0C0B0A0908)


@pytest.mark.parametrize("n,k,t,h", [(40960, 2000, 5, 13), (40000, 2000, 5, 13), (3000, 300, 12, 8)])
def test_check_matches_oracle(ctx, n, k, t, h):
    import torch

    from cuot_amd.ferret import FerretCOT

    print(f"ferret consistency n={n} t={t} h={h}", flush=True)
    fc = FerretCOT(ctx, n=n, k=k, t=t, h=h, seed=SEED)
    fc.set_malicious(True)
    ref = po.FerretSession(n, k, t, h, np.array(SEED, dtype=np.uint64), nthreads=8)
    ref.set_malicious(True)
    for rnd in range(2):
        fc.extend_f2k()
        torch.cuda.synchronize()
        s, r = ref.extend()
        assert np.array_equal(to_np(fc.out_s), s) and np.array_equal(to_np(fc.out_r), r), rnd
        ok, xp, da, db = fc.check_result()
        rok, rxp, rda, rdb = ref.check_result()
        assert ok and rok, rnd
        assert xp == bytes(rxp) and da == rda and db == rdb, rnd
    fc.close()


def test_corrupted_leaf_fails(ctx):
    import torch

    from cuot_amd import CuotError, _lib
    from cuot_amd.ferret import FerretCOT

    print("ferret consistency, corrupted leaf", flush=True)
    n, k, t, h = 40960, 2000, 5, 13
    fc = FerretCOT(ctx, n=n, k=k, t=t, h=h, seed=SEED)
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    args = (fc.handle, fc.out_s.data_ptr(), fc.out_r.data_ptr())
    assert lib.cuot_ferret_session_extend_stages(*args, 4, st) != 0  # check before any MPCOT
    with pytest.raises(CuotError):  # no result before a check ran
        fc.check_result()
    assert lib.cuot_ferret_session_extend_stages(*args, 1, st) == 0
    assert lib.cuot_ferret_session_extend_stages(*args, 4, st) == 0
    ok, _, da, db = fc.check_result()
    assert ok and da == db
    fc.out_r[8192 + 77, 1] ^= 1 << 40  # one receiver leaf of tree 1
    assert lib.cuot_ferret_session_extend_stages(*args, 4, st) == 0
    ok2, _, da2, db2 = fc.check_result()
    assert not ok2 and da2 == da and db2 != db
    fc.out_r[8192 + 77, 1] ^= 1 << 40
    assert lib.cuot_ferret_session_extend_stages(*args, 2, st) == 0
    torch.cuda.synchronize()
    ref = po.FerretSession(n, k, t, h, np.array(SEED, dtype=np.uint64), nthreads=8)
    s, r = ref.extend()
    assert np.array_equal(to_np(fc.out_s), s) and np.array_equal(to_np(fc.out_r), r)
    # the next extend's MPCOT without a check discards the last result: the
    # previous extend's digests are not reported as the current one's
    assert lib.cuot_ferret_session_extend_stages(*args, 1, st) == 0
    with pytest.raises(CuotError):
        fc.check_result()
    assert lib.cuot_ferret_session_extend_stages(*args, 2, st) == 0
    with pytest.raises(CuotError):
        fc.check_result()
    fc.close()
