"""GPU parity of the iterating Ferret session (cuot_ferret_session_*, the
FerretCOT extend loop): three extends byte-identical to the oracle's session
(both parties, including the pre-OT hash, the level-sum OTs, LPN with
kk = bootstrap + 128 and the bootstrap refresh), and the FerretCOT.rcot mirror
serving requests across extends with w = v ^ lsb(w)*Delta."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import to_np

pytestmark = pytest.mark.gpu

SEED = (0x0706050403020100, 0x0F0E0D0C0B0A0908)


@pytest.mark.parametrize("n,k,t,h,ntasks", [(40000, 2000, 5, 13, 1), (100000, 4546, 13, 13, 2),
                                            (30000, 1500, 30, 10, 1)])
def test_session_matches_oracle(ctx, n, k, t, h, ntasks):
    import torch

    from cuot_amd.ferret import FerretCOT

    print(f"ferret session n={n} k={k} t={t} h={h}", flush=True)
    fc = FerretCOT(ctx, n=n, k=k, t=t, h=h, seed=SEED, ntasks=ntasks)
    ref = po.FerretSession(n, k, t, h, np.array(SEED, dtype=np.uint64), ntasks=ntasks, nthreads=8)
    assert fc.M == ref.M and fc.ot_limit == ref.ot_limit
    assert np.array_equal(to_np(fc.Delta)[0], ref.delta)
    for rnd in range(3):
        fc.extend_f2k()
        torch.cuda.synchronize()
        s, r = ref.extend()
        assert np.array_equal(to_np(fc.out_s), s), f"sender round {rnd}"
        assert np.array_equal(to_np(fc.out_r), r), f"receiver round {rnd}"
    assert fc.rounds == 3
    fc.close()


def test_rcot_spans_extends(ctx):
    import torch

    from cuot_amd.ferret import FerretCOT

    fc = FerretCOT(ctx, n=40000, k=2000, t=5, h=13, seed=SEED)
    num = fc.ot_limit + 1000
    s, r = fc.rcot(num)
    torch.cuda.synchronize()
    assert s.shape == (num, 2) and fc.rounds == 2 and fc.silent_ot_left() == fc.ot_limit - 1000
    d = fc.Delta[0]
    lsb = (r[:, 0] & 1).to(torch.bool)
    assert torch.equal(torch.where(lsb.unsqueeze(1), s ^ d, s), r)
    ref = po.FerretSession(40000, 2000, 5, 13, np.array(SEED, dtype=np.uint64), nthreads=8)
    s0, _ = ref.extend()
    s1, _ = ref.extend()
    exp = np.concatenate([s0[: ref.ot_limit], s1[:1000]])
    assert np.array_equal(to_np(s), exp)
    fc.close()
