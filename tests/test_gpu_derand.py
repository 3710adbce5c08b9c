"""GPU parity of online derandomisation (cuot_cot_online / cuot_cot_to_ot /
cuot_cot_to_rot) on a Ferret session's COTs: byte-identical to the oracle,
with M = K ^ b*Delta, OT outputs m_r and ROT outputs data_r."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import to_np

pytestmark = pytest.mark.gpu

SEED = (0x0706050403020100, 0x0F0E0D0C0B0A0908)


@pytest.mark.parametrize("num,key_start", [(1, 0), (1000, 0), (30001, 24)])
def test_derand_matches_oracle(ctx, num, key_start):
    import torch

    from cuot_amd.ferret import FerretCOT

    print(f"derand num={num}", flush=True)
    g = torch.Generator().manual_seed(num)
    fc = FerretCOT(ctx, n=40000, k=2000, t=5, h=13, seed=SEED)
    delta = to_np(fc.Delta)[0]
    b = torch.randint(0, 2, (num,), dtype=torch.uint8, generator=g).cuda()
    m0 = torch.randint(-2**62, 2**62, (num, 2), dtype=torch.int64, generator=g).cuda()
    m1 = torch.randint(-2**62, 2**62, (num, 2), dtype=torch.int64, generator=g).cuda()
    s = np.array([0x0123456789ABCDEF, 0x0FEDCBA987654321], dtype=np.uint64)
    # chosen-bit COTs
    ref = FerretCOT(ctx, n=40000, k=2000, t=5, h=13, seed=SEED)  # same COT stream, kept random
    ks0, mr0 = ref.rcot(num)
    ks, mr, bo = fc.send_recv_cot(b)
    torch.cuda.synchronize()
    K2, bo_ref = po.cot_online(to_np(ks0), to_np(mr0), to_np(b), delta)
    assert np.array_equal(to_np(bo), bo_ref) and np.array_equal(to_np(ks), K2)
    assert torch.equal(mr, torch.where(b.bool().unsqueeze(1), ks ^ fc.Delta[0], ks))
    # chosen-message OT on the next COTs
    ks1, mr1 = ref.rcot(num)
    K3, _ = po.cot_online(to_np(ks1), to_np(mr1), to_np(b), delta)
    pad, out = fc.send_recv_ot(m0, m1, b, s, key_start)
    torch.cuda.synchronize()
    pad_ref, out_ref = po.cot_to_ot(K3, to_np(mr1), delta, s, key_start, to_np(b), to_np(m0), to_np(m1))
    assert np.array_equal(to_np(pad), pad_ref) and np.array_equal(to_np(out), out_ref)
    assert torch.equal(out, torch.where(b.bool().unsqueeze(1), m1, m0))
    # random OT
    ks2, mr2 = ref.rcot(num)
    K4, _ = po.cot_online(to_np(ks2), to_np(mr2), to_np(b), delta)
    d0, d1, outr = fc.send_recv_rot(b, s, key_start)
    torch.cuda.synchronize()
    r0, r1, rr = po.cot_to_rot(K4, to_np(mr2), delta, s, key_start)
    assert np.array_equal(to_np(d0), r0) and np.array_equal(to_np(d1), r1) and np.array_equal(to_np(outr), rr)
    assert torch.equal(outr, torch.where(b.bool().unsqueeze(1), d1, d0))
    fc.close()
    ref.close()
