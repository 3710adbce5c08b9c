"""GPU parity of the Ferret session's state formats (emp-ot byte layouts of
ferret_cot.hpp:215-247 and 309-347): the assembled blobs equal the oracle's
layout of the oracle session's Delta and bootstrap, a session resumed from
blobs or from pre-OT files continues byte-identically to the oracle, bad
blobs and files are refused like emp-ot refuses them, the two-call extend
equals the one-call extend, and the event log records SeedExp / LPN."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import to_np

pytestmark = pytest.mark.gpu

SEED = (0x0706050403020100, 0x0F0E0D0C0B0A0908)
P = dict(n=40000, k=2000, t=5, h=13)


def _pair(ctx, rounds):
    import torch

    from cuot_amd.ferret import FerretCOT

    fc = FerretCOT(ctx, seed=SEED, **P)
    ref = po.FerretSession(P["n"], P["k"], P["t"], P["h"], np.array(SEED, dtype=np.uint64), nthreads=8)
    for _ in range(rounds):
        fc.extend_f2k()
        ref.extend()
    torch.cuda.synchronize()
    return fc, ref


def test_state_blobs_match_oracle_layout(ctx):
    from cuot_amd.ferret import ALICE, BOB

    print("ferret state blobs", flush=True)
    fc, ref = _pair(ctx, 2)
    bs, br = ref.bootstrap()
    a, b = fc.assemble_state(ALICE), fc.assemble_state(BOB)
    assert len(a) == fc.state_size() == 48 + 16 * ref.M
    assert a == po.ferret_state_blob(1, P["n"], P["t"], P["k"], ref.delta, bs)
    assert b == po.ferret_state_blob(2, P["n"], P["t"], P["k"], np.zeros(2, np.uint64), br)
    fc.close()


def test_resume_from_state(ctx):
    import torch

    from cuot_amd import CuotError
    from cuot_amd.ferret import ALICE, BOB, FerretCOT

    print("ferret resume from state", flush=True)
    fc, ref = _pair(ctx, 2)
    a, b = fc.assemble_state(ALICE), fc.assemble_state(BOB)
    fc2 = FerretCOT.disassemble_state(ctx, a, b, seed=SEED, round=2, **P)
    assert fc2.rounds == 2
    assert np.array_equal(to_np(fc2.Delta)[0], ref.delta)
    for rnd in range(2):
        fc.extend_f2k()
        fc2.extend_f2k()
        torch.cuda.synchronize()
        s, r = ref.extend()
        assert np.array_equal(to_np(fc2.out_s), s) and np.array_equal(to_np(fc2.out_r), r), rnd
        assert torch.equal(fc.out_s, fc2.out_s) and torch.equal(fc.out_r, fc2.out_r)
    # emp-ot's disassemble_state refuses another party or other (n, t, k)
    with pytest.raises(CuotError):
        FerretCOT.disassemble_state(ctx, b, a, seed=SEED, round=2, **P)
    with pytest.raises(CuotError):
        FerretCOT.disassemble_state(ctx, a, b, seed=SEED, round=2, **dict(P, k=P["k"] + 1))
    fc.close()
    fc2.close()


def test_pre_ot_files(ctx, tmp_path):
    import torch

    from cuot_amd import CuotError
    from cuot_amd.ferret import ALICE, BOB, FerretCOT

    print("ferret pre-OT files", flush=True)
    fc, ref = _pair(ctx, 1)
    fa, fb = str(tmp_path / "pre_send"), str(tmp_path / "pre_recv")
    fc.write_pre_data128_to_file(ALICE, fa)
    fc.write_pre_data128_to_file(BOB, fb)
    bs, br = ref.bootstrap()
    assert open(fa, "rb").read() == po.ferret_pre_file(1, P["n"], P["t"], P["k"], ref.delta, bs)
    assert open(fb, "rb").read() == po.ferret_pre_file(2, P["n"], P["t"], P["k"], None, br)
    with pytest.raises(CuotError):  # "wrong party"
        FerretCOT.read_pre_data128_from_file(ctx, fb, fa, seed=SEED, round=1, **P)
    with pytest.raises(CuotError):  # "wrong parameters"
        FerretCOT.read_pre_data128_from_file(ctx, fa, fb, seed=SEED, round=1, **dict(P, t=P["t"] + 1))
    assert os.path.exists(fa) and os.path.exists(fb)
    fc2 = FerretCOT.read_pre_data128_from_file(ctx, fa, fb, seed=SEED, round=1, **P)
    assert not os.path.exists(fa) and not os.path.exists(fb)  # consumed, as emp-ot removes them
    fc2.extend_f2k()
    torch.cuda.synchronize()
    s, r = ref.extend()
    assert np.array_equal(to_np(fc2.out_s), s) and np.array_equal(to_np(fc2.out_r), r)
    fc.close()
    fc2.close()


def test_staged_extend_and_log(ctx, tmp_path):
    import torch

    from cuot_amd import _lib
    from cuot_amd.ferret import FerretCOT
    from cuot_amd.logger import Log, Recver, Sender

    print("ferret staged extend + log", flush=True)
    fc, ref = _pair(ctx, 0)
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    args = (fc.handle, fc.out_s.data_ptr(), fc.out_r.data_ptr())
    assert lib.cuot_ferret_session_extend_stages(*args, 2, st) != 0  # LPN before MPCOT
    Log.reset()
    files = [str(tmp_path / "send.txt"), str(tmp_path / "recv.txt")]
    Log.open(Sender, files[0], True, 2)
    Log.open(Recver, files[1], True, 2)
    try:
        for rnd in range(2):
            fc.extend_f2k()  # staged under the log
            torch.cuda.synchronize()
            s, r = ref.extend()
            assert np.array_equal(to_np(fc.out_s), s) and np.array_equal(to_np(fc.out_r), r), rnd
    finally:
        Log.close(Sender)
        Log.close(Recver)
        Log.reset()
    for f in files:
        lines = open(f).read().splitlines()
        assert lines[:6] == ["0 CudaInit", "1 BaseOT", "2 SeedExp", "3 LPN", "4 Neural", "--------------------"]
        body = [ln.split()[:2] for ln in lines[6:14]]
        assert body == [["s", "2"], ["e", "2"], ["s", "3"], ["e", "3"]] * 2
        assert [ln.split()[:2] for ln in lines[14:]] == [["t", str(e)] for e in range(5)] + \
            [["m", str(e)] for e in range(5)]
        assert float(lines[14 + 2].split()[2]) > 0 and float(lines[14 + 3].split()[2]) > 0
    fc.close()


def test_state_and_files_with_n_pre_above_bootstrap(ctx, tmp_path):
    """emp-ot's blob and files carry param.n_pre blocks of ot_pre_data
    (ferret_cot.hpp:226,238,318,346), more than the M the extend reads: with
    n_pre > M the M bootstrap COTs are followed by zeros up to n_pre, the sizes
    are emp-ot's state_size (48 + 16 n_pre) and file length, and sessions read
    back from both continue byte-identically to the oracle."""
    import torch

    from cuot_amd import CuotError
    from cuot_amd.ferret import ALICE, BOB, FerretCOT

    print("ferret state with n_pre > M", flush=True)
    fc, ref = _pair(ctx, 1)
    M = ref.M
    n_pre = M + 77
    # a blob of the n_pre = M session is too short for a session with more pre-OT blocks
    with pytest.raises(CuotError):
        FerretCOT.disassemble_state(ctx, fc.assemble_state(ALICE), fc.assemble_state(BOB), seed=SEED, round=1,
                                    n_pre=n_pre, **P)
    fcp = FerretCOT(ctx, seed=SEED, n_pre=n_pre, **P)
    fcp.extend_f2k()
    torch.cuda.synchronize()
    assert np.array_equal(to_np(fcp.out_s), to_np(fc.out_s)) and np.array_equal(to_np(fcp.out_r), to_np(fc.out_r))
    a, b = fcp.assemble_state(ALICE), fcp.assemble_state(BOB)
    assert len(a) == fcp.state_size() == 48 + 16 * n_pre
    bs, br = ref.bootstrap()
    pad = np.zeros((n_pre - M, 2), np.uint64)
    assert a == po.ferret_state_blob(1, P["n"], P["t"], P["k"], ref.delta, np.concatenate([bs, pad]))
    assert b == po.ferret_state_blob(2, P["n"], P["t"], P["k"], np.zeros(2, np.uint64), np.concatenate([br, pad]))
    fa, fb = str(tmp_path / "send"), str(tmp_path / "recv")
    fcp.write_pre_data128_to_file(ALICE, fa)
    fcp.write_pre_data128_to_file(BOB, fb)
    assert open(fa, "rb").read() == po.ferret_pre_file(1, P["n"], P["t"], P["k"], ref.delta, np.concatenate([bs, pad]))
    assert os.path.getsize(fb) == 8 + 24 + 16 * n_pre
    fcf = FerretCOT.read_pre_data128_from_file(ctx, fa, fb, seed=SEED, round=1, n_pre=n_pre, **P)
    fcs = FerretCOT.disassemble_state(ctx, a, b, seed=SEED, round=1, n_pre=n_pre, **P)
    fcf.extend_f2k()
    fcs.extend_f2k()
    torch.cuda.synchronize()
    s, r = ref.extend()
    for x in (fcf, fcs):
        assert np.array_equal(to_np(x.out_s), s) and np.array_equal(to_np(x.out_r), r)
    # n_pre below M is emp-ot's "LPN parameter not matched"
    with pytest.raises(CuotError):
        FerretCOT(ctx, seed=SEED, n_pre=M - 1, **P)
    for x in (fc, fcp, fcf, fcs):
        x.close()
