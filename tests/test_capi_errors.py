"""Failure handling that the host forces without a GPU: malformed instance
parameters are refused by the size queries of the C ABI (0 workspace bytes,
-1 state size) before any device work, and every refusal sets a message
(the reference throws or calls emp's error() for these,
ferret/emp-ot/emp-ot/ferret/constants.h:14-19, gpu/gpu_matrix.cu:44-46)."""
import ctypes as C
import os

import pytest

SEED = (0x0706050403020100, 0x0F0E0D0C0B0A0908)


@pytest.fixture(scope="module")
def lib():
    from cuot_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    return _lib.load()


def _sot(n, t, h):
    import cuot_amd

    return cuot_amd.SotConfig(n=n, t=t, h=h, reserved=0, k0=cuot_amd.to_block((1, 0)), k1=cuot_amd.to_block((2, 0)),
                              seed=cuot_amd.to_block(SEED))


def _ferret(n, k, t, h, ntasks=1, n_pre=0):
    import cuot_amd

    return cuot_amd.FerretConfig(n=n, k=k, t=t, h=h, ntasks=ntasks, seed=cuot_amd.to_block(SEED), n_pre=n_pre)


def test_sot_workspace_refuses_malformed(lib):
    assert lib.cuot_sot_workspace_bytes(C.byref(_sot(1 << 20, 256, 13))) > 0
    for n, t, h in [(1 << 20, 255, 13),   # t * 2^h < 2n: the trees do not cover the leaves
                    (0, 16, 7),           # n < 1
                    (1 << 10, 16, 0),     # h < 1
                    (1 << 10, 1, 17),     # h > 16
                    (1 << 10, 0, 7)]:     # t < 1
        assert lib.cuot_sot_workspace_bytes(C.byref(_sot(n, t, h))) == 0, (n, t, h)
    assert lib.cuot_sot_workspace_bytes(None) == 0


def test_ferret_workspace_refuses_malformed(lib):
    assert lib.cuot_ferret_workspace_bytes(C.byref(_ferret(10168320, 452160, 1270, 13))) > 0
    for n, k, t, h, nt in [(10168320, 452160, 1241, 13, 1),  # t * 2^h < n
                           (20000, 20000, 3, 13, 1),         # k >= n
                           (20000, 0, 3, 13, 1),             # k < 1
                           (20000, 1000, 3, 13, 0),          # ntasks < 1
                           (20000, 1000, 3, 17, 1)]:         # h > 16
        assert lib.cuot_ferret_workspace_bytes(C.byref(_ferret(n, k, t, h, nt))) == 0, (n, k, t, h, nt)


def test_ferret_state_size_refuses_malformed(lib):
    n, k, t, h = 40000, 2000, 5, 13
    M = k + h * t + 128
    assert lib.cuot_ferret_state_size(C.byref(_ferret(n, k, t, h))) == 48 + 16 * M
    assert lib.cuot_ferret_state_size(C.byref(_ferret(n, k, t, h, n_pre=M + 5))) == 48 + 16 * (M + 5)
    # emp-ot's PrimalLPNParameter: n_pre >= k + t * log_bin_sz + 128 (constants.h:14-19)
    assert lib.cuot_ferret_state_size(C.byref(_ferret(n, k, t, h, n_pre=M - 1))) == -1
    # the bootstrap must leave outputs: M < n
    assert lib.cuot_ferret_state_size(C.byref(_ferret(M, k, t, h))) == -1
    assert lib.cuot_ferret_state_size(None) == -1


def test_knobs_refuse_bad_values_and_set_the_crossover(lib):
    """Knob validation on the host: out-of-range values are refused with a
    message and leave the knobs unchanged; the matrix-core crossover follows
    qc_mfma_max (0: the built-in value, -1: never, capped at 2^16)."""
    from cuot_amd import _lib, tune

    try:
        base = int(lib.cuot_qc_mfma_crossover())
        assert 17 <= base <= 1 << 16
        for bad in (dict(qc_mfma_max=-2), dict(aes_impl=3), dict(qc_radix=9), dict(qc_cols=3)):
            k = _lib.KnobValues(**{f: 0 for f, _ in _lib.KnobValues._fields_})
            for f, v in bad.items():
                setattr(k, f, v)
            assert lib.cuot_set_knobs(C.byref(k)) == 1
            assert lib.cuot_last_error()
            assert int(lib.cuot_qc_mfma_crossover()) == base
        tune.set_knobs(qc_mfma_max=-1)
        assert int(lib.cuot_qc_mfma_crossover()) == 0
        tune.set_knobs(qc_mfma_max=30000)
        assert int(lib.cuot_qc_mfma_crossover()) == 30000
        tune.set_knobs(qc_mfma_max=1 << 20)
        assert int(lib.cuot_qc_mfma_crossover()) == 1 << 16
        tune.set_knobs(aes_impl=2)
        assert tune.get_knobs()["aes_impl"] == 2
    finally:
        tune.set_knobs()
    assert int(lib.cuot_qc_mfma_crossover()) == base


def test_ggm_bitsliced_work_sizes(lib):
    """The bitsliced GGM work queries: two levels of t * 2^(h-1) blocks for the
    sender; the receiver adds t zero roots, t * h * 2 level sums and t
    punctured indices; malformed trees give 0."""
    import cuot_amd

    B = cuot_amd.to_block
    for t, h in [(1, 1), (3, 5), (1024, 15), (7, 16)]:
        p = cuot_amd.GgmParams(t=t, h=h, mode=0, n_out=t << h, k0=B((1, 0)), k1=B((2, 0)))
        s = lib.cuot_ggm_sender_bs_work_bytes(C.byref(p))
        assert s == 2 * (t << (h - 1)) * 16, (t, h)
        assert lib.cuot_ggm_recver_bs_work_bytes(C.byref(p)) == s + 16 * t + 32 * t * h + 8 * t, (t, h)
    for t, h, mode in [(0, 5, 0), (3, 0, 0), (3, 17, 0), (3, 5, 7)]:
        p = cuot_amd.GgmParams(t=t, h=h, mode=mode, n_out=0, k0=B((1, 0)), k1=B((2, 0)))
        assert lib.cuot_ggm_sender_bs_work_bytes(C.byref(p)) == 0, (t, h, mode)
        assert lib.cuot_ggm_recver_bs_work_bytes(C.byref(p)) == 0, (t, h, mode)
