"""The oracle's iterating Ferret (FerretCOT::extend_f2k restated): every
round's outputs satisfy w = v ^ lsb(w)*Delta, the bootstrap after a round is
that round's last M outputs (ferret_cot.hpp:100), rounds differ, and the
correlation-robust hash matches its definition H(x) = AES0(s(x)) ^ s(x)
computed independently from the FIPS-pinned AES."""
import numpy as np

from oracle import pyoracle as po


def _relation(s, r, d):
    lsb = (r[:, 0] & np.uint64(1)).astype(bool)
    exp = s.copy()
    exp[lsb] ^= d
    return np.array_equal(exp, r)


def test_session_rounds_relation_and_refresh():
    fs = po.FerretSession(40000, 2000, 5, 13, (0x1234, 0x5678), nthreads=4)
    assert fs.M == 2000 + 13 * 5 + 128 and fs.ot_limit == 40000 - fs.M
    prev = None
    for _ in range(3):
        s, r = fs.extend()
        assert _relation(s, r, fs.delta)
        bs, br = fs.bootstrap()
        assert np.array_equal(bs, s[fs.ot_limit:]) and np.array_equal(br, r[fs.ot_limit:])
        assert not np.any(bs[:, 0] & np.uint64(1))  # sender bootstrap LSBs stay 0
        if prev is not None:
            assert not np.array_equal(prev, s)
        prev = s


def test_ccrh_definition():
    rng = np.random.default_rng(4)
    x = rng.integers(0, 2**63, size=(64, 2), dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=(64, 2),
                                                                                          dtype=np.uint64)
    sig = np.empty_like(x)
    sig[:, 0] = x[:, 1]          # lo' = hi
    sig[:, 1] = x[:, 1] ^ x[:, 0]  # hi' = hi ^ lo
    exp = po.aes_ecb(np.zeros(2, dtype=np.uint64), sig) ^ sig
    assert np.array_equal(po.ccrh(x), exp)
