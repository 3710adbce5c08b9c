"""Pin the oracle's AES-128 against FIPS-197 vectors, OpenSSL and itself."""
import ctypes
import ctypes.util

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import rand_blocks, rng


def _b(hexs):
    return po.block_from_bytes(bytes.fromhex(hexs))


def test_fips197_appendix_c1():
    key = po.block_from_bytes(bytes(range(16)))
    for impl in (0, 1):
        ct = po.aes_ecb(key, _b("00112233445566778899aabbccddeeff"), impl)
        assert po.block_to_bytes(ct).hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"


def test_fips197_appendix_a1_key_schedule():
    rk = po.aes_expand(_b("2b7e151628aed2a6abf7158809cf4f3c"))
    words = [rk[4 * i:4 * i + 4].tobytes().hex() for i in range(44)]
    # FIPS-197 A.1: w4, w5, w43.  (The reference's Rcon offset would give
    # w4 = 2cfafe17.)
    assert words[4] == "a0fafe17"
    assert words[5] == "88542cb1"
    assert words[43] == "b6630ca6"


def test_fips197_appendix_b():
    ct = po.aes_ecb(_b("2b7e151628aed2a6abf7158809cf4f3c"), _b("3243f6a8885a308d313198a2e0370734"))
    assert po.block_to_bytes(ct).hex() == "3925841d02dc09fbdc118597196a0b32"


def test_portable_matches_aesni():
    if not po.have_aesni():
        pytest.skip("no AES-NI on this host")
    r = rng(1)
    for _ in range(4):
        key = rand_blocks(r, 1)
        x = rand_blocks(r, 257)
        assert np.array_equal(po.aes_ecb(key, x, 0), po.aes_ecb(key, x, 1))


def _openssl():
    name = ctypes.util.find_library("crypto")
    if not name:
        return None
    try:
        L = ctypes.CDLL(name)
        L.EVP_CIPHER_CTX_new.restype = ctypes.c_void_p
        L.EVP_aes_128_ecb.restype = ctypes.c_void_p
        L.EVP_EncryptInit_ex.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                         ctypes.c_char_p]
        L.EVP_CIPHER_CTX_set_padding.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.EVP_EncryptUpdate.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                        ctypes.c_char_p, ctypes.c_int]
        L.EVP_CIPHER_CTX_free.argtypes = [ctypes.c_void_p]
        return L
    except OSError:
        return None


def test_matches_openssl_ecb():
    L = _openssl()
    if L is None:
        pytest.skip("libcrypto not available")
    r = rng(2)
    key = rand_blocks(r, 1)
    x = rand_blocks(r, 64)
    c = L.EVP_CIPHER_CTX_new()
    assert L.EVP_EncryptInit_ex(c, L.EVP_aes_128_ecb(), None, po.block_to_bytes(key), None) == 1
    L.EVP_CIPHER_CTX_set_padding(c, 0)
    out = ctypes.create_string_buffer(64 * 16)
    ol = ctypes.c_int(0)
    assert L.EVP_EncryptUpdate(c, out, ctypes.byref(ol), po.block_to_bytes(x), 64 * 16) == 1
    L.EVP_CIPHER_CTX_free(c)
    assert out.raw[:ol.value] == po.block_to_bytes(po.aes_ecb(key, x))


def test_prg_is_ctr_of_counter_blocks():
    r = rng(3)
    seed = rand_blocks(r, 1)
    out = po.prg(seed, 7, 100, 9)
    ctr = np.zeros((9, 2), dtype=np.uint64)
    ctr[:, 0] = np.arange(100, 109, dtype=np.uint64)
    ctr[:, 1] = 7
    assert np.array_equal(out, po.aes_ecb(seed, ctr))


def test_matches_reference_cryptotools_aes():
    """The reference's vendored cryptoTools AES (compiled into oracle/_ref from
    the reference's own sources) agrees with the oracle at random keys."""
    import ctypes
    import os

    path = os.path.join(os.path.dirname(po.__file__), "_ref", "libref_aes.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L = ctypes.CDLL(path)
    L.ref_aes_ecb.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    r = rng(11)
    for _ in range(5):
        key = rand_blocks(r, 1)
        x = rand_blocks(r, 100)
        out = np.zeros_like(x)
        L.ref_aes_ecb(key.ctypes.data, x.ctypes.data, out.ctypes.data, 100)
        assert np.array_equal(out, po.aes_ecb(key, x))
