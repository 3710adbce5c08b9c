"""The bit-row basis conversion used inside the QC encoder, compiled for the
host with g++: both the generic and the compile-time specialised forms agree
with a direct element-level restatement of the subspace-polynomial division
(every level, forward and inverse) and round-trip up to 2^20-bit rows."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def test_conversion_host_build():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "conv_check")
        subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "conv_host_check.cpp")],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout
    assert "conversion ok" in out


def test_gf32_bitsliced_multiply_host_build():
    """gf32_mulc_top (uniform two-bit and one-bit steps, per-lane steps) equals
    scalar GF(2^32) multiplication on every row."""
    with tempfile.TemporaryDirectory() as d:
        for two_bit in (0, 1):
            exe = os.path.join(d, f"gf32_check{two_bit}")
            subprocess.run(["g++", "-O2", "-std=c++17", f"-DGF32_TWO_BIT={two_bit}", "-o", exe,
                            os.path.join(HERE, "gf32_host_check.cpp")], check=True)
            out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout
            assert "gf32 top ok" in out
