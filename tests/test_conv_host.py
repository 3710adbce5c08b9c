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
    """The tower GF(2^32) of the QC encoder is a field (axioms, x^(2^32-1) = 1,
    subfield and theta degree), its chunk embedding reproduces carry-less chunk
    products, its Cantor basis satisfies beta_i^2 + beta_i = beta_{i-1}, and the
    bitsliced products (subfield and general constants, wave-uniform and
    per-lane) and plane maps equal the scalar ones on every row."""
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "gf32_check")
        subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "gf32_host_check.cpp")], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=300).stdout
        assert "gf32 tower ok" in out
