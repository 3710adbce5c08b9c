"""The reference-side dev_layer shim (cuot_amd/shims/dev_layer_shim.cpp)
compiles against the reference's own headers (dev_layer.h, gpu/gpu_matrix.h)
and links against libcuot_amd.so, so that a caller compiled against
dev_layer.h -- as MpcotReg and LpnF2 are -- resolves all four entry points
through it.  Nothing is run.  Skipped where the reference tree, g++ or the
built library is absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
FERRET_DIR = os.path.join(REF, "ferret/emp-ot/emp-ot/ferret")
GPU_DIR = os.path.join(REF, "gpu")
SHIM = os.path.join(ROOT, "cuot_amd", "shims", "dev_layer_shim.cpp")
LIBDIR = os.path.join(ROOT, "cuot_amd", "lib")

# a caller that uses the four functions exactly as declared by the reference
CALLER = r"""
#include "dev_layer.h"
int main(int argc, char**) {
  if (argc > 99) {  // never taken: only the references to the symbols matter
    cuda_setdev(0);
    cuda_mpcot_sender(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 13, nullptr, 1);
    cuda_mpcot_recver(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 13, nullptr, 1);
    cuda_primal_lpn(Sender, nullptr, 10, nullptr, 0, nullptr, nullptr, nullptr, nullptr, nullptr, 1);
  }
  return 0;
}
"""

needs = pytest.mark.skipif(not (os.path.isdir(FERRET_DIR) and os.path.isdir(GPU_DIR) and shutil.which("g++")),
                           reason="reference headers or g++ not available")


@needs
def test_shim_compiles_against_reference_headers():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", f"-I{FERRET_DIR}",
                        f"-I{GPU_DIR}", f"-I{os.path.join(ROOT, 'include')}", SHIM], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@needs
def test_reference_caller_links_through_shim(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libcuot_amd.so")):
        pytest.skip("library not built")
    caller = tmp_path / "caller.cpp"
    caller.write_text(CALLER)
    exe = tmp_path / "caller"
    cmd = ["g++", "-std=c++17", f"-I{FERRET_DIR}", f"-I{GPU_DIR}", f"-I{os.path.join(ROOT, 'include')}",
           str(caller), SHIM, f"-L{LIBDIR}", "-lcuot_amd", "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
