"""The C++ interface headers (include/ed/*.hpp) compile with hipcc against include/edhip.h and
link to libedhip.so (no GPU needed to build)."""
import os
import shutil
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "exact_diagonalization_for_fermionic_systems_with_gpu_amd")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_cpp_driver_compiles_and_links():
    if not os.path.exists(os.path.join(PKG, "libedhip.so")):
        pytest.skip("libedhip.so not built")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "drv")
        r = subprocess.run([HIPCC, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"),
                            os.path.join(REPO, "tests", "cpp", "kagome12_driver.cpp"), "-o", exe, "-L", PKG,
                            "-ledhip"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        assert os.path.getsize(exe) > 0
