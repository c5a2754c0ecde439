"""The reference-named C++ shims (include/tfg_ref) compile against types with
the reference's layouts and link against libtfg_hip.so's C ABI."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"), reason="no toolchain")
def test_reference_shims_compile():
    src = os.path.join(ROOT, "tests", "cpp", "shim_check.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-Wall", "-Wno-unused-function", src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
