"""The reference-named C++ surface (include/gwn/gwn.hpp) compiles with g++ and links against
libgwn_hip.so; its host-decided calls behave like the reference's (no GPU needed)."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "smallgwn_amd")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_cpp_surface_compiles_links_and_validates(tmp_path):
    if not os.path.exists(os.path.join(LIB_DIR, "libgwn_hip.so")):
        pytest.fail("libgwn_hip.so is not built")
    exe = tmp_path / "gwn_surface_check"
    cmd = [
        "g++", "-std=c++20", "-O1", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
        "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
        os.path.join(ROOT, "tests", "cpp", "gwn_surface_check.cpp"),
        "-L", LIB_DIR, "-lgwn_hip", f"-Wl,-rpath,{LIB_DIR}",
        "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe),
    ]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all checks passed" in out.stdout
