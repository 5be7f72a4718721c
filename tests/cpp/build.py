"""Builds the test-side native libraries next to their sources (tests/cpp/lib*.so).

- libgwn_surface_gpu.so: the reference-named C++ surface (include/gwn/gwn.hpp) driving a GPU
  workflow through libgwn_hip.so (g++, no device code of its own).
- libgwn_point_api.so: kernels calling the device point API (include/gwn/gwn_device.hpp), built
  with hipcc and the library's own flags (gfx950, correctly rounded float division and square
  root, packed FP32 off).

The GPU tests load them with ctypes; __graft_entry__.build() calls build() after the library.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
LIB_DIR = os.path.join(ROOT, "smallgwn_amd")
SURFACE_SO = os.path.join(HERE, "libgwn_surface_gpu.so")
POINT_SO = os.path.join(HERE, "libgwn_point_api.so")


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")


def build() -> list[str]:
    tmp = SURFACE_SO + ".tmp"
    _run([
        "g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__",
        "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
        os.path.join(HERE, "gwn_surface_gpu.cpp"),
        "-L", LIB_DIR, "-lgwn_hip", "-Wl,-rpath,$ORIGIN/../../smallgwn_amd",
        "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", tmp,
    ])
    os.replace(tmp, SURFACE_SO)
    from smallgwn_amd import build as lib_build

    tmp = POINT_SO + ".tmp"
    _run([
        lib_build.hipcc(), *lib_build.COMMON_FLAGS, "-shared", os.path.join(HERE, "gwn_point_api.hip"),
        "-L", LIB_DIR, "-lgwn_hip", "-Wl,-rpath,$ORIGIN/../../smallgwn_amd", "-o", tmp,
    ])
    os.replace(tmp, POINT_SO)
    return [SURFACE_SO, POINT_SO]


if __name__ == "__main__":
    for p in build():
        print(f"[build] {p}")
