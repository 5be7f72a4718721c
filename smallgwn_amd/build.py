"""Compile libgwn_hip.so (HIP for gfx950) in-tree with hipcc.

Every .hip file under csrc/ is compiled to an object in smallgwn_amd/_build/ (in parallel)
and linked into smallgwn_amd/libgwn_hip.so.  No GPU is needed: hipcc cross-compiles.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(HERE, "libgwn_hip.so")
ARCH = os.environ.get("GWN_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # IEEE-correct float division and square root (moment parity with the CPU oracle).
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # Packed FP32 instructions stay disabled on this platform.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src)[:-4] + ".o")
    cmd = [hipcc(), *COMMON_FLAGS, "-c", src, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {os.path.basename(src)}:\n{proc.stdout}\n{proc.stderr}")
    warnings = "\n".join(
        line for line in proc.stderr.splitlines()
        if "warning" in line and "packed-fp32-ops" not in line and "target-feature" not in line
    )
    if warnings:
        print(f"[build] {os.path.basename(src)}:\n{warnings}", file=sys.stderr)
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
    """Compile every source and link libgwn_hip.so; returns the library path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(_compile, srcs))
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"[build] {LIB_PATH} ({len(objs)} objects)")
    return LIB_PATH


if __name__ == "__main__":
    build()
