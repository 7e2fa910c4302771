"""Build the HIP step library for gfx950 in-tree.

The library ``forge2d_amd/libforge2d_amd.so`` is the product: C-ABI entry
points (``include/forge2d_amd.h``) over the fused world-step kernel. Floating
point is kept IEEE-exact so results match the reference's x86-64 SSE2 build:
no contraction into FMA, correctly rounded division and square root, denormals
kept. ``-packed-fp32-ops`` is disabled for every HIP translation unit.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libforge2d_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("F2D_OFFLOAD_ARCH", "gfx950")

FP_FLAGS = [
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
] + FP_FLAGS

SOURCES = ["step.hip", "api.cpp"]


def build(verbose: bool = False) -> str:
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        out = os.path.join(build_dir, src + ".o")
        cmd = [HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out] + HIP_FLAGS
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(out)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
