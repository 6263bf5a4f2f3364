"""Build the gfx950 library (libvp8g.so) in-tree with hipcc.

Usage: python oxideav-vp8_amd/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libvp8g.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 instruction selection stays off for every HIP translation unit.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

SOURCES = ["capi.hip", "arnr_kernels.hip", "mc_kernels.hip", "search_kernels.hip", "proposer_kernels.hip",
           "proposer_host.hip", "synth_kernels.hip", "synth_host.hip", "tuning.cpp"]


def build(verbose: bool = False) -> Path:
    srcs = [str(CSRC / s) for s in SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", str(tmp), *srcs, "-ldl"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(f"built {LIB}")
    sys.exit(0)
