"""Build the gfx950 library (libvp8g.so) in-tree with hipcc.

Every translation unit is compiled in parallel into a temporary directory,
then linked.  Usage: python oxideav-vp8_amd/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
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
           "proposer_host.hip", "synth_kernels.hip", "synth_host.hip", "tuning.cpp", "inter_search_kernels.hip", "split_kernels.hip",
           "stream_kernels.hip", "lf_kernels.hip", "lf_host.hip"]


def build(verbose: bool = False) -> Path:
    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with tempfile.TemporaryDirectory(prefix="vp8g_build_") as tmpdir:
        objs = [Path(tmpdir) / (Path(s).stem + ".o") for s in SOURCES]

        def compile_one(i: int) -> None:
            src = CSRC / SOURCES[i]
            lang = ["-x", "hip"] if src.suffix == ".cpp" else []
            cmd = [HIPCC, *HIP_FLAGS, *lang, "-c", str(src), "-o", str(objs[i])]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)

        with ThreadPoolExecutor(jobs) as pool:
            list(pool.map(compile_one, range(len(SOURCES))))
        tmp = LIB.with_suffix(".so.tmp")
        cmd = [HIPCC, "--offload-arch=gfx950", "-fPIC", "-shared", "-o", str(tmp), *map(str, objs), "-ldl"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(f"built {LIB}")
    sys.exit(0)
