"""Compiles the HIP answer engine into an in-tree shared library for gfx950.

    python -m altopir_public_amd.build        # -> altopir_public_amd/libaltopir_hip.so

The library exports the C ABI of include/altopir_engine.h.  It is built with hipcc
directly (no cmake/ninja): the kernels and the engine's host side, the serving loop and its TCP
front end (host code only).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libaltopir_hip.so"
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are kept out of every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    return "hipcc"


def build(verbose: bool = False) -> Path:
    srcs = [str(CSRC / "alto_engine.hip"), str(CSRC / "alto_serve.hip"), str(CSRC / "alto_net.cpp")]
    tmp = LIB.with_suffix(".so.tmp")
    # the serving loop issues its uploads to the DMA engines through the HSA runtime
    rocm_lib = Path(hipcc()).resolve().parent.parent / "lib"
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{ROOT / 'include'}", "-o", str(tmp), *srcs, f"-L{rocm_lib}", "-lhsa-runtime64"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
    print(LIB)
