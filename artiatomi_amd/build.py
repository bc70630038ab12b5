"""Builds libartia.so (gfx950) in-tree with hipcc.

Usage: python -m artiatomi_amd.build   (or artiatomi_amd.build.build_library())
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libartia.so"

SOURCES = [
    "context.hip",
    "geometry.cpp",
    "formats.cpp",
    "fp.hip",
    "bp.hip",
    "image.hip",
    "spectral.hip",
    "subvol.hip",
    "adjoint.hip",
    "refine.hip",
    "sart.hip",
    "prep.hip",
]

# -ffp-contract=off: the reference-numerics kernels keep the reference's float operation order;
# explicit fmaf is used where a fused multiply-add is intended.
# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def build_library(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIP_FLAGS, "-I", str(PKG.parent / "include"), "-o", str(tmp)]
    cmd += [str(CSRC / s) for s in SOURCES]
    cmd += ["-lhipfft"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
