"""Builds libartia.so (gfx950) in-tree with hipcc.

Usage: python -m artiatomi_amd.build   (or artiatomi_amd.build.build_library())
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libartia.so"
OBJ = PKG / "_build"

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
    "rotate.hip",
    "sart.hip",
    "prep.hip",
]

# -ffp-contract=off: the reference-numerics kernels keep the reference's float operation order;
# explicit fmaf is used where a fused multiply-add is intended.
# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
COMPILE_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
LINK_FLAGS = ["--offload-arch=gfx950", "-shared", "-fPIC"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def _compile(src: Path, obj: Path) -> tuple:
    cmd = [hipcc(), *COMPILE_FLAGS, "-I", str(PKG.parent / "include"), "-c", "-o", str(obj), str(src)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    return src.name, res.returncode, res.stdout + res.stderr, " ".join(cmd)


def build_library(verbose: bool = False, jobs: int = 0, incremental: bool = False) -> Path:
    """Compiles every source to an object under artiatomi_amd/_build (in parallel), then links
    libartia.so.  incremental=True recompiles only sources newer than their object, or all of them
    when a header is newer than the oldest object."""
    OBJ.mkdir(exist_ok=True)
    todo = []
    headers = max(p.stat().st_mtime for p in list(CSRC.glob("*.hpp")) + list((PKG.parent / "include").glob("*.h")))
    for s in SOURCES:
        src, obj = CSRC / s, OBJ / (s + ".o")
        if (not incremental or not obj.exists() or obj.stat().st_mtime < src.stat().st_mtime
                or obj.stat().st_mtime < headers):
            todo.append((src, obj))
    jobs = jobs or min(len(todo), os.cpu_count() or 4, 16) or 1
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for name, rc, out, cmd in pool.map(lambda a: _compile(*a), todo):
            if verbose:
                print(cmd, flush=True)
            if rc != 0:
                sys.stderr.write(out)
                raise RuntimeError(f"hipcc failed on {name} with exit code {rc}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *LINK_FLAGS, "-o", str(tmp)] + [str(OBJ / (s + ".o")) for s in SOURCES] + ["-lhipfft"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc link failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
