"""The C-ABI library loads without a GPU and exports every entry point include/artia.h declares;
host-only entry points work and reject bad arguments with a status (no exceptions)."""
import ctypes as C
import re
from pathlib import Path

import artiatomi_amd as A
from artiatomi_amd import _abi

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for hdr in (ROOT / "include").glob("*.h"):
        text = hdr.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(artia_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "artia_bp" in names and "artia_fp_march" in names and "artia_sart_run" in names


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_every_declared_symbol():
    assert set(declared_functions()) <= set(_abi.SIGNATURES)


def test_host_entry_points_validate_arguments():
    lib = _abi.load()
    out = A.TiltGeom()
    tp = A.TiltParams(0, 0, 1, 0, 0)
    gp = A.GlobalParams(0, 0, 0, 0, 0)
    assert lib.artia_geom_tilt(C.byref(tp), C.byref(gp), 0, 10, 1.0, 0.0, C.byref(out)) != 0
    assert lib.artia_geom_tilt(None, C.byref(gp), 10, 10, 1.0, 0.0, C.byref(out)) != 0
    assert lib.artia_geom_volume(0, 1, 1, _abi.farr([1, 1, 1]), _abi.farr([0, 0, 0]), C.byref(A.VolGeom())) != 0
    assert lib.artia_projection_index_list(_abi.farr([0.0]), 0, (C.c_int * 1)()) != 0
    assert lib.artia_bp(None, None, None, None, None, 1, 1, 1, 1.0, 1, 0.0, 1.0, 0, None) != 0
    assert lib.artia_version().decode().startswith("artia")


def test_cpp_functor_adapter_compiles(tmp_path):
    """include/artia_kernels.hpp (the Kernels.h-shaped C++ adapter) compiles against artia.h."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    assert gxx, "g++ is required"
    src = tmp_path / "adapter.cpp"
    src.write_text('#include "artia_kernels.hpp"\n'
                   "int main() { artia::Context* c = nullptr; (void)c; return 0; }\n")
    res = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", str(ROOT / "include"),
                          str(src)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
