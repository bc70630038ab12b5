"""ctypes view of the CPU oracle (oracle/liboracle.so) for tests, smoke() and the CPU baseline.

The oracle's geometry carriers have the same memory layout as the product's artia_tilt_geom /
artia_vol_geom / artia_tilt_params / artia_global_params / artia_sart_params, so the same
ctypes structure classes describe both.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from artiatomi_amd._abi import GlobalParams, SartParams, TiltGeom, TiltParams, VolGeom

ROOT = Path(__file__).resolve().parents[1]


class CtfConsts(C.Structure):
    """Oracle CtfConsts: c_cs, c_voltage, c_pixelsize, c_pixelcount, c_freqStepSize,
    c_phaseContrast, c_ampContrast."""

    _fields_ = [("cs_mm", C.c_float), ("voltage_kV", C.c_float), ("pixelsize_m", C.c_float),
                ("pixelcount", C.c_float), ("freq_step", C.c_float), ("phase_contrast", C.c_float),
                ("amp_contrast", C.c_float)]


class CtfSlabParams(C.Structure):
    """Oracle CtfSlabParams: CTFSliceThickness, pixel size, VolumeShift.z, IgnoreZShiftForCTF,
    SwitchCTFDirectionForIMOD."""

    _fields_ = [("slice_thickness_nm", C.c_float), ("pixelsize_nm", C.c_float), ("volume_shift_z", C.c_float),
                ("ignore_z_shift", C.c_int), ("switch_direction", C.c_int)]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"

_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not ORACLE_SO.exists():
        subprocess.run(["bash", str(ROOT / "oracle" / "build.sh")], check=True)
    lib = C.CDLL(str(ORACLE_SO))
    P = C.c_void_p
    for suf in ("f32", "f64"):
        getattr(lib, f"oracle_march_{suf}").argtypes = [P, P, P, C.c_int, C.c_int, C.c_size_t, P, P, P]
        getattr(lib, f"oracle_slicer_{suf}").argtypes = [P, P, P, C.c_int, C.c_int, C.c_size_t, P, C.c_double,
                                                          C.c_double, P]
        getattr(lib, f"oracle_vol_traversal_{suf}").argtypes = [P, P, C.c_int, C.c_int, C.c_size_t, P, P]
        getattr(lib, f"oracle_compare_{suf}").argtypes = [C.c_int, C.c_int, C.c_size_t, P, P, P, C.c_float, P, P,
                                                           C.c_float]
        getattr(lib, f"oracle_crop_border_{suf}").argtypes = [C.c_int, C.c_int, C.c_size_t, P, P, P, P]
        getattr(lib, f"oracle_back_project_{suf}").argtypes = [P, P, P, P, C.c_int, C.c_int, C.c_size_t, C.c_float,
                                                                C.c_int, C.c_float, C.c_float, C.c_int]
        getattr(lib, f"oracle_make_square_{suf}").argtypes = [C.c_int, C.c_int, C.c_size_t, P, C.c_int, P, C.c_int,
                                                               C.c_int, C.c_int, C.c_int]
        getattr(lib, f"oracle_sart_{suf}").argtypes = [P, P, P, C.c_int, P, P, C.c_int, C.c_int, P, P, C.c_int, P]
        getattr(lib, f"oracle_march_adjoint_{suf}").argtypes = [P, P, P, C.c_int, C.c_int, C.c_size_t, P, P,
                                                                 C.c_double]
        getattr(lib, f"oracle_march_shifted_{suf}").argtypes = [P, P, P, C.c_int, C.c_int, C.c_size_t, P, P,
                                                                 C.c_double, C.c_double]
    lib.oracle_real_length.restype = C.c_float
    lib.oracle_ctf_wavelength.restype = C.c_float
    lib.oracle_ctf_wavelength.argtypes = [C.c_float]
    lib.oracle_ctf_consts.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, C.c_float, P]
    lib.oracle_defocus_distances.argtypes = [P, P, C.c_int, C.c_int, P, P]
    lib.oracle_defocus_min_max.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, P, P, P]
    for suf in ("f32", "f64"):
        getattr(lib, f"oracle_ctf_apply_{suf}").argtypes = [P, C.c_int, C.c_size_t, P, C.c_float, C.c_float,
                                                             C.c_float, C.c_int, C.c_int, C.c_float, P]
        getattr(lib, f"oracle_wbp_weight_{suf}").argtypes = [P, C.c_int, C.c_size_t, C.c_float, C.c_int]
        getattr(lib, f"oracle_fourier_filter_{suf}").argtypes = [P, C.c_int, C.c_size_t, C.c_float, C.c_float,
                                                                  C.c_float, C.c_float]
        getattr(lib, f"oracle_dose_weighting_{suf}").argtypes = [P, C.c_int, C.c_size_t, C.c_float, C.c_float]
    lib.oracle_set_threads.argtypes = [C.c_int]
    lib.oracle_get_threads.restype = C.c_int
    _lib = lib
    return lib


def _suffix(dtype) -> str:
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def ref(o):
    return C.byref(o)


def tilt_geometry(tilt_deg, W, H, psi_deg=0.0, mag=1.0, shift=(0.0, 0.0), glob=None, mag_aniso=(1.0, 0.0)):
    g = TiltGeom()
    tp = TiltParams(tilt_deg, psi_deg, mag, shift[0], shift[1])
    gp = glob if glob is not None else GlobalParams(0, 0, 0, 0, 0)
    load().oracle_tilt_geometry(ref(tp), ref(gp), W, H, C.c_float(mag_aniso[0]), C.c_float(mag_aniso[1]), ref(g))
    return g


def volume_geometry(nx, ny, nz, vs=(1.0, 1.0, 1.0), shift=(0.0, 0.0, 0.0)):
    v = VolGeom()
    load().oracle_volume_geometry(nx, ny, nz, (C.c_float * 3)(*vs), (C.c_float * 3)(*shift), ref(v))
    return v


def subvolume_geometry(n, vs_sub, vs_full, full, pos, shift=(0.0, 0.0, 0.0)):
    v = VolGeom()
    load().oracle_subvolume_geometry(n, C.c_float(vs_sub), (C.c_float * 3)(*vs_full), ref(full),
                                     (C.c_float * 3)(*pos), (C.c_float * 3)(*shift), ref(v))
    return v


def hit_points(full, g):
    out = (C.c_int * 8)()
    load().oracle_hit_points(ref(full), ref(g), out)
    return list(out)


def real_length(full, g) -> float:
    return float(load().oracle_real_length(ref(full), ref(g)))


def index_list(tilts):
    n = len(tilts)
    out = (C.c_int * n)()
    load().oracle_index_list((C.c_float * n)(*tilts), n, out)
    return list(out)


def glibc_rand(seed, count):
    out = (C.c_int * count)()
    load().oracle_glibc_rand(C.c_uint(seed), count, out)
    return list(out)


def march(vg, tg, vol, W, H, dtype=np.float32, roi=None, proj=None):
    vol = np.ascontiguousarray(vol, dtype=dtype)
    proj = np.zeros((H, W), dtype) if proj is None else np.ascontiguousarray(proj, dtype=dtype).copy()
    dist = np.zeros((H, W), dtype)
    r = (C.c_int * 4)(*(roi or (0, 0, W, H)))
    getattr(load(), f"oracle_march_{_suffix(dtype)}")(ref(vg), ref(tg), _p(vol), W, H, W, _p(proj), _p(dist), r)
    return proj, dist


def march_adjoint(vg, tg, resid, dtype=np.float64, roi=None, scale=1.0):
    """scale * P^T resid for march's projection output P."""
    H, W = resid.shape
    resid = np.ascontiguousarray(resid, dtype=dtype)
    grad = np.zeros((vg.nz, vg.ny, vg.nx), dtype)
    r = (C.c_int * 4)(*(roi or (0, 0, W, H)))
    getattr(load(), f"oracle_march_adjoint_{_suffix(dtype)}")(ref(vg), ref(tg), _p(resid), W, H, W, _p(grad), r,
                                                              float(scale))
    return grad


def march_shifted(vg, tg, vol, W, H, sx, sy, dtype=np.float64, roi=None):
    """march's projection with the tilt shift changed by (sx, sy) pixels in T precision."""
    vol = np.ascontiguousarray(vol, dtype=dtype)
    proj = np.zeros((H, W), dtype)
    r = (C.c_int * 4)(*(roi or (0, 0, W, H)))
    getattr(load(), f"oracle_march_shifted_{_suffix(dtype)}")(ref(vg), ref(tg), _p(vol), W, H, W, _p(proj), r,
                                                              float(sx), float(sy))
    return proj


def slicer(vg, tg, vol, W, H, tmin, tmax, dtype=np.float32, roi=None):
    vol = np.ascontiguousarray(vol, dtype=dtype)
    proj = np.zeros((H, W), dtype)
    r = (C.c_int * 4)(*(roi or (0, 0, W, H)))
    getattr(load(), f"oracle_slicer_{_suffix(dtype)}")(ref(vg), ref(tg), _p(vol), W, H, W, _p(proj), float(tmin),
                                                       float(tmax), r)
    return proj


def vol_traversal(vg, tg, W, H, dtype=np.float32, roi=None):
    dist = np.zeros((H, W), dtype)
    r = (C.c_int * 4)(*(roi or (0, 0, W, H)))
    getattr(load(), f"oracle_vol_traversal_{_suffix(dtype)}")(ref(vg), ref(tg), W, H, W, _p(dist), r)
    return dist


def compare(real, virt, dist, real_len, crop=(0, 0, 0, 0), crop_dim=(0, 0, 0, 0), scale=1.0, dtype=np.float32):
    H, W = virt.shape
    real = np.ascontiguousarray(real, dtype=dtype)
    out = np.ascontiguousarray(virt, dtype=dtype).copy()
    dist = np.ascontiguousarray(dist, dtype=dtype)
    getattr(load(), f"oracle_compare_{_suffix(dtype)}")(W, H, W, _p(real), _p(out), _p(dist), C.c_float(real_len),
                                                        (C.c_float * 4)(*crop), (C.c_float * 4)(*crop_dim),
                                                        C.c_float(scale))
    return out


def crop_border(img, points, cut=(0, 0), dim=(0, 0), dtype=np.float32):
    H, W = img.shape
    out = np.ascontiguousarray(img, dtype=dtype).copy()
    getattr(load(), f"oracle_crop_border_{_suffix(dtype)}")(W, H, W, _p(out), (C.c_float * 2)(*cut),
                                                            (C.c_float * 2)(*dim), (C.c_int * 8)(*points))
    return out


def back_project(vg, tg, vol, img, lam, os_=1, dist_min=0.0, dist_max=9999999999999.0, filt=0, dtype=np.float32):
    H, W = img.shape
    out = np.ascontiguousarray(vol, dtype=dtype).copy()
    img = np.ascontiguousarray(img, dtype=dtype)
    getattr(load(), f"oracle_back_project_{_suffix(dtype)}")(ref(vg), ref(tg), _p(out), _p(img), W, H, W,
                                                             C.c_float(lam), os_, C.c_float(dist_min),
                                                             C.c_float(dist_max), filt)
    return out


def make_square(img, N, bx, by, mirror_y=False, fill_zero=True, dtype=np.float32):
    H, W = img.shape
    img = np.ascontiguousarray(img, dtype=dtype)
    out = np.zeros((N, N), dtype)
    getattr(load(), f"oracle_make_square_{_suffix(dtype)}")(W, H, W, _p(img), N, _p(out), bx, by, int(mirror_y),
                                                            int(fill_zero))
    return out


def sart(vg, full, geoms, index_list_, projections, vol, params: SartParams, iterations, empty=True,
         dtype=np.float32):
    n = len(geoms)
    T, H, W = projections.shape
    assert T == n
    tg = (TiltGeom * n)(*geoms)
    il = (C.c_int * n)(*index_list_)
    proj = np.ascontiguousarray(projections, dtype=dtype)
    out = np.ascontiguousarray(vol, dtype=dtype).copy()
    flag = C.c_int(1 if empty else 0)
    getattr(load(), f"oracle_sart_{_suffix(dtype)}")(ref(vg), ref(full), tg, n, il, _p(proj), W, H, _p(out),
                                                     ref(params), iterations, C.byref(flag))
    return out, bool(flag.value)


def set_threads(n: int) -> None:
    load().oracle_set_threads(n)


# ------------------------------------------------------------------ CTF / WBP weighting

def ctf_wavelength(kV: float) -> float:
    return load().oracle_ctf_wavelength(C.c_float(kV))


def ctf_consts(cs_mm, kV, ps_nm, N, amp=0.0) -> CtfConsts:
    k = CtfConsts()
    load().oracle_ctf_consts(C.c_float(cs_mm), C.c_float(kV), C.c_float(ps_nm), int(N), C.c_float(amp), ref(k))
    return k


def _interleaved(spec, dtype):
    a = np.ascontiguousarray(spec)
    out = np.empty(a.shape + (2,), dtype)
    out[..., 0] = a.real
    out[..., 1] = a.imag
    return out


def ctf_apply(spec, consts: CtfConsts, dmin_nm, dmax_nm, angle_deg, for_fp=True, phase_flip=False, wiener=0.1,
              beta=(200.0, 0.0, 0.008, 0.0), dtype=np.float64):
    """CTF multiply of an (N, N/2+1) half spectrum; returns a complex array."""
    N = spec.shape[0]
    assert spec.shape == (N, N // 2 + 1)
    buf = _interleaved(spec, dtype)
    getattr(load(), f"oracle_ctf_apply_{_suffix(dtype)}")(_p(buf), N, N // 2 + 1, ref(consts), C.c_float(dmin_nm),
                                                          C.c_float(dmax_nm), C.c_float(angle_deg), int(for_fp),
                                                          int(phase_flip), C.c_float(wiener),
                                                          (C.c_float * 4)(*beta))
    return buf[..., 0] + 1j * buf[..., 1]


def wbp_weight(spec, psi_deg, method, dtype=np.float64):
    N = spec.shape[0]
    assert spec.shape == (N, N // 2 + 1)
    buf = _interleaved(spec, dtype)
    getattr(load(), f"oracle_wbp_weight_{_suffix(dtype)}")(_p(buf), N, N // 2 + 1, C.c_float(psi_deg), int(method))
    return buf[..., 0] + 1j * buf[..., 1]


def fourier_filter(spec, lp, hp, lps, hps, dtype=np.float64):
    N = spec.shape[0]
    assert spec.shape == (N, N // 2 + 1)
    buf = _interleaved(spec, dtype)
    getattr(load(), f"oracle_fourier_filter_{_suffix(dtype)}")(_p(buf), N, N // 2 + 1, C.c_float(lp), C.c_float(hp),
                                                               C.c_float(lps), C.c_float(hps))
    return buf[..., 0] + 1j * buf[..., 1]


def dose_weighting(spec, dose, pixelsize_A, dtype=np.float64):
    N = spec.shape[0]
    assert spec.shape == (N, N // 2 + 1)
    buf = _interleaved(spec, dtype)
    getattr(load(), f"oracle_dose_weighting_{_suffix(dtype)}")(_p(buf), N, N // 2 + 1, C.c_float(dose),
                                                               C.c_float(pixelsize_A))
    return buf[..., 0] + 1j * buf[..., 1]


def defocus_distances(full, g, W, H):
    a, b = C.c_float(), C.c_float()
    load().oracle_defocus_distances(ref(full), ref(g), W, H, C.byref(a), C.byref(b))
    return a.value, b.value


def defocus_min_max(ray, tilt_deg, dmin_nm, dmax_nm, slab: CtfSlabParams):
    a, b = C.c_float(), C.c_float()
    load().oracle_defocus_min_max(C.c_float(ray), C.c_float(tilt_deg), C.c_float(dmin_nm), C.c_float(dmax_nm),
                                  ref(slab), C.byref(a), C.byref(b))
    return a.value, b.value
