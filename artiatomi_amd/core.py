"""Host-side mirror of the EmSART reconstruction-core operator interface.

`ArtiaContext` exposes one method per reference functor of src/EmSART/Kernels.h:37-245
(FPKernel -> fp_march, SlicerKernel -> fp_slicer, VolTravLengthKernel -> vol_traversal,
CompKernel -> compare, CropBorderKernel -> crop_border, BPKernel -> bp,
CopyToSquareKernel -> make_square, CTFKernel -> ctf_apply, WbpWeightingKernel -> wbp_weight,
plus the FFT slab filters of Reconstructor.cpp -> ctf_slab_filter / wbp_filter) over the C ABI in include/artia.h, taking torch CUDA
tensors as device memory.  Like the reference functors, each call returns the elapsed
milliseconds when `timed=True` (the call then synchronises); otherwise it returns None and
does not block.  Errors raise `ArtiaError` carrying the library's message (the reference
throws CudaException, CudaException.h:34-36).

Geometry helpers mirror Projection / Volume / Reconstructor host code through the ABI.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

from . import _abi
from ._abi import (CtfParams, GlobalParams, SartParams, Subvol, SubvolTuning, TiltDefocus, TiltGeom, TiltParams,
                   VolGeom)


class ArtiaError(RuntimeError):
    pass


def _check(status: int, ctx=None, what: str = "") -> None:
    if status != _abi.ARTIA_OK:
        msg = ""
        if ctx is not None:
            msg = _abi.load().artia_last_error(ctx).decode(errors="replace")
        raise ArtiaError(f"{what} failed with status {status}: {msg}")


# ------------------------------------------------------------------ geometry

def tilt_geometry(tilt_deg: float, width: int, height: int, psi_deg: float = 0.0, mag: float = 1.0,
                  shift: Sequence[float] = (0.0, 0.0), glob: Optional[GlobalParams] = None,
                  mag_aniso: Sequence[float] = (1.0, 0.0)) -> TiltGeom:
    """Projection::GetPosition/GetPixelU,VPitch/GetNormalVector/GetDetectorMatrix for one tilt."""
    lib = _abi.load()
    tp = TiltParams(tilt_deg, psi_deg, mag, shift[0], shift[1])
    gp = glob if glob is not None else GlobalParams(0.0, 0.0, 0.0, 0, 0.0)
    out = TiltGeom()
    _check(lib.artia_geom_tilt(C.byref(tp), C.byref(gp), int(width), int(height), float(mag_aniso[0]),
                               float(mag_aniso[1]), C.byref(out)), what="artia_geom_tilt")
    return out


def volume_geometry(nx: int, ny: int, nz: int, voxel_size: Sequence[float] = (1.0, 1.0, 1.0),
                    shift: Sequence[float] = (0.0, 0.0, 0.0)) -> VolGeom:
    """Volume::PositionInSpace for the whole volume."""
    lib = _abi.load()
    out = VolGeom()
    _check(lib.artia_geom_volume(nx, ny, nz, _abi.farr(voxel_size, 3), _abi.farr(shift, 3), C.byref(out)),
           what="artia_geom_volume")
    return out


def subvolume_geometry(n: int, voxel_size_sub: float, voxel_size_full: Sequence[float], full: VolGeom,
                       position: Sequence[float], shift: Sequence[float] = (0.0, 0.0, 0.0)) -> VolGeom:
    """Volume::PositionInSpace for a motive-list sub-volume (1-based position)."""
    lib = _abi.load()
    out = VolGeom()
    _check(lib.artia_geom_subvolume(n, voxel_size_sub, _abi.farr(voxel_size_full, 3), C.byref(full),
                                    _abi.farr(position, 3), _abi.farr(shift, 3), C.byref(out)),
           what="artia_geom_subvolume")
    return out


def subvolume_tilt_geometry(tilt_deg: float, width: int, height: int, psi_deg: float = 0.0, mag: float = 1.0,
                            shift: Sequence[float] = (0.0, 0.0), extra_shift: Sequence[float] = (0.0, 0.0),
                            glob: Optional[GlobalParams] = None) -> TiltGeom:
    """Tilt geometry seen by one sub-volume: the marker shift plus the sub-volume's extra shift,
    added in float as Projection::GetPosition does (Projection.cpp:183-184)."""
    import numpy as np

    sx = float(np.float32(shift[0]) + np.float32(extra_shift[0]))
    sy = float(np.float32(shift[1]) + np.float32(extra_shift[1]))
    return tilt_geometry(tilt_deg, width, height, psi_deg=psi_deg, mag=mag, shift=(sx, sy), glob=glob)


def footprint_window(vg: VolGeom, tg: TiltGeom, width: int, height: int, margin: int = 0) -> tuple:
    """Detector window (x0, y0, w, h) holding the projection of the volume's bounding box at tilt
    geometry tg, widened by `margin` pixels and clipped to the width x height detector: the eight
    box corners through the detector matrix (as BackProjectionSquareOS.cu:246-253 maps a point to
    a pixel) and the inverse anisotropic magnification.  (0, 0, 0, 0) when nothing is left."""
    import math

    M = list(tg.det_matrix)
    Ai = list(tg.mag_aniso_inv)
    xs, ys = [], []
    for c in range(8):
        p = [vg.bbox_max[i] if (c >> i) & 1 else vg.bbox_min[i] for i in range(3)]
        a = M[0] * p[0] + M[1] * p[1] + M[2] * p[2] + M[3]
        b = M[4] * p[0] + M[5] * p[1] + M[6] * p[2] + M[7]
        xs.append(Ai[0] * a + Ai[1] * b + Ai[2])
        ys.append(Ai[3] * a + Ai[4] * b + Ai[5])
    x0 = max(0, int(math.floor(min(xs))) - margin)
    y0 = max(0, int(math.floor(min(ys))) - margin)
    x1 = min(width, int(math.ceil(max(xs))) + 1 + margin)
    y1 = min(height, int(math.ceil(max(ys))) + 1 + margin)
    if x1 <= x0 or y1 <= y0:
        return (0, 0, 0, 0)
    return (x0, y0, x1 - x0, y1 - y0)


def hitpoints(full: VolGeom, tg: TiltGeom) -> list:
    lib = _abi.load()
    out = (C.c_int * 8)()
    _check(lib.artia_geom_hitpoints(C.byref(full), C.byref(tg), out), what="artia_geom_hitpoints")
    return list(out)


def real_length(full: VolGeom, tg: TiltGeom) -> float:
    lib = _abi.load()
    out = C.c_float()
    _check(lib.artia_geom_real_length(C.byref(full), C.byref(tg), C.byref(out)), what="artia_geom_real_length")
    return out.value


def projection_index_list(tilts_deg: Sequence[float]) -> list:
    lib = _abi.load()
    n = len(tilts_deg)
    out = (C.c_int * n)()
    _check(lib.artia_projection_index_list(_abi.farr(tilts_deg), n, out), what="artia_projection_index_list")
    return list(out)


def defocus_distances(full: VolGeom, tg: TiltGeom) -> tuple:
    """[t_in, t_out) of the CTF slab loop along the tilt normal (Reconstructor.cpp:54-173)."""
    a, b = C.c_float(), C.c_float()
    _check(_abi.load().artia_geom_defocus_distances(C.byref(full), C.byref(tg), C.byref(a), C.byref(b)),
           what="artia_geom_defocus_distances")
    return a.value, b.value


def defocus_minmax(ray: float, d: TiltDefocus, p: CtfParams) -> tuple:
    """Defocus (nm) of the slab starting at `ray` (Reconstructor.cpp:178-201)."""
    a, b = C.c_float(), C.c_float()
    _check(_abi.load().artia_geom_defocus_minmax(float(ray), C.byref(d), C.byref(p), C.byref(a), C.byref(b)),
           what="artia_geom_defocus_minmax")
    return a.value, b.value


def ctf_params(cs_mm: float = 2.7, voltage_kV: float = 300.0, pixelsize_nm: float = 1.0, amp_contrast: float = 0.0,
               wiener: float = 0.1, beta_fac=(200.0, 0.0, 0.008, 0.0), phase_flip_only: bool = False,
               slice_thickness_nm: float = 50.0, volume_shift_z: float = 0.0, ignore_z_shift: bool = False,
               switch_direction: bool = False) -> CtfParams:
    """CTF settings; the defaults are EmSART's (CTFSliceThickness 50 nm, CTFBetaFac
    (200, 0, 0.008, 0), WienerFilterNoiseLevel 0.1, amplitude contrast 0)."""
    return CtfParams(cs_mm, voltage_kV, pixelsize_nm, amp_contrast, wiener, _abi.farr(beta_fac, 4),
                     int(bool(phase_flip_only)), slice_thickness_nm, volume_shift_z, int(bool(ignore_z_shift)),
                     int(bool(switch_direction)))


@dataclass
class Acquisition:
    """Per-tilt geometry of a tilt series, as EmSART builds it from the marker file."""

    width: int
    height: int
    tilts_deg: list
    psis_deg: list
    shifts: list
    geoms: list = field(default_factory=list)
    hitpoints: list = field(default_factory=list)
    real_lengths: list = field(default_factory=list)
    index_list: list = field(default_factory=list)
    slab_ranges: list = field(default_factory=list)

    @classmethod
    def build(cls, full: VolGeom, width: int, height: int, tilts_deg, psis_deg=None, shifts=None,
              glob: Optional[GlobalParams] = None) -> "Acquisition":
        n = len(tilts_deg)
        psis = list(psis_deg) if psis_deg is not None else [0.0] * n
        sh = [tuple(s) for s in shifts] if shifts is not None else [(0.0, 0.0)] * n
        acq = cls(width, height, [float(t) for t in tilts_deg], [float(p) for p in psis], sh)
        for t, p, s in zip(acq.tilts_deg, psis, sh):
            g = tilt_geometry(t, width, height, psi_deg=p, shift=s, glob=glob)
            acq.geoms.append(g)
            acq.hitpoints.append(hitpoints(full, g))
            acq.real_lengths.append(real_length(full, g))
        acq.index_list = projection_index_list(acq.tilts_deg)
        acq.slab_ranges = [defocus_distances(full, g) for g in acq.geoms]
        return acq

    def c_arrays(self):
        n = len(self.geoms)
        tg = (TiltGeom * n)(*self.geoms)
        hp = (C.c_int * (8 * n))(*[v for h in self.hitpoints for v in h])
        rl = (C.c_float * n)(*self.real_lengths)
        il = (C.c_int * n)(*self.index_list)
        return tg, hp, rl, il


def sart_params(lambda_: float = 0.3, sirt_count: int = 1, oversampling: int = 1, crop=(0, 0, 0, 0),
                crop_dim=(0, 0, 0, 0), cut_length=(0, 0), dim_length=(0, 0), proj_scale: float = 1.0) -> SartParams:
    return SartParams(sirt_count, lambda_, oversampling, _abi.farr(crop, 4), _abi.farr(crop_dim, 4),
                      _abi.farr(cut_length, 2), _abi.farr(dim_length, 2), proj_scale)


def marker_tilt_params(marker_data, proj: int, ref_marker: int = 0):
    """MarkerFile values of projection proj (tilt, psi, magnification, shifts) at the reference
    marker, and whether the projection is good (CheckIfProjIndexIsGood).  marker_data: the marker
    file's float array, shape (markers, projections, items)."""
    import numpy as np

    a = np.ascontiguousarray(marker_data, np.float32)
    if a.ndim != 3:
        raise ArtiaError("marker data must be (markers, projections, items)")
    tp, good = TiltParams(), C.c_int()
    lib = _abi.load()
    st = lib.artia_marker_tilt_params(a.ctypes.data_as(C.POINTER(C.c_float)), a.shape[2], a.shape[1], a.shape[0],
                                      ref_marker, proj, C.byref(tp), C.byref(good))
    _check(st, what="artia_marker_tilt_params")
    return tp, bool(good.value)


def ctf_file_defocus(ctf_data, index: int, tilt_deg: float) -> TiltDefocus:
    """CtfFile defocus of projection index: ctf_data is the CTF file's float array, shape (rows,
    projections)."""
    import numpy as np

    a = np.ascontiguousarray(ctf_data, np.float32)
    if a.ndim != 2:
        raise ArtiaError("CTF data must be (rows, projections)")
    d = TiltDefocus()
    st = _abi.load().artia_ctf_file_defocus(a.ctypes.data_as(C.POINTER(C.c_float)), a.shape[1], a.shape[0], index,
                                            float(tilt_deg), C.byref(d))
    _check(st, what="artia_ctf_file_defocus")
    return d


def _as_f32(a, ndim: int, what: str):
    import numpy as np

    a = np.ascontiguousarray(a, np.float32)
    if a.ndim != ndim:
        raise ArtiaError(f"{what} must have {ndim} dimensions")
    return a


def shift_file_get(shift_data, proj: int, motive: int) -> tuple:
    """ShiftFile::operator(): the (negated) extra shift of motive at projection proj.  shift_data is
    the shift file's float array, shape (2, projections, motives)."""
    a = _as_f32(shift_data, 3, "shift file data")
    out = (C.c_float * 2)()
    st = _abi.load().artia_shift_file_get(a.ctypes.data_as(C.POINTER(C.c_float)), a.shape[2], a.shape[1], a.shape[0],
                                          proj, motive, out)
    _check(st, what="artia_shift_file_get")
    return float(out[0]), float(out[1])


def shift_file_set(shift_data, proj: int, motive: int, value) -> None:
    """ShiftFile::SetValue into a writable float32 array of shape (2, projections, motives)."""
    import numpy as np

    if not (isinstance(shift_data, np.ndarray) and shift_data.dtype == np.float32 and shift_data.ndim == 3
            and shift_data.flags.c_contiguous and shift_data.flags.writeable):
        raise ArtiaError("shift file data must be a writable C-contiguous float32 array (2, projections, motives)")
    st = _abi.load().artia_shift_file_set(shift_data.ctypes.data_as(C.POINTER(C.c_float)), shift_data.shape[2],
                                          shift_data.shape[1], shift_data.shape[0], proj, motive,
                                          _abi.farr(value, 2))
    _check(st, what="artia_shift_file_set")


def rot_matrix(phi: float, psi: float, theta: float):
    """computeRotMat's float matrix (artia_rot_matrix): 3 x 3 numpy float32, R = Rz(psi) Rx(theta) Rz(phi)."""
    import numpy as np

    R = (C.c_float * 9)()
    _abi.load().artia_rot_matrix(float(phi), float(psi), float(theta), R)
    return np.array(list(R), dtype=np.float32).reshape(3, 3)


class MotiveList:
    """A motive list (MotiveListe.h:33-106) over the file's float array of shape (particles, 20),
    with the position and shift binnings EmSART's ScaleMotivelistPosition/-Shift apply."""

    GROUP_MODES = {"by_class": _abi.ARTIA_GROUP_BY_CLASS, "max_dist": _abi.ARTIA_GROUP_MAX_DIST,
                   "max_count": _abi.ARTIA_GROUP_MAX_COUNT}

    def __init__(self, data, binning_click: float = 1.0, binning_shift: float = 1.0):
        self.data = _as_f32(data, 2, "motive list data")
        if self.data.shape[1] != 20:
            raise ArtiaError("a motive list row has 20 floats")
        self.click, self.shift = float(binning_click), float(binning_shift)
        self._p = self.data.ctypes.data_as(C.POINTER(C.c_float))

    @classmethod
    def from_file(cls, path, binning_click: float = 1.0, binning_shift: float = 1.0) -> "MotiveList":
        info, data = read_file(path, "em")
        return cls(data.reshape(info.ny, info.nx), binning_click, binning_shift)

    def __len__(self) -> int:
        return self.data.shape[0]

    def __getitem__(self, index: int) -> "_abi.Motive":
        m = _abi.Motive()
        _check(_abi.load().artia_motive_get(self._p, 20, len(self), int(index), self.click, self.shift, C.byref(m)),
               what="artia_motive_get")
        return m

    def group_count(self, mode: str) -> int:
        n = C.c_int()
        _check(_abi.load().artia_motive_group_count(self._p, 20, len(self), self.GROUP_MODES[mode], C.byref(n)),
               what="artia_motive_group_count")
        return n.value

    def group(self, mode: str, group: int, max_distance: float = 0.0, group_size: int = 0) -> list:
        """Row indices of the particles of group `group` (GetNeighbours), the group's first particle first."""
        out = (C.c_int * len(self))()
        n = C.c_int()
        _check(_abi.load().artia_motive_group(self._p, 20, len(self), self.GROUP_MODES[mode], int(group),
                                              float(max_distance), int(group_size), self.click, self.shift, out,
                                              len(self), C.byref(n)), what="artia_motive_group")
        return list(out[:n.value])


def subvolume_job(full: VolGeom, acq: "Acquisition", motives: MotiveList, shift_data, size: int,
                  voxel_size_sub: float, voxel_size_full: Sequence[float] = (1.0, 1.0, 1.0),
                  max_shift: float = 1e30, glob: Optional[GlobalParams] = None):
    """EmSARTSubVolumes' inputs for the sub-volume WBP (EmSartSubVolumes.cpp:611-662), from the motive
    list and the shift file: per motive the sub-volume geometry (PositionInSpace with the motive's
    position and shift), per (motive, tilt) the tilt geometry with the shift file's extra shift,
    and whether the pair is backprojected (|extra shift| < MaxShift - 0.5).  Returns (volume
    geometries, tilt geometries [s * ntilts + t], active flags [s * ntilts + t])."""
    import numpy as np

    shifts = _as_f32(shift_data, 3, "shift file data") if shift_data is not None else None
    vgs, geoms, active = [], [], []
    for s in range(len(motives)):
        m = motives[s]
        vgs.append(subvolume_geometry(size, voxel_size_sub, voxel_size_full, full, (m.x, m.y, m.z),
                                      (m.x_shift, m.y_shift, m.z_shift)))
        for t in range(len(acq.geoms)):
            es = shift_file_get(shifts, t, s) if shifts is not None else (0.0, 0.0)
            ex, ey = np.float32(es[0]), np.float32(es[1])
            active.append(bool(np.sqrt(ex * ex + ey * ey) < np.float32(max_shift) - np.float32(0.5)))
            geoms.append(subvolume_tilt_geometry(acq.tilts_deg[t], acq.width, acq.height, psi_deg=acq.psis_deg[t],
                                                 shift=acq.shifts[t], extra_shift=(float(ex), float(ey)), glob=glob))
    return vgs, geoms, active


FILE_FORMATS = {"em": _abi.ARTIA_FILE_EM, "mrc": _abi.ARTIA_FILE_MRC}
# numpy element types of the EM DataType / MRC MODE codes that have one
_EM_NP = {1: "u1", 2: "i2", 4: "i4", 5: "f4", 8: "c8", 9: "f8", 10: "f2"}
_MRC_NP = {0: "u1", 1: "i2", 2: "f4", 4: "c8", 6: "u2", 12: "f2", 64: "f2"}


def parse_file_header(fmt: str, header: bytes) -> "_abi.FileInfo":
    """Sizes, element type and data offset of an EM or MRC file from its first bytes (512 / 1024)."""
    info = _abi.FileInfo()
    buf = C.create_string_buffer(bytes(header), len(header))
    st = _abi.load().artia_file_parse_header(FILE_FORMATS[fmt], buf, len(header), C.byref(info))
    _check(st, what="artia_file_parse_header")
    return info


def read_file(path, fmt: Optional[str] = None):
    """(FileInfo, array) of an EM or MRC stack or volume: the data as a read-only numpy memory map of
    shape (nz, ny, nx) in the file's byte order.  fmt defaults to the file's suffix (.em, else MRC)."""
    import numpy as np

    fmt = fmt or ("em" if str(path).lower().endswith(".em") else "mrc")
    with open(path, "rb") as f:
        head = f.read(1024 if fmt == "mrc" else 512)
    info = parse_file_header(fmt, head)
    code = (_EM_NP if fmt == "em" else _MRC_NP).get(info.raw_type)
    if code is None:
        raise ArtiaError(f"{fmt} element type {info.raw_type} has no array type")
    dt = np.dtype(code).newbyteorder(">" if info.byte_swapped else "<")
    data = np.memmap(path, dtype=dt, mode="r", offset=info.data_offset, shape=(info.nz, info.ny, info.nx))
    return info, data


def volume_file_header(fmt: str, nx: int, ny: int, nz: int, pixel_size_nm: float, voxel_size=(1.0, 1.0, 1.0),
                       half: bool = False) -> bytes:
    """The header EmSART writes before a reconstructed volume (EmSart.cpp:717-759)."""
    out = C.create_string_buffer(1024)
    n = C.c_size_t()
    st = _abi.load().artia_file_volume_header(FILE_FORMATS[fmt], int(bool(half)), nx, ny, nz, float(pixel_size_nm),
                                              _abi.farr(voxel_size, 3), out, 1024, C.byref(n))
    _check(st, what="artia_file_volume_header")
    return out.raw[:n.value]


def prep_params(dtype: str = "float32", mirror_y: bool = False, correct_bad_pixels: bool = True,
                bad_pixel_value: float = 10.0, skip_filter: bool = False, lp: float = 800.0, lps: float = 200.0,
                hp: float = 0.0, hps: float = 0.0, dose_weighting: bool = False, pixelsize_nm: float = 1.0,
                wbp: bool = False, wbp_method: str = "ramp", down_weight_tilts: bool = False,
                normalization: str = "mean") -> "_abi.PrepParams":
    """Projection preparation settings; the defaults are EmSART's (fourFilterLP 800, LPS 200, HP 0,
    HPS 0, BadPixelValue 10, CorrectBadPixels on, ProjectionNormalization mean, WBPFilter ramp)."""
    if dtype not in _abi.DTYPES:
        raise ArtiaError(f"unknown projection data type {dtype!r}")
    return _abi.PrepParams(_abi.DTYPES[dtype], int(bool(mirror_y)), int(bool(correct_bad_pixels)), bad_pixel_value,
                           int(bool(skip_filter)), lp, lps, hp, hps, int(bool(dose_weighting)), pixelsize_nm,
                           int(bool(wbp)), _abi.WBP_METHODS[wbp_method], int(bool(down_weight_tilts)),
                           _abi.NORMALIZATIONS[normalization])


# ------------------------------------------------------------------ device operators

def _cptr(t, name: str, n: int) -> int:
    """Device pointer of an (n, n/2+1) complex64 half spectrum."""
    import torch

    if (not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.complex64 or not t.is_contiguous()
            or tuple(t.shape) != (n, n // 2 + 1)):
        raise ArtiaError(f"{name} must be a contiguous ({n}, {n // 2 + 1}) complex64 CUDA tensor")
    return t.data_ptr()


def _ptr(t, name: str, numel: Optional[int] = None) -> int:
    import torch

    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise ArtiaError(f"{name} must be a contiguous float32 CUDA tensor")
    if numel is not None and t.numel() < numel:
        raise ArtiaError(f"{name} holds {t.numel()} elements, needs {numel}")
    return t.data_ptr()


def _vol(t, vg: VolGeom, name: str = "vol"):
    """A volume tensor's pointer and entry-point suffix: contiguous float32 (fp32 storage) or float16
    (half storage, the reference's FP16Volume) CUDA tensor of nx * ny * nz voxels."""
    import torch

    if isinstance(t, torch.Tensor) and t.dtype == torch.float16:
        if not t.is_cuda or not t.is_contiguous() or t.numel() < vg.nx * vg.ny * vg.nz:
            raise ArtiaError(f"{name} must be a contiguous CUDA tensor of {vg.nx * vg.ny * vg.nz} voxels")
        return t.data_ptr(), "_f16"
    return _ptr(t, name, vg.nx * vg.ny * vg.nz), ""


class ArtiaContext:
    """One libartia context on one GPU, bound to torch's current stream of that device.  Volume
    arguments of the projectors, the backprojector and the SART drivers are float32 or float16
    tensors; float16 selects the half-storage entry points."""

    def __init__(self, device: int = 0, numerics: str = "exact"):
        import torch

        self.lib = _abi.load()
        self.device = device
        h = C.c_void_p()
        _check(self.lib.artia_create(device, C.byref(h)), what="artia_create")
        self.h = h
        stream = torch.cuda.current_stream(device)
        _check(self.lib.artia_set_stream(self.h, C.c_void_p(stream.cuda_stream)), self.h, "artia_set_stream")
        self.set_numerics(numerics)

    def set_numerics(self, numerics: str) -> None:
        mode = {"exact": _abi.ARTIA_NUMERICS_EXACT, "reference": _abi.ARTIA_NUMERICS_REFERENCE}[numerics]
        _check(self.lib.artia_set_numerics(self.h, mode), self.h, "artia_set_numerics")
        self.numerics = numerics

    def set_fp_staging(self, enable: bool) -> None:
        """LDS-staged exact forward projection where it applies (default) or direct reads; both
        give bit-identical projections."""
        _check(self.lib.artia_set_fp_staging(self.h, int(bool(enable))), self.h, "artia_set_fp_staging")

    def set_stack_limit(self, nbytes: int) -> None:
        """Bytes the CTF driver may use for its per-tilt slab stacks (0: the built-in caps)."""
        _check(self.lib.artia_set_stack_limit(self.h, int(nbytes)), self.h, "artia_set_stack_limit")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.artia_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        _check(self.lib.artia_sync(self.h), self.h, "artia_sync")

    def _call(self, name, *args, timed: bool = False):
        ms = C.c_float() if timed else None
        st = getattr(self.lib, name)(self.h, *args, C.byref(ms) if timed else None)
        _check(st, self.h, name)
        return ms.value if timed else None

    @staticmethod
    def _roi(roi):
        return None if roi is None else _abi.iarr(roi, 4)

    def fp_march(self, vg: VolGeom, tg: TiltGeom, vol, proj, dist, roi=None, timed=False):
        H, W = proj.shape
        vp, sfx = _vol(vol, vg)
        return self._call("artia_fp_march" + sfx, C.byref(vg), C.byref(tg), vp, W, H, W, _ptr(proj, "proj"),
                          _ptr(dist, "dist", W * H), self._roi(roi), timed=timed)

    def fp_slicer(self, vg: VolGeom, tg: TiltGeom, vol, proj, tmin, tmax, roi=None, timed=False):
        H, W = proj.shape
        vp, sfx = _vol(vol, vg)
        return self._call("artia_fp_slicer" + sfx, C.byref(vg), C.byref(tg), vp, W, H, W, _ptr(proj, "proj"),
                          float(tmin), float(tmax), self._roi(roi), timed=timed)

    def vol_traversal(self, vg: VolGeom, tg: TiltGeom, dist, roi=None, timed=False):
        H, W = dist.shape
        return self._call("artia_vol_traversal", C.byref(vg), C.byref(tg), W, H, W, _ptr(dist, "dist"),
                          self._roi(roi), timed=timed)

    def compare(self, real, virt, dist, real_length, crop=(0, 0, 0, 0), crop_dim=(0, 0, 0, 0), scale=1.0,
                timed=False):
        H, W = virt.shape
        return self._call("artia_compare", W, H, W, _ptr(real, "real", W * H), _ptr(virt, "virt"),
                          _ptr(dist, "dist", W * H), float(real_length), _abi.farr(crop, 4),
                          _abi.farr(crop_dim, 4), float(scale), timed=timed)

    def fp_residual(self, vg: VolGeom, tg: TiltGeom, vol, real, residual, real_length, crop=(0, 0, 0, 0),
                    crop_dim=(0, 0, 0, 0), scale=1.0, points=None, cut_length=(0, 0), dim_length=(0, 0),
                    dist=None, volume_is_empty=False, timed=False):
        """residual = cropBorder(compare(real, FP(vol), mean chord)) in the forward projector's
        epilogue (artia_fp_residual); points=None skips cropBorder.  dist (H, W) receives the chord
        for an empty volume or under reference numerics."""
        H, W = residual.shape
        if vol is None:
            vp, sfx = None, ""
        else:
            vp, sfx = _vol(vol, vg)
        return self._call("artia_fp_residual" + sfx, C.byref(vg), C.byref(tg), vp, int(bool(volume_is_empty)), W, H, W,
                          _ptr(real, "real", W * H), float(real_length), _abi.farr(crop, 4), _abi.farr(crop_dim, 4),
                          float(scale), _abi.farr(cut_length, 2), _abi.farr(dim_length, 2),
                          None if points is None else _abi.iarr(points, 8), _ptr(residual, "residual"),
                          None if dist is None else _ptr(dist, "dist", W * H), timed=timed)

    def crop_border(self, img, points, cut_length=(0, 0), dim_length=(0, 0), timed=False):
        H, W = img.shape
        return self._call("artia_crop_border", W, H, W, _ptr(img, "img"), _abi.farr(cut_length, 2),
                          _abi.farr(dim_length, 2), _abi.iarr(points, 8), timed=timed)

    def bp(self, vg: VolGeom, tg: TiltGeom, vol, img, lambda_, oversampling=1, dist_min=0.0,
           dist_max=9999999999999.0, filter="point", timed=False):
        H, W = img.shape
        f = {"point": _abi.ARTIA_FILTER_POINT, "linear": _abi.ARTIA_FILTER_LINEAR}[filter]
        vp, sfx = _vol(vol, vg)
        return self._call("artia_bp" + sfx, C.byref(vg), C.byref(tg), vp, _ptr(img, "img"), W, H, W, float(lambda_),
                          int(oversampling), float(dist_min), float(dist_max), f, timed=timed)

    def volume_f16_to_f32(self, vol, out, z0=0, count=None, timed=False):
        """convertVolumeFP16ToFP32: out[k] = -vol[z0 + k] as float32 for k < count (all slices from z0
        by default); vol (nz, ny, nx) float16, out (count, ny, nx) float32."""
        import torch

        if not (isinstance(vol, torch.Tensor) and vol.is_cuda and vol.dtype == torch.float16 and vol.is_contiguous()
                and vol.dim() == 3):
            raise ArtiaError("vol must be a contiguous (nz, ny, nx) float16 CUDA tensor")
        nz, ny, nx = vol.shape
        count = nz - z0 if count is None else int(count)
        return self._call("artia_volume_f16_to_f32", C.c_void_p(vol.data_ptr()), nx, ny, nz, int(z0), count,
                          _ptr(out, "out", max(count, 0) * ny * nx), nx, timed=timed)

    def make_square(self, img, out, border_x, border_y, mirror_y=False, fill_zero=True, timed=False):
        H, W = img.shape
        N = out.shape[0]
        return self._call("artia_make_square", W, H, W, _ptr(img, "img"), N, _ptr(out, "out", N * N),
                          int(border_x), int(border_y), int(bool(mirror_y)), int(bool(fill_zero)), timed=timed)

    def ctf_apply(self, spec, params: CtfParams, defocus_min_nm, defocus_max_nm, angle_deg, for_fp=True,
                  timed=False):
        """In-place CTF multiply (for_fp) or Wiener inverse on an (n, n/2+1) half spectrum."""
        n = spec.shape[0]
        return self._call("artia_ctf_apply", _cptr(spec, "spectrum", n), n, n // 2 + 1, C.byref(params),
                          float(defocus_min_nm), float(defocus_max_nm), float(angle_deg), int(bool(for_fp)),
                          timed=timed)

    def wbp_weight(self, spec, psi_deg, method="exact", timed=False):
        """In-place WBP weighting of an (n, n/2+1) half spectrum."""
        n = spec.shape[0]
        return self._call("artia_wbp_weight", _cptr(spec, "spectrum", n), n, n // 2 + 1, float(psi_deg),
                          _abi.WBP_METHODS[method], timed=timed)

    def ctf_slab_filter(self, img, out, corners, params: CtfParams, defocus_min_nm, defocus_max_nm, angle_deg,
                        for_fp=True, accumulate=False, cut_length=(0, 0), dim_length=(0, 0), timed=False):
        """cropBorder(img) in place, then out (+)= cropBorder(iFFT(ctf * FFT(square(img))) / n^2)."""
        H, W = img.shape
        return self._call("artia_ctf_slab_filter", _ptr(img, "img"), W, H, W, _ptr(out, "out", W * H), W,
                          int(bool(accumulate)), _abi.iarr(corners, 8), _abi.farr(cut_length, 2),
                          _abi.farr(dim_length, 2), C.byref(params), float(defocus_min_nm), float(defocus_max_nm),
                          float(angle_deg), int(bool(for_fp)), timed=timed)

    def wbp_filter(self, img, out, psi_deg, method="exact", timed=False):
        """out = iFFT(w(psi) * FFT(mirror_square(img))) / n^2, cropped to the image."""
        H, W = img.shape
        return self._call("artia_wbp_filter", _ptr(img, "img"), W, H, W, _ptr(out, "out", W * H), W,
                          float(psi_deg), _abi.WBP_METHODS[method], timed=timed)

    def prepare_projections(self, pp, images, out, tilt_deg=None, psi_deg=None, dose=None, timed=False):
        """PrepareProjection over a stack: images (n, H, W) CUDA tensor of pp's data type, out (n, H, W)
        float32.  Returns (mean, std, bad_pixels) per projection (numpy), and the milliseconds when
        timed."""
        import numpy as np
        import torch

        tdt = {0: torch.int16, 1: torch.uint16, 2: torch.int32, 3: torch.uint32, 4: torch.float32}[pp.dtype]
        if (not isinstance(images, torch.Tensor) or not images.is_cuda or images.dtype != tdt
                or not images.is_contiguous() or images.dim() != 3):
            raise ArtiaError(f"images must be a contiguous (n, H, W) {tdt} CUDA tensor")
        n, H, W = images.shape
        _ptr(out, "out", n * H * W)

        def arr(v, what):
            if v is None:
                return None
            a = np.ascontiguousarray(np.asarray(v, np.float32).reshape(-1))
            if a.size != n:
                raise ArtiaError(f"{what} needs {n} values")
            return a

        tl, ps, ds = arr(tilt_deg, "tilt_deg"), arr(psi_deg, "psi_deg"), arr(dose, "dose")
        mean, std = np.zeros(n, np.float32), np.zeros(n, np.float32)
        bad = np.zeros(n, np.int32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None
        ms = C.c_float()
        st = self.lib.artia_prepare_projections(self.h, C.byref(pp), C.c_void_p(images.data_ptr()), n, W, H, W,
                                                fp(tl), fp(ps), fp(ds), C.c_void_p(out.data_ptr()), W, fp(mean),
                                                fp(std), bad.ctypes.data_as(C.c_void_p),
                                                C.byref(ms) if timed else None)
        _check(st, self.h, "artia_prepare_projections")
        return (mean, std, bad, ms.value) if timed else (mean, std, bad)

    def fp_adjoint(self, vg: VolGeom, tg: TiltGeom, resid, vol_grad, scale=1.0, roi=None, timed=False):
        """vol_grad += scale * P^T resid (exact transpose of the exact-numerics forward projector)."""
        H, W = resid.shape
        return self._call("artia_fp_adjoint", C.byref(vg), C.byref(tg), _ptr(resid, "resid"), W, H, W,
                          _ptr(vol_grad, "vol_grad", vg.nx * vg.ny * vg.nz), self._roi(roi), float(scale),
                          timed=timed)

    def fp_shift_grad(self, vg: VolGeom, tg: TiltGeom, vol, grad_proj, out, roi=None, timed=False):
        """out (2 float64 on the device) += sum grad_proj * dP/d(shift_x, shift_y)."""
        import torch

        H, W = grad_proj.shape
        if not (isinstance(out, torch.Tensor) and out.is_cuda and out.dtype == torch.float64 and out.numel() >= 2):
            raise ArtiaError("out must be a float64 CUDA tensor of 2 elements")
        return self._call("artia_fp_shift_grad", C.byref(vg), C.byref(tg), _ptr(vol, "vol", vg.nx * vg.ny * vg.nz),
                          _ptr(grad_proj, "grad_proj"), W, H, W, self._roi(roi), out.data_ptr(), timed=timed)

    def profile(self, enable: bool = True) -> None:
        """Enables (and resets) per-kernel-class event timing of driver launches."""
        _check(self.lib.artia_profile_enable(self.h, int(enable)), self.h, "artia_profile_enable")

    def profile_read(self) -> dict:
        """{kernel class: (total device ms, launches)} accumulated since profile(True)."""
        out = {}
        for name, k in _abi.PROF_KERNELS.items():
            ms, n = C.c_double(), C.c_longlong()
            _check(self.lib.artia_profile_read(self.h, k, C.byref(ms), C.byref(n)), self.h, "artia_profile_read")
            out[name] = (ms.value, n.value)
        return out

    def fp_ctf(self, vg: VolGeom, tg: TiltGeom, vol, slab_range, defocus: TiltDefocus, ctf: CtfParams, proj, dist,
               cut_length=(0, 0), dim_length=(0, 0), volume_is_empty=False, timed=False):
        """ForwardProjectionCTF for one tilt (artia_fp_ctf): proj += the CTF-filtered slab projections
        over slab_range = (t_in, t_out); dist = the centre-ray traversal length."""
        H, W = proj.shape
        if vol is None:
            vp, sfx = None, ""
        else:
            vp, sfx = _vol(vol, vg)
        return self._call("artia_fp_ctf" + sfx, C.byref(vg), C.byref(tg), vp, int(bool(volume_is_empty)), W, H, W,
                          float(slab_range[0]), float(slab_range[1]), C.byref(defocus), C.byref(ctf),
                          _abi.farr(cut_length, 2), _abi.farr(dim_length, 2), _ptr(proj, "proj"),
                          _ptr(dist, "dist", W * H), timed=timed)

    def bp_ctf(self, vg: VolGeom, tg: TiltGeom, vol, residual, points, slab_range, defocus: TiltDefocus,
               ctf: CtfParams, lambda_, oversampling=1, cut_length=(0, 0), dim_length=(0, 0), timed=False):
        """BackProjectionCTF for one tilt (artia_bp_ctf): every slab of slab_range Wiener-filtered and
        backprojected inside its distance gate; the residual is left cropped once per slab."""
        H, W = residual.shape
        vp, sfx = _vol(vol, vg)
        return self._call("artia_bp_ctf" + sfx, C.byref(vg), C.byref(tg), vp, _ptr(residual, "residual"), W, H, W,
                          _abi.iarr(points, 8), float(slab_range[0]), float(slab_range[1]), C.byref(defocus),
                          C.byref(ctf), _abi.farr(cut_length, 2), _abi.farr(dim_length, 2), float(lambda_),
                          int(oversampling), timed=timed)

    def rot3d(self, vol, angles, out=None, timed=False):
        """The size^3 reference cube rotated per motive (artia_rot3d, RotKernel / rot3d): angles is a
        sequence of (phi, psi, theta) in degrees; returns out (n, size, size, size) float32."""
        import torch

        S = int(vol.shape[0])
        if tuple(vol.shape) != (S, S, S):
            raise ArtiaError(f"rot3d takes a cube, got shape {tuple(vol.shape)}")
        flat = [float(a) for triple in angles for a in triple]
        if len(flat) % 3:
            raise ArtiaError("rot3d angles are (phi, psi, theta) triples")
        n = len(flat) // 3
        if out is None:
            out = torch.empty((n, S, S, S), device=vol.device, dtype=torch.float32)
        ms = self._call("artia_rot3d", _ptr(vol, "vol", S ** 3), S, _abi.farr(flat) if n else None, n,
                        _ptr(out, "out", n * S ** 3), timed=timed)
        return (out, ms) if timed else out

    def sart_run(self, vg: VolGeom, acq: Acquisition, projections, vol, params: SartParams, iterations=1,
                 volume_is_empty: bool = True, timed=False, noalloc=False):
        """Runs `iterations` SART iterations; returns (ms or None, volume_is_empty).  noalloc=True takes
        the allocation-free entry point (fp32 volumes; its workspace is the library's fixed per-GPU
        array)."""
        n = len(acq.geoms)
        tg, hp, rl, il = acq.c_arrays()
        empty = C.c_int(1 if volume_is_empty else 0)
        ms = C.c_float() if timed else None
        vp, sfx = _vol(vol, vg)
        if noalloc:
            if sfx:
                raise ArtiaError("the allocation-free SART entry point takes float32 volumes")
            sfx = "_noalloc"
        st = getattr(self.lib, "artia_sart_run" + sfx)(
            self.h, C.byref(vg), tg, hp, rl, n, il, _ptr(projections, "projections", n * acq.width * acq.height),
            acq.width, acq.height, vp, C.byref(params), int(iterations), C.byref(empty),
            C.byref(ms) if timed else None)
        _check(st, self.h, "artia_sart_run" + sfx)
        return (ms.value if timed else None), bool(empty.value)

    def sart_ctf_run(self, vg: VolGeom, acq: Acquisition, defocus, projections, vol, params: SartParams,
                     ctf: CtfParams, iterations=1, volume_is_empty: bool = True, timed=False):
        """SART with 3D-CTF slab correction; `defocus` is one TiltDefocus per tilt.  Returns
        (ms or None, volume_is_empty)."""
        n = len(acq.geoms)
        if len(defocus) != n:
            raise ArtiaError(f"expected {n} TiltDefocus entries, got {len(defocus)}")
        tg, hp, rl, il = acq.c_arrays()
        sr = (C.c_float * (2 * n))(*[v for r in acq.slab_ranges for v in r])
        df = (TiltDefocus * n)(*defocus)
        empty = C.c_int(1 if volume_is_empty else 0)
        ms = C.c_float() if timed else None
        vp, sfx = _vol(vol, vg)
        st = getattr(self.lib, "artia_sart_ctf_run" + sfx)(
            self.h, C.byref(vg), tg, hp, rl, sr, df, n, il,
            _ptr(projections, "projections", n * acq.width * acq.height), acq.width, acq.height, vp, C.byref(params),
            C.byref(ctf), int(iterations), C.byref(empty), C.byref(ms) if timed else None)
        _check(st, self.h, "artia_sart_ctf_run" + sfx)
        return (ms.value if timed else None), bool(empty.value)


# ------------------------------------------------------------------ sub-volume WBP

def _subvol_arrays(subs, geoms, active, ntilts):
    """subs: [(VolGeom, tensor)], geoms: nsub lists of ntilts TiltGeom, active: nsub x ntilts or None."""
    import numpy as np

    nsub = len(subs)
    sv = (Subvol * nsub)(*[Subvol(g, C.c_void_p(_ptr(v, "sub-volume", g.nx * g.ny * g.nz))) for g, v in subs])
    flat = [g for row in geoms for g in row]
    if len(flat) != nsub * ntilts:
        raise ArtiaError(f"expected {nsub} x {ntilts} tilt geometries, got {len(flat)}")
    tg = (TiltGeom * len(flat))(*flat)
    act = None
    if active is not None:
        a = np.ascontiguousarray(np.asarray(active, dtype=np.uint8).reshape(nsub, ntilts))
        act = (C.c_ubyte * a.size).from_buffer_copy(a.tobytes())
    return sv, tg, act


def _ctf_arrays(ctf, slab_ranges, defocus, ntilts):
    if ctf is None:
        return None, None, None
    if slab_ranges is None or defocus is None or len(slab_ranges) != ntilts or len(defocus) != ntilts:
        raise ArtiaError("CTF sub-volume WBP needs one slab range and one TiltDefocus per tilt")
    sr = (C.c_float * (2 * ntilts))(*[v for r in slab_ranges for v in r])
    df = (TiltDefocus * ntilts)(*defocus)
    return C.byref(ctf), sr, df


def subvol_wbp_plain(ctx: "ArtiaContext", subs, geoms, images, oversampling=1, active=None, ctf=None,
                     slab_ranges=None, defocus=None, timed=False):
    """The reference's sub-volume schedule: one BP launch per active (sub-volume, tilt) [and slab]."""
    T, H, W = images.shape
    sv, tg, act = _subvol_arrays(subs, geoms, active, T)
    cp, sr, df = _ctf_arrays(ctf, slab_ranges, defocus, T)
    ms = C.c_float() if timed else None
    st = ctx.lib.artia_subvol_wbp_plain(ctx.h, sv, len(subs), tg, act, T, _ptr(images, "images", T * H * W), W, H, W,
                                        int(oversampling), cp, sr, df, C.byref(ms) if timed else None)
    _check(st, ctx.h, "artia_subvol_wbp_plain")
    return ms.value if timed else None


class SubvolPlan:
    """Batched sub-volume WBP (one persistent launch, graph-replayed); see artia_subvol_plan_create."""

    def __init__(self, ctx: "ArtiaContext", subs, geoms, images, oversampling=1, active=None, ctf=None,
                 slab_ranges=None, defocus=None, tuning: Optional[SubvolTuning] = None):
        T, H, W = images.shape
        self.ctx = ctx
        self._keep = (subs, images)   # the plan holds raw pointers into these tensors
        sv, tg, act = _subvol_arrays(subs, geoms, active, T)
        cp, sr, df = _ctf_arrays(ctf, slab_ranges, defocus, T)
        h = C.c_void_p()
        st = ctx.lib.artia_subvol_plan_create(ctx.h, sv, len(subs), tg, act, T, _ptr(images, "images", T * H * W), W,
                                              H, W, int(oversampling), cp, sr, df,
                                              C.byref(tuning) if tuning is not None else None, C.byref(h))
        _check(st, ctx.h, "artia_subvol_plan_create")
        self.h = h

    def execute(self, timed=False):
        ms = C.c_float() if timed else None
        _check(self.ctx.lib.artia_subvol_plan_execute(self.h, C.byref(ms) if timed else None), self.ctx.h,
               "artia_subvol_plan_execute")
        return ms.value if timed else None

    @property
    def tuning(self) -> SubvolTuning:
        t = SubvolTuning()
        _check(self.ctx.lib.artia_subvol_plan_tuning(self.h, C.byref(t)), self.ctx.h, "artia_subvol_plan_tuning")
        return t

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.artia_subvol_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _tptr(t, name: str, dtype, numel: int) -> int:
    """Device pointer of a contiguous CUDA tensor of `dtype` holding at least `numel` elements."""
    import torch

    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ArtiaError(f"{name} must be a contiguous {dtype} CUDA tensor")
    if t.numel() < numel:
        raise ArtiaError(f"{name} holds {t.numel()} elements, needs {numel}")
    return t.data_ptr()


class PairSet:
    """All (sub-volume, tilt) pairs of a refinement step, each operation one launch over all of
    them (artia_pairs_*).  Sub-volumes share one flat float32 buffer; window images are packed
    in pair order into flat buffers of `n_pixels` floats; shifts are (npairs, 2) float64.

    volumes: sequence of (VolGeom, float offset into the volume buffer); pairs: sequence of
    (volume index, x0, y0, w, h) detector windows; geoms: one TiltGeom per pair."""

    def __init__(self, ctx: "ArtiaContext", volumes, pairs, geoms):
        import torch  # noqa: F401  (device buffers are torch tensors)

        self.ctx = ctx
        nv, n = len(volumes), len(pairs)
        if len(geoms) != n:
            raise ArtiaError(f"{n} pairs need {n} geometries, got {len(geoms)}")
        vol_arr = (_abi.PairVolume * nv)()
        self.vol_floats = 0
        self.volumes = []
        for k, (vg, off) in enumerate(volumes):
            vol_arr[k].geom = vg
            vol_arr[k].offset = int(off)
            self.volumes.append((vg, int(off)))
            self.vol_floats = max(self.vol_floats, int(off) + vg.nx * vg.ny * vg.nz)
        pair_arr = (_abi.Pair * n)()
        self.windows = []
        off = 0
        for p, (vi, x0, y0, w, h) in enumerate(pairs):
            q = pair_arr[p]
            q.volume, q.x0, q.y0, q.w, q.h, q.offset = int(vi), int(x0), int(y0), int(w), int(h), off
            self.windows.append((int(vi), int(x0), int(y0), int(w), int(h), off))
            off += int(w) * int(h)
        self.n_pixels = off
        self.npairs = n
        geo_arr = (TiltGeom * n)(*geoms)
        h = C.c_void_p()
        _check(ctx.lib.artia_pairs_create(ctx.h, vol_arr, nv, pair_arr, geo_arr, n, C.byref(h)), ctx.h,
               "artia_pairs_create")
        self.h = h

    def _shifts(self, shifts):
        import torch

        if shifts is None:
            return None
        return _tptr(shifts, "shifts", torch.float64, 2 * self.npairs)

    def _run(self, name, *args, timed=False):
        ms = C.c_float() if timed else None
        _check(getattr(self.ctx.lib, name)(self.h, *args, C.byref(ms) if timed else None), self.ctx.h, name)
        return ms.value if timed else None

    def window(self, buf, p: int):
        """(h, w) view of pair p's window in a packed image buffer."""
        _, _, _, w, h, off = self.windows[p]
        return buf[off:off + w * h].view(h, w)

    def project(self, volumes, shifts=None, out=None, timed=False):
        """Packed window images P_s volumes (a new buffer unless `out` is given)."""
        import torch

        if out is None:
            out = torch.empty(self.n_pixels, device=volumes.device, dtype=torch.float32)
        ms = self._run("artia_pairs_project", _ptr(volumes, "volumes", self.vol_floats), self._shifts(shifts),
                       _ptr(out, "images", self.n_pixels), timed=timed)
        return (out, ms) if timed else out

    def loss_grad(self, volumes, shifts, targets, resid=None, normal=False, timed=False):
        """Per pair sum (P_s - targets)^2 and its shift gradient, (npairs,) and (npairs, 2) float64;
        with normal=True also J^T J as (npairs, 3) float64 (xx, xy, yy), J = dP_s/ds."""
        import torch

        loss = torch.empty(self.npairs, device=volumes.device, dtype=torch.float64)
        grad = torch.empty((self.npairs, 2), device=volumes.device, dtype=torch.float64)
        jtj = torch.empty((self.npairs, 3), device=volumes.device, dtype=torch.float64) if normal else None
        ms = self._run("artia_pairs_loss_grad", _ptr(volumes, "volumes", self.vol_floats), self._shifts(shifts),
                       _ptr(targets, "targets", self.n_pixels), loss.data_ptr(), grad.data_ptr(),
                       None if jtj is None else jtj.data_ptr(),
                       None if resid is None else _ptr(resid, "resid", self.n_pixels), timed=timed)
        res = (loss, grad, jtj) if normal else (loss, grad)
        return res + (ms,) if timed else res

    def shift_grad(self, volumes, shifts, weights, timed=False):
        """Per pair sum weights * dP_s/ds, (npairs, 2) float64."""
        import torch

        grad = torch.empty((self.npairs, 2), device=volumes.device, dtype=torch.float64)
        ms = self._run("artia_pairs_shift_grad", _ptr(volumes, "volumes", self.vol_floats), self._shifts(shifts),
                       _ptr(weights, "weights", self.n_pixels), grad.data_ptr(), timed=timed)
        return (grad, ms) if timed else grad

    def adjoint(self, shifts, images, vol_grads, scale=1.0, timed=False):
        """vol_grads += scale * P_s^T images."""
        return self._run("artia_pairs_adjoint", self._shifts(shifts), _ptr(images, "images", self.n_pixels),
                         float(scale), _ptr(vol_grads, "vol_grads", self.vol_floats), timed=timed)

    def cc_search(self, models, targets, square: int, max_shift: int, want_map=False, timed=False,
                  multi_peak=False):
        """GetDisplacement's search per pair: (peak (npairs, 2) int32, cc (npairs,) float32, map or None).
        multi_peak=True takes MultiPeakDetection's peak (the qualifying local maximum nearest to zero)."""
        import torch

        dev = models.device
        peak = torch.empty((self.npairs, 2), device=dev, dtype=torch.int32)
        cc = torch.empty(self.npairs, device=dev, dtype=torch.float32)
        side = 2 * int(max_shift) + 1
        ccmap = torch.empty((self.npairs, side, side), device=dev, dtype=torch.float32) if want_map else None
        ms = self._run("artia_pairs_cc_search", _ptr(models, "models", self.n_pixels),
                       _ptr(targets, "targets", self.n_pixels), int(square), int(max_shift), int(bool(multi_peak)),
                       peak.data_ptr(),
                       cc.data_ptr(), None if ccmap is None else ccmap.data_ptr(), timed=timed)
        return (peak, cc, ccmap, ms) if timed else (peak, cc, ccmap)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.artia_pairs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
