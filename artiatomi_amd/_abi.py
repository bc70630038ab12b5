"""ctypes binding of include/artia.h (libartia.so, built for gfx950).

The library is loaded from this package directory.  A missing library is an error: there is
no CPU or PyTorch fallback for any operator.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libartia.so"

ARTIA_OK = 0
ARTIA_NUMERICS_EXACT = 0
ARTIA_NUMERICS_REFERENCE = 1
ARTIA_FILTER_POINT = 0
ARTIA_FILTER_LINEAR = 1
PROF_KERNELS = {"fp": 0, "bp": 1, "compare": 2, "crop": 3, "fill": 4, "fft": 5}
WBP_METHODS = {"ramp": 0, "exact": 1, "contrast2": 2, "contrast10": 3, "contrast30": 4}

_F3 = C.c_float * 3


class TiltGeom(C.Structure):
    """artia_tilt_geom: c_detektor, c_uPitch, c_vPitch, c_projNorm, c_DetectorMatrix,
    c_magAniso, c_magAnisoInv (DeviceVariables.cuh:13-34)."""

    _fields_ = [
        ("d", _F3), ("u", _F3), ("v", _F3), ("n", _F3),
        ("det_matrix", C.c_float * 16),
        ("mag_aniso", C.c_float * 9), ("mag_aniso_inv", C.c_float * 9),
    ]


class VolGeom(C.Structure):
    """artia_vol_geom: c_bBoxMin/Max, c_volumeBBoxRcp, c_volumeDim, c_voxelSize, zShift."""

    _fields_ = [
        ("bbox_min", _F3), ("bbox_max", _F3), ("bbox_rcp", _F3), ("dim", _F3),
        ("voxel_size", _F3), ("z_shift", C.c_float),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
    ]


class TiltParams(C.Structure):
    """artia_tilt_params: MarkerFile items of one projection."""

    _fields_ = [("tilt_deg", C.c_float), ("psi_deg", C.c_float), ("mag", C.c_float),
                ("shift_x", C.c_float), ("shift_y", C.c_float)]


class GlobalParams(C.Structure):
    """artia_global_params: AddTiltAngle, AddTiltXAngle, PhiAngle, UseFixPsiAngle, PsiAngle."""

    _fields_ = [("add_tilt_deg", C.c_float), ("add_tilt_x_deg", C.c_float), ("phi_deg", C.c_float),
                ("use_fix_psi", C.c_int), ("fix_psi_deg", C.c_float)]


class SartParams(C.Structure):
    """artia_sart_params: SIRTCount, Lambda, OverSampling, Crop, CropDim, CutLength,
    DimLength, ProjectionScaleFactor."""

    _fields_ = [("sirt_count", C.c_int), ("lambda_", C.c_float), ("oversampling", C.c_int),
                ("crop", C.c_float * 4), ("crop_dim", C.c_float * 4),
                ("cut_length", C.c_float * 2), ("dim_length", C.c_float * 2),
                ("proj_scale", C.c_float)]


class CtfParams(C.Structure):
    """artia_ctf_params: Cs, Voltage, pixel size, amplitude contrast, WienerFilterNoiseLevel,
    CTFBetaFac, PhaseFlipOnly, CTFSliceThickness, VolumeShift.z, IgnoreZShiftForCTF,
    SwitchCTFDirectionForIMOD."""

    _fields_ = [("cs_mm", C.c_float), ("voltage_kV", C.c_float), ("pixelsize_nm", C.c_float),
                ("amp_contrast", C.c_float), ("wiener", C.c_float), ("beta_fac", C.c_float * 4),
                ("phase_flip_only", C.c_int), ("slice_thickness_nm", C.c_float),
                ("volume_shift_z", C.c_float), ("ignore_z_shift", C.c_int), ("switch_direction", C.c_int)]


DTYPES = {"int16": 0, "uint16": 1, "int32": 2, "uint32": 3, "float32": 4}
NORMALIZATIONS = {"mean": 0, "std": 1, "none": 2}


class PrepParams(C.Structure):
    """artia_prep_params: data type, FRM_DM4, CorrectBadPixels, BadPixelValue, SkipFilter,
    fourFilterLP/LPS/HP/HPS, DoseWeighting, pixel size, WBP, WBPFilter, DownWeightTiltsForWBP,
    ProjectionNormalization."""

    _fields_ = [("dtype", C.c_int), ("mirror_y", C.c_int), ("correct_bad_pixels", C.c_int),
                ("bad_pixel_value", C.c_float), ("skip_filter", C.c_int), ("lp", C.c_float), ("lps", C.c_float),
                ("hp", C.c_float), ("hps", C.c_float), ("dose_weighting", C.c_int), ("pixelsize_nm", C.c_float),
                ("wbp", C.c_int), ("wbp_method", C.c_int), ("down_weight_tilts", C.c_int),
                ("normalization", C.c_int)]


class TiltDefocus(C.Structure):
    """artia_tilt_defocus: CtfFile min/max defocus (nm), astigmatism angle (deg), tilt angle."""

    _fields_ = [("defocus_min_nm", C.c_float), ("defocus_max_nm", C.c_float),
                ("astig_angle_deg", C.c_float), ("tilt_deg", C.c_float)]


class FileInfo(C.Structure):
    """artia_file_info: element type, sizes and data offset of an EM or MRC file."""

    _fields_ = [("dtype", C.c_int), ("raw_type", C.c_int), ("elem_bytes", C.c_int), ("nx", C.c_int),
                ("ny", C.c_int), ("nz", C.c_int), ("data_offset", C.c_size_t), ("pixel_size_nm", C.c_float),
                ("byte_swapped", C.c_int)]


ARTIA_FILE_EM = 0
ARTIA_FILE_MRC = 1


class Motive(C.Structure):
    """artia_motive: one motive-list row (struct motive, MotiveListe.h:81-106), binnings applied."""

    _fields_ = [(n, C.c_float) for n in (
        "cc_coeff", "x_coord_2d", "y_coord_2d", "part_nr", "tomo_nr", "part_nr_in_tomo", "wedge_idx", "x", "y", "z",
        "x_shift", "y_shift", "z_shift", "x_shift_before", "y_shift_before", "z_shift_before", "phi", "psi", "theta",
        "class_no")]


ARTIA_GROUP_BY_CLASS = 0
ARTIA_GROUP_MAX_DIST = 1
ARTIA_GROUP_MAX_COUNT = 2


class Subvol(C.Structure):
    """artia_subvol: sub-volume placement and its device volume."""

    _fields_ = [("geom", VolGeom), ("vol", C.c_void_p)]


class SubvolTuning(C.Structure):
    """artia_subvol_tuning of the batched sub-volume WBP."""

    _fields_ = [("autotune", C.c_int), ("tile_policy", C.c_int), ("quads_per_thread", C.c_int),
                ("blocks_per_cu", C.c_int)]


class PairVolume(C.Structure):
    """artia_pair_volume: a sub-volume's placement and its float offset in the volume buffer."""

    _fields_ = [("geom", VolGeom), ("offset", C.c_longlong)]


class Pair(C.Structure):
    """artia_pair: sub-volume index, detector window and the window image's float offset."""

    _fields_ = [("volume", C.c_int), ("x0", C.c_int), ("y0", C.c_int), ("w", C.c_int), ("h", C.c_int),
                ("offset", C.c_longlong)]


_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_SZ = C.c_size_t
_PF = C.POINTER(C.c_float)
_PI = C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "artia_create": (_I, [_I, C.POINTER(_P)]),
    "artia_destroy": (_I, [_P]),
    "artia_set_stream": (_I, [_P, _P]),
    "artia_get_stream": (_I, [_P, C.POINTER(_P)]),
    "artia_set_numerics": (_I, [_P, _I]),
    "artia_set_fp_staging": (_I, [_P, _I]),
    "artia_set_stack_limit": (_I, [_P, _SZ]),
    "artia_sync": (_I, [_P]),
    "artia_last_error": (C.c_char_p, [_P]),
    "artia_version": (C.c_char_p, []),
    "artia_geom_tilt": (_I, [C.POINTER(TiltParams), C.POINTER(GlobalParams), _I, _I, _F, _F,
                             C.POINTER(TiltGeom)]),
    "artia_geom_volume": (_I, [_I, _I, _I, _PF, _PF, C.POINTER(VolGeom)]),
    "artia_geom_subvolume": (_I, [_I, _F, _PF, C.POINTER(VolGeom), _PF, _PF, C.POINTER(VolGeom)]),
    "artia_geom_hitpoints": (_I, [C.POINTER(VolGeom), C.POINTER(TiltGeom), _PI]),
    "artia_geom_real_length": (_I, [C.POINTER(VolGeom), C.POINTER(TiltGeom), _PF]),
    "artia_projection_index_list": (_I, [_PF, _I, _PI]),
    "artia_marker_tilt_params": (_I, [_PF, _I, _I, _I, _I, _I, C.POINTER(TiltParams), _PI]),
    "artia_ctf_file_defocus": (_I, [_PF, _I, _I, _I, _F, C.POINTER(TiltDefocus)]),
    "artia_geom_defocus_distances": (_I, [C.POINTER(VolGeom), C.POINTER(TiltGeom), _PF, _PF]),
    "artia_geom_defocus_minmax": (_I, [_F, C.POINTER(TiltDefocus), C.POINTER(CtfParams), _PF, _PF]),
    "artia_fp_march": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _I, _I, _SZ, _P, _P, _PI, _PF]),
    "artia_fp_slicer": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _I, _I, _SZ, _P, _F, _F, _PI, _PF]),
    "artia_vol_traversal": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _I, _I, _SZ, _P, _PI, _PF]),
    "artia_compare": (_I, [_P, _I, _I, _SZ, _P, _P, _P, _F, _PF, _PF, _F, _PF]),
    "artia_fp_residual": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _I, _I, _I, _SZ, _P, _F, _PF, _PF, _F,
                               _PF, _PF, _PI, _P, _P, _PF]),
    "artia_fp_residual_f16": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _I, _I, _I, _SZ, _P, _F, _PF, _PF,
                                   _F, _PF, _PF, _PI, _P, _P, _PF]),
    "artia_crop_border": (_I, [_P, _I, _I, _SZ, _P, _PF, _PF, _PI, _PF]),
    "artia_bp": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _P, _I, _I, _SZ, _F, _I, _F, _F, _I, _PF]),
    "artia_make_square": (_I, [_P, _I, _I, _SZ, _P, _I, _P, _I, _I, _I, _I, _PF]),
    "artia_fp_march_f16": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _I, _I, _SZ, _P, _P, _PI, _PF]),
    "artia_fp_slicer_f16": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _I, _I, _SZ, _P, _F, _F, _PI,
                                 _PF]),
    "artia_bp_f16": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _P, _I, _I, _SZ, _F, _I, _F, _F, _I,
                          _PF]),
    "artia_volume_f16_to_f32": (_I, [_P, _P, _I, _I, _I, _I, _I, _P, _SZ, _PF]),
    "artia_file_parse_header": (_I, [_I, _P, _SZ, C.POINTER(FileInfo)]),
    "artia_file_volume_header": (_I, [_I, _I, _I, _I, _I, _F, _PF, _P, _SZ, C.POINTER(_SZ)]),
    "artia_shift_file_get": (_I, [_PF, _I, _I, _I, _I, _I, _PF]),
    "artia_shift_file_set": (_I, [_PF, _I, _I, _I, _I, _I, _PF]),
    "artia_motive_get": (_I, [_PF, _I, _I, _I, _F, _F, C.POINTER(Motive)]),
    "artia_motive_group_count": (_I, [_PF, _I, _I, _I, _PI]),
    "artia_motive_group": (_I, [_PF, _I, _I, _I, _I, _F, _I, _F, _F, _PI, _I, _PI]),
    "artia_ctf_apply": (_I, [_P, _P, _I, _SZ, C.POINTER(CtfParams), _F, _F, _F, _I, _PF]),
    "artia_wbp_weight": (_I, [_P, _P, _I, _SZ, _F, _I, _PF]),
    "artia_ctf_slab_filter": (_I, [_P, _P, _I, _I, _SZ, _P, _SZ, _I, _PI, _PF, _PF, C.POINTER(CtfParams), _F, _F,
                                   _F, _I, _PF]),
    "artia_wbp_filter": (_I, [_P, _P, _I, _I, _SZ, _P, _SZ, _F, _I, _PF]),
    "artia_prepare_projections": (_I, [_P, C.POINTER(PrepParams), _P, _I, _I, _I, _SZ, _P, _P, _P, _P, _SZ, _P, _P,
                                       _P, _PF]),
    "artia_subvol_wbp_plain": (_I, [_P, C.POINTER(Subvol), _I, C.POINTER(TiltGeom), _P, _I, _P, _I, _I, _SZ, _I,
                                    C.POINTER(CtfParams), _PF, C.POINTER(TiltDefocus), _PF]),
    "artia_subvol_plan_create": (_I, [_P, C.POINTER(Subvol), _I, C.POINTER(TiltGeom), _P, _I, _P, _I, _I, _SZ, _I,
                                      C.POINTER(CtfParams), _PF, C.POINTER(TiltDefocus), C.POINTER(SubvolTuning),
                                      C.POINTER(_P)]),
    "artia_subvol_plan_execute": (_I, [_P, _PF]),
    "artia_subvol_plan_tuning": (_I, [_P, C.POINTER(SubvolTuning)]),
    "artia_subvol_plan_destroy": (_I, [_P]),
    "artia_fp_adjoint": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _I, _I, _SZ, _P, _PI, _F, _PF]),
    "artia_fp_shift_grad": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _P, _I, _I, _SZ, _PI, _P, _PF]),
    "artia_pairs_create": (_I, [_P, C.POINTER(PairVolume), _I, C.POINTER(Pair), C.POINTER(TiltGeom), _I,
                                C.POINTER(_P)]),
    "artia_pairs_destroy": (_I, [_P]),
    "artia_pairs_project": (_I, [_P, _P, _P, _P, _PF]),
    "artia_pairs_loss_grad": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _PF]),
    "artia_pairs_shift_grad": (_I, [_P, _P, _P, _P, _P, _PF]),
    "artia_pairs_adjoint": (_I, [_P, _P, _P, _F, _P, _PF]),
    "artia_pairs_cc_search": (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P, _PF]),
    "artia_rot_matrix": (None, [_F, _F, _F, _PF]),
    "artia_rot3d": (_I, [_P, _P, _I, _PF, _I, _P, _PF]),
    "artia_profile_enable": (_I, [_P, _I]),
    "artia_profile_read": (_I, [_P, _I, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "artia_fp_ctf": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _I, _I, _I, _SZ, _F, _F,
                          C.POINTER(TiltDefocus), C.POINTER(CtfParams), _PF, _PF, _P, _P, _PF]),
    "artia_fp_ctf_f16": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _I, _I, _I, _SZ, _F, _F,
                              C.POINTER(TiltDefocus), C.POINTER(CtfParams), _PF, _PF, _P, _P, _PF]),
    "artia_bp_ctf": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _P, _I, _I, _SZ, _PI, _F, _F,
                          C.POINTER(TiltDefocus), C.POINTER(CtfParams), _PF, _PF, _F, _I, _PF]),
    "artia_bp_ctf_f16": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _P, _P, _I, _I, _SZ, _PI, _F, _F,
                              C.POINTER(TiltDefocus), C.POINTER(CtfParams), _PF, _PF, _F, _I, _PF]),
    "artia_sart_run": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _PI, _PF, _I, _PI, _P, _I, _I, _P,
                            C.POINTER(SartParams), _I, _PI, _PF]),
    "artia_sart_ctf_run": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _PI, _PF, _PF, C.POINTER(TiltDefocus),
                                _I, _PI, _P, _I, _I, _P, C.POINTER(SartParams), C.POINTER(CtfParams), _I, _PI,
                                _PF]),
    "artia_sart_run_noalloc": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _PI, _PF, _I, _PI, _P, _I, _I, _P,
                                    C.POINTER(SartParams), _I, _PI, _PF]),
    "artia_sart_run_f16": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _PI, _PF, _I, _PI, _P, _I, _I, _P,
                                C.POINTER(SartParams), _I, _PI, _PF]),
    "artia_sart_ctf_run_f16": (_I, [_P, C.POINTER(VolGeom), C.POINTER(TiltGeom), _PI, _PF, _PF,
                                    C.POINTER(TiltDefocus), _I, _PI, _P, _I, _I, _P, C.POINTER(SartParams),
                                    C.POINTER(CtfParams), _I, _PI, _PF]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libartia.so and declares every entry point; raises if the library is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m artiatomi_amd.build` "
            "(there is no fallback implementation)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def farr(values, n=None):
    vals = list(values)
    if n is not None and len(vals) != n:
        raise ValueError(f"expected {n} values, got {len(vals)}")
    return (C.c_float * len(vals))(*vals)


def iarr(values, n=None):
    vals = list(values)
    if n is not None and len(vals) != n:
        raise ValueError(f"expected {n} values, got {len(vals)}")
    return (C.c_int * len(vals))(*vals)
