"""artiatomi_amd - EmSART's reconstruction core (forward projection, residual, SART
backprojection, 3D-CTF slab correction, WBP weighting) as gfx950 HIP kernels behind the C ABI
of include/artia.h.  This package is the Python view of that library; the compute path is the
HIP library (libartia.so) and nothing else."""
from ._abi import (CtfParams, FileInfo, GlobalParams, SartParams, Subvol, SubvolTuning, TiltDefocus,  # noqa: F401
                   TiltGeom, TiltParams, VolGeom, load)
from .core import (  # noqa: F401
    Acquisition,
    ArtiaContext,
    ArtiaError,
    PairSet,
    SubvolPlan,
    ctf_params,
    prep_params,
    marker_tilt_params,
    ctf_file_defocus,
    MotiveList,
    shift_file_get,
    shift_file_set,
    subvolume_job,
    parse_file_header,
    read_file,
    volume_file_header,
    defocus_distances,
    defocus_minmax,
    footprint_window,
    hitpoints,
    projection_index_list,
    real_length,
    rot_matrix,
    sart_params,
    subvol_wbp_plain,
    subvolume_geometry,
    subvolume_tilt_geometry,
    tilt_geometry,
    volume_geometry,
)

__all__ = [
    "Acquisition", "ArtiaContext", "ArtiaError", "CtfParams", "PairSet", "Subvol", "SubvolPlan", "SubvolTuning",
    "subvol_wbp_plain", "subvolume_tilt_geometry", "GlobalParams", "TiltDefocus",
    "ctf_params", "defocus_distances", "defocus_minmax", "footprint_window", "SartParams", "TiltGeom", "TiltParams",
    "VolGeom", "hitpoints", "load", "projection_index_list", "real_length", "sart_params",
    "subvolume_geometry", "tilt_geometry", "volume_geometry", "FileInfo", "parse_file_header", "read_file",
    "volume_file_header", "marker_tilt_params", "ctf_file_defocus", "prep_params", "MotiveList", "shift_file_get",
    "shift_file_set", "subvolume_job", "rot_matrix",
]
