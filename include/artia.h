/*
 * artia.h - C ABI of the MI355X (gfx950) EmSART reconstruction core.
 *
 * The entry points replace the reference's operator boundary, the C++ functor classes of
 * src/EmSART/Kernels.h:37-245 plus the cuFFT/NPP glue between them inside
 * src/EmSART/Reconstructor.cpp, with plain-pointer C functions:
 *   - no exceptions cross the ABI: every call returns an artia_status, and
 *     artia_last_error() holds the message (the C++ adapter in artia_kernels.hpp rethrows,
 *     keeping the reference's exception convention of CudaException.h:34-36);
 *   - geometry is passed per call in artia_tilt_geom / artia_vol_geom instead of the
 *     module __constant__ globals of DeviceVariables.cuh:13-34 that SetConstantValues
 *     (Kernels.cpp:949-1132) uploads one cuMemcpyHtoD at a time;
 *   - images and volumes are caller-owned device pointers: fp32, row-major, x fastest.
 *     Projections use a pitch counted in elements; volumes are dense [z][y][x];
 *   - every launch is stream-ordered on the context's stream.  When `ms` is non-NULL the
 *     call synchronises and returns the elapsed milliseconds, which is the reference
 *     functors' "float operator()" contract (Kernels.cpp:65-86).  When `ms` is NULL the
 *     call does not block.
 *
 * Numerics (artia_set_numerics):
 *   ARTIA_NUMERICS_REFERENCE  kernels evaluate the reference kernel bodies in the reference's
 *                             float operation order (IEEE division and square root, no FMA
 *                             contraction, exact-float texture filtering).
 *   ARTIA_NUMERICS_EXACT      (default) kernels use numerically stable formulations of the same
 *                             algorithm, built for throughput, which track the double-precision
 *                             evaluation of the reference algorithm.
 * A context is used by one host thread at a time; several contexts may coexist.
 */
#ifndef ARTIA_H
#define ARTIA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct artia_ctx artia_ctx;

typedef enum {
    ARTIA_OK = 0,
    ARTIA_ERR_INVALID_ARGUMENT = -1,
    ARTIA_ERR_DEVICE = -2,
    ARTIA_ERR_OUT_OF_MEMORY = -3,
    ARTIA_ERR_UNSUPPORTED = -4,
    ARTIA_ERR_FFT = -5
} artia_status;

typedef enum { ARTIA_NUMERICS_EXACT = 0, ARTIA_NUMERICS_REFERENCE = 1 } artia_numerics;

/* BP source-image sampling: POINT without CTF, LINEAR with CTF (Reconstructor.cpp:318-327) */
typedef enum { ARTIA_FILTER_POINT = 0, ARTIA_FILTER_LINEAR = 1 } artia_filter;

/* Per-tilt detector geometry: the values of c_detektor, c_uPitch, c_vPitch, c_projNorm,
 * c_DetectorMatrix, c_magAniso, c_magAnisoInv (DeviceVariables.cuh:13-34). */
typedef struct {
    float d[3];
    float u[3];
    float v[3];
    float n[3];
    float det_matrix[16]; /* row-major 4x4 inverse of [u; v; n; d] */
    float mag_aniso[9];   /* row-major 3x3, forward projection */
    float mag_aniso_inv[9];
} artia_tilt_geom;

/* Volume placement: c_bBoxMin, c_bBoxMax, c_volumeBBoxRcp, c_volumeDim, c_voxelSize,
 * c_zShiftForPartialVolume, plus the extent of the stored array. */
typedef struct {
    float bbox_min[3];
    float bbox_max[3];
    float bbox_rcp[3];
    float dim[3];
    float voxel_size[3];
    float z_shift;
    int nx, ny, nz;
} artia_vol_geom;

/* Per-tilt acquisition values (MarkerFile items, MarkerFile.h:34-46) */
typedef struct {
    float tilt_deg;
    float psi_deg;
    float mag;
    float shift_x;
    float shift_y;
} artia_tilt_params;

/* Global config keys entering the rotation (Config.cpp:303-406) */
typedef struct {
    float add_tilt_deg;
    float add_tilt_x_deg;
    float phi_deg;
    int use_fix_psi;
    float fix_psi_deg;
} artia_global_params;

/* SART schedule parameters (EmSart.cpp:560-675, Config.cpp:303-406) */
typedef struct {
    int sirt_count;       /* ordered-subset size S */
    float lambda;         /* relaxation; each BP uses lambda / S */
    int oversampling;     /* BP sub-pixel sampling factor */
    float crop[4];        /* Crop (compare border) */
    float crop_dim[4];    /* CropDim */
    float cut_length[2];  /* CutLength (cropBorder) */
    float dim_length[2];  /* DimLength */
    float proj_scale;     /* ProjectionScaleFactor */
} artia_sart_params;

/* 3D-CTF settings: the CTF kernel constants (SetConstantValues(CTFKernel...),
 * Kernels.cpp:1104-1132), the functor arguments (Kernels.cpp:395-401) and the slab keys of
 * Reconstructor::GetDefocusMinMax (Reconstructor.cpp:178-201). */
typedef struct {
    float cs_mm;              /* Cs */
    float voltage_kV;         /* Voltage */
    float pixelsize_nm;       /* projection pixel size */
    float amp_contrast;       /* amplitude contrast; the reference hard-codes 0 (Kernels.cpp:1111) */
    float wiener;             /* WienerFilterNoiseLevel */
    float beta_fac[4];        /* CTFBetaFac: low-frequency radius, envelope coefficients */
    int phase_flip_only;      /* PhaseFlipOnly */
    float slice_thickness_nm; /* CTFSliceThickness */
    float volume_shift_z;     /* VolumeShift.z (voxels) */
    int ignore_z_shift;       /* IgnoreZShiftForCTF */
    int switch_direction;     /* SwitchCTFDirectionForIMOD */
} artia_ctf_params;

/* Per-tilt defocus (CtfFile rows: min, max, astigmatism angle; CtfFile.cpp:37-56) and the
 * tilt angle entering the z-shift correction (marker tilt + AddTiltAngle). */
typedef struct {
    float defocus_min_nm;
    float defocus_max_nm;
    float astig_angle_deg;
    float tilt_deg;
} artia_tilt_defocus;

/* ---------------------------------------------------------------- context */
artia_status artia_create(int device, artia_ctx** out);
artia_status artia_destroy(artia_ctx* ctx);
artia_status artia_set_stream(artia_ctx* ctx, void* hip_stream);
artia_status artia_get_stream(artia_ctx* ctx, void** hip_stream);
artia_status artia_set_numerics(artia_ctx* ctx, int numerics);
/* Exact forward projection (march and slicer), the reference-numerics march, and the exact
 * adjoint: enable != 0 (the default) walks each tile's rays through LDS-staged slabs of the volume
 * where the geometry allows it, 0 always reads (or, for the adjoint, atomically adds to) the volume
 * directly.  Projections are bit-identical either way; adjoint sums differ only in the order of
 * their float atomics.  Pair
 * sets take the setting in force when they are created.  No reference counterpart (a tuning
 * choice). */
artia_status artia_set_fp_staging(artia_ctx* ctx, int enable);
/* Bytes the CTF driver (artia_sart_ctf_run) may use for its per-tilt slab stacks: the filtered slab
 * images of a tilt and the forward slabs' sub-ray sums.  0 (the default) allows 32 GiB of images and
 * 16 GiB of sums.  Within a smaller limit the forward slabs go in several volume passes, or slab by
 * slab, and without room for the images every slab is a separate launch, as in the reference's loop;
 * the results are the same.  A CTF sub-volume plan created afterwards takes the same limit for the
 * filtered slab images of its chunk of tilts (0: 16 GiB). */
artia_status artia_set_stack_limit(artia_ctx* ctx, size_t bytes);
artia_status artia_sync(artia_ctx* ctx);
const char* artia_last_error(artia_ctx* ctx);
const char* artia_version(void);

/* ---------------------------------------------------------------- host geometry
 * Projection::RotateMatrix, GetPosition, GetPixelUPitch, GetPixelVPitch, GetNormalVector,
 * GetDetectorMatrix (Projection.cpp:80-404) and Reconstructor::GetMagAnistropyMatrix
 * (Reconstructor.cpp:386-426). */
artia_status artia_geom_tilt(const artia_tilt_params* tp, const artia_global_params* gp, int width,
                             int height, float mag_aniso_amount, float mag_aniso_angle_deg,
                             artia_tilt_geom* out);
/* Volume::PositionInSpace for the whole volume (Volume.cpp:270-286) */
artia_status artia_geom_volume(int nx, int ny, int nz, const float voxel_size[3],
                               const float volume_shift[3], artia_vol_geom* out);
/* Volume::PositionInSpace for a cubic motive-list sub-volume (Volume.cpp:288-306) */
artia_status artia_geom_subvolume(int n, float voxel_size_sub, const float voxel_size_full[3],
                                  const artia_vol_geom* full, const float position[3],
                                  const float shift[3], artia_vol_geom* out);
/* Projection::ComputeHitPoints (Projection.cpp:850-1049): int corners p1..p4 as 8 ints */
artia_status artia_geom_hitpoints(const artia_vol_geom* full, const artia_tilt_geom* g, int out[8]);
/* Reconstructor::Compare's volume traversal length (Reconstructor.cpp:1390-1393) */
artia_status artia_geom_real_length(const artia_vol_geom* full, const artia_tilt_geom* g, float* out);
/* Projection::CreateProjectionIndexList(PLT_RANDOM_START_ZERO_TILT) (Projection.cpp:561-596) */
artia_status artia_projection_index_list(const float* tilt_deg, int count, int* out);
/* MarkerFile::operator() (MarkerFile.cpp:1216-1221): the values of projection proj at marker
 * ref_marker in a marker file's float data, data[marker * dim_x * dim_y + proj * dim_x + item]
 * (dim_x items per projection, at least 10; dim_y projections), read as EmSART reads them: tilt
 * MFI_TiltAngle, psi MFI_RotationPsi, magnification MFI_Magnifiaction, shifts MFI_X_Shift and
 * MFI_Y_Shift (MarkerFile.h:34-46).  *good = CheckIfProjIndexIsGood (MarkerFile.cpp:143-146):
 * both coordinates of the reference marker above -1.  dim_z is the marker count
 * (MarkerFile::GetMarkerCount); ref_marker must lie below it. */
artia_status artia_marker_tilt_params(const float* data, int dim_x, int dim_y, int dim_z, int ref_marker,
                                      int proj, artia_tilt_params* out, int* good);
/* CtfFile (CtfFile.cpp:37-56): projection index's defocus from a CTF file's float data
 * (dim_x columns, dim_y >= 5 rows): minimum row 1, maximum row 2 (nm), astigmatism angle row 4 in
 * radians, returned in degrees; tilt_deg is stored as given. */
artia_status artia_ctf_file_defocus(const float* data, int dim_x, int dim_y, int index, float tilt_deg,
                                    artia_tilt_defocus* out);

/* Reconstructor::GetDefocusDistances (Reconstructor.cpp:54-173): the range [t_in, t_out) of
 * the CTF slab loop along the tilt normal, in pixels */
artia_status artia_geom_defocus_distances(const artia_vol_geom* full, const artia_tilt_geom* g,
                                          float* t_in, float* t_out);
/* Reconstructor::GetDefocusMinMax (Reconstructor.cpp:178-201): defocus of the slab at `ray` */
artia_status artia_geom_defocus_minmax(float ray, const artia_tilt_defocus* d, const artia_ctf_params* p,
                                       float* defocus_min_nm, float* defocus_max_nm);

/* ---------------------------------------------------------------- operators */
/* FPKernel / march (Kernels.cpp:41-87, ForwardProjectionRayMarcher_TL.cu:57-160):
 * proj += ray integral, dist = mean chord, over pixels in roi = {x0, y0, x1, y1}
 * (roi NULL = whole image). */
artia_status artia_fp_march(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                            const float* vol, int width, int height, size_t pitch, float* proj,
                            float* dist, const int* roi, float* ms);
/* SlicerKernel / slicer (Kernels.cpp:103-156, ForwardProjectionSlicer.cu:56-220) */
artia_status artia_fp_slicer(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                             const float* vol, int width, int height, size_t pitch, float* proj,
                             float tmin, float tmax, const int* roi, float* ms);
/* VolTravLengthKernel / volTraversalLength (ForwardProjectionSlicer.cu:223-288) */
artia_status artia_vol_traversal(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                                 int width, int height, size_t pitch, float* dist, const int* roi,
                                 float* ms);
/* Reconstructor::ForwardProjectionNoCTF, Reconstructor::Compare and the cropBorder that
 * BackProjectionNoCTF applies first (Reconstructor.cpp:429-490, 1386-1405, 871-880), fused into the
 * forward projector's epilogue: residual = cropBorder(compare(real, FP(vol), mean chord)), with the
 * projection taken from zero (the reference zeroes proj_d first).  The projection and the mean
 * chord never reach memory; the residual equals fill + artia_fp_march + artia_compare +
 * artia_crop_border bit for bit.  hitpoints == NULL skips cropBorder (its corners as
 * artia_geom_hitpoints gives them, cut_length and dim_length as artia_crop_border takes them).
 * volume_is_empty != 0 takes EmSART's first-projection rule (EmSart.cpp:598-601): no projection,
 * the centre-ray traversal length (artia_vol_traversal) as the chord; vol may then be NULL.  dist
 * (W x H, same pitch) is written with the chord only for an empty volume or under reference
 * numerics, which run the separate launches; it may be NULL otherwise. */
artia_status artia_fp_residual(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, const float* vol,
                               int volume_is_empty, int width, int height, size_t pitch, const float* real,
                               float real_length, const float crop[4], const float crop_dim[4], float proj_scale,
                               const float cut_length[2], const float dim_length[2], const int hitpoints[8],
                               float* residual, float* dist, float* ms);
/* CompKernel / compare (Kernels.cpp:216-270, Compare.cu:49-130): virt <- residual */
artia_status artia_compare(artia_ctx* ctx, int width, int height, size_t pitch, const float* real,
                           float* virt, const float* dist, float real_length, const float crop[4],
                           const float crop_dim[4], float proj_scale, float* ms);
/* CropBorderKernel / cropBorder (Kernels.cpp:276-320, Compare.cu:170-277); p = p1..p4 */
artia_status artia_crop_border(artia_ctx* ctx, int width, int height, size_t pitch, float* img,
                               const float cut_length[2], const float dim_length[2], const int p[8],
                               float* ms);
/* BPKernel / backProjection (Kernels.cpp:339-378, BackProjectionSquareOS.cu:263-782):
 * vol += lambda * chord-weighted mean of img over each voxel's sub-rays; quads of 4 voxels
 * along x (nx % 4 == 0) whose first corner distance lies in [dist_min, dist_max). */
artia_status artia_bp(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, float* vol,
                      const float* img, int width, int height, size_t pitch, float lambda,
                      int oversampling, float dist_min, float dist_max, int filter, float* ms);
/* CopyToSquareKernel / makeSquare (Kernels.cpp:804-830, CopyToSquare.cu:37-92) */
artia_status artia_make_square(artia_ctx* ctx, int width, int height, size_t pitch, const float* in,
                               int n, float* out, int border_x, int border_y, int mirror_y,
                               int fill_zero, float* ms);

/* CTFKernel / ctf (Kernels.cpp:395-435, ctf.cu:80-240) on an R2C half spectrum of an n x n
 * image: (n/2+1) x n complex64 values, row pitch in complex elements.  for_fp selects the
 * forward multiply by -sin(chi)*E, else the Wiener inverse; phase_flip_only flips the sign
 * where sin(chi) >= 0.  Defocus in nm, astigmatism angle in degrees. */
artia_status artia_ctf_apply(artia_ctx* ctx, void* spectrum, int n, size_t pitch, const artia_ctf_params* p,
                             float defocus_min_nm, float defocus_max_nm, float angle_deg, int for_fp,
                             float* ms);
/* WbpWeightingKernel / wbpWeighting (Kernels.cpp:453-488, wbpWeighting.cu:48-144) on an R2C half
 * spectrum; method 0 ramp, 1 exact (no weighting), 2/3/4 contrast 2/10/30 */
artia_status artia_wbp_weight(artia_ctx* ctx, void* spectrum, int n, size_t pitch, float psi_deg,
                              int method, float* ms);
/* One CTF slab filter (Reconstructor.cpp:557-583 forward, :952-965 backward), n = max(W, H):
 * cropBorder(img, corners) in place -> makeSquare(zero fill) -> R2C -> ctf -> C2R ->
 * out = cropBorder(square / n^2) (accumulate = 0) or out += it (accumulate = 1).  With
 * corners == NULL neither crop is applied (the sub-volume path, Reconstructor.cpp:1006-1012).
 * The FFT plans and n x n buffers are cached in the context. */
artia_status artia_ctf_slab_filter(artia_ctx* ctx, float* img, int width, int height, size_t pitch,
                                   float* out, size_t out_pitch, int accumulate, const int corners[8],
                                   const float cut_length[2], const float dim_length[2],
                                   const artia_ctf_params* p, float defocus_min_nm, float defocus_max_nm,
                                   float angle_deg, int for_fp, float* ms);
/* WBP projection filter (Reconstructor.cpp:1271-1319): makeSquare(mirror) -> R2C ->
 * wbpWeighting(psi, method) -> C2R -> out = square / n^2 cropped to W x H */
artia_status artia_wbp_filter(artia_ctx* ctx, const float* img, int width, int height, size_t pitch,
                              float* out, size_t out_pitch, float psi_deg, int method, float* ms);

/* ---------------------------------------------------------------- projection preparation
 * PrepareProjection (Reconstructor.cpp:1158-1383, called per projection by EmSart.cpp:480-545)
 * over a stack of n projections in one call.  Per projection:
 *   1. convert to float (nppiConvert_*32f), square it with mirror padding (CopyToSquare, y flipped
 *      first for DM4 stacks), mean of the square;
 *   2. bad pixels: square values > bad_pixel_value * mean are set to the mean (when
 *      correct_bad_pixels), and counted;
 *   3. SART (wbp == 0): unless skip_filter, R2C -> band-pass (fourierFilter) [-> doseWeighting] ->
 *      C2R; crop / n^2 (n = max(W, H)); then out = (img - m) / -s, with m and the standard
 *      deviation of that image (of its central W/2 x H/2 for ARTIA_NORM_NONE) and s = std, mean or 1
 *      for ARTIA_NORM_STD, MEAN, NONE;
 *   4. WBP (wbp == 1): the normalisation of step 3 first, on the raw cropped image (NONE: m = 0,
 *      s = 1; s divided by cos(tilt) with down_weight_tilts), mirror-squared again, R2C ->
 *      [band-pass] [-> doseWeighting] -> wbpWeighting(psi, wbp_method) -> C2R; out = crop / n^2.
 * images: device, n images of height rows of `pitch` elements of dtype, image i at
 * i * height * pitch elements.  out: device, n images of height x out_pitch floats.  tilt_deg,
 * psi_deg, dose (AccumulatedDose, used with dose_weighting; may be NULL otherwise): host arrays of
 * n.  mean, std (the statistics the reference reports: of the filtered image for SART, of the raw
 * image for WBP) and bad_pixels: host arrays of n, filled before return (the call synchronises).
 * Standard deviations are population deviations (division by the pixel count). */
typedef enum {
    ARTIA_DT_SHORT = 0,
    ARTIA_DT_USHORT = 1,
    ARTIA_DT_INT = 2,
    ARTIA_DT_UINT = 3,
    ARTIA_DT_FLOAT = 4
} artia_dtype;
typedef enum { ARTIA_NORM_MEAN = 0, ARTIA_NORM_STD = 1, ARTIA_NORM_NONE = 2 } artia_norm;
typedef struct {
    int dtype;                 /* artia_dtype */
    int mirror_y;              /* FRM_DM4 read mode */
    int correct_bad_pixels;    /* CorrectBadPixels */
    float bad_pixel_value;     /* BadPixelValue (times the square's mean) */
    int skip_filter;           /* SkipFilter */
    float lp, lps, hp, hps;    /* fourFilterLP, fourFilterLPS, fourFilterHP, fourFilterHPS */
    int dose_weighting;        /* DoseWeighting */
    float pixelsize_nm;        /* doseWeighting's pixel size (passed in Angstrom, x10) */
    int wbp;                   /* WBP */
    int wbp_method;            /* WBPFilter: as artia_wbp_weight */
    int down_weight_tilts;     /* DownWeightTiltsForWBP */
    int normalization;         /* artia_norm (ProjectionNormalization) */
} artia_prep_params;
artia_status artia_prepare_projections(artia_ctx* ctx, const artia_prep_params* pp, const void* images, int n,
                                       int width, int height, size_t pitch, const float* tilt_deg,
                                       const float* psi_deg, const float* dose, float* out, size_t out_pitch,
                                       float* mean, float* std, int* bad_pixels, float* ms);

/* ---------------------------------------------------------------- adjoint and gradients
 * No reference counterpart (the reference's backprojector is SART's normalised update, not the
 * transpose of its ray marcher); for gradient-based refinement of volumes and shifts.  P is the
 * forward projector of artia_fp_march in ARTIA_NUMERICS_EXACT:
 *   (P x)[pixel] = 0.25 * step * sum over the pixel's 4 sub-rays and samples of trilinear(x, q).
 * The volume needs at least 2 voxels per axis.  roi = {x0, y0, x1, y1} or NULL for the image. */
/* vol_grad += scale * P^T resid: the exact transpose of P (same sub-rays, samples and trilinear
 * weights, recomputed), accumulated with device float atomics. */
artia_status artia_fp_adjoint(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                              const float* resid, int width, int height, size_t pitch, float* vol_grad,
                              const int* roi, float scale, float* ms);
/* out[0..1] += sum_pixels grad_proj[pixel] * d(P vol)[pixel] / d(shift_x, shift_y), the tilt's 2D
 * shift as in artia_tilt_params (the detector origin moves by -shift_x*u - shift_y*v,
 * Projection.cpp:183-195).  out is a device array of 2 doubles. */
artia_status artia_fp_shift_grad(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                                 const float* vol, const float* grad_proj, int width, int height, size_t pitch,
                                 const int* roi, double* out, float* ms);

/* ---------------------------------------------------------------- batched refinement (pair sets)
 * EmSARTRefine projects each motive's sub-volume at each tilt (ForwardProjectionROI,
 * EmSartRefine.cpp:809-858) and searches that tilt's 2D shift by cross-correlation
 * (GetDisplacement, Reconstructor.cpp:1746-1959), one call sequence per (sub-volume, tilt).  A
 * pair set keeps all (sub-volume, tilt) pairs of such a step on the device and runs each
 * operation below over all of them in one launch (one workgroup per pair, the pairs of one
 * sub-volume kept on one XCD).  No reference counterpart as an interface.
 *
 * Sub-volumes live in one device buffer, each a dense [z][y][x] fp32 array at its own float
 * offset, with 2 .. 2^24 - 1 voxels per axis, below 2^30 voxels and ny * nz < 2^24.  Pair p
 * projects sub-volume pairs[p].volume through geoms[p] (artia_geom_tilt of its tilt, with the
 * sub-volume's own shifts added to the marker shift) into the detector window
 * x0 <= x < x0 + w, y0 <= y < y0 + h; window images are packed row-major (w floats a row) at
 * pairs[p].offset floats into image buffers, which must hold every window.  Each operation takes
 * `shifts`: 2 doubles per pair on the device, or NULL for zero.  The refinement shift s moves the
 * pair's detector origin by -s_x u - s_y v, applied in double to every sub-ray origin (as the
 * marker shifts enter the origin, Projection.cpp:183-195), so that P_s[x] = P_0[x - s].  P is the
 * exact projector of artia_fp_march restricted to the window,
 *   (P vol)[pixel] = (float)(0.25 * step * sum over the pixel's 4 sub-rays and their samples),
 * bit-identical to artia_fp_march's projection (added to a zero image) at s = 0. */
typedef struct {
    artia_vol_geom geom;  /* sub-volume placement */
    long long offset;     /* float offset of its array in the volume buffer */
} artia_pair_volume;
typedef struct {
    int volume;           /* index into the pair set's sub-volumes */
    int x0, y0, w, h;     /* detector window */
    long long offset;     /* float offset of the window image in image buffers */
} artia_pair;
typedef struct artia_pairs artia_pairs;
artia_status artia_pairs_create(artia_ctx* ctx, const artia_pair_volume* volumes, int nvol, const artia_pair* pairs,
                                const artia_tilt_geom* geoms, int npairs, artia_pairs** out);
artia_status artia_pairs_destroy(artia_pairs* ps);
/* images[window of p] = P_s(p) volumes */
artia_status artia_pairs_project(artia_pairs* ps, const float* volumes, const double* shifts, float* images,
                                 float* ms);
/* Per pair, with r = P_s volumes - targets over its window and J = dP_s/ds (w h x 2):
 * loss[p] = sum r^2, grad[2p..2p+1] = d loss / d s = 2 J^T r, and (normal non-NULL)
 * normal[3p..3p+2] = the Gauss-Newton matrix J^T J as (xx, xy, yy); all double, overwritten.
 * resid (or NULL) receives r in float.  One pass: every cell a sub-ray crosses is read once for
 * the value and the shift derivative, both summed in closed form over the cell's samples. */
artia_status artia_pairs_loss_grad(artia_pairs* ps, const float* volumes, const double* shifts, const float* targets,
                                   double* loss, double* grad, double* normal, float* resid, float* ms);
/* grad[2p], grad[2p+1] = sum over the window of weights * dP_s/ds (double, overwritten) */
artia_status artia_pairs_shift_grad(artia_pairs* ps, const float* volumes, const double* shifts,
                                    const float* weights, double* grad, float* ms);
/* vol_grads += scale * P_s^T images: every pair's window scattered into its sub-volume's
 * gradient (same layout as the volume buffer) by the transpose of P's per-cell closed form, with
 * device float atomics. */
artia_status artia_pairs_adjoint(artia_pairs* ps, const double* shifts, const float* images, float scale,
                                 float* vol_grads, float* ms);
/* GetDisplacement's shift search (Reconstructor.cpp:1746-1959) per pair over its window.  A is the
 * pair's target window, B its model window; both count as zero elsewhere in the reference's
 * N x N square (N = `square`, max(width, height)).  With the square's means and
 * SumA = sum (A - mean A)^2 (likewise B), the normalised correlation
 *   CC(r) = (sum_x A(x) B(x + r) - sum A sum B / N^2) / sqrt(SumA SumB)
 * is searched over the integer shifts with sqrtf(rx^2 + ry^2) <= max_shift (maxShift,
 * wbpWeighting.cu:265-294); the largest wins, ties to the first in the square's row-major order
 * (shifts wrapped modulo N, nppiMaxIndx).  It equals the reference's FFT correlation when the
 * window lies in the square with w, h <= N - 2 max_shift.  peak[2p], peak[2p+1] = (rx, ry): the
 * target matches the model moved by -r, so the shift estimate is s = -r.  Both are -1000 (the
 * reference's failure value) when the best CC is <= 0 or SumA SumB is 0.  cc_value[p] = the best
 * CC; cc_map (or NULL) receives (2 max_shift + 1)^2 floats per pair, [ry + max_shift][rx +
 * max_shift], 0 outside the disk.  0 <= max_shift <= 32, square >= 2 max_shift + 1.
 * multi_peak != 0 is GetDisplacement's MultiPeakDetection (Reconstructor.cpp:1887-1935): the peak is
 * instead the qualifying local maximum nearest to zero shift.  On the float map (0 outside the disk)
 * a pixel qualifies when it is above its left, right, upper, lower and lower-diagonal neighbours,
 * at least equal to its upper-diagonal ones (neighbours wrap around the square; findPeak,
 * wbpWeighting.cu:329-386) and at least 0.9f times the maximum; the winner maximises
 * 1 / (sqrtf(rx^2 + ry^2) + 0.0001f) (maxShiftWeighted, wbpWeighting.cu:295-327), ties to the first
 * in the square's row-major order.  cc_value and the failure test stay those of the maximum. */
artia_status artia_pairs_cc_search(artia_pairs* ps, const float* models, const float* targets, int square,
                                   int max_shift, int multi_peak, int* peak, float* cc_value, float* cc_map,
                                   float* ms);

/* ---------------------------------------------------------------- motive rotation */
/* RotKernel::computeRotMat (Kernels.cpp:1227-1327): R = Rz(psi) Rx(theta) Rz(phi) from the
 * motive list's phi, psi, theta in degrees, row-major R[3 i + j] = rotMat[i][j], in float; the
 * sixteen angles 0, 30, 45, 60, .., 330 take tabulated exact cosines and sines. */
void artia_rot_matrix(float phi, float psi, float theta, float R[9]);
/* The reference cube rotated per motive (EmSartRefine.cpp:805-865 with RotKernel::operator(),
 * Kernels.cpp:1338-1381, and rot3d, wbpWeighting.cu:388-405), n rotations in one call:
 * out + r size^3, [z][y][x], = in sampled at c + R_r^T ((x, y, z) - c), c = (float)(size / 2),
 * trilinear with indices clamped to [0, size - 1] (the texture's CLAMP addressing), exact-float
 * weights.  vol and out are device pointers; angles is a host array of n (phi, psi, theta)
 * triples.  1 <= size <= 1024. */
artia_status artia_rot3d(artia_ctx* ctx, const float* vol, int size, const float* angles, int n, float* out,
                         float* ms);

/* ---------------------------------------------------------------- per-tilt 3D-CTF steps
 * Reconstructor::ForwardProjectionCTF (Reconstructor.cpp:496-644) for one tilt: for every slab
 * ray = slab_start, slab_start + T/ps, ... < slab_end (T = ctf->slice_thickness_nm,
 * ps = ctf->pixelsize_nm, stepped in float; slab_start/slab_end as artia_geom_defocus_distances
 * gives them), the slicer image of [ray, ray + T/ps), cropBorder with the image corners and
 * cut_length/dim_length, the CTF multiply at the slab's defocus (artia_geom_defocus_minmax with
 * *defocus), cropBorder again, added into proj in slab order: proj accumulates, as proj_d does
 * after ResetProjectionsDevice.  dist (W x H, same pitch) is overwritten with the centre-ray
 * traversal length (volTraversalLength).  volume_is_empty != 0: only dist, and vol may be NULL.
 * Exact numerics run every slab of the tilt in one volume pass and filter the slabs in batched
 * FFTs; the result equals the per-slab loop bit for bit. */
artia_status artia_fp_ctf(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, const float* vol,
                          int volume_is_empty, int width, int height, size_t pitch, float slab_start, float slab_end,
                          const artia_tilt_defocus* defocus, const artia_ctf_params* ctf, const float cut_length[2],
                          const float dim_length[2], float* proj, float* dist, float* ms);
/* Reconstructor::BackProjectionCTF (Reconstructor.cpp:920-969) for one tilt: for every slab (as
 * artia_fp_ctf steps them), cropBorder of the residual in place with hitpoints (artia_geom_hitpoints),
 * the Wiener (or phase-flip) filter at the slab's defocus, cropBorder, and the LINEAR-sampled
 * backprojection with `lambda` (Lambda / SIRTCount) gated to the slab's distances.  On return the
 * residual is cropped once per slab, as the reference's loop leaves proj_d.  Exact numerics filter
 * the slabs in batched FFTs into a stack and backproject all of them in one launch, bit for bit
 * the per-slab loop. */
artia_status artia_bp_ctf(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, float* vol,
                          float* residual, int width, int height, size_t pitch, const int hitpoints[8],
                          float slab_start, float slab_end, const artia_tilt_defocus* defocus,
                          const artia_ctf_params* ctf, const float cut_length[2], const float dim_length[2],
                          float lambda, int oversampling, float* ms);

/* ---------------------------------------------------------------- drivers
 * EmSART's ordered-subset SART loop (EmSart.cpp:579-675) without CTF, device resident:
 * for every subset, per tilt FP (volTraversalLength while the volume is empty) -> compare
 * against projections[tilt] -> residual kept on the device; then per tilt cropBorder with
 * hitpoints[8*tilt] -> BP with lambda / S.  tilts/hitpoints/real_lengths/index_list are host
 * arrays of length ntilts; projections is a device array of ntilts * height * width floats
 * (dense rows).  *volume_is_empty carries EmSART's volumeIsEmpty flag across calls. */
artia_status artia_sart_run(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                            const int* hitpoints, const float* real_lengths, int ntilts,
                            const int* index_list, const float* projections, int width, int height,
                            float* vol, const artia_sart_params* sp, int iterations,
                            int* volume_is_empty, float* ms);
/* artia_sart_run without any allocation, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It has no counterpart
 * in the reference.  Same arguments and result as artia_sart_run (bit for bit); the (1 + SIRTCount)
 * projection-sized workspace images come from a 64 MiB device array the library holds on every
 * GPU it runs on for the life of the process, so (1 + sirt_count) * width * height must not exceed
 * 16 Mi floats (2048 x 2048 with SIRTCount <= 3; ARTIA_ERR_INVALID_ARGUMENT beyond).  Calls that
 * overlap on one GPU share that array and race; kernel profiling (artia_profile_enable) must be
 * off. */
artia_status artia_sart_run_noalloc(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                                    const int* hitpoints, const float* real_lengths, int ntilts,
                                    const int* index_list, const float* projections, int width, int height,
                                    float* vol, const artia_sart_params* sp, int iterations, int* volume_is_empty,
                                    float* ms);
/* The same loop with 3D-CTF correction (Reconstructor.cpp:496-645 forward, :920-969 backward):
 * per tilt, slabs ray = t_in, t_in + T/ps, ... < t_out (slab_range[2*tilt] = t_in, t_out from
 * artia_geom_defocus_distances), each sliced, CTF-filtered and summed into the forward
 * projection, and each slab of the residual Wiener-filtered and backprojected with LINEAR
 * sampling inside its distance gate.  defocus[tilt] gives each tilt's defocus. */
artia_status artia_sart_ctf_run(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                                const int* hitpoints, const float* real_lengths, const float* slab_range,
                                const artia_tilt_defocus* defocus, int ntilts, const int* index_list,
                                const float* projections, int width, int height, float* vol,
                                const artia_sart_params* sp, const artia_ctf_params* ctf, int iterations,
                                int* volume_is_empty, float* ms);

/* ---------------------------------------------------------------- half volume storage
 * EmSART's FP16Volume mode (EmSart.cpp:354-374, Config FP16Volume): the volume is stored as IEEE
 * binary16 (uint16_t bits, dense [z][y][x], half the bytes of fp32).  Forward projection filters the
 * half voxels as floats (the reference's texture over a half-float CUDA array); backprojection is
 * BPKernel(modBP, fp16 = true), backProjectionFP16 (BackProjectionSquareOS.cu:785-1290): the fp32
 * kernel's arithmetic between a read of the quad's half voxels and a write rounded to nearest even.
 * Each call behaves as its fp32 namesake on the float values of the half voxels, so a half volume
 * projects bit-identically to the fp32 volume holding the same values, and a backprojection into a
 * half volume equals the fp32 backprojection rounded to half.  The volume pointer is 8-byte
 * aligned. */
artia_status artia_fp_march_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                                const uint16_t* vol, int width, int height, size_t pitch, float* proj,
                                float* dist, const int* roi, float* ms);
artia_status artia_fp_slicer_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                                 const uint16_t* vol, int width, int height, size_t pitch, float* proj,
                                 float tmin, float tmax, const int* roi, float* ms);
artia_status artia_fp_residual_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                                   const uint16_t* vol, int volume_is_empty, int width, int height, size_t pitch,
                                   const float* real, float real_length, const float crop[4], const float crop_dim[4],
                                   float proj_scale, const float cut_length[2], const float dim_length[2],
                                   const int hitpoints[8], float* residual, float* dist, float* ms);
artia_status artia_bp_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, uint16_t* vol,
                          const float* img, int width, int height, size_t pitch, float lambda,
                          int oversampling, float dist_min, float dist_max, int filter, float* ms);
/* artia_sart_run / artia_sart_ctf_run over a half volume (EmSart.cpp:599-667 with FP16Volume) */
artia_status artia_sart_run_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                                const int* hitpoints, const float* real_lengths, int ntilts,
                                const int* index_list, const float* projections, int width, int height,
                                uint16_t* vol, const artia_sart_params* sp, int iterations,
                                int* volume_is_empty, float* ms);
artia_status artia_sart_ctf_run_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                                    const int* hitpoints, const float* real_lengths, const float* slab_range,
                                    const artia_tilt_defocus* defocus, int ntilts, const int* index_list,
                                    const float* projections, int width, int height, uint16_t* vol,
                                    const artia_sart_params* sp, const artia_ctf_params* ctf, int iterations,
                                    int* volume_is_empty, float* ms);
artia_status artia_fp_ctf_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, const uint16_t* vol,
                              int volume_is_empty, int width, int height, size_t pitch, float slab_start,
                              float slab_end, const artia_tilt_defocus* defocus, const artia_ctf_params* ctf,
                              const float cut_length[2], const float dim_length[2], float* proj, float* dist,
                              float* ms);
artia_status artia_bp_ctf_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, uint16_t* vol,
                              float* residual, int width, int height, size_t pitch, const int hitpoints[8],
                              float slab_start, float slab_end, const artia_tilt_defocus* defocus,
                              const artia_ctf_params* ctf, const float cut_length[2], const float dim_length[2],
                              float lambda, int oversampling, float* ms);
/* ConvVolKernel / convertVolumeFP16ToFP32 (Kernels.cpp:1134-1180, BackProjectionSquareOS.cu:1294-1323),
 * as Reconstructor::ConvertVolumeFP16 (Reconstructor.cpp:1734-1745) uses it slice by slice for the
 * output file: slices z0 .. z0+count-1 of an nx x ny x nz half volume as negated floats (EmSART
 * writes volumes with this sign; the fp32 path negates with Volume::Invert, EmSart.cpp:712), slice
 * z0+k at out + k * ny * out_pitch.  (The reference's convertVolume3DFP16ToFP32,
 * BackProjectionSquareOS.cu:1325-1355, stores every slice into the same plane; count > 1 here
 * gives the slices one plane each.) */
artia_status artia_volume_f16_to_f32(artia_ctx* ctx, const uint16_t* vol, int nx, int ny, int nz, int z0,
                                     int count, float* out, size_t out_pitch, float* ms);

/* ---------------------------------------------------------------- file headers
 * EM (EmTools/FileIO/EmHeader.h, EmFile.cpp) and MRC (mrcHeader.h, MRCFile.cpp) stacks and
 * volumes.  artia_file_parse_header reads a file's first bytes (EM: 512, MRC: 1024) and reports
 * what a reader needs to hand the data to the device: the sizes, the element type and where the
 * data starts (MRC: after NEXT extended-header bytes).  dtype is the artia_dtype the projection
 * preparation converts from (EM short/int/float, MRC modes 1, 2, 6), -1 for the other types
 * (bytes, complex, double, half, RGB); raw_type is the header's own code (EM DataType, MRC MODE).
 * EM files must have PC machine coding (EmFile.cpp:75-78; else ARTIA_ERR_UNSUPPORTED).  An MRC
 * header whose NX reads above 32000 is big-endian (MRCFile.cpp:84-88): its fields are swapped and
 * byte_swapped = 1 tells the caller to swap the data.  pixel_size_nm: MRC Xlen / NX / 10
 * (MRCFile::GetPixelsize, MRCFile.cpp:585-591, 0 when unset), EM 1 (TiltSeries::GetPixelSize).
 * artia_file_volume_header writes the header EmSART writes before a reconstructed volume
 * (EmSart.cpp:717-759): MRC mode 2 (64 for a half volume), Xlen/Ylen/Zlen = pixel size x dim x
 * voxel size x 10, axes x, y, z, all else 0; or EM with PC coding, DataType float (10 for half) and
 * ObjectPixelSize = (int)(pixel size x voxel size x 1000).  *header_bytes = 1024 (MRC) or 512 (EM).
 * Host-only calls; no context. */
typedef enum { ARTIA_FILE_EM = 0, ARTIA_FILE_MRC = 1 } artia_file_format;
typedef struct {
    int dtype;            /* artia_dtype, or -1 */
    int raw_type;         /* EM DataType or MRC MODE */
    int elem_bytes;       /* bytes per element */
    int nx, ny, nz;
    size_t data_offset;   /* bytes from the file's start to the first element */
    float pixel_size_nm;
    int byte_swapped;     /* MRC: the file is big-endian */
} artia_file_info;
artia_status artia_file_parse_header(int format, const void* header, size_t size, artia_file_info* out);
artia_status artia_file_volume_header(int format, int half, int nx, int ny, int nz, float pixel_size_nm,
                                      const float voxel_size[3], void* out, size_t capacity,
                                      size_t* header_bytes);

/* ShiftFile (EmTools/FileIO/ShiftFile.cpp:68-84): an EM float file of dim_x motives by dim_y
 * projections by dim_z = 2 components, component c of (proj, motive) at
 * data[c * dim_x * dim_y + proj * dim_x + motive].  artia_shift_file_get is operator(): both
 * components negated (the shift EmSARTSubVolumes applies, EmSartSubVolumes.cpp:651-653);
 * artia_shift_file_set is SetValue: stored as given. */
artia_status artia_shift_file_get(const float* data, int dim_x, int dim_y, int dim_z, int proj, int motive,
                                  float out[2]);
artia_status artia_shift_file_set(float* data, int dim_x, int dim_y, int dim_z, int proj, int motive,
                                  const float value[2]);

/* MotiveList (EmTools/FileIO/MotiveListe.h:33-106, MotiveList.cpp): an EM float file of
 * dim_x = 20 columns (struct motive, in this order) by dim_y particles.  artia_motive_get is GetAt:
 * the row of particle `index`, with the five coordinates scaled by binning_click and the three
 * shifts by binning_shift (ScaleMotivelistPosition, ScaleMotivelistShift). */
typedef struct {
    float cc_coeff, x_coord_2d, y_coord_2d, part_nr, tomo_nr, part_nr_in_tomo, wedge_idx;
    float x, y, z;                      /* x_Coord, y_Coord, z_Coord: position, voxels, from 1 */
    float x_shift, y_shift, z_shift;
    float x_shift_before, y_shift_before, z_shift_before;
    float phi, psi, theta, class_no;
} artia_motive;
artia_status artia_motive_get(const float* data, int dim_x, int dim_y, int index, float binning_click,
                              float binning_shift, artia_motive* out);
/* The particle groups EmSARTRefine iterates over (EmSartRefine.cpp:697-710):
 *   ARTIA_GROUP_BY_CLASS  MotiveList::GetNeighbours(int groupNr) (MotiveList.cpp:202-219): group g
 *                         is the particles whose class equals the g-th distinct class number in
 *                         file order (class numbers are kept as int, MotiveList.cpp:57-73);
 *   ARTIA_GROUP_MAX_DIST  GetNeighbours(index, maxDist) (:180-193): particle g first, then every
 *                         other particle within max_distance (scaled coordinates), in file order;
 *   ARTIA_GROUP_MAX_COUNT GetNeighbours(index, count) (:147-178): the group_size nearest particles
 *                         by (distance, index), particle g itself first.
 * artia_motive_group_count is GetGroupCount (:221-233).  artia_motive_group writes the group's
 * particle indices (global rows) to out (capacity entries) and their number to *count. */
typedef enum { ARTIA_GROUP_BY_CLASS = 0, ARTIA_GROUP_MAX_DIST = 1, ARTIA_GROUP_MAX_COUNT = 2 } artia_group_mode;
artia_status artia_motive_group_count(const float* data, int dim_x, int dim_y, int mode, int* count);
artia_status artia_motive_group(const float* data, int dim_x, int dim_y, int mode, int group, float max_distance,
                                int group_size, float binning_click, float binning_shift, int* out, int capacity,
                                int* count);

/* ---------------------------------------------------------------- sub-volume WBP
 * EmSARTSubVolumes (EmSartSubVolumes.cpp:611-707): every motive-list sub-volume is reconstructed
 * by backprojecting each tilt's WBP-filtered projection (artia_wbp_filter) with lambda 1 through
 * the tilt's geometry shifted by the pair's extra shift (Projection.cpp:183-184).
 *   geoms[s * ntilts + t]   geometry of tilt t for sub-volume s (marker shift + extra shift);
 *   active[s * ntilts + t]  nonzero to backproject the pair (EmSARTSubVolumes skips pairs with
 *                           |extra shift| >= MaxShift - 0.5); NULL = all active;
 *   images                  device, ntilts images of height rows at `pitch` floats, in the order
 *                           the tilts are processed (the reference's index list order);
 *   ctf (optional)          with slab_range[2 t] (artia_geom_defocus_distances of the full
 *                           volume) and defocus[t]: each tilt is filtered per defocus slab without
 *                           border crop and backprojected with LINEAR sampling inside the slab gate
 *                           (Reconstructor.cpp:975-1031); NULL = POINT sampling, gate [0, 1e13). */
typedef struct {
    artia_vol_geom geom; /* artia_geom_subvolume; geom.nx % 4 == 0 */
    float* vol;          /* device, dense [z][y][x] fp32, 16-byte aligned; accumulated into */
} artia_subvol;

/* The reference's schedule (Reconstructor.cpp:894-916, 975-1031): one backprojection launch per
 * active (sub-volume, tilt) pair and slab. */
artia_status artia_subvol_wbp_plain(artia_ctx* ctx, const artia_subvol* subs, int nsub,
                                    const artia_tilt_geom* geoms, const unsigned char* active, int ntilts,
                                    const float* images, int width, int height, size_t pitch,
                                    int oversampling, const artia_ctf_params* ctf, const float* slab_range,
                                    const artia_tilt_defocus* defocus, float* ms);

/* Batched form of the same computation.  No reference counterpart: one persistent launch over a
 * device work queue of (sub-volume, voxel tile) items sorted by size, each thread accumulating all
 * tilts in registers; captured once in a hipGraph (without CTF) and replayed by execute.  With CTF
 * the tilts run in chunks whose filtered slab images fit together (artia_set_stack_limit bytes, by
 * default 16 GiB, at most half the free device memory, at least one tilt): each tilt of a chunk is
 * transformed once and filtered per slab, then one persistent launch adds the chunk's tilts.  Results
 * equal artia_subvol_wbp_plain's bit for bit in ARTIA_NUMERICS_EXACT (the only numerics it
 * supports).  The plan keeps the pointers it was given; they must stay valid until destroy.
 * tuning NULL: the plan times the candidate tile shapes, quads per thread and persistent block
 * counts on a sample of the queue (sub-volume contents are restored) and keeps the fastest. */
typedef struct {
    int autotune;         /* set by the plan when its tuning came from the autotuner */
    int tile_policy;      /* 0: widest x tile (up to 16 quads x 4 rows), 1: 8-row tiles */
    int quads_per_thread; /* 1 or 2 (stacked along z) */
    int blocks_per_cu;    /* persistent blocks per compute unit, 1..16 */
} artia_subvol_tuning;
typedef struct artia_subvol_plan artia_subvol_plan;
artia_status artia_subvol_plan_create(artia_ctx* ctx, const artia_subvol* subs, int nsub,
                                      const artia_tilt_geom* geoms, const unsigned char* active, int ntilts,
                                      const float* images, int width, int height, size_t pitch,
                                      int oversampling, const artia_ctf_params* ctf, const float* slab_range,
                                      const artia_tilt_defocus* defocus, const artia_subvol_tuning* tuning,
                                      artia_subvol_plan** out);
artia_status artia_subvol_plan_execute(artia_subvol_plan* plan, float* ms);
artia_status artia_subvol_plan_tuning(const artia_subvol_plan* plan, artia_subvol_tuning* out);
artia_status artia_subvol_plan_destroy(artia_subvol_plan* plan);

/* ---------------------------------------------------------------- profiling
 * With profiling enabled, artia_sart_run brackets each kernel launch with HIP events on the
 * context stream (no extra synchronisation inside the run) and accumulates, per kernel class,
 * the summed device time and the launch count; enabling resets the counters.  This replaces
 * the per-functor millisecond return values of Kernels.cpp:65-86, which the reference computes
 * and drops. */
typedef enum {
    ARTIA_PROF_FP = 0,        /* march / volTraversalLength */
    ARTIA_PROF_BP = 1,        /* backProjection */
    ARTIA_PROF_COMPARE = 2,   /* compare */
    ARTIA_PROF_CROP = 3,      /* cropBorder */
    ARTIA_PROF_FILL = 4,      /* projection / distance resets */
    ARTIA_PROF_FFT = 5,       /* CTF / WBP spectral filters (FFT pair, multiply, crop) */
    ARTIA_PROF_COUNT = 6
} artia_prof_kernel;
artia_status artia_profile_enable(artia_ctx* ctx, int enable);
artia_status artia_profile_read(artia_ctx* ctx, int kernel, double* total_ms, long long* launches);

#ifdef __cplusplus
}
#endif

#endif /* ARTIA_H */
