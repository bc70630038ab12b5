"""Oracle compositions for the Fourier-space paths: the CTF slab filter, the WBP projection filter
and EmSART's SART loop with 3D-CTF correction, built from the oracle's kernels (cropBorder,
makeSquare, ctf, wbpWeighting, slicer, volTraversalLength, compare, backProjection) and numpy's
FFT.  numpy's irfft2 equals the reference's unnormalised C2R followed by the division by N^2
(Reconstructor.cpp:578-579).  TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import numpy as np

import oracle_lib as O


def slab_filter(img, corners, consts, dmin, dmax, angle, for_fp, cut=(0, 0), dim=(0, 0), phase_flip=False,
                wiener=0.1, beta=(200.0, 0.0, 0.008, 0.0), dtype=np.float64):
    """cropBorder(img) -> makeSquare(zero) -> R2C -> ctf -> C2R / N^2 -> crop to W x H ->
    cropBorder (no crops when corners is None).  Returns (cropped img, filtered image)."""
    H, W = img.shape
    N = max(W, H)
    bx, by = (N - W) // 2, (N - H) // 2
    a = O.crop_border(img, corners, cut, dim, dtype=dtype) if corners is not None else np.asarray(img, dtype)
    sq = O.make_square(a, N, bx, by, mirror_y=False, fill_zero=True, dtype=dtype)
    X = np.fft.rfft2(sq)
    X = O.ctf_apply(X, consts, dmin, dmax, angle, for_fp=for_fp, phase_flip=phase_flip, wiener=wiener,
                    beta=beta, dtype=dtype)
    r = np.fft.irfft2(X, s=(N, N)).astype(dtype)
    out = np.ascontiguousarray(r[by:by + H, bx:bx + W])
    if corners is None:   # the sub-volume path filters without border crops
        return a, out
    return a, O.crop_border(out, corners, cut, dim, dtype=dtype)


def wbp_filter(img, psi, method, dtype=np.float64):
    """makeSquare(mirror) -> R2C -> wbpWeighting -> C2R / N^2 -> crop to W x H."""
    H, W = img.shape
    N = max(W, H)
    bx, by = (N - W) // 2, (N - H) // 2
    sq = O.make_square(img, N, bx, by, mirror_y=False, fill_zero=False, dtype=dtype)
    X = O.wbp_weight(np.fft.rfft2(sq), psi, method, dtype=dtype)
    r = np.fft.irfft2(X, s=(N, N)).astype(dtype)
    return np.ascontiguousarray(r[by:by + H, bx:bx + W])


def slab_rays(t_in, t_out, step):
    """The reference's float slab loop `for (float ray = t_in; ray < t_out; ray += step)`."""
    rays = []
    ray = np.float32(t_in)
    step = np.float32(step)
    while ray < np.float32(t_out):
        rays.append(float(ray))
        ray = np.float32(ray + step)
    return rays


def sart_ctf(vg, full, geoms, hitpoints, real_lengths, slab_ranges, defocus, index_list, projections, vol, sp,
             ctf, iterations, empty=True, dtype=np.float64):
    """EmSART's loop (EmSart.cpp:579-675) with ForwardProjectionCTF (Reconstructor.cpp:496-645) and
    BackProjectionCTF (:920-969).  `defocus` holds (dmin_nm, dmax_nm, astig_deg, tilt_deg) per
    tilt; `ctf` is an artia_ctf_params."""
    T, H, W = projections.shape
    N = max(W, H)
    consts = O.ctf_consts(ctf.cs_mm, ctf.voltage_kV, ctf.pixelsize_nm, N, ctf.amp_contrast)
    slab = O.CtfSlabParams(ctf.slice_thickness_nm, ctf.pixelsize_nm, ctf.volume_shift_z, ctf.ignore_z_shift,
                           ctf.switch_direction)
    step = float(np.float32(ctf.slice_thickness_nm) / np.float32(ctf.pixelsize_nm))
    beta = tuple(ctf.beta_fac)
    frame = (0, H - 1, W - 1, H - 1, 0, 0, W - 1, 0)
    cut, dimL = tuple(sp.cut_length), tuple(sp.dim_length)
    S = sp.sirt_count
    lam = float(np.float32(sp.lambda_) / np.float32(S))
    vol = np.ascontiguousarray(vol, dtype=dtype).copy()
    kw = dict(phase_flip=bool(ctf.phase_flip_only), wiener=ctf.wiener, beta=beta, dtype=dtype)
    for _ in range(iterations):
        for step_i in range((T + S - 1) // S):
            residuals = {}
            for i in range(S):
                if step_i * S + i >= T:
                    continue
                idx = index_list[step_i * S + i]
                g = geoms[idx]
                dmin0, dmax0, ang, tilt = defocus[idx]
                proj = np.zeros((H, W), dtype)
                if not empty:
                    for ray in slab_rays(*slab_ranges[idx], step):
                        dmin, dmax = O.defocus_min_max(ray, tilt, dmin0, dmax0, slab)
                        dist = O.slicer(vg, g, vol, W, H, ray, float(np.float32(ray) + np.float32(step)),
                                        dtype=dtype)
                        _, f = slab_filter(dist, frame, consts, dmin, dmax, ang, True, cut, dimL, **kw)
                        proj = (proj + f).astype(dtype)
                dist = O.vol_traversal(vg, g, W, H, dtype=dtype)
                residuals[i] = O.compare(projections[idx], proj, dist, real_lengths[idx], tuple(sp.crop),
                                         tuple(sp.crop_dim), sp.proj_scale, dtype=dtype)
            for i in range(S):
                if step_i * S + i >= T:
                    continue
                empty = False
                idx = index_list[step_i * S + i]
                g = geoms[idx]
                dmin0, dmax0, ang, tilt = defocus[idx]
                res = residuals[i]
                for ray in slab_rays(*slab_ranges[idx], step):
                    dmin, dmax = O.defocus_min_max(ray, tilt, dmin0, dmax0, slab)
                    res, f = slab_filter(res, hitpoints[idx], consts, dmin, dmax, ang, False, cut, dimL, **kw)
                    vol = O.back_project(vg, g, vol, f, lam, sp.oversampling, ray,
                                         float(np.float32(ray) + np.float32(step)), filt=1, dtype=dtype)
    return vol, empty


def subvol_wbp(subs, geoms, images, os_, active=None, ctf=None, slab_ranges=None, defocus=None, dtype=np.float64):
    """EmSARTSubVolumes' loop (EmSartSubVolumes.cpp:611-707): per tilt, every active sub-volume gets
    a lambda-1 backprojection of the tilt's image; with CTF, per defocus slab the image is filtered
    without crops (Reconstructor.cpp:975-1031) and backprojected with LINEAR sampling in the gate.
    subs: [(vol_geom, volume array)]; returns the updated volumes."""
    T, H, W = images.shape
    vols = [np.ascontiguousarray(v, dtype=dtype).copy() for _, v in subs]
    if ctf is not None:
        N = max(W, H)
        consts = O.ctf_consts(ctf.cs_mm, ctf.voltage_kV, ctf.pixelsize_nm, N, ctf.amp_contrast)
        slab = O.CtfSlabParams(ctf.slice_thickness_nm, ctf.pixelsize_nm, ctf.volume_shift_z, ctf.ignore_z_shift,
                               ctf.switch_direction)
        step = float(np.float32(ctf.slice_thickness_nm) / np.float32(ctf.pixelsize_nm))
    for t in range(T):
        if ctf is None:
            for s, (vg, _) in enumerate(subs):
                if active is not None and not active[s][t]:
                    continue
                vols[s] = O.back_project(vg, geoms[s][t], vols[s], images[t], 1.0, os_, 0.0, 9999999999999.0,
                                         filt=0, dtype=dtype)
            continue
        dmin0, dmax0, ang, tilt = defocus[t]
        for ray in slab_rays(*slab_ranges[t], step):
            dmin, dmax = O.defocus_min_max(ray, tilt, dmin0, dmax0, slab)
            _, f = slab_filter(images[t].astype(dtype), None, consts, dmin, dmax, ang, False,
                               phase_flip=bool(ctf.phase_flip_only), wiener=ctf.wiener, beta=tuple(ctf.beta_fac),
                               dtype=dtype)
            for s, (vg, _) in enumerate(subs):
                if active is not None and not active[s][t]:
                    continue
                vols[s] = O.back_project(vg, geoms[s][t], vols[s], f, 1.0, os_, ray,
                                         float(np.float32(ray) + np.float32(step)), filt=1, dtype=dtype)
    return vols
