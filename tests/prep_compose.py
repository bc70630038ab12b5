"""Oracle composition of PrepareProjection (Reconstructor.cpp:1158-1383) from the oracle's
makeSquare, fourierFilter, doseWeighting and wbpWeighting kernels and numpy's FFT, with NPP's
statistics restated: nppiMean as the double mean, nppiMean_StdDev as the double mean and the
population standard deviation, nppiCompareC/nppiSum_8u/nppiSet_C1MR as a mask count and
replacement, nppiSubC/nppiDivC as float subtraction and true division.  numpy's irfft2 equals the
reference's unnormalised C2R followed by the division by N^2.  TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import numpy as np

import oracle_lib as O

f32 = np.float32


def _stats(a):
    a = np.asarray(a, np.float64)
    m = float(a.mean())
    return m, float(np.sqrt(((a - m) ** 2).mean()))


def prepare(img, *, mirror_y=False, correct_bad_pixels=True, bad_pixel_value=10.0, skip_filter=False, lp=800.0,
            lps=200.0, hp=0.0, hps=0.0, dose_weighting=False, pixelsize_nm=1.0, wbp=False, wbp_method="ramp",
            down_weight_tilts=False, normalization="mean", tilt=0.0, psi=0.0, dose=0.0, dtype=np.float64):
    """One projection (H, W) of any integer or float type.  Returns (out, mean, std, bad_pixels) as
    the reference leaves them: out in float32 (cast from dtype), mean and std the statistics it
    reports.  dtype is the arithmetic of the spectral steps (float64: the FP64 oracle)."""
    H, W = img.shape
    N = max(W, H)
    bx, by = (N - W) // 2, (N - H) // 2
    real = np.asarray(img).astype(np.float32)                       # nppiConvert_*32f
    sq = O.make_square(real, N, bx, by, mirror_y=mirror_y, fill_zero=False, dtype=np.float32)
    meanValue = f32(_stats(sq)[0])                                  # nppiMean over the square
    bad = 0
    if correct_bad_pixels:
        mask = sq > f32(f32(bad_pixel_value) * meanValue)
        bad = int(mask.sum())
        sq[mask] = meanValue
    method = {"ramp": 0, "exact": 1, "contrast2": 2, "contrast10": 3, "contrast30": 4}[wbp_method]
    stdValue = f32(0)
    if wbp:
        us = sq[by:by + H, bx:bx + W] / f32(1.0)
        m, s = _stats(us)
        meanValue, stdValue = f32(m), f32(s)
        std_hf, mean_hf = stdValue, meanValue
        if normalization == "mean":
            std_hf = meanValue
        if normalization == "none":
            std_hf, mean_hf = f32(1), f32(0)
        if down_weight_tilts:
            arg = f32(float(f32(abs(f32(tilt))) / f32(180.0)) * np.pi)
            std_hf = f32(std_hf / f32(np.cos(arg)))
        rp = ((us - mean_hf).astype(np.float32) / f32(-std_hf)).astype(np.float32)
        sq = O.make_square(rp, N, bx, by, mirror_y=False, fill_zero=False, dtype=np.float32)
    if not skip_filter or wbp:
        X = np.fft.rfft2(np.asarray(sq, dtype))
        if not skip_filter:
            X = O.fourier_filter(X, lp, hp, lps, hps, dtype=dtype)
        if dose_weighting:
            X = O.dose_weighting(X, dose, f32(f32(pixelsize_nm) * f32(10.0)), dtype=dtype)
        if wbp:
            X = O.wbp_weight(X, psi, method, dtype=dtype)
        us = np.fft.irfft2(X, s=(N, N))[by:by + H, bx:bx + W]
    else:
        us = sq[by:by + H, bx:bx + W] / f32(1.0)
    if wbp:
        return np.asarray(us, np.float32), float(meanValue), float(stdValue), bad
    # SART: statistics of the filtered image (its centre for IMOD-aligned stacks), normalised
    roi = us[H // 4:H // 4 + H // 2, W // 4:W // 4 + W // 2] if normalization == "none" else us
    m, s = _stats(roi)
    meanValue, stdValue = f32(m), f32(s)
    std_hf, mean_hf = stdValue, meanValue
    if normalization == "mean":
        std_hf = meanValue
    if normalization == "none":
        std_hf = f32(1)
    out = (np.asarray(us, dtype) - dtype(mean_hf)) / dtype(-std_hf)
    return np.asarray(out, np.float32), float(meanValue), float(stdValue), bad
