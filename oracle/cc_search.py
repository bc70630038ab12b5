"""CPU oracle of EmSARTRefine's shift search (Reconstructor::GetDisplacement).

TEST INFRASTRUCTURE ONLY.  Nothing in the product (artiatomi_amd/) may import or call this
module; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, and only as
the checker.

Restated in float64 numpy, step for step, from src/EmSART/Reconstructor.cpp:1746-1959 (paths
relative to the uermel/Artiatomi source tree):
  1. cts(proj_d -> projSquare_d): the measured-minus-rest image placed in the N x N square
     (N = max(width, height)) (:1769);
  2. nppiMean, nppiSubC: subtract its mean over the square (:1775-1779); nppiSqr + nppiSum: SumA,
     the sum of squares of the mean-free image (:1781-1786);
  3. cufftExecR2C (:1790);
  4. the same for the model projection projSubVols_d: mean, SumB, R2C (:1795-1814);
  5. conjMul: conj(F_A) * F_B (wbpWeighting.cu:244-261, called at :1817);
  6. cufftExecC2R, unnormalised (:1819), then nppiDivC by N^2 sqrt(SumA SumB) (:1829);
  7. maxShift: zero every pixel whose wrapped distance sqrtf(mx^2 + my^2) > maxShift
     (wbpWeighting.cu:265-294, called at :1864);
  8. nppiMaxIndx over the square: the first maximum in row-major order (:1866);
  9. wrap the index to a signed shift (x > N/2 -> x - N) and report (-1000, -1000) when the
     maximum is <= 0 (:1938-1956).
With MultiPeakDetection (:1887-1935) step 8 is followed by:
  8a. findPeak (wbpWeighting.cu:329-386) on the masked float map: a pixel is a peak when it
      exceeds its left, right, upper and lower neighbours and its lower diagonal neighbours
      strictly, its upper diagonal neighbours or equals them (the reference's <= for the two
      pixels of row y - 1), and reaches maxVal * 0.9f (neighbours wrap around the square);
  8b. nppiSet 1 everywhere, then 0 where the inverted mask is set: 1 at the peaks, 0 elsewhere;
  8c. maxShiftWeighted (wbpWeighting.cu:295-327): 0 outside the disk, value / (dist + 0.0001f)
      inside (in float);
  8d. nppiMaxIndx again; the shift comes from this index, the failure test and the reported CC
      value from step 8's maximum.
The cuFFT pair computes, up to float rounding, the circular correlation
CC(r) = sum_x A'(x) B'(x + r) / sqrt(SumA SumB) of the mean-free images, which is what this
restatement evaluates (in double, through numpy's FFT).

Pinned by tests/test_oracle_cc.py: a brute-force circular correlation on small squares, exact
recovery of circular shifts (CC = 1), the failure value for an anti-correlated target, the disk
mask, and the row-major tie rule.
"""
from __future__ import annotations

import numpy as np


def to_square(img: np.ndarray, N: int, x0: int = 0, y0: int = 0) -> np.ndarray:
    """The image (or a detector window at (x0, y0)) in an N x N square, zero elsewhere."""
    sq = np.zeros((N, N), dtype=np.float64)
    h, w = img.shape
    sq[y0:y0 + h, x0:x0 + w] = img
    return sq


def multi_peak_index(cc32: np.ndarray, max_val: float, max_shift: int) -> tuple:
    """Steps 8a-8d on the masked float32 map: (x, y) square index of the nearest qualifying peak."""
    N = cc32.shape[0]
    v = cc32
    L, R = np.roll(v, 1, axis=1), np.roll(v, -1, axis=1)       # [y][x - 1], [y][x + 1]
    U, D = np.roll(v, 1, axis=0), np.roll(v, -1, axis=0)       # [y - 1][x], [y + 1][x]
    UL, UR = np.roll(U, 1, axis=1), np.roll(U, -1, axis=1)     # [y - 1][x - 1], [y - 1][x + 1]
    DL, DR = np.roll(D, 1, axis=1), np.roll(D, -1, axis=1)     # [y + 1][x - 1], [y + 1][x + 1]
    thr = np.float32(np.float32(max_val) * np.float32(0.9))
    peak = (L < v) & (R < v) & (U < v) & (D < v) & (UL <= v) & (DL < v) & (UR <= v) & (DR < v) & (v >= thr)
    idx = np.arange(N)
    m = np.where(idx > N // 2, N - idx, idx).astype(np.float32)
    dist = np.sqrt(m[None, :] * m[None, :] + m[:, None] * m[:, None], dtype=np.float32)
    w = np.where(peak, np.float32(1.0), np.float32(0.0))
    w = np.where(dist > np.float32(max_shift), np.float32(0.0), (w / (dist + np.float32(0.0001))).astype(np.float32))
    y, x = divmod(int(np.argmax(w)), N)
    return x, y


def get_displacement(A_sq: np.ndarray, B_sq: np.ndarray, max_shift: int, multi_peak: bool = False):
    """Steps 2-9 on two N x N squares: A the measured-minus-rest image, B the model.  Returns
    (rx, ry, max value, masked correlation map); multi_peak takes steps 8a-8d."""
    N = A_sq.shape[0]
    A = A_sq - A_sq.mean()
    B = B_sq - B_sq.mean()
    sumA = float((A * A).sum())
    sumB = float((B * B).sum())
    FA = np.fft.rfft2(A)
    FB = np.fft.rfft2(B)
    cc = np.fft.irfft2(np.conj(FA) * FB, s=(N, N)) * (N * N)   # cuFFT's C2R is unnormalised
    cc = cc / (N * N * np.sqrt(sumA * sumB))
    idx = np.arange(N)
    m = np.where(idx > N // 2, N - idx, idx).astype(np.float32)
    dist = np.sqrt(m[None, :] ** 2 + m[:, None] ** 2, dtype=np.float32)
    cc = np.where(dist > np.float32(max_shift), 0.0, cc)
    flat = int(np.argmax(cc))            # first maximum in row-major order
    y, x = divmod(flat, N)
    val = float(cc[y, x])
    if multi_peak:
        # the float map of the reference; its maximum is step 8's value
        x, y = multi_peak_index(cc.astype(np.float32), float(np.float32(val)), max_shift)
    rx = x - N if x > N // 2 else x
    ry = y - N if y > N // 2 else y
    if val <= 0:
        rx = ry = -1000
    return rx, ry, val, cc


def cc_map_of(cc: np.ndarray, max_shift: int) -> np.ndarray:
    """The (2 M + 1)^2 neighbourhood of shift 0 of a wrapped correlation map, [ry + M][rx + M]."""
    M = max_shift
    r = np.arange(-M, M + 1)
    return cc[np.ix_(r % cc.shape[0], r % cc.shape[1])]
