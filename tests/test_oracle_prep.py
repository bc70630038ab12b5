"""Pins the projection-preparation oracle (fourierFilter, doseWeighting and the PrepareProjection
composition) against values the mathematics fixes: the band-pass's plateau and Gaussian edges, the
dose filter's exponential, an FFT pass with a unit filter, the sign and scale of the
normalisations, the bad-pixel count and replacement, and the WBP order (normalise, then filter)."""
import numpy as np
import pytest

import oracle_lib as O
import prep_compose as P


def unit_spectrum(N):
    return np.ones((N, N // 2 + 1), np.complex128)


def freq_dist(N):
    y = np.arange(N, dtype=np.float64)
    y = np.where(y > N * 0.5, -(N - y), y)
    x = np.arange(N // 2 + 1, dtype=np.float64)
    return np.sqrt(x[None, :] ** 2 + y[:, None] ** 2)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fourier_filter_low_pass_edge(dtype):
    N = 64
    f = O.fourier_filter(unit_spectrum(N), 20.0, 0.0, 4.0, 0.0, dtype=dtype).real
    d = freq_dist(N)
    # lp - lps = 16: plateau up to 16, then exp(-(d - 16)^2 / 32) while above 0.001, else 0
    g = np.exp(-((d - 16.0) ** 2) / 32.0)
    want = np.where(d <= 16.0, 1.0, np.where(g > 0.001, g, 0.0))
    tol = 1e-6 if dtype == np.float32 else 1e-12
    assert np.abs(f - want).max() < tol
    assert f[0, 20] == pytest.approx(np.exp(-0.5), rel=1e-6)


def test_fourier_filter_high_pass_edge_and_default_low_pass():
    N = 64
    f = O.fourier_filter(unit_spectrum(N), 0.0, 5.0, 0.0, 2.0).real
    d = freq_dist(N)
    # lp <= 0: pass up to N/2 - 1 = 31; hp + hps = 7: zero below 7 except the Gaussian rise
    low = (d <= 31).astype(np.float64)
    g = np.exp(-((d - 7.0) ** 2) / 8.0)
    want = np.where(d >= 7.0, low, np.where(g > 0.001, g, low * 0.0))
    assert np.abs(f - want).max() < 1e-12
    assert f[0, 5] == pytest.approx(np.exp(-0.5), rel=1e-12)


def test_fourier_filter_pure_gaussian():
    # lp == lps and no high pass: the kernel's special case exp(-d^2 / (2 lps^2))
    N = 32
    f = O.fourier_filter(unit_spectrum(N), 3.0, 0.0, 3.0, 0.0).real
    assert np.abs(f - np.exp(-freq_dist(N) ** 2 / 18.0)).max() < 1e-12


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dose_weighting(dtype):
    N, dose, pixA = 64, 3.0, 2.0
    f = O.dose_weighting(unit_spectrum(N), dose, pixA, dtype=dtype).real
    want = np.exp(-dose * freq_dist(N) / (32 / pixA))
    assert np.abs(f - want).max() < (1e-6 if dtype == np.float32 else 1e-12)


def rand_img(H, W, seed, lo=100.0, hi=200.0):
    return np.random.default_rng(seed).uniform(lo, hi, (H, W)).astype(np.float32)


@pytest.mark.parametrize("norm", ["std", "mean", "none"])
def test_sart_skip_filter_normalisation(norm):
    img = rand_img(40, 56, 1)
    out, m, s, bad = P.prepare(img, skip_filter=True, normalization=norm, correct_bad_pixels=False)
    roi = img[10:30, 14:42] if norm == "none" else img
    mm, ss = roi.astype(np.float64).mean(), roi.astype(np.float64).std()
    assert bad == 0
    assert m == pytest.approx(mm, rel=1e-6) and s == pytest.approx(ss, rel=1e-6)
    div = {"std": ss, "mean": mm, "none": 1.0}[norm]
    # the reference's sign: (img - mean) / -s
    assert np.allclose(out, (img - mm) / -div, rtol=1e-5, atol=1e-5 * abs((img - mm) / div).max())


def test_unit_band_pass_is_identity():
    img = rand_img(48, 40, 2)
    a, *_ = P.prepare(img, skip_filter=True, correct_bad_pixels=False, normalization="std")
    # lp = 4N with no edge: the filter is 1 on the whole spectrum; the FFT pair returns the image
    b, *_ = P.prepare(img, lp=4 * 48, lps=0.0, correct_bad_pixels=False, normalization="std")
    assert np.abs(a - b).max() < 1e-5


def test_bad_pixels_replaced_by_square_mean():
    img = np.full((32, 32), 1.0, np.float32)
    img[16, 16] = 500.0
    img[5, 7] = 300.0
    out, m, s, bad = P.prepare(img, skip_filter=True, bad_pixel_value=10.0, normalization="none")
    assert bad == 2
    sq_mean = np.float32((32 * 32 - 2 + 800.0) / (32 * 32))
    # normalisation 'none' subtracts the centre's mean only; both replaced pixels hold the square's mean
    centre = np.full((16, 16), 1.0)
    centre[8, 8] = sq_mean
    assert out[16, 16] == pytest.approx(-(sq_mean - centre.mean()), abs=1e-6)
    assert out[5, 7] == pytest.approx(-(sq_mean - centre.mean()), abs=1e-6)


def test_wbp_normalises_before_filtering():
    img = rand_img(40, 40, 3)
    mm, ss = img.astype(np.float64).mean(), img.astype(np.float64).std()
    # exact weighting (identity) with a unit band-pass: the output is the normalised raw image
    out, m, s, bad = P.prepare(img, wbp=True, wbp_method="exact", lp=200.0, lps=0.0, normalization="std",
                               correct_bad_pixels=False, psi=3.0)
    assert m == pytest.approx(mm, rel=1e-6) and s == pytest.approx(ss, rel=1e-6)
    assert np.allclose(out, (img - mm) / -ss, atol=2e-5)
    # down-weighting divides s by cos(tilt): the output scales by cos(tilt)
    out2, *_ = P.prepare(img, wbp=True, wbp_method="exact", lp=200.0, lps=0.0, normalization="std",
                         correct_bad_pixels=False, down_weight_tilts=True, tilt=-60.0)
    assert np.allclose(out2, out * 0.5, atol=2e-5)


def test_zero_dose_is_identity():
    img = rand_img(32, 48, 4)
    a, *_ = P.prepare(img, lp=20.0, lps=5.0, correct_bad_pixels=False)
    b, *_ = P.prepare(img, lp=20.0, lps=5.0, correct_bad_pixels=False, dose_weighting=True, dose=0.0,
                      pixelsize_nm=0.3)
    assert np.array_equal(a, b)
