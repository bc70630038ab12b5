"""Pins the oracle's CTF, WBP-weighting and CTF-slab geometry functions (ctf.cu:80-240,
wbpWeighting.cu:48-144, Reconstructor.cpp:54-201) with closed forms written independently of
the kernel's formulation, hand-computed known answers, and bit-equality with the product's host
restatement of the slab geometry."""
import math

import numpy as np
import pytest

import oracle_lib as O
import artiatomi_amd as A
from artiatomi_amd import _abi

H_REF, C_REF, E_REF = 6.63e-34, 3.0e8, 1.602e-19   # the kernel's rounded constants (ctf.cu:47-63)


def test_wavelength_closed_form():
    # relativistic electron wavelength lambda = h c / sqrt(eV (2 E0 + eV)), energies in eV
    for kv in (80.0, 120.0, 200.0, 300.0):
        ev = kv * 1e3
        lam = H_REF * C_REF / math.sqrt(ev * (2 * 511e3 + ev) * E_REF * E_REF)
        assert O.ctf_wavelength(kv) == pytest.approx(lam, rel=2e-7)
    # CODATA value at 300 kV is 1.9687 pm; the rounded constants shift it by about +0.14%
    assert O.ctf_wavelength(300.0) == pytest.approx(1.96875e-12, rel=3e-3)


def _freq_grid(N, ps_nm):
    """Spatial frequencies (1/m) of the R2C half spectrum; the Nyquist row y = N/2 is kept at
    +N/2 (the kernel folds only y > N/2), which matters only for the astigmatic angle."""
    step = 1.0 / (N * ps_nm * 1e-9)
    y = np.arange(N)
    ky = (np.where(y > N // 2, y - N, y) * step)[:, None]
    kx = (np.arange(N // 2 + 1) * step)[None, :]
    return kx + 0 * ky, ky + 0 * kx


def _sin_chi(N, ps_nm, cs_mm, kv, dz_m, amp=0.0):
    """Phase-contrast transfer sin(chi + phi) of the standard weak-phase CTF with the reference's
    sign convention (m = -chi), phi = asin(amplitude contrast)."""
    kx, ky = _freq_grid(N, ps_nm)
    k2 = kx * kx + ky * ky
    lam = O.ctf_wavelength(kv)
    cs = cs_mm * 1e-3
    chi = math.pi * lam * k2 * (dz_m - 0.5 * cs * lam * lam * k2)
    return np.sin(-chi + math.asin(amp))


@pytest.mark.parametrize("amp", [0.0, 0.07])
def test_ctf_forward_closed_form(amp):
    N, ps, cs, kv, dz = 64, 0.25, 2.7, 300.0, 3000.0
    k = O.ctf_consts(cs, kv, ps, N, amp)
    X = np.ones((N, N // 2 + 1), complex)
    Y = O.ctf_apply(X, k, dz, dz, 0.0, for_fp=True, beta=(0.0, 0.0, 0.0, 0.0), dtype=np.float64)
    s = _sin_chi(N, ps, cs, kv, dz * 1e-9, amp)
    val = np.where(np.abs(s) < 1e-4, np.where(s >= 0, 1e-4, -1e-4), s)
    expect = -val
    expect[0, 0] = 1.0   # radius 0 is not beyond betaFac.x = 0: untouched
    # the kernel rounds lambda, Cs, the frequency step and the defocus to float before the phase
    np.testing.assert_allclose(Y.real, expect, atol=2e-5)
    assert np.all(Y.imag == 0)
    # the float evaluation tracks the double one
    Yf = O.ctf_apply(X, k, dz, dz, 0.0, for_fp=True, beta=(0.0, 0.0, 0.0, 0.0), dtype=np.float32)
    assert np.max(np.abs(Yf - Y)) < 1e-4


def test_ctf_envelope_and_low_frequency_disk():
    N, ps, cs, kv, dz = 64, 0.25, 2.7, 300.0, 4000.0
    k = O.ctf_consts(cs, kv, ps, N)
    beta = (5.0, 1.0, 0.5, 0.02)
    X = np.ones((N, N // 2 + 1), complex)
    Y = O.ctf_apply(X, k, dz, dz, 0.0, for_fp=True, beta=beta, dtype=np.float64).real
    kx, ky = _freq_grid(N, ps)
    kk = np.sqrt(kx * kx + ky * ky)
    l = kk / 1e8
    env = np.maximum(np.exp(-beta[1] * l - beta[2] * l * l - beta[3] * l ** 3), 0.01)
    v = _sin_chi(N, ps, cs, kv, dz * 1e-9) * env
    v = np.where(np.abs(v) < 1e-4, np.where(v >= 0, 1e-4, -1e-4), v)
    radius = np.hypot(np.fft.rfftfreq(N, 1.0 / N)[None, :], np.fft.fftfreq(N, 1.0 / N)[:, None])
    expect = np.where(radius > beta[0], -v, 1.0)
    np.testing.assert_allclose(Y, expect, atol=2e-5)
    assert np.all(Y[radius <= beta[0]] == 1.0)


def test_ctf_astigmatism():
    # defocus = d0 + (1 - cos 2(alpha - angle)) (d1 - d0) = d0 + 2 sin^2(alpha - angle) (d1 - d0)
    N, ps, cs, kv, d0, d1, ang = 64, 0.25, 2.7, 300.0, 3000.0, 3300.0, 25.0
    k = O.ctf_consts(cs, kv, ps, N)
    X = np.ones((N, N // 2 + 1), complex)
    Y = O.ctf_apply(X, k, d0, d1, ang, for_fp=True, beta=(0.0, 0.0, 0.0, 0.0), dtype=np.float64).real
    kx, ky = _freq_grid(N, ps)
    alpha = np.where(kx == 0, math.pi / 2, np.arctan2(ky, kx))
    dz = d0 + 2 * np.sin(alpha - math.radians(ang)) ** 2 * (d1 - d0)
    lam = O.ctf_wavelength(kv)
    k2 = kx * kx + ky * ky
    s = np.sin(-math.pi * lam * k2 * (dz * 1e-9 - 0.5 * cs * 1e-3 * lam * lam * k2))
    v = np.where(np.abs(s) < 1e-4, np.where(s >= 0, 1e-4, -1e-4), s)
    expect = -v
    expect[0, 0] = 1.0
    np.testing.assert_allclose(Y, expect, atol=5e-5)


def test_ctf_wiener_and_phase_flip():
    N, ps, cs, kv, dz, W = 64, 0.25, 2.7, 300.0, 3500.0, 0.1
    k = O.ctf_consts(cs, kv, ps, N)
    X = np.ones((N, N // 2 + 1), complex)
    s = _sin_chi(N, ps, cs, kv, dz * 1e-9)
    B = O.ctf_apply(X, k, dz, dz, 0.0, for_fp=False, wiener=W, beta=(0.0, 0.0, 0.0, 0.0), dtype=np.float64).real
    expect = -s / (s * s + W)
    expect[0, 0] = 1.0
    np.testing.assert_allclose(B, expect, atol=2e-4)
    # forward then Wiener: s^2 / (s^2 + W) -> 1 where |s| >> sqrt(W)
    F = O.ctf_apply(X, k, dz, dz, 0.0, for_fp=True, beta=(0.0, 0.0, 0.0, 0.0), dtype=np.float64)
    FB = O.ctf_apply(F, k, dz, dz, 0.0, for_fp=False, wiener=W, beta=(0.0, 0.0, 0.0, 0.0), dtype=np.float64).real
    strong = np.abs(s) > 0.9
    assert strong.sum() > 50
    assert np.all(FB[strong] > 0.88) and np.all(FB[strong] < 1.0)
    # phase flip: -1 exactly where sin >= 0, +1 elsewhere, no envelope
    P = O.ctf_apply(X, k, dz, dz, 0.0, phase_flip=True, beta=(0.0, 0.0, 0.0, 0.0), dtype=np.float64).real
    assert set(np.unique(P)) <= {-1.0, 1.0}
    clear = np.abs(s) > 1e-3
    np.testing.assert_array_equal(P[clear], np.where(s[clear] >= 0, -1.0, 1.0))


@pytest.mark.parametrize("psi", [0.0, 90.0, 30.0, -47.0])
def test_wbp_ramp_closed_form(psi):
    N = 64
    X = np.ones((N, N // 2 + 1), complex)
    Y = O.wbp_weight(X, psi, 0, dtype=np.float64).real
    # frequency vector (kx, ky) in pixels rotated by -psi; the weight is |ky'| / N / 2
    a = math.radians(-psi)
    kx = np.arange(N // 2 + 1)[None, :]
    ky = np.where(np.arange(N) > N / 2, np.arange(N) - N, np.arange(N))[:, None]
    ky_rot = math.sin(a) * kx + math.cos(a) * ky
    np.testing.assert_allclose(Y, np.minimum(np.abs(ky_rot) / N / 2, 1.0), atol=1e-6)
    Yf = O.wbp_weight(X, psi, 0, dtype=np.float32).real
    np.testing.assert_allclose(Yf, Y, atol=1e-6)


def test_wbp_exact_and_contrast():
    N = 64
    rng = np.random.default_rng(3)
    X = rng.standard_normal((N, N // 2 + 1)) + 1j * rng.standard_normal((N, N // 2 + 1))
    np.testing.assert_array_equal(O.wbp_weight(X, 17.0, 1), X)   # FM_EXACT has no case: weight 1
    ones = np.ones((N, N // 2 + 1), complex)
    coeff = {2: (1.000654227857550, 0.006008581017124, 0.004159659493151, 0.000975903396538),
             3: (1.003784816598589, 0.029016377161629, 0.019582940715148, 0.004559409669984),
             4: (0.999884616010943, 0.000307646262566, 0.004742915272196, 0.004806551368900)}
    for m, p in coeff.items():
        Y = O.wbp_weight(ones, 0.0, m).real
        assert Y[0, 0] == 0.0 and Y[0, 5] == 0.0          # ky = 0: length 0
        assert np.all((Y >= 0) & (Y <= 1))
        L = math.log(0.25)                                  # Nyquist row: |ky| = N/2 -> 0.25
        w = min(max(p[0] + p[1] * L + p[2] * L * L + p[3] * L ** 3, 0.0), 1.0)
        assert Y[N // 2, 3] == pytest.approx(w, abs=1e-6)
        # weight grows with |ky| towards 1 on every column
        col = Y[1:N // 2 + 1, 0]
        assert np.all(np.diff(col) >= -1e-12)


def _ctf(**kw):
    return A.ctf_params(**kw)


def test_defocus_distances_tilt0_and_extent():
    full = O.volume_geometry(64, 48, 16)
    W, H = 64, 64
    g0 = O.tilt_geometry(0.0, W, H)
    t_in, t_out = O.defocus_distances(full, g0, W, H)
    dz = g0.d[2]
    assert t_in == pytest.approx(dz - full.bbox_max[2], abs=1e-3)
    assert t_out == pytest.approx(dz - full.bbox_min[2], abs=1e-3)
    for tilt, psi in [(30.0, 0.0), (-55.0, 3.0), (12.0, -4.0)]:
        g = O.tilt_geometry(tilt, W, H, psi_deg=psi, shift=(2.0, -3.0))
        a, b = O.defocus_distances(full, g, W, H)
        extent = sum(abs(g.n[i]) * (full.bbox_max[i] - full.bbox_min[i]) for i in range(3))
        assert b - a == pytest.approx(extent, rel=1e-4)
        # the product's host restatement is bit-identical
        vg = A.volume_geometry(64, 48, 16)
        pg = A.tilt_geometry(tilt, W, H, psi_deg=psi, shift=(2.0, -3.0))
        assert A.defocus_distances(vg, pg) == (a, b)


def test_defocus_min_max_known_answers():
    slab = O.CtfSlabParams(4.0, 0.25, 0.0, 0, 0)
    # ray 10100 px -> 100 px from the detector plane -> round(100*0.25*4)/4 + 4/2 = 27 nm
    assert O.defocus_min_max(10100.0, 0.0, 3000.0, 3200.0, slab) == (3027.0, 3227.0)
    assert O.defocus_min_max(10100.6, 0.0, 3000.0, 3200.0, slab) == (3027.25, 3227.25)
    slab_z = O.CtfSlabParams(4.0, 0.25, 8.0, 0, 0)        # minus 8 * 0.25 * cos(60 deg) = 1 nm
    a, b = O.defocus_min_max(10100.0, 60.0, 3000.0, 3200.0, slab_z)
    assert a == pytest.approx(3026.0, abs=1e-3) and b == pytest.approx(3226.0, abs=1e-3)
    slab_ign = O.CtfSlabParams(4.0, 0.25, 8.0, 1, 0)
    assert O.defocus_min_max(10100.0, 60.0, 3000.0, 3200.0, slab_ign) == (3027.0, 3227.0)
    slab_sw = O.CtfSlabParams(4.0, 0.25, 0.0, 0, 1)
    assert O.defocus_min_max(10100.0, 0.0, 3000.0, 3200.0, slab_sw) == (2973.0, 3173.0)


def test_defocus_min_max_product_matches_oracle():
    rng = np.random.default_rng(11)
    for _ in range(200):
        ray = float(np.float32(rng.uniform(9800, 10300)))
        tilt = float(np.float32(rng.uniform(-60, 60)))
        d0, d1 = (float(np.float32(v)) for v in rng.uniform(2000, 5000, 2))
        T, ps, vz = float(np.float32(rng.uniform(2, 60))), float(np.float32(rng.uniform(0.1, 1.5))), \
            float(np.float32(rng.uniform(-20, 20)))
        ign, sw = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        expect = O.defocus_min_max(ray, tilt, d0, d1, O.CtfSlabParams(T, ps, vz, ign, sw))
        p = _ctf(pixelsize_nm=ps, slice_thickness_nm=T, volume_shift_z=vz, ignore_z_shift=bool(ign),
                 switch_direction=bool(sw))
        got = A.defocus_minmax(ray, _abi.TiltDefocus(d0, d1, 0.0, tilt), p)
        assert got == expect
