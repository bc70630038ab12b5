"""GPU parity of projection preparation (artia_prepare_projections, PrepareProjection of
Reconstructor.cpp:1158-1383) against the oracle composition (prep_compose.prepare).

Gates: the prepared images within 1e-5 relative L2 of the FP64 oracle (single-precision FFT and
float filters), the reported means and standard deviations within 1e-5 relative, the bad-pixel
counts equal.  A stack prepared in one call equals its projections prepared one at a time, bit
for bit (chunks of the batched FFT do not change a projection's result).
"""
import numpy as np
import pytest
import torch

import artiatomi_amd as A
import prep_compose as P
from conftest import rel_l2

pytestmark = pytest.mark.gpu

NP_TYPES = {"int16": np.int16, "uint16": np.uint16, "int32": np.int32, "uint32": np.uint32, "float32": np.float32}


@pytest.fixture(scope="module")
def ctx(gpu):
    c = A.ArtiaContext(0)
    yield c
    c.close()


def stack(n, H, W, dtype, seed, hot=3):
    """Counts-like images with a smooth gradient, noise and a few hot pixels."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    base = 800.0 + 3.0 * xx - 2.0 * yy
    imgs = []
    for i in range(n):
        a = base + 40.0 * i + rng.normal(0.0, 25.0, (H, W))
        for _ in range(hot):
            a[rng.integers(0, H), rng.integers(0, W)] = 60000.0 if dtype in ("uint16", "uint32") else 30000.0
        imgs.append(np.clip(np.rint(a), 0, 65535) if dtype != "float32" else a)
    return np.stack(imgs).astype(NP_TYPES[dtype])


CASES = {
    "sart_std_float": dict(dtype="float32", normalization="std", lp=30.0, lps=6.0),
    "sart_mean_uint16_hp": dict(dtype="uint16", normalization="mean", lp=0.0, lps=0.0, hp=4.0, hps=2.0),
    "sart_none_int16_dm4": dict(dtype="int16", normalization="none", mirror_y=True, lp=25.0, lps=5.0),
    "sart_skip_int32": dict(dtype="int32", normalization="std", skip_filter=True),
    "sart_dose_uint32": dict(dtype="uint32", normalization="mean", lp=40.0, lps=8.0, dose_weighting=True,
                             pixelsize_nm=0.25),
    "wbp_ramp_downweight": dict(dtype="float32", normalization="std", wbp=True, wbp_method="ramp",
                                down_weight_tilts=True, lp=35.0, lps=5.0),
    "wbp_exact_skip_mean": dict(dtype="uint16", normalization="mean", wbp=True, wbp_method="exact",
                                skip_filter=True),
    "wbp_contrast10_none_dose": dict(dtype="float32", normalization="none", wbp=True, wbp_method="contrast10",
                                     dose_weighting=True, pixelsize_nm=0.3, lp=30.0, lps=4.0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_prepare_parity(ctx, name):
    kw = CASES[name]
    n, H, W = 5, 72, 88
    imgs = stack(n, H, W, kw["dtype"], seed=list(CASES).index(name) + 1)
    tilts = np.linspace(-57.0, 54.0, n).astype(np.float32)
    psis = np.linspace(-4.0, 3.0, n).astype(np.float32)
    # doseWeighting's exp(-dose * |k| N/2 / pixelsize) at these doses keeps image content; much larger
    # doses leave a nearly constant image whose normalised residue is all cancellation
    doses = np.linspace(0.05, 2.0, n).astype(np.float32)
    pp = A.prep_params(**kw)
    d_in = torch.from_numpy(imgs).cuda()
    out = torch.zeros((n, H, W), device="cuda")
    mean, std, bad = ctx.prepare_projections(pp, d_in, out, tilt_deg=tilts, psi_deg=psis, dose=doses)
    got = out.cpu().numpy()
    okw = {k: v for k, v in kw.items() if k != "dtype"}
    worst = 0.0
    for i in range(n):
        ref, m, s, b = P.prepare(imgs[i], tilt=float(tilts[i]), psi=float(psis[i]), dose=float(doses[i]),
                                 dtype=np.float64, **okw)
        e = rel_l2(got[i], ref)
        worst = max(worst, e)
        assert b == bad[i], (i, b, bad[i])
        assert mean[i] == pytest.approx(m, rel=1e-5, abs=1e-6), (i, mean[i], m)
        assert std[i] == pytest.approx(s, rel=1e-5, abs=1e-6), (i, std[i], s)
        assert e < 1e-5, (i, e)
    print(f"prepare {name}: worst rel L2 {worst:.2e}, bad pixels {bad.tolist()}")


def test_stack_equals_one_at_a_time(ctx):
    # 7 projections: batched chunks of 4 + 2 + 1 against seven single-projection calls
    n, H, W = 7, 64, 80
    imgs = stack(n, H, W, "uint16", seed=11)
    pp = A.prep_params(dtype="uint16", normalization="std", lp=20.0, lps=4.0, dose_weighting=True,
                       pixelsize_nm=0.2)
    doses = np.linspace(1.0, 40.0, n).astype(np.float32)
    d_in = torch.from_numpy(imgs).cuda()
    out = torch.zeros((n, H, W), device="cuda")
    mean, std, bad = ctx.prepare_projections(pp, d_in, out, dose=doses)
    for i in range(n):
        one = torch.zeros((1, H, W), device="cuda")
        m1, s1, b1 = ctx.prepare_projections(pp, d_in[i:i + 1].contiguous(), one, dose=doses[i:i + 1])
        assert torch.equal(one[0], out[i]), i
        assert m1[0] == mean[i] and s1[0] == std[i] and b1[0] == bad[i]


def test_prepare_rejects_bad_arguments(ctx):
    imgs = torch.zeros((2, 16, 16), device="cuda")
    out = torch.zeros((2, 16, 16), device="cuda")
    with pytest.raises(A.ArtiaError):   # WBP needs psi angles
        ctx.prepare_projections(A.prep_params(wbp=True), imgs, out)
    with pytest.raises(A.ArtiaError):   # dose weighting needs doses
        ctx.prepare_projections(A.prep_params(dose_weighting=True), imgs, out)
    with pytest.raises(A.ArtiaError):   # the tensor's type must match the declared one
        ctx.prepare_projections(A.prep_params(dtype="int16"), imgs, out)
