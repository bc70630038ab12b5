"""The fused forward-projection residual (artia_fp_residual) against the separate operator launches
it replaces, through the C ABI: zero the projection, artia_fp_march, artia_compare, then
artia_crop_border (Reconstructor.cpp:429-490, 1386-1405, 871-880).  The residual must equal that
composition bit for bit, for both numerics, staged and direct walks, fp32 and half volumes, with and
without the border crop, and for EmSART's empty-volume first projection.
"""
import numpy as np
import pytest
import torch

import artiatomi_amd as A

pytestmark = pytest.mark.gpu

W, H = 72, 64
GEOMS = [(0.0, 0.0, (0.0, 0.0)), (31.0, 2.5, (1.5, -2.0)), (-56.0, -4.0, (-3.0, 1.0))]


@pytest.fixture(scope="module")
def ctx(gpu):
    c = A.ArtiaContext(0)
    yield c
    c.close()


def composed(ctx, vg, g, vol, real, L, crop, crop_dim, scale, pts, cut, dim, empty=False):
    proj = torch.zeros((H, W), device="cuda")
    dist = torch.zeros((H, W), device="cuda")
    if empty:
        ctx.vol_traversal(vg, g, dist)
    else:
        ctx.fp_march(vg, g, vol, proj, dist)
    ctx.compare(real, proj, dist, L, crop=crop, crop_dim=crop_dim, scale=scale)
    if pts is not None:
        ctx.crop_border(proj, pts, cut_length=cut, dim_length=dim)
    return proj


@pytest.mark.parametrize("numerics", ["exact", "reference"])
@pytest.mark.parametrize("staged", [True, False])
@pytest.mark.parametrize("half", [False, True])
def test_fp_residual_equals_composition(ctx, numerics, staged, half):
    vg = A.volume_geometry(48, 44, 16)
    gen = torch.Generator(device="cuda").manual_seed(3)
    vol = torch.rand((16, 44, 48), generator=gen, device="cuda")
    if half:
        vol = vol.half()
    real = torch.rand((H, W), generator=gen, device="cuda") * 20.0
    ctx.set_numerics(numerics)
    ctx.set_fp_staging(staged)
    try:
        for tilt, psi, shift in GEOMS:
            g = A.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
            L = A.real_length(vg, g)
            pts = A.hitpoints(vg, g)
            for crop, crop_dim, cut, dim, use_pts in (((0, 0, 0, 0), (0, 0, 0, 0), (0, 0), (0, 0), False),
                                                      ((2, 3, 4, 5), (6, 5, 4, 3), (2, 3), (5, 7), True)):
                print(f"fp_residual {numerics} staged={staged} half={half} tilt={tilt} crop={use_pts}", flush=True)
                p = pts if use_pts else None
                ref = composed(ctx, vg, g, vol, real, L, crop, crop_dim, 1.3, p, cut, dim)
                res = torch.full((H, W), float("nan"), device="cuda")
                dist = torch.zeros((H, W), device="cuda")
                ctx.fp_residual(vg, g, vol, real, res, L, crop=crop, crop_dim=crop_dim, scale=1.3, points=p,
                                cut_length=cut, dim_length=dim, dist=dist)
                torch.cuda.synchronize()
                assert torch.equal(res, ref)
                assert float(ref.abs().max()) > 0
    finally:
        ctx.set_numerics("exact")
        ctx.set_fp_staging(True)


def test_fp_residual_empty_volume(ctx):
    vg = A.volume_geometry(40, 40, 12)
    real = torch.rand((H, W), device="cuda")
    g = A.tilt_geometry(22.0, W, H, psi_deg=1.0)
    L = A.real_length(vg, g)
    pts = A.hitpoints(vg, g)
    ref = composed(ctx, vg, g, None, real, L, (1, 1, 1, 1), (2, 2, 2, 2), 1.0, pts, (1, 1), (3, 3), empty=True)
    res = torch.empty((H, W), device="cuda")
    dist = torch.empty((H, W), device="cuda")
    ctx.fp_residual(vg, g, None, real, res, L, crop=(1, 1, 1, 1), crop_dim=(2, 2, 2, 2), points=pts,
                    cut_length=(1, 1), dim_length=(3, 3), dist=dist, volume_is_empty=True)
    torch.cuda.synchronize()
    assert torch.equal(res, ref)


def test_fp_residual_needs_dist_when_empty(ctx):
    vg = A.volume_geometry(40, 40, 12)
    g = A.tilt_geometry(0.0, W, H)
    res = torch.empty((H, W), device="cuda")
    with pytest.raises(A.ArtiaError):
        ctx.fp_residual(vg, g, None, torch.rand((H, W), device="cuda"), res, 1.0, volume_is_empty=True)
