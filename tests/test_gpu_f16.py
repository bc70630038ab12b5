"""Half volume storage (EmSART's FP16Volume mode, EmSart.cpp:354-374) through the C ABI.

The reference's FP16 path is the fp32 path on the half voxels' float values: the forward projectors
filter a half-float texture as floats, and backProjectionFP16 (BackProjectionSquareOS.cu:785-1290)
is backProjection between __half2float reads and __float2half_rn writes.  The gates follow from
that, bit for bit:
  - a projection of a half volume equals the projection of the fp32 volume holding its values;
  - a backprojection into a half volume equals the fp32 backprojection rounded to half;
  - the half SART drivers equal the loops composed from the half operators, slab by slab for CTF;
  - the slice export is -(float)voxel (convertVolumeFP16ToFP32).
The fp32 operators these rest on are pinned against the oracle in test_gpu_parity.py,
test_gpu_sart.py and test_gpu_ctf.py.
"""
import numpy as np
import pytest
import torch

import artiatomi_amd as A
import ctf_compose as CC
from artiatomi_amd import phantom
from conftest import rel_l2
from test_gpu_ctf import ctf_case, per_slab_sart_ctf
from test_gpu_sart import make_case

pytestmark = pytest.mark.gpu

GEOMS = [
    (0.0, 0.0, (0.0, 0.0)),
    (27.0, 0.0, (1.3, -0.7)),
    (-58.0, 3.1, (-2.0, 1.5)),
]


@pytest.fixture(scope="module")
def ctx(gpu):
    c = A.ArtiaContext(0)
    yield c
    c.close()


def half_volume(dims, seed, scale=1.0):
    nx, ny, nz = dims
    v = np.random.default_rng(seed).random((nz, ny, nx)).astype(np.float32) * np.float32(scale)
    h = torch.from_numpy(v).cuda().half()
    return h, h.float()


@pytest.mark.parametrize("numerics", ["exact", "reference"])
@pytest.mark.parametrize("tilt,psi,shift", GEOMS)
def test_march_and_slicer_f16_equal_fp32_values(ctx, numerics, tilt, psi, shift):
    W, H, dims = 64, 56, (48, 40, 16)
    vg = A.volume_geometry(*dims)
    g = A.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
    vh, vf = half_volume(dims, 1)
    ctx.set_numerics(numerics)
    try:
        for staging in (True, False):
            ctx.set_fp_staging(staging)
            print(f"march f16 {numerics} tilt={tilt} staging={staging}", flush=True)
            out = []
            for vol in (vh, vf):
                proj = torch.full((H, W), 0.5, device="cuda")
                dist = torch.zeros((H, W), device="cuda")
                ctx.fp_march(vg, g, vol, proj, dist)
                sl = torch.zeros((H, W), device="cuda")
                ctx.fp_slicer(vg, g, vol, sl, 9990.0, 10004.0)   # a slab through the volume's middle
                out.append((proj, dist, sl))
            torch.cuda.synchronize()
            assert float(out[1][0].abs().max()) > 1.0 and float(out[1][2].abs().max()) > 0.0
            for a, b in zip(*out):
                assert torch.equal(a, b)
    finally:
        ctx.set_fp_staging(True)
        ctx.set_numerics("exact")


@pytest.mark.parametrize("numerics", ["exact", "reference"])
@pytest.mark.parametrize("os_,filt", [(1, "point"), (2, "point"), (3, "point"), (2, "linear")])
@pytest.mark.parametrize("tilt,psi,shift", GEOMS)
def test_bp_f16_is_fp32_rounded(ctx, numerics, os_, filt, tilt, psi, shift):
    W, H, dims = 64, 56, (48, 40, 16)
    vg = A.volume_geometry(*dims)
    g = A.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
    img = torch.from_numpy(np.random.default_rng(2).normal(0, 1, (H, W)).astype(np.float32)).cuda()
    vh, vf = half_volume(dims, 3, scale=0.25)
    ctx.set_numerics(numerics)
    try:
        print(f"bp f16 {numerics} tilt={tilt} os={os_} {filt}", flush=True)
        # a distance gate that keeps part of the volume: gated quads stay as they are
        for dmin, dmax in ((0.0, 9999999999999.0), (9994.0, 10001.0)):
            before = vf.clone()
            ctx.bp(vg, g, vh, img, 0.7, oversampling=os_, dist_min=dmin, dist_max=dmax, filter=filt)
            ctx.bp(vg, g, vf, img, 0.7, oversampling=os_, dist_min=dmin, dist_max=dmax, filter=filt)
            torch.cuda.synchronize()
            assert not torch.equal(vf, before)
            assert torch.equal(vh, vf.half())
            vf = vh.float()
    finally:
        ctx.set_numerics("exact")


def plain_sart(ctx, vg, acq, d_proj, vol, sp, iterations):
    """EmSART's loop (EmSart.cpp:579-675) for S = 1 through the per-operator entry points."""
    H, W = acq.height, acq.width
    proj = torch.zeros((H, W), device="cuda")
    dist = torch.zeros((H, W), device="cuda")
    lam = float(np.float32(sp.lambda_) / np.float32(sp.sirt_count))
    empty = True
    for _ in range(iterations):
        for idx in acq.index_list:
            g = acq.geoms[idx]
            proj.zero_()
            dist.zero_()
            if empty:
                ctx.vol_traversal(vg, g, dist)
            else:
                ctx.fp_march(vg, g, vol, proj, dist)
            ctx.compare(d_proj[idx], proj, dist, acq.real_lengths[idx], tuple(sp.crop), tuple(sp.crop_dim),
                        sp.proj_scale)
            empty = False
            ctx.crop_border(proj, acq.hitpoints[idx], tuple(sp.cut_length), tuple(sp.dim_length))
            ctx.bp(vg, g, vol, proj, lam, oversampling=sp.oversampling)
    torch.cuda.synchronize()


@pytest.mark.parametrize("numerics", ["exact", "reference"])
def test_sart_f16_equals_operator_loop(gpu, numerics):
    dims, W, H = (40, 44, 12), 52, 48
    vg, acq, projs = make_case(dims, W, H, phantom.tilt_series(-60, 60, 9), seed=5, psi_range=4.0,
                               shift_range=2.0)
    sp = A.sart_params(lambda_=0.4, oversampling=2, crop=(3, 3, 3, 3), crop_dim=(2, 2, 2, 2), cut_length=(1, 1),
                       dim_length=(2, 2))
    nx, ny, nz = dims
    ctx = A.ArtiaContext(0, numerics=numerics)
    try:
        d_proj = torch.from_numpy(projs).cuda()
        vh = torch.zeros((nz, ny, nx), device="cuda", dtype=torch.float16)
        _, empty = ctx.sart_run(vg, acq, d_proj, vh, sp, iterations=2)
        ref = torch.zeros_like(vh)
        plain_sart(ctx, vg, acq, d_proj, ref, sp, iterations=2)
        assert not empty and float(ref.float().abs().max()) > 0
        assert torch.equal(vh, ref)
        # half storage tracks the fp32 reconstruction to about the half rounding
        vf = torch.zeros((nz, ny, nx), device="cuda")
        ctx.sart_run(vg, acq, d_proj, vf, sp, iterations=2)
        err = rel_l2(vh.float().cpu().numpy(), vf.cpu().numpy())
        print(f"sart f16 {numerics}: rel L2 to fp32 storage {err:.2e}", flush=True)
        assert err < 5e-3
    finally:
        ctx.close()


def test_sart_ctf_f16_equals_per_slab_composition(gpu):
    dims, vg, acq, projs, defocus, ctf = ctf_case(seed=4)
    nx, ny, nz = dims
    sp = A.sart_params(lambda_=0.3, oversampling=2, cut_length=(1, 1), dim_length=(2, 2))
    tdef = [A.TiltDefocus(*d) for d in defocus]
    ctx = A.ArtiaContext(0)
    try:
        d_proj = torch.from_numpy(projs).cuda()
        vh = torch.zeros((nz, ny, nx), device="cuda", dtype=torch.float16)
        ctx.sart_ctf_run(vg, acq, tdef, d_proj, vh, sp, ctf, iterations=2)
        ref = torch.zeros_like(vh)
        per_slab_sart_ctf(ctx, vg, acq, tdef, d_proj, ref, sp, ctf, iterations=2)
        assert float(ref.float().abs().max()) > 0
        assert torch.equal(vh, ref)
        # and under a stack limit that sends the forward slabs through several volume passes
        npix = acq.width * acq.height
        step = np.float32(ctf.slice_thickness_nm) / np.float32(ctf.pixelsize_nm)
        max_slabs = max(len(CC.slab_rays(a, b, step)) for a, b in acq.slab_ranges)
        ctx.set_stack_limit(max_slabs * npix * 4 + 32 * npix + 32 * npix * max(1, max_slabs // 3))
        v2 = torch.zeros_like(vh)
        ctx.sart_ctf_run(vg, acq, tdef, d_proj, v2, sp, ctf, iterations=2)
        torch.cuda.synchronize()
        assert torch.equal(v2, vh)
    finally:
        ctx.close()


def test_volume_f16_to_f32(ctx):
    vh, vf = half_volume((36, 20, 9), 6, scale=100.0)
    out = torch.full((4, 20, 36), 7.0, device="cuda")
    ctx.volume_f16_to_f32(vh, out, z0=3, count=4)
    torch.cuda.synchronize()
    assert torch.equal(out, -vf[3:7])
    with pytest.raises(A.ArtiaError):
        ctx.volume_f16_to_f32(vh, out, z0=6, count=4)


def test_f16_rejects_misaligned_volume(ctx):
    dims, W, H = (48, 40, 16), 64, 56
    vg = A.volume_geometry(*dims)
    g = A.tilt_geometry(10.0, W, H)
    buf = torch.zeros(48 * 40 * 16 + 4, device="cuda", dtype=torch.float16)
    img = torch.zeros((H, W), device="cuda")
    with pytest.raises(A.ArtiaError):   # a quad of half voxels is one 8-byte access
        ctx.bp(vg, g, buf[2:2 + 48 * 40 * 16], img, 1.0)
    ctx.bp(vg, g, buf[4:4 + 48 * 40 * 16], img, 1.0)
