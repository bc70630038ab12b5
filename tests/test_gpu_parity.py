"""GPU parity of every operator through the C ABI against the CPU oracle, on seeded inputs.

reference numerics: the kernels evaluate the reference kernel bodies in float operation order,
  so they must match the float oracle to <= 1e-6 relative L2 (bit-exact except where a
  transcendental - expf in compare/cropBorder - may differ by one ulp between libm and ocml).
exact numerics: the throughput kernels must match the double-precision oracle to <= 1e-5 relative
  L2 per operator (the SART gate is in test_gpu_sart.py).
"""
import numpy as np
import pytest
import torch

import artiatomi_amd as A
import oracle_lib as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

GEOMS = [
    (0.0, 0.0, (0.0, 0.0)),
    (2.0, 0.0, (0.0, 0.0)),
    (27.0, 0.0, (1.3, -0.7)),
    (-58.0, 3.1, (-2.0, 1.5)),
    (44.0, -4.6, (4.2, -3.3)),
]


@pytest.fixture(scope="module")
def ctx(gpu):
    c = A.ArtiaContext(0)
    yield c
    c.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def setup(tilt, psi, shift, dims=(48, 40, 16), W=64, H=56, seed=0):
    vg = A.volume_geometry(*dims)
    g = A.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
    nx, ny, nz = dims
    vol = np.random.default_rng(seed).random((nz, ny, nx)).astype(np.float32)
    return vg, g, vol


@pytest.mark.parametrize("tilt,psi,shift", GEOMS)
def test_march_parity(ctx, tilt, psi, shift):
    W, H = 64, 56
    vg, g, vol = setup(tilt, psi, shift, W=W, H=H)
    print(f"march tilt={tilt} psi={psi}", flush=True)
    base = np.random.default_rng(1).random((H, W)).astype(np.float32)
    for numerics, dtype, tol in (("reference", np.float32, 1e-6), ("exact", np.float64, 1e-5)):
        ctx.set_numerics(numerics)
        proj, dist = dev(base), dev(np.zeros((H, W)))
        ctx.fp_march(vg, g, dev(vol), proj, dist)
        po, do = O.march(vg, g, vol.astype(dtype), W, H, dtype=dtype, proj=base.astype(dtype))
        assert rel_l2(host(proj), po) <= tol, numerics
        assert rel_l2(host(dist), do) <= tol, numerics


@pytest.mark.parametrize("tilt,psi,shift", GEOMS[2:4])
def test_march_roi_and_slicer_parity(ctx, tilt, psi, shift):
    W, H = 64, 56
    vg, g, vol = setup(tilt, psi, shift, W=W, H=H)
    roi = (5, 7, 50, 41)
    tmin, tmax = 9990.0, 10004.0
    for numerics, dtype, tol in (("reference", np.float32, 1e-6), ("exact", np.float64, 1e-5)):
        ctx.set_numerics(numerics)
        proj, dist = dev(np.zeros((H, W))), dev(np.zeros((H, W)))
        ctx.fp_march(vg, g, dev(vol), proj, dist, roi=roi)
        po, do = O.march(vg, g, vol.astype(dtype), W, H, dtype=dtype, roi=roi)
        assert rel_l2(host(proj), po) <= tol
        sl = dev(np.zeros((H, W)))
        ctx.fp_slicer(vg, g, dev(vol), sl, tmin, tmax)
        so = O.slicer(vg, g, vol.astype(dtype), W, H, tmin, tmax, dtype=dtype)
        assert rel_l2(host(sl), so) <= tol


@pytest.mark.parametrize("tilt,psi,shift", GEOMS)
def test_vol_traversal_parity(ctx, tilt, psi, shift):
    W, H = 64, 56
    vg, g, _ = setup(tilt, psi, shift, W=W, H=H)
    for numerics, dtype, tol in (("reference", np.float32, 0.0), ("exact", np.float64, 1e-7)):
        ctx.set_numerics(numerics)
        d = dev(np.zeros((H, W)))
        ctx.vol_traversal(vg, g, d)
        ref = O.vol_traversal(vg, g, W, H, dtype=dtype)
        assert rel_l2(host(d), ref) <= tol


def test_compare_and_crop_parity(ctx):
    W, H = 70, 50
    r = np.random.default_rng(3)
    real, virt = r.random((H, W)).astype(np.float32), r.random((H, W)).astype(np.float32)
    dist = (r.random((H, W)) * 3).astype(np.float32)
    crop, crop_dim = (2, 3, 4, 5), (6, 5, 4, 3)
    v = dev(virt)
    ctx.compare(dev(real), v, dev(dist), 17.5, crop=crop, crop_dim=crop_dim, scale=1.7)
    ref = O.compare(real, virt, dist, 17.5, crop=crop, crop_dim=crop_dim, scale=1.7)
    assert rel_l2(host(v), ref) <= 1e-6
    pts = [3, 44, 66, 47, 1, 4, 60, 2]
    img = r.random((H, W)).astype(np.float32)
    t = dev(img)
    ctx.crop_border(t, pts, cut_length=(2, 3), dim_length=(5, 7))
    refc = O.crop_border(img, pts, cut=(2, 3), dim=(5, 7))
    assert rel_l2(host(t), refc) <= 1e-6


def test_make_square_parity(ctx):
    img = np.random.default_rng(4).random((40, 64)).astype(np.float32)
    for fill in (True, False):
        for mirror in (False, True):
            out = dev(np.zeros((64, 64)))
            ctx.make_square(dev(img), out, 0, 12, mirror_y=mirror, fill_zero=fill)
            ref = O.make_square(img, 64, 0, 12, mirror_y=mirror, fill_zero=fill)
            assert np.array_equal(host(out), ref)


@pytest.mark.parametrize("tilt,psi,shift", GEOMS)
@pytest.mark.parametrize("os_", [1, 2])
def test_bp_parity(ctx, tilt, psi, shift, os_):
    W, H = 64, 56
    vg, g, vol = setup(tilt, psi, shift, W=W, H=H)
    print(f"bp tilt={tilt} psi={psi} os={os_}", flush=True)
    img = (np.random.default_rng(5).random((H, W)) - 0.5).astype(np.float32)
    for numerics, dtype, tol in (("reference", np.float32, 1e-6), ("exact", np.float64, 1e-5)):
        ctx.set_numerics(numerics)
        for filt in ("point", "linear"):
            v = dev(vol)
            ctx.bp(vg, g, v, dev(img), 0.37, oversampling=os_, filter=filt)
            ref = O.back_project(vg, g, vol.astype(dtype), img.astype(dtype), 0.37, os_=os_,
                                 filt=0 if filt == "point" else 1, dtype=dtype)
            got = host(v)
            assert rel_l2(got - vol, ref - vol) <= tol, (numerics, filt)


def test_bp_slab_gate_parity(ctx):
    W, H = 64, 56
    vg, g, vol = setup(33.0, 1.0, (0.5, 0.5), W=W, H=H)
    img = np.random.default_rng(6).random((H, W)).astype(np.float32)
    for numerics, dtype in (("reference", np.float32), ("exact", np.float64)):
        ctx.set_numerics(numerics)
        v = dev(vol)
        ctx.bp(vg, g, v, dev(img), 1.0, oversampling=2, dist_min=9995.0, dist_max=10003.0)
        ref = O.back_project(vg, g, vol.astype(dtype), img.astype(dtype), 1.0, os_=2, dist_min=9995.0,
                             dist_max=10003.0, dtype=dtype)
        got = host(v)
        changed = np.abs(ref - vol) > 0
        assert 0.05 < changed.mean() < 0.95
        assert rel_l2(got - vol, ref - vol) <= (1e-6 if numerics == "reference" else 1e-5)


def test_bp_rejects_bad_volume_shape(ctx):
    vg = A.volume_geometry(30, 8, 4)   # nx not a multiple of 4
    g = A.tilt_geometry(0.0, 32, 32)
    with pytest.raises(A.ArtiaError):
        ctx.bp(vg, g, dev(np.zeros((4, 8, 30))), dev(np.zeros((32, 32))), 1.0)
