"""GPU check that the LDS-staged exact forward projector (artia_set_fp_staging(ctx, 1), the default)
and the direct-read one give bit-identical projections and distance images, for march and slicer,
across tilts (including steep ones that fall back to direct reads), tilt-axis rotations, shifts,
magnification anisotropy, regions of interest, a z-shifted sub-volume, and a whole SART iteration."""
import numpy as np
import pytest
import torch

import artiatomi_amd as A

pytestmark = pytest.mark.gpu

CASES = [
    # (nx, ny, nz, W, H, tilt, psi, shift, roi)
    (64, 48, 20, 72, 64, 0.0, 0.0, (0.0, 0.0), None),
    (64, 48, 20, 72, 64, 17.0, 3.0, (1.3, -2.2), None),
    (64, 48, 20, 72, 64, -33.0, -4.5, (-0.7, 0.4), (5, 3, 61, 50)),
    (64, 48, 20, 72, 64, 45.0, 0.0, (0.0, 0.0), None),
    (64, 48, 20, 72, 64, 58.0, 2.0, (2.0, 1.0), None),
    (64, 48, 20, 72, 64, -60.0, -5.0, (0.0, 0.0), None),
    (64, 48, 20, 72, 64, 72.0, 1.0, (0.0, 0.0), None),
    (40, 56, 36, 64, 64, 38.0, 88.0, (0.5, 0.5), None),
    (48, 48, 48, 40, 40, -24.0, 30.0, (0.0, 0.0), (0, 0, 33, 27)),
]


@pytest.fixture(scope="module")
def ctx(gpu):
    c = A.ArtiaContext(0)
    yield c
    c.set_fp_staging(True)
    c.close()


def project_both(ctx, fn):
    out = []
    for staged in (True, False):
        ctx.set_fp_staging(staged)
        out.append(fn())
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("nx,ny,nz,W,H,tilt,psi,shift,roi", CASES)
def test_march_staged_equals_direct(ctx, nx, ny, nz, W, H, tilt, psi, shift, roi):
    print(f"march staged vs direct: {nx}x{ny}x{nz} tilt={tilt} psi={psi} roi={roi}", flush=True)
    vg = A.volume_geometry(nx, ny, nz)
    g = A.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
    gen = torch.Generator(device="cuda").manual_seed(nx + int(tilt))
    vol = torch.rand((nz, ny, nx), generator=gen, device="cuda")

    def run():
        proj = torch.zeros((H, W), device="cuda")
        dist = torch.zeros((H, W), device="cuda")
        ctx.fp_march(vg, g, vol, proj, dist, roi=roi)
        return proj, dist

    (p1, d1), (p2, d2) = project_both(ctx, run)
    assert float(p1.abs().max()) > 0
    assert torch.equal(p1, p2)
    assert torch.equal(d1, d2)


@pytest.mark.parametrize("tilt", [0.0, 29.0, -51.0])
def test_slicer_staged_equals_direct(ctx, tilt):
    vg = A.volume_geometry(48, 40, 24)
    W = H = 56
    g = A.tilt_geometry(tilt, W, H, psi_deg=1.5)
    vol = torch.rand((24, 40, 48), generator=torch.Generator(device="cuda").manual_seed(5), device="cuda")
    t0, t1 = A.defocus_distances(vg, g)
    tm = 0.5 * (t0 + t1)

    def run():
        proj = torch.zeros((H, W), device="cuda")
        ctx.fp_slicer(vg, g, vol, proj, t0 - 1.0, tm)
        return proj

    p1, p2 = project_both(ctx, run)
    assert float(p1.abs().max()) > 0
    assert torch.equal(p1, p2)


def test_subvolume_with_zshift_staged_equals_direct(ctx):
    full = A.volume_geometry(96, 96, 40)
    sv = A.subvolume_geometry(24, 1.0, (1.0, 1.0, 1.0), full, (40, 52, 18), shift=(0.0, 0.0, 3.5))
    W = H = 96
    vol = torch.rand((24, 24, 24), generator=torch.Generator(device="cuda").manual_seed(6), device="cuda")
    for tilt in (-42.0, 11.0):
        g = A.tilt_geometry(tilt, W, H, psi_deg=-2.0)

        def run():
            proj = torch.zeros((H, W), device="cuda")
            dist = torch.zeros((H, W), device="cuda")
            ctx.fp_march(sv, g, vol, proj, dist)
            return proj

        p1, p2 = project_both(ctx, run)
        assert float(p1.abs().max()) > 0
        assert torch.equal(p1, p2)


def test_sart_iteration_staged_equals_direct(ctx):
    vg = A.volume_geometry(48, 48, 16)
    W = H = 56
    tilts = [-54.0, -27.0, 0.0, 27.0, 54.0]
    acq = A.Acquisition.build(vg, W, H, tilts, psis_deg=[1.0, -2.0, 0.5, 3.0, -1.0])
    gen = torch.Generator(device="cuda").manual_seed(7)
    projs = torch.rand((len(tilts), H, W), generator=gen, device="cuda")
    vols = []
    for staged in (True, False):
        ctx.set_fp_staging(staged)
        vol = torch.zeros((16, 48, 48), device="cuda")
        ctx.sart_run(vg, acq, projs, vol, A.sart_params(lambda_=0.3, oversampling=2), iterations=2)
        vols.append(vol)
    torch.cuda.synchronize()
    assert float(vols[0].abs().max()) > 0
    assert torch.equal(vols[0], vols[1])


def test_mag_anisotropy_staged_equals_direct(ctx):
    vg = A.volume_geometry(56, 48, 24)
    W, H = 64, 64
    g = A.tilt_geometry(-37.0, W, H, psi_deg=4.0, mag_aniso=(1.04, 25.0))
    vol = torch.rand((24, 48, 56), generator=torch.Generator(device="cuda").manual_seed(8), device="cuda")

    def run():
        proj = torch.zeros((H, W), device="cuda")
        dist = torch.zeros((H, W), device="cuda")
        ctx.fp_march(vg, g, vol, proj, dist)
        return proj

    p1, p2 = project_both(ctx, run)
    assert float(p1.abs().max()) > 0
    assert torch.equal(p1, p2)


@pytest.mark.parametrize("nx,ny,nz,W,H,tilt,psi,shift,roi", CASES)
def test_reference_march_staged_equals_direct(ctx, nx, ny, nz, W, H, tilt, psi, shift, roi):
    """Reference numerics: the LDS-staged march takes the reference kernel's samples, order and float
    arithmetic, so it leaves the same bits as the direct-read kernel."""
    print(f"reference march staged vs direct: {nx}x{ny}x{nz} tilt={tilt} psi={psi} roi={roi}", flush=True)
    vg = A.volume_geometry(nx, ny, nz)
    g = A.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
    gen = torch.Generator(device="cuda").manual_seed(3 * nx + int(tilt))
    vol = torch.rand((nz, ny, nx), generator=gen, device="cuda")

    def run():
        proj = torch.zeros((H, W), device="cuda")
        dist = torch.zeros((H, W), device="cuda")
        ctx.fp_march(vg, g, vol, proj, dist, roi=roi)
        return proj, dist

    ctx.set_numerics("reference")
    try:
        (p1, d1), (p2, d2) = project_both(ctx, run)
    finally:
        ctx.set_numerics("exact")
    assert float(p1.abs().max()) > 0
    assert torch.equal(p1, p2)
    assert torch.equal(d1, d2)


def test_reference_sart_iteration_staged_equals_direct(ctx):
    vg = A.volume_geometry(48, 48, 16)
    sv = A.subvolume_geometry(24, 1.0, (1.0, 1.0, 1.0), A.volume_geometry(96, 96, 40), (40, 52, 18),
                              shift=(0.0, 0.0, 3.5))
    W = H = 56
    tilts = [-54.0, -27.0, 0.0, 27.0, 54.0]
    acq = A.Acquisition.build(vg, W, H, tilts, psis_deg=[1.0, -2.0, 0.5, 3.0, -1.0])
    gen = torch.Generator(device="cuda").manual_seed(11)
    projs = torch.rand((len(tilts), H, W), generator=gen, device="cuda")
    vols = []
    ctx.set_numerics("reference")
    try:
        for staged in (True, False):
            ctx.set_fp_staging(staged)
            vol = torch.zeros((16, 48, 48), device="cuda")
            ctx.sart_run(vg, acq, projs, vol, A.sart_params(lambda_=0.3, oversampling=2), iterations=2)
            vols.append(vol)
        # a z-shifted sub-volume
        g = A.tilt_geometry(-42.0, 96, 96, psi_deg=-2.0)
        svol = torch.rand((24, 24, 24), generator=gen, device="cuda")
        outs = []
        for staged in (True, False):
            ctx.set_fp_staging(staged)
            proj = torch.zeros((96, 96), device="cuda")
            dist = torch.zeros((96, 96), device="cuda")
            ctx.fp_march(sv, g, svol, proj, dist)
            outs.append(proj)
    finally:
        ctx.set_numerics("exact")
        ctx.set_fp_staging(True)
    torch.cuda.synchronize()
    assert float(vols[0].abs().max()) > 0
    assert torch.equal(vols[0], vols[1])
    assert float(outs[0].abs().max()) > 0
    assert torch.equal(outs[0], outs[1])
