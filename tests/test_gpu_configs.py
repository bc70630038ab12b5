"""The two benchmark configurations small enough to check whole against the oracle.

Plain SART, 256 x 256 x 64 (the reference's CPU-runnable case): a seeded phantom of 50 ellipsoids
(semi-axes 4-40 voxels, density U[0.1, 1]); 41 tilts -60..60 step 3 deg of 256 x 256, made by the
FP64 forward projector plus Gaussian noise sigma = 0.05 * mean; oversampling 1, ordered subset
size 1, lambda 0.3, two iterations, no CTF, no border crop (EmSart.cpp:579-675):
  - reference numerics against the float oracle after two iterations, <= 1e-6 relative L2;
  - exact numerics against the FP64 oracle, <= 1e-5 after one iteration and <= 1e-4 after two.

Adjoint / gradient form, 512 x 512 x 128 field, 41 tilts -60..60 step 3 deg of 512 x 512, 500
zero-margin 64^3 sub-volumes (Gaussian blobs below 1e-6 at the faces) at seeded positions, one
(sub-volume, tilt) pair each, detector shifts U[-3, 3] px:
  - dot-product test of the whole-volume projector and adjoint over the 41 tilts, <= 1e-6;
  - dot-product test of the pair-set projector and adjoint with per-pair shifts, <= 1e-6;
  - the fused launch's shift gradient against FP64 central differences of the oracle's loss on
    sampled pairs: h = 1e-6 px within 1e-5, h = 1e-3 px within 1e-3 (the wider step straddles
    the trilinear kinks of the discretised projector, a few 1e-4 of truncation);
  - Levenberg-Marquardt on the fused launch recovers all 20 500 shifts within 1e-3 px.
"""
import numpy as np
import pytest
import torch

import artiatomi_amd as A
import oracle_lib as O
from artiatomi_amd import autograd as AG
from artiatomi_amd import phantom
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_sart_256x256x64_41tilts_two_iterations(gpu):
    dims, W, H = (256, 256, 64), 256, 256
    nx, ny, nz = dims
    tilts = phantom.tilt_series(-60, 60, 3)
    assert len(tilts) == 41
    vg = A.volume_geometry(*dims)
    acq = A.Acquisition.build(vg, W, H, tilts)
    truth = phantom.phantom(dims, 50, seed=1).numpy().astype(np.float64)
    print("sart 256x256x64: FP64 projections", flush=True)
    projs = np.stack([O.march(vg, g, truth, W, H, dtype=np.float64)[0] for g in acq.geoms])
    projs = (projs + phantom.gaussian_noise(projs.shape, 0.05 * float(projs.mean()), 2)).astype(np.float32)
    sp = A.sart_params(lambda_=0.3, sirt_count=1, oversampling=1)
    ctx = A.ArtiaContext(0)
    try:
        d_proj = torch.from_numpy(projs).cuda()
        print("sart 256x256x64: reference numerics, two iterations, float oracle", flush=True)
        ctx.set_numerics("reference")
        vol = torch.zeros((nz, ny, nx), device="cuda")
        ctx.sart_run(vg, acq, d_proj, vol, sp, iterations=2)
        ref32, _ = O.sart(vg, vg, acq.geoms, acq.index_list, projs, np.zeros((nz, ny, nx), np.float32), sp, 2,
                          dtype=np.float32)
        e32 = rel_l2(vol.cpu().numpy(), ref32)
        print(f"  reference numerics vs float oracle after 2 iterations: {e32:.2e}", flush=True)
        assert e32 <= 1e-6
        print("sart 256x256x64: exact numerics, FP64 oracle", flush=True)
        ctx.set_numerics("exact")
        vol = torch.zeros((nz, ny, nx), device="cuda")
        _, empty = ctx.sart_run(vg, acq, d_proj, vol, sp, iterations=1)
        ref1, empty64 = O.sart(vg, vg, acq.geoms, acq.index_list, projs, np.zeros((nz, ny, nx)), sp, 1,
                               dtype=np.float64)
        e1 = rel_l2(vol.cpu().numpy(), ref1)
        print(f"  exact numerics vs FP64 oracle after 1 iteration: {e1:.2e}", flush=True)
        assert e1 <= 1e-5
        ctx.sart_run(vg, acq, d_proj, vol, sp, iterations=1, volume_is_empty=empty)
        ref2, _ = O.sart(vg, vg, acq.geoms, acq.index_list, projs, ref1, sp, 1, empty=empty64, dtype=np.float64)
        e2 = rel_l2(vol.cpu().numpy(), ref2)
        print(f"  exact numerics vs FP64 oracle after 2 iterations: {e2:.2e}; float vs FP64 oracle "
              f"{rel_l2(ref32, ref2):.2e}", flush=True)
        assert e2 <= 1e-4
        # the reconstruction moved towards the phantom
        assert rel_l2(ref2, truth) < rel_l2(ref1, truth) < 1.0
    finally:
        ctx.close()


FIELD, DET, EDGE, NSUB = (512, 512, 128), 512, 64, 500


@pytest.fixture(scope="module")
def pairs512(gpu):
    rng = np.random.default_rng(9)
    FX, FY, FZ = FIELD
    tilts = phantom.tilt_series(-60, 60, 3)
    full = A.volume_geometry(*FIELD)
    ctx = A.ArtiaContext(0)
    vols = phantom.blob_subvolumes(EDGE, NSUB, 9)
    base = [A.tilt_geometry(t, DET, DET) for t in tilts]
    subs, pairs, geoms = [], [], []
    for k in range(NSUB):
        pos = rng.uniform([EDGE / 2] * 3, [FX - EDGE / 2, FY - EDGE / 2, FZ - EDGE / 2]) + 1.0
        vg = A.subvolume_geometry(EDGE, 1.0, (1.0, 1.0, 1.0), full, pos.tolist())
        subs.append((vg, k * EDGE ** 3))
        for g in base:
            x0, y0, w, h = A.footprint_window(vg, g, DET, DET, 8)
            pairs.append((k, x0, y0, w, h))
            geoms.append(g)
    ps = A.PairSet(ctx, subs, pairs, geoms)
    s_true = rng.uniform(-3.0, 3.0, (ps.npairs, 2))
    yield dict(ctx=ctx, ps=ps, vols=vols, subs=subs, pairs=pairs, geoms=geoms, base=base, full=full,
               s_true=s_true, rng=rng)
    ps.close()
    ctx.close()


def test_dot_product_512x512x128_41tilts(pairs512):
    c = pairs512
    ctx, full = c["ctx"], c["full"]
    FX, FY, FZ = FIELD
    gen = torch.Generator(device="cuda").manual_seed(31)
    x = torch.randn((FZ, FY, FX), generator=gen, device="cuda")
    proj = torch.empty((DET, DET), device="cuda")
    dist = torch.empty((DET, DET), device="cuda")
    pty = torch.zeros_like(x)
    lhs = 0.0
    for g in c["base"]:
        y = torch.randn((DET, DET), generator=gen, device="cuda")
        proj.zero_()
        ctx.fp_march(full, g, x, proj, dist)
        ctx.fp_adjoint(full, g, y, pty)    # accumulates P_t^T y_t over the tilts
        lhs += float((proj.double() * y.double()).sum())
    rhs = float((x.double() * pty.double()).sum())
    err = abs(lhs - rhs) / abs(lhs)
    print(f"dot test 512x512x128, 41 tilts: <Px,y> {lhs:.9e}, <x,P^T y> {rhs:.9e}, rel {err:.2e}", flush=True)
    assert err <= 1e-6


def test_pair_dot_product_with_shifts(pairs512):
    c = pairs512
    ps = c["ps"]
    gen = torch.Generator(device="cuda").manual_seed(32)
    s = torch.from_numpy(c["s_true"]).cuda()
    x = torch.randn(c["vols"].shape, generator=gen, device="cuda")
    y = torch.randn((ps.n_pixels,), generator=gen, device="cuda")
    px = ps.project(x, s)
    pty = torch.zeros_like(x)
    ps.adjoint(s, y, pty)
    lhs = float((px.double() * y.double()).sum())
    rhs = float((x.double() * pty.double()).sum())
    err = abs(lhs - rhs) / abs(lhs)
    print(f"pair-set dot test, {ps.npairs} pairs with shifts: rel {err:.2e}", flush=True)
    assert err <= 1e-6


def test_shift_gradient_vs_fp64_central_differences(pairs512):
    c = pairs512
    ps = c["ps"]
    rng = np.random.default_rng(33)
    s_true = torch.from_numpy(c["s_true"]).cuda()
    targets = ps.project(c["vols"], s_true)
    s = s_true + torch.from_numpy(rng.uniform(-0.7, 0.7, (ps.npairs, 2))).cuda()
    _, grad = ps.loss_grad(c["vols"], s.contiguous(), targets)
    torch.cuda.synchronize()
    gh, sh, th = grad.cpu().numpy(), s.cpu().numpy(), targets.cpu().numpy().astype(np.float64)
    T = len(c["base"])
    sample = [0, 20, 40, 41 * 137 + 5, 41 * 311 + 33, 41 * 499 + 12]
    worst = {1e-6: 0.0, 1e-3: 0.0}
    for p in sample:
        k, x0, y0, w, h, off = ps.windows[p]
        vg, voff = c["subs"][k]
        v64 = c["vols"][voff:voff + EDGE ** 3].view(EDGE, EDGE, EDGE).cpu().numpy().astype(np.float64)
        b = th[off:off + w * h].reshape(h, w)
        roi = (x0, y0, x0 + w, y0 + h)
        for hstep in (1e-6, 1e-3):
            fd = []
            for ex, ey in ((1, 0), (0, 1)):
                ls = []
                for sg in (1, -1):
                    P = O.march_shifted(vg, c["geoms"][p], v64, DET, DET, sh[p, 0] + sg * hstep * ex,
                                        sh[p, 1] + sg * hstep * ey, roi=roi)
                    ls.append(((P[y0:y0 + h, x0:x0 + w] - b) ** 2).sum())
                fd.append((ls[0] - ls[1]) / (2 * hstep))
            fd = np.array(fd)
            e = np.abs(gh[p] - fd).max() / np.abs(fd).max()
            print(f"pair {p} (sub-volume {k}, tilt {p % T}): grad {gh[p]}, FP64 FD h={hstep:g} {fd}, rel {e:.2e}",
                  flush=True)
            worst[hstep] = max(worst[hstep], e)
    assert worst[1e-6] <= 1e-5
    assert worst[1e-3] <= 1e-3


def test_refinement_recovers_shifts_512(pairs512):
    c = pairs512
    ps = c["ps"]
    s_true = torch.from_numpy(c["s_true"]).cuda()
    targets = ps.project(c["vols"], s_true)
    s = AG.refine_pair_shifts(c["vols"], targets, ps, steps=12, method="gn")
    err = (s - s_true).abs().max().item()
    print(f"Levenberg-Marquardt, 12 evaluations, {ps.npairs} pairs: max shift error {err:.2e} px", flush=True)
    assert err <= 1e-3
