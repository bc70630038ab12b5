"""Parity at the benchmark sizes, where the oracle cannot run whole.

SART workload (1024 x 1024 x 256 volume, 1024^2 detector, exact numerics):
  - forward projection at three tilts against the FP64 oracle on 24 x 24-pixel windows (the
    detector's centre, the edge of the volume's shadow, a corner), within 1e-5 relative L2;
  - backprojection (OS 2) against the FP64 oracle on two-slice slabs at the bottom face, the middle
    and the top face of the volume (the oracle runs on the slab alone: a voxel's update depends only
    on its own position and the image).  A sub-ray parallel to a voxel face (every tilt about y has
    a beam direction whose y component is zero or a float residue, 3.7e-9 at 47 deg) that lies
    within rounding of that face is inside or outside the whole voxel column by rounding alone, in
    any arithmetic: the reference's own float evaluation moves such voxels by up to 3x against
    FP64.  Those voxels are found geometrically, in FP64 on the host (tie_voxels: a sub-ray through
    the voxel's cross-section within 1e-6 voxels of a face it runs parallel to, which covers the
    float rounding of a sub-ray's offset, ~5e-7); every other voxel agrees with the oracle within
    1e-5 of the slab's largest update, and within 1e-5 relative L2 over the slab.
CTF workload (2048 x 2048 x 512 volume, 2048^2 detector, 4 nm slabs at 0.25 nm/px): one CTF-SART
pass over two tilts from a non-empty volume, stacked forward slabs and backprojection, equals the
reference's per-slab loop through the operator entry points bit for bit.
Sub-volume workload (4096^2 projections, 4096 x 4096 x 1024 field, 3D CTF): the batched plan equals
the per-call schedule bit for bit on 40 cubes of the target mix over three tilts.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import artiatomi_amd as A
import oracle_lib as O
from conftest import rel_l2
from test_gpu_ctf import per_slab_sart_ctf

pytestmark = pytest.mark.gpu

DIMS, DET = (1024, 1024, 256), 1024


@pytest.fixture(scope="module")
def cfg2(gpu):
    ctx = A.ArtiaContext(0)
    gen = torch.Generator(device="cuda").manual_seed(11)
    nx, ny, nz = DIMS
    vol = torch.rand((nz, ny, nx), generator=gen, device="cuda")
    host = vol.cpu().numpy().astype(np.float64)
    yield ctx, A.volume_geometry(*DIMS), vol, host
    ctx.close()


@pytest.mark.parametrize("tilt,psi,shift", [(0.0, 0.0, (0.0, 0.0)), (37.0, 2.5, (3.0, -2.0)),
                                            (-58.0, -4.0, (-1.5, 4.0))])
def test_march_windows_full_size(cfg2, tilt, psi, shift):
    ctx, vg, vol, host = cfg2
    g = A.tilt_geometry(tilt, DET, DET, psi_deg=psi, shift=shift)
    proj = torch.zeros((DET, DET), device="cuda")
    dist = torch.zeros((DET, DET), device="cuda")
    ctx.fp_march(vg, g, vol, proj, dist)
    p, d = proj.cpu().numpy(), dist.cpu().numpy()
    for x0, y0 in ((500, 500), (16, 500), (1000, 1000)):
        roi = (x0, y0, x0 + 24, y0 + 24)
        print(f"march 1024^3/4 tilt={tilt} window {roi}", flush=True)
        po, do = O.march(vg, g, host, DET, DET, dtype=np.float64, roi=roi)
        sl = np.s_[y0:y0 + 24, x0:x0 + 24]
        if np.abs(po[sl]).max() == 0:
            assert np.abs(p[sl]).max() == 0
            continue
        assert rel_l2(p[sl], po[sl]) <= 1e-5
        assert rel_l2(d[sl], do[sl]) <= 1e-5


def tie_voxels(vg, g, os_, eps):
    """Voxels of vg crossed by a sub-ray of the OS lattice (detector points ((i + 0.5)/OS,
    (j + 0.5)/OS)) that runs parallel to one of the voxel's faces (|n_c| < 1e-6) and lies within
    eps voxels of it, in FP64.  Rays are d + X u + Y v + t n (the tilt's pitch vectors, as the
    backprojector's sub-rays); "crosses" means a positive chord through the voxel's extent along the
    other axes.  Returns a boolean [nz, ny, nx] mask."""
    d, u, v, n = (np.array(a, dtype=np.float64) for a in (g.d, g.u, g.v, g.n))
    n = n / np.linalg.norm(n)
    s = np.array(vg.voxel_size, dtype=np.float64)
    binv = np.linalg.inv(np.stack([u, v, n], axis=1))
    par = [c for c in range(3) if abs(n[c]) < 1e-6]
    nonpar = [c for c in range(3) if c not in par]
    tol = eps + 3.0 * float(s.max()) * max([abs(n[c]) for c in par] + [0.0])   # drift along the ray
    z, y, x = np.meshgrid(np.arange(vg.nz), np.arange(vg.ny), np.arange(vg.nx), indexing="ij")
    corner = np.stack([vg.bbox_min[c] + (x, y, z)[c] * s[c] for c in range(3)], axis=-1) - d
    pu = corner @ binv[0]
    pv = corner @ binv[1]
    # the voxel's projected box relative to its corner's projection, from its 8 corners
    offs = np.array([[(k >> 2) & 1, (k >> 1) & 1, k & 1] for k in range(8)], dtype=np.float64) * s
    du, dv = offs @ binv[0], offs @ binv[1]
    ni = int(np.ceil((du.max() - du.min()) * os_)) + 3
    nj = int(np.ceil((dv.max() - dv.min()) * os_)) + 3
    i0 = np.ceil((pu + du.min()) * os_ - 0.5) - 1
    j0 = np.ceil((pv + dv.min()) * os_ - 0.5) - 1
    mask = np.zeros(pu.shape, dtype=bool)
    for di in range(ni):
        a = (i0 + di + 0.5) / os_ - pu
        for dj in range(nj):
            b = (j0 + dj + 0.5) / os_ - pv
            w = {c: a * u[c] + b * v[c] for c in range(3)}
            tin = np.full(pu.shape, -np.inf)
            tout = np.full(pu.shape, np.inf)
            for c in nonpar:
                t0, t1 = -w[c] / n[c], (s[c] - w[c]) / n[c]
                tin = np.maximum(tin, np.minimum(t0, t1))
                tout = np.minimum(tout, np.maximum(t0, t1))
            crosses = tout > tin
            for c in par:
                near = (np.abs(w[c]) < tol) | (np.abs(w[c] - s[c]) < tol)
                for c2 in par:
                    if c2 != c:
                        near &= (w[c2] > -tol) & (w[c2] < s[c2] + tol)
                mask |= near & crosses
    return mask


@pytest.mark.parametrize("tilt,psi,shift", [(0.0, 0.0, (0.0, 0.0)), (-47.0, 3.0, (2.0, 1.0))])
def test_bp_slabs_full_size(cfg2, tilt, psi, shift):
    ctx, vg, _, _ = cfg2
    nx, ny, nz = DIMS
    g = A.tilt_geometry(tilt, DET, DET, psi_deg=psi, shift=shift)
    img = np.random.default_rng(5).normal(0.0, 1.0, (DET, DET)).astype(np.float32)
    vol = torch.zeros((nz, ny, nx), device="cuda")
    ctx.bp(vg, g, vol, torch.from_numpy(img).cuda(), 0.3, oversampling=2)
    torch.cuda.synchronize()
    for z0 in (0, nz // 2 - 1, nz - 2):
        sub = A.VolGeom.from_buffer_copy(vg)
        sub.bbox_min[2] = vg.bbox_min[2] + z0 * vg.voxel_size[2]
        assert float(np.float32(sub.bbox_min[2])) == vg.bbox_min[2] + z0 * vg.voxel_size[2]   # exact
        sub.nz = 2
        print(f"bp 1024^3/4 tilt={tilt} slab z={z0}", flush=True)
        ref = O.back_project(sub, g, np.zeros((2, ny, nx)), img, 0.3, os_=2, dtype=np.float64)
        got = vol[z0:z0 + 2].cpu().numpy().astype(np.float64)
        assert np.abs(ref).max() > 0
        tie = tie_voxels(sub, g, 2, 1e-6)
        bad = (np.abs(got - ref) > 1e-5 * np.abs(ref).max()) & ~tie
        print(f"  {int(tie.sum())} of {tie.size} voxels have a sub-ray within 1e-6 of a parallel face; "
              f"{int(bad.sum())} others off by > 1e-5 of the max; rel L2 of the others "
              f"{rel_l2(got[~tie], ref[~tie]):.2e}", flush=True)
        assert tie.sum() <= 1e-4 * tie.size
        assert not bad.any(), np.argwhere(bad)[:10]
        assert rel_l2(got[~tie], ref[~tie]) <= 1e-5


def test_sart_ctf_two_tilts_full_size_equals_per_slab(gpu):
    nx, ny, nz = 2048, 2048, 512
    W = H = 2048
    vg = A.volume_geometry(nx, ny, nz)
    tilts = [-36.0, 44.0]
    acq = A.Acquisition.build(vg, W, H, tilts, psis_deg=[1.5, -2.0], shifts=[(2.0, -1.0), (-3.0, 0.5)])
    ctf = A.ctf_params(pixelsize_nm=0.25, slice_thickness_nm=4.0, amp_contrast=0.07)
    tdef = [A.TiltDefocus(3500.0, 3700.0, 30.0, tilts[0]), A.TiltDefocus(4200.0, 4400.0, -65.0, tilts[1])]
    sp = A.sart_params(lambda_=0.3, oversampling=2, cut_length=(2, 2), dim_length=(4, 4))
    ctx = A.ArtiaContext(0)
    try:
        gen = torch.Generator(device="cuda").manual_seed(21)
        start = torch.rand((nz, ny, nx), generator=gen, device="cuda")
        d_proj = torch.randn((2, H, W), generator=gen, device="cuda") * 50.0
        mid = start[nz // 2].clone()
        print("ctf 2048^2x512: stacked driver", flush=True)
        vol = start.clone()
        _, empty = ctx.sart_ctf_run(vg, acq, tdef, d_proj, vol, sp, ctf, iterations=1, volume_is_empty=False)
        print("ctf 2048^2x512: per-slab loop", flush=True)
        ref = start
        per_slab_sart_ctf(ctx, vg, acq, tdef, d_proj, ref, sp, ctf, iterations=1, volume_is_empty=False)
        assert not empty
        assert not torch.equal(vol[nz // 2], mid)   # the two tilts changed the volume
        assert torch.equal(vol, ref)
    finally:
        ctx.close()


def test_subvol_ctf_4096_batched_equals_plain(gpu):
    """Sub-volume WBP at the target detector size: 40 cubes of 32/64/96/128 voxels (4:3:2:1) in a
    4096 x 4096 x 1024 field, three 4096^2 projections, per-(cube, tilt) shifts, 3D CTF with 16-px
    slabs; the batched plan equals the reference's per-call schedule bit for bit."""
    rng = np.random.default_rng(17)
    W = H = 4096
    full = A.volume_geometry(4096, 4096, 1024)
    tilts = [-50.0, 0.0, 38.0]
    sizes = [32] * 16 + [64] * 12 + [96] * 8 + [128] * 4
    subs = []
    for n in sizes:
        pos = rng.uniform([n, n, n / 2], [4096 - n, 4096 - n, 1024 - n / 2])
        g = A.subvolume_geometry(n, 1.0, (1.0, 1.0, 1.0), full, pos.tolist())
        subs.append((g, torch.rand((n, n, n), device="cuda")))
    extra = rng.uniform(-4.0, 4.0, (len(sizes), len(tilts), 2))
    geoms = [[A.subvolume_tilt_geometry(t, W, H, extra_shift=extra[s, i]) for i, t in enumerate(tilts)]
             for s in range(len(sizes))]
    images = torch.randn((len(tilts), H, W), device="cuda")
    ctf = A.ctf_params(pixelsize_nm=0.25, slice_thickness_nm=4.0, amp_contrast=0.07)
    base = [A.tilt_geometry(t, W, H) for t in tilts]
    kw = dict(ctf=ctf, slab_ranges=[A.defocus_distances(full, g) for g in base],
              defocus=[A.TiltDefocus(3000.0 + 500 * i, 3200.0 + 500 * i, 20.0 * i, t) for i, t in enumerate(tilts)])
    ctx = A.ArtiaContext(0)
    try:
        plain = [(g, v.clone()) for g, v in subs]
        print("subvol 4096^2 ctf: plain", flush=True)
        A.subvol_wbp_plain(ctx, plain, geoms, images, 2, **kw)
        print("subvol 4096^2 ctf: batched plan", flush=True)
        batched = [(g, v.clone()) for g, v in subs]
        plan = A.SubvolPlan(ctx, batched, geoms, images, 2, **kw)
        plan.execute()
        torch.cuda.synchronize()
        plan.close()
        changed = 0
        for (_, a), (_, b), (_, v0) in zip(batched, plain, subs):
            assert torch.equal(a, b)
            changed += int(not torch.equal(a, v0))
        assert changed == len(sizes)
    finally:
        ctx.close()
