"""GPU parity of the Fourier-space path through the C ABI: the CTF multiply / Wiener inverse /
phase flip and the WBP weighting against the oracle kernels, the fused slab and WBP filters
against numpy-FFT compositions of oracle kernels, and the SART loop with 3D-CTF correction
against the oracle's CTF SART composition.

Gates (relative L2): exact numerics vs the double oracle <= 1e-6 for the spectrum multiplies,
<= 1e-5 for the FFT filters (single-precision FFT) and for one CTF SART iteration, <= 1e-4 after
two; reference numerics vs the float oracle <= 1e-5 (the FFT libraries round differently).
"""
import numpy as np
import pytest
import torch

import artiatomi_amd as A
import ctf_compose as CC
import oracle_lib as O
from artiatomi_amd import phantom
from conftest import rel_l2

pytestmark = pytest.mark.gpu


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


def spectrum(N, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((N, N // 2 + 1)) + 1j * rng.standard_normal((N, N // 2 + 1))).astype(np.complex64)


CTF_CASES = [
    # N, for_fp, phase_flip, amp, dmin, dmax, angle, beta, wiener
    (64, True, False, 0.0, 3000.0, 3000.0, 0.0, (2.0, 0.0, 0.008, 0.0), 0.1),
    (96, True, False, 0.07, 2800.0, 3300.0, 33.0, (3.0, 0.5, 0.01, 0.001), 0.1),
    (64, False, False, 0.0, 4100.0, 4300.0, -20.0, (2.0, 0.0, 0.008, 0.0), 0.1),
    (80, False, False, 0.07, 3000.0, 3200.0, 71.0, (0.0, 0.0, 0.008, 0.0), 0.02),
    (64, True, True, 0.0, 3500.0, 3600.0, 12.0, (2.0, 0.0, 0.008, 0.0), 0.1),
]


@pytest.mark.parametrize("case", range(len(CTF_CASES)))
def test_ctf_apply_parity(ctx, case):
    N, fp, pf, amp, d0, d1, ang, beta, w = CTF_CASES[case]
    print(f"ctf case {case}", flush=True)
    p = A.ctf_params(pixelsize_nm=0.25, amp_contrast=amp, wiener=w, beta_fac=beta, phase_flip_only=pf)
    k = O.ctf_consts(2.7, 300.0, 0.25, N, amp)
    X = spectrum(N, case)
    for numerics, dtype, tol in (("reference", np.float32, 1e-5), ("exact", np.float64, 1e-6)):
        ctx.set_numerics(numerics)
        d = torch.from_numpy(X).cuda()
        ctx.ctf_apply(d, p, d0, d1, ang, for_fp=fp)
        expect = O.ctf_apply(X, k, d0, d1, ang, for_fp=fp, phase_flip=pf, wiener=w, beta=beta, dtype=dtype)
        got = host(d)
        assert rel_l2(got.real, expect.real) <= tol and rel_l2(got.imag, expect.imag) <= tol, numerics
    ctx.set_numerics("exact")


@pytest.mark.parametrize("method", ["ramp", "exact", "contrast2", "contrast10", "contrast30"])
def test_wbp_weight_parity(ctx, method):
    m = A._abi.WBP_METHODS[method]
    for N, psi in ((64, 0.0), (96, 37.5), (80, -84.0)):
        X = spectrum(N, N)
        for numerics, dtype in (("reference", np.float32), ("exact", np.float64)):
            ctx.set_numerics(numerics)
            d = torch.from_numpy(X).cuda()
            ctx.wbp_weight(d, psi, method)
            expect = O.wbp_weight(X, psi, m, dtype=dtype)
            got = host(d)
            # the reference's __sinf/__cosf of the angle are evaluated as sinf/cosf here
            assert rel_l2(got.real, expect.real) <= 1e-6 and rel_l2(got.imag, expect.imag) <= 1e-6, (numerics, N)
    ctx.set_numerics("exact")


@pytest.mark.parametrize("for_fp,accumulate", [(True, True), (False, False)])
def test_ctf_slab_filter_parity(ctx, for_fp, accumulate):
    W, H = 64, 48
    rng = np.random.default_rng(7)
    img = rng.random((H, W)).astype(np.float32)
    base = rng.random((H, W)).astype(np.float32)
    vg = A.volume_geometry(48, 40, 12)
    g = A.tilt_geometry(30.0, W, H, psi_deg=2.0, shift=(1.0, -2.0))
    corners = A.hitpoints(vg, g) if not for_fp else [0, H - 1, W - 1, H - 1, 0, 0, W - 1, 0]
    cut, dim = (2.0, 3.0), (4.0, 2.0)
    p = A.ctf_params(pixelsize_nm=0.5, beta_fac=(2.0, 0.0, 0.008, 0.0), amp_contrast=0.07)
    k = O.ctf_consts(2.7, 300.0, 0.5, max(W, H), 0.07)
    for numerics, dtype, tol in (("reference", np.float32, 1e-5), ("exact", np.float64, 1e-5)):
        ctx.set_numerics(numerics)
        d_img, d_out = dev(img), dev(base)
        ctx.ctf_slab_filter(d_img, d_out, corners, p, 3000.0, 3150.0, 20.0, for_fp=for_fp, accumulate=accumulate,
                            cut_length=cut, dim_length=dim)
        a, f = CC.slab_filter(img.astype(dtype), corners, k, 3000.0, 3150.0, 20.0, for_fp, cut, dim,
                              beta=tuple(p.beta_fac), dtype=dtype)
        expect = base.astype(np.float64) + f if accumulate else f
        assert rel_l2(host(d_img), a) <= 1e-6, numerics
        assert rel_l2(host(d_out), expect) <= tol, numerics
    ctx.set_numerics("exact")


@pytest.mark.parametrize("method", ["ramp", "exact", "contrast10"])
def test_wbp_filter_parity(ctx, method):
    W, H = 64, 40
    img = np.random.default_rng(9).random((H, W)).astype(np.float32)
    m = A._abi.WBP_METHODS[method]
    for numerics, dtype in (("reference", np.float32), ("exact", np.float64)):
        ctx.set_numerics(numerics)
        out = dev(np.zeros((H, W)))
        ctx.wbp_filter(dev(img), out, -3.5, method)
        expect = CC.wbp_filter(img.astype(dtype), -3.5, m, dtype=dtype)
        assert rel_l2(host(out), expect) <= 1e-5, numerics
    ctx.set_numerics("exact")


def test_ctf_filter_rejects_bad_arguments(ctx):
    img = dev(np.zeros((8, 8)))
    p = A.ctf_params()
    bad = A.ctf_params(voltage_kV=0.0)
    with pytest.raises(A.ArtiaError):
        ctx.ctf_slab_filter(img, img, [0, 7, 7, 7, 0, 0, 7, 0], bad, 3000, 3000, 0)
    with pytest.raises(A.ArtiaError):   # filter method 9 does not exist
        ctx._call("artia_wbp_filter", img.data_ptr(), 8, 8, 8, img.data_ptr(), 8, 0.0, 9)
    ctx.ctf_slab_filter(img, img, [0, 7, 7, 7, 0, 0, 7, 0], p, 3000, 3000, 0)


def ctf_case(seed=4):
    dims, W, H = (40, 40, 12), 48, 48
    tilts = phantom.tilt_series(-60, 60, 20)
    rng = np.random.default_rng(seed)
    psis = rng.uniform(-3, 3, len(tilts)).tolist()
    shifts = rng.uniform(-2, 2, (len(tilts), 2)).tolist()
    vg = A.volume_geometry(*dims)
    acq = A.Acquisition.build(vg, W, H, tilts, psis_deg=psis, shifts=shifts)
    truth = phantom.rasterize(dims, phantom.make_shapes(dims, 10, seed, axes=(2.0, 9.0))).numpy()
    projs = np.stack([O.march(vg, g, truth.astype(np.float64), W, H, dtype=np.float64)[0] for g in acq.geoms])
    projs = (projs + phantom.gaussian_noise(projs.shape, 0.05 * float(projs.mean()), seed + 1)).astype(np.float32)
    defocus = [(float(d), float(d) + 150.0, float(a), float(t))
               for d, a, t in zip(rng.uniform(2500, 4000, len(tilts)), rng.uniform(-90, 90, len(tilts)), tilts)]
    # 3.1 nm / 0.5 nm = 6.2 px per slab: not a multiple of the 0.15 px march step, so no sample
    # lands on a slab boundary (where rounding alone decides which slab takes it)
    ctf = A.ctf_params(pixelsize_nm=0.5, slice_thickness_nm=3.1, beta_fac=(2.0, 0.0, 0.008, 0.0),
                       amp_contrast=0.07, volume_shift_z=2.0)
    return dims, vg, acq, projs, defocus, ctf


def test_sart_ctf_parity(gpu):
    dims, vg, acq, projs, defocus, ctf = ctf_case()
    nx, ny, nz = dims
    sp = A.sart_params(lambda_=0.3, oversampling=2, cut_length=(1, 1), dim_length=(2, 2))
    tdef = [A.TiltDefocus(*d) for d in defocus]
    args = (vg, vg, acq.geoms, acq.hitpoints, acq.real_lengths, acq.slab_ranges, defocus, acq.index_list, projs)
    ctx = A.ArtiaContext(0)
    try:
        d_proj = torch.from_numpy(projs).cuda()
        for numerics, dtype, gates in (("exact", np.float64, ((1, 1e-5), (2, 1e-4))),
                                       ("reference", np.float32, ((2, 1e-5),))):
            ctx.set_numerics(numerics)
            for iters, tol in gates:
                print(f"sart ctf {numerics} iterations={iters}", flush=True)
                vol = torch.zeros((nz, ny, nx), device="cuda")
                _, empty = ctx.sart_ctf_run(vg, acq, tdef, d_proj, vol, sp, ctf, iterations=iters)
                ref, ref_empty = CC.sart_ctf(*args, np.zeros((nz, ny, nx)), sp, ctf, iters, dtype=dtype)
                err = rel_l2(host(vol), ref)
                print(f"  rel_l2 = {err:.3e}", flush=True)
                assert not empty and not ref_empty
                assert err <= tol, (numerics, iters, err)
    finally:
        ctx.close()


def per_slab_sart_ctf(ctx, vg, acq, tdef, d_proj, vol, sp, ctf, iterations, volume_is_empty=True):
    """The reference's per-slab CTF-SART loop (Reconstructor.cpp:496-645 forward, :920-969 backward)
    driven through the per-operator entry points, S = 1: per tilt, slab by slab, fp_slicer ->
    ctf_slab_filter (accumulated), compare, then per slab ctf_slab_filter -> gated bp.  vol (float32
    or float16) is updated in place."""
    f32 = np.float32
    H, W = d_proj.shape[1:]
    step = f32(ctf.slice_thickness_nm) / f32(ctf.pixelsize_nm)
    frame = (0, H - 1, W - 1, H - 1, 0, 0, W - 1, 0)
    proj = torch.zeros((H, W), device="cuda")
    dist = torch.zeros((H, W), device="cuda")
    cut, dim = tuple(sp.cut_length), tuple(sp.dim_length)

    def slabs(idx):
        ray = f32(acq.slab_ranges[idx][0])
        while ray < f32(acq.slab_ranges[idx][1]):
            yield ray, f32(ray + step)
            ray = f32(ray + step)

    empty = volume_is_empty
    for _ in range(iterations):
        for idx in acq.index_list:
            g, d = acq.geoms[idx], tdef[idx]
            proj.zero_()
            dist.zero_()
            if not empty:
                for a, b in slabs(idx):
                    dmin, dmax = A.defocus_minmax(float(a), d, ctf)
                    dist.zero_()
                    ctx.fp_slicer(vg, g, vol, dist, float(a), float(b))
                    ctx.ctf_slab_filter(dist, proj, frame, ctf, dmin, dmax, d.astig_angle_deg, for_fp=True,
                                        accumulate=True, cut_length=cut, dim_length=dim)
                dist.zero_()
            ctx.vol_traversal(vg, g, dist)
            ctx.compare(d_proj[idx], proj, dist, acq.real_lengths[idx], tuple(sp.crop), tuple(sp.crop_dim),
                        sp.proj_scale)
            empty = False
            for a, b in slabs(idx):
                dmin, dmax = A.defocus_minmax(float(a), d, ctf)
                ctx.ctf_slab_filter(proj, dist, acq.hitpoints[idx], ctf, dmin, dmax, d.astig_angle_deg,
                                    for_fp=False, accumulate=False, cut_length=cut, dim_length=dim)
                ctx.bp(vg, g, vol, dist, float(f32(sp.lambda_)), oversampling=sp.oversampling, dist_min=float(a),
                       dist_max=float(b), filter="linear")
    torch.cuda.synchronize()


def test_sart_ctf_equals_per_slab_composition(gpu):
    """artia_sart_ctf_run backprojects a tilt's CTF slabs in one launch over a stack of the filtered
    slab images; driving the reference's per-slab loop (Reconstructor.cpp:920-969) through the
    per-operator entry points (gated artia_bp per slab) gives the same volume bit for bit."""
    dims, vg, acq, projs, defocus, ctf = ctf_case()
    nx, ny, nz = dims
    sp = A.sart_params(lambda_=0.3, oversampling=2, cut_length=(1, 1), dim_length=(2, 2))
    tdef = [A.TiltDefocus(*d) for d in defocus]
    ctx = A.ArtiaContext(0)
    try:
        d_proj = torch.from_numpy(projs).cuda()
        vol = torch.zeros((nz, ny, nx), device="cuda")
        ctx.sart_ctf_run(vg, acq, tdef, d_proj, vol, sp, ctf, iterations=2)
        ref = torch.zeros_like(vol)
        per_slab_sart_ctf(ctx, vg, acq, tdef, d_proj, ref, sp, ctf, iterations=2)
        assert float(ref.abs().max()) > 0
        assert torch.equal(vol, ref)
    finally:
        ctx.close()


def test_sart_ctf_same_under_stack_limits(gpu):
    """Within a smaller stack limit the driver takes its other schedules: the forward slabs in
    several volume passes, slab by slab, and without room for the slab images one launch per slab.
    Every schedule gives the same volume bit for bit."""
    dims, vg, acq, projs, defocus, ctf = ctf_case(seed=7)
    nx, ny, nz = dims
    H, W = projs.shape[1:]
    npix = H * W
    sp = A.sart_params(lambda_=0.3, oversampling=2, cut_length=(1, 1), dim_length=(2, 2))
    tdef = [A.TiltDefocus(*d) for d in defocus]
    step = np.float32(ctf.slice_thickness_nm) / np.float32(ctf.pixelsize_nm)
    max_slabs = max(len(CC.slab_rays(a, b, step)) for a, b in acq.slab_ranges)
    stack = max_slabs * npix * 4
    limits = {
        "default": 0,
        "three passes or more": stack + 32 * npix + 32 * npix * max(1, max_slabs // 3),
        "slab by slab": stack + 64 * npix - 1,
        "no stack": stack - 1,
    }
    ctx = A.ArtiaContext(0)
    try:
        d_proj = torch.from_numpy(projs).cuda()
        vols = {}
        for name, lim in limits.items():
            ctx.set_stack_limit(lim)
            vol = torch.zeros((nz, ny, nx), device="cuda")
            ctx.sart_ctf_run(vg, acq, tdef, d_proj, vol, sp, ctf, iterations=2)
            torch.cuda.synchronize()
            vols[name] = vol
            print(f"stack limit {name} ({lim} bytes, {max_slabs} slabs): max |vol| {float(vol.abs().max()):.4g}",
                  flush=True)
        for name in limits:
            assert torch.equal(vols[name], vols["default"]), name
    finally:
        ctx.close()


@pytest.mark.parametrize("numerics", ["exact", "reference"])
def test_per_tilt_ctf_steps_equal_per_slab_loop(gpu, numerics):
    """artia_fp_ctf and artia_bp_ctf (Reconstructor::ForwardProjectionCTF / BackProjectionCTF for one
    tilt, Reconstructor.cpp:496-644, 920-969) against the reference's per-slab loop driven through
    the operator entry points: projection, chord image, volume update and the residual left cropped
    once per slab agree bit for bit.  Exact numerics take the stacked forms (one volume pass, batched
    filters, one backprojection launch), reference numerics the per-slab launches."""
    dims, vg, acq, projs, defocus, ctf = ctf_case(seed=5)
    nx, ny, nz = dims
    H, W = projs.shape[1:]
    f32 = np.float32
    step = f32(ctf.slice_thickness_nm) / f32(ctf.pixelsize_nm)
    frame = (0, H - 1, W - 1, H - 1, 0, 0, W - 1, 0)
    cut, dim = (1.0, 1.0), (2.0, 2.0)
    tdef = [A.TiltDefocus(*d) for d in defocus]
    gen = torch.Generator(device="cuda").manual_seed(11)
    vol0 = torch.rand((nz, ny, nx), generator=gen, device="cuda")
    ctx = A.ArtiaContext(0, numerics=numerics)
    try:
        for idx in acq.index_list[:4]:
            print(f"per-tilt ctf {numerics} tilt {idx}", flush=True)
            g, d, sr = acq.geoms[idx], tdef[idx], acq.slab_ranges[idx]
            rays = CC.slab_rays(sr[0], sr[1], step)
            assert len(rays) > 1
            proj, dist = torch.zeros((H, W), device="cuda"), torch.full((H, W), 7.0, device="cuda")
            ctx.fp_ctf(vg, g, vol0, sr, d, ctf, proj, dist, cut_length=cut, dim_length=dim)
            proj_r, dist_r = torch.zeros((H, W), device="cuda"), torch.zeros((H, W), device="cuda")
            for a in rays:
                dmin, dmax = A.defocus_minmax(float(a), d, ctf)
                dist_r.zero_()
                ctx.fp_slicer(vg, g, vol0, dist_r, float(a), float(f32(a + step)))
                ctx.ctf_slab_filter(dist_r, proj_r, frame, ctf, dmin, dmax, d.astig_angle_deg, for_fp=True,
                                    accumulate=True, cut_length=cut, dim_length=dim)
            dist_r.zero_()
            ctx.vol_traversal(vg, g, dist_r)
            torch.cuda.synchronize()
            assert float(proj.abs().max()) > 0
            assert torch.equal(proj, proj_r)
            assert torch.equal(dist, dist_r)
            # empty volume: only the chord image
            dist_e = torch.zeros((H, W), device="cuda")
            proj_e = torch.zeros((H, W), device="cuda")
            ctx.fp_ctf(vg, g, None, sr, d, ctf, proj_e, dist_e, volume_is_empty=True)
            torch.cuda.synchronize()
            assert torch.equal(dist_e, dist_r) and float(proj_e.abs().max()) == 0.0

            res = torch.randn((H, W), generator=gen, device="cuda")
            res_r = res.clone()
            vol_a, vol_b = vol0.clone(), vol0.clone()
            lam = float(f32(0.3))
            ctx.bp_ctf(vg, g, vol_a, res, acq.hitpoints[idx], sr, d, ctf, lam, oversampling=2, cut_length=cut,
                       dim_length=dim)
            tmp = torch.zeros((H, W), device="cuda")
            for a in rays:
                dmin, dmax = A.defocus_minmax(float(a), d, ctf)
                ctx.ctf_slab_filter(res_r, tmp, acq.hitpoints[idx], ctf, dmin, dmax, d.astig_angle_deg,
                                    for_fp=False, accumulate=False, cut_length=cut, dim_length=dim)
                ctx.bp(vg, g, vol_b, tmp, lam, oversampling=2, dist_min=float(a), dist_max=float(f32(a + step)),
                       filter="linear")
            torch.cuda.synchronize()
            assert not torch.equal(vol_a, vol0)
            assert torch.equal(vol_a, vol_b)
            assert torch.equal(res, res_r)
    finally:
        ctx.close()
