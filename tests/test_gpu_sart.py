"""SART (EmSART ordered-subset loop, no CTF) on the GPU through artia_sart_run against the CPU
oracle's SART driver on the same seeded inputs.

Gates: reference numerics vs the float oracle <= 1e-6 relative L2 after 2 iterations;
exact numerics vs the double oracle <= 1e-5 after one iteration and <= 1e-4 after two.
"""
import numpy as np
import pytest
import torch

import artiatomi_amd as A
import oracle_lib as O
from artiatomi_amd import phantom
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def make_case(dims, W, H, tilts, seed, psi_range=0.0, shift_range=0.0, n_ell=12):
    rng = np.random.default_rng(seed)
    psis = rng.uniform(-psi_range, psi_range, len(tilts)) if psi_range else np.zeros(len(tilts))
    shifts = rng.uniform(-shift_range, shift_range, (len(tilts), 2)) if shift_range else np.zeros((len(tilts), 2))
    vg = A.volume_geometry(*dims)
    acq = A.Acquisition.build(vg, W, H, tilts, psis_deg=psis.tolist(), shifts=shifts.tolist())
    shapes = phantom.make_shapes(dims, n_ell, seed, axes=(2.0, 10.0))
    truth = phantom.rasterize(dims, shapes).numpy()
    projs = np.stack([O.march(vg, g, truth.astype(np.float64), W, H, dtype=np.float64)[0] for g in acq.geoms])
    projs = projs + phantom.gaussian_noise(projs.shape, 0.05 * float(projs.mean()), seed + 1)
    return vg, acq, projs.astype(np.float32)


CASES = {
    "plain": dict(dims=(48, 48, 12), W=56, H=56, tilts=phantom.tilt_series(-60, 60, 10), seed=3, os=1,
                  crop={}),
    "psi_shift_os2_border": dict(dims=(40, 44, 12), W=52, H=48, tilts=phantom.tilt_series(-60, 60, 12), seed=5,
                                 os=2, psi_range=5.0, shift_range=3.0,
                                 crop=dict(crop=(4, 4, 4, 4), crop_dim=(3, 3, 3, 3), cut_length=(2, 2),
                                           dim_length=(3, 3))),
    "subset3": dict(dims=(32, 32, 8), W=40, H=40, tilts=phantom.tilt_series(-60, 60, 15), seed=8, os=2,
                    sirt=3, crop={}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_sart_parity(gpu, name):
    c = CASES[name]
    print(f"sart case {name}", flush=True)
    vg, acq, projs = make_case(c["dims"], c["W"], c["H"], c["tilts"], c["seed"], c.get("psi_range", 0.0),
                               c.get("shift_range", 0.0))
    sp = A.sart_params(lambda_=0.3, sirt_count=c.get("sirt", 1), oversampling=c["os"], **c["crop"])
    nx, ny, nz = c["dims"]
    ctx = A.ArtiaContext(0)
    try:
        d_proj = torch.from_numpy(projs).cuda()
        # reference numerics, two iterations, against the float oracle
        ctx.set_numerics("reference")
        vol = torch.zeros((nz, ny, nx), device="cuda")
        _, empty = ctx.sart_run(vg, acq, d_proj, vol, sp, iterations=2)
        ref32, _ = O.sart(vg, vg, acq.geoms, acq.index_list, projs, np.zeros((nz, ny, nx), np.float32), sp, 2,
                          dtype=np.float32)
        torch.cuda.synchronize()
        assert not empty
        e32 = rel_l2(vol.cpu().numpy(), ref32)
        assert e32 <= 1e-6, e32
        # exact numerics against the double oracle, one then two iterations
        ctx.set_numerics("exact")
        vol = torch.zeros((nz, ny, nx), device="cuda")
        _, empty = ctx.sart_run(vg, acq, d_proj, vol, sp, iterations=1)
        ref64, e64flag = O.sart(vg, vg, acq.geoms, acq.index_list, projs, np.zeros((nz, ny, nx)), sp, 1,
                                dtype=np.float64)
        torch.cuda.synchronize()
        e1 = rel_l2(vol.cpu().numpy(), ref64)
        assert e1 <= 1e-5, e1
        ctx.sart_run(vg, acq, d_proj, vol, sp, iterations=1, volume_is_empty=empty)
        ref64b, _ = O.sart(vg, vg, acq.geoms, acq.index_list, projs, ref64, sp, 1, empty=e64flag, dtype=np.float64)
        torch.cuda.synchronize()
        e2 = rel_l2(vol.cpu().numpy(), ref64b)
        assert e2 <= 1e-4, e2
        print(f"  ref-vs-f32 {e32:.2e}  exact-vs-f64 it1 {e1:.2e} it2 {e2:.2e}  "
              f"f32-vs-f64 {rel_l2(ref32, ref64b):.2e}", flush=True)
    finally:
        ctx.close()


def per_launch_sart(ctx, vg, acq, d_proj, vol, sp, S, lam, empty, border):
    """One SART iteration of EmSART's ordered-subset loop (EmSart.cpp:579-675) composed from the
    operator entry points: per subset member the fused residual (artia_fp_residual, or the
    traversal length while the volume is empty), then one artia_bp launch per member."""
    H, W = d_proj.shape[1:]
    T = len(acq.geoms)
    dist = torch.zeros((H, W), device="cuda")
    for step in range((T + S - 1) // S):
        members = acq.index_list[step * S:(step + 1) * S]
        res = []
        for idx in members:
            r = torch.empty((H, W), device="cuda")
            ctx.fp_residual(vg, acq.geoms[idx], None if empty else vol, d_proj[idx], r, acq.real_lengths[idx],
                            crop=border.get("crop", (0, 0, 0, 0)), crop_dim=border.get("crop_dim", (0, 0, 0, 0)),
                            points=acq.hitpoints[idx], cut_length=border.get("cut_length", (0, 0)),
                            dim_length=border.get("dim_length", (0, 0)), dist=dist, volume_is_empty=empty)
            res.append(r)
        for idx, r in zip(members, res):
            ctx.bp(vg, acq.geoms[idx], vol, r, lam, oversampling=sp.oversampling)
        empty = False
    return empty


@pytest.mark.parametrize("sirt,os_", [(3, 2), (5, 1), (1, 2)])
def test_sart_subset_equals_per_launch_loop(gpu, sirt, os_):
    """The driver backprojects up to four tilts of a subset per launch, adding their updates in
    registers; the volume equals one backprojection launch per tilt, bit for bit."""
    border = dict(crop=(2, 2, 2, 2), crop_dim=(3, 3, 3, 3), cut_length=(1, 1), dim_length=(3, 3))
    vg, acq, projs = make_case((40, 36, 12), 48, 44, phantom.tilt_series(-57, 60, 9), 21, psi_range=4.0,
                               shift_range=2.0)
    sp = A.sart_params(lambda_=0.3, sirt_count=sirt, oversampling=os_, **border)
    ctx = A.ArtiaContext(0)
    try:
        d_proj = torch.from_numpy(projs).cuda()
        vol = torch.zeros((12, 36, 40), device="cuda")
        _, empty = ctx.sart_run(vg, acq, d_proj, vol, sp, iterations=2)
        ref = torch.zeros_like(vol)
        lam = float(np.float32(0.3) / np.float32(sirt))   # Lambda / SIRTCount in float
        e = per_launch_sart(ctx, vg, acq, d_proj, ref, sp, sirt, lam, True, border)
        per_launch_sart(ctx, vg, acq, d_proj, ref, sp, sirt, lam, e, border)
        torch.cuda.synchronize()
        assert not empty
        assert float(vol.abs().max()) > 0
        assert torch.equal(vol, ref)
    finally:
        ctx.close()
