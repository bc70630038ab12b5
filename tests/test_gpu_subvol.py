"""GPU parity of the sub-volume WBP (EmSARTSubVolumes' loop) through the C ABI.

plain form: against the oracle's composition of lambda-1 backprojections (EmSartSubVolumes.cpp:
  611-707, Reconstructor.cpp:894-1034) - exact numerics <= 1e-5 relative L2 against FP64,
  reference numerics <= 1e-6 against the float oracle.
batched form: bit-identical to the plain form in exact numerics, for the autotuned and for fixed
  tunings, with and without CTF, including repeated graph replays; with CTF also for stack budgets
  that split the tilts into chunks of about two tilts and of one tilt.
"""
import numpy as np
import pytest
import torch

import artiatomi_amd as A
import ctf_compose as CC
from artiatomi_amd import phantom
from conftest import rel_l2

pytestmark = pytest.mark.gpu

SIZES = [24, 16, 32, 12, 24, 16, 32, 8]


def make_case(seed=21, ctf=False):
    rng = np.random.default_rng(seed)
    W = H = 128
    full = A.volume_geometry(128, 128, 32)
    tilts = phantom.tilt_series(-60, 60, 15)
    psis = rng.uniform(-3, 3, len(tilts))
    shifts = rng.uniform(-2, 2, (len(tilts), 2))
    subs_geo, subs_vol = [], []
    for n in SIZES:
        pos = rng.uniform([n, n, n / 2], [128 - n, 128 - n, 32 - n / 2]) + 1.0
        subs_geo.append(A.subvolume_geometry(n, 1.0, (1.0, 1.0, 1.0), full, pos.tolist()))
        subs_vol.append(rng.random((n, n, n)).astype(np.float32))
    extra = rng.uniform(-2.2, 2.2, (len(SIZES), len(tilts), 2))
    active = np.hypot(extra[..., 0], extra[..., 1]) < 2.5 - 0.5      # |shift| < MaxShift - 0.5
    geoms = [[A.subvolume_tilt_geometry(t, W, H, psi_deg=p, shift=sh, extra_shift=extra[s, i])
              for i, (t, p, sh) in enumerate(zip(tilts, psis, shifts))] for s in range(len(SIZES))]
    images = rng.random((len(tilts), H, W)).astype(np.float32)
    out = dict(full=full, subs_geo=subs_geo, subs_vol=subs_vol, geoms=geoms, active=active, images=images)
    if ctf:
        out["ctf"] = A.ctf_params(pixelsize_nm=0.5, slice_thickness_nm=3.1, beta_fac=(2.0, 0.0, 0.008, 0.0),
                                  amp_contrast=0.07)
        base = [A.tilt_geometry(t, W, H, psi_deg=p, shift=sh) for t, p, sh in zip(tilts, psis, shifts)]
        out["slab_ranges"] = [A.defocus_distances(full, g) for g in base]
        out["defocus"] = [(float(d), float(d) + 200.0, float(a), float(t))
                          for d, a, t in zip(rng.uniform(2500, 4500, len(tilts)), rng.uniform(-90, 90, len(tilts)),
                                             tilts)]
    return out


def device_subs(c):
    return [(g, torch.from_numpy(v.copy()).cuda()) for g, v in zip(c["subs_geo"], c["subs_vol"])]


def ctf_kwargs(c):
    if "ctf" not in c:
        return {}
    return dict(ctf=c["ctf"], slab_ranges=c["slab_ranges"], defocus=[A.TiltDefocus(*d) for d in c["defocus"]])


def run_plain(ctx, c, os_):
    subs = device_subs(c)
    A.subvol_wbp_plain(ctx, subs, c["geoms"], torch.from_numpy(c["images"]).cuda(), os_, active=c["active"],
                       **ctf_kwargs(c))
    torch.cuda.synchronize()
    return [v.cpu().numpy() for _, v in subs]


def oracle(c, os_, dtype):
    subs = [(g, v.astype(dtype)) for g, v in zip(c["subs_geo"], c["subs_vol"])]
    ctf = c.get("ctf")
    return CC.subvol_wbp(subs, c["geoms"], c["images"], os_, active=c["active"], ctf=ctf,
                         slab_ranges=c.get("slab_ranges"), defocus=c.get("defocus"), dtype=dtype)


def all_rel(got, ref):
    return rel_l2(np.concatenate([g.ravel() for g in got]), np.concatenate([r.ravel() for r in ref]))


@pytest.mark.parametrize("ctf,os_", [(False, 1), (False, 2), (True, 2)])
def test_subvol_plain_parity(gpu, ctf, os_):
    c = make_case(ctf=ctf)
    ctx = A.ArtiaContext(0)
    try:
        for numerics, dtype, tol in (("exact", np.float64, 1e-5), ("reference", np.float32, 1e-6)):
            print(f"subvol plain ctf={ctf} os={os_} {numerics}", flush=True)
            ctx.set_numerics(numerics)
            got = run_plain(ctx, c, os_)
            err = all_rel(got, oracle(c, os_, dtype))
            print(f"  rel_l2 = {err:.3e}", flush=True)
            assert err <= tol, (numerics, err)
    finally:
        ctx.close()


TUNINGS = [None, (0, 0, 1, 2), (0, 1, 2, 4), (0, 0, 2, 8), (0, 1, 1, 1)]


@pytest.mark.parametrize("ctf", [False, True])
def test_subvol_batched_equals_plain(gpu, ctf):
    c = make_case(seed=5, ctf=ctf)
    os_ = 2
    ctx = A.ArtiaContext(0)
    try:
        images = torch.from_numpy(c["images"]).cuda()
        plain1 = run_plain(ctx, c, os_)
        # plain applied twice: the reference's accumulate semantics across two runs
        subs2 = [(g, torch.from_numpy(v).cuda()) for g, v in zip(c["subs_geo"], plain1)]
        A.subvol_wbp_plain(ctx, subs2, c["geoms"], images, os_, active=c["active"], **ctf_kwargs(c))
        torch.cuda.synchronize()
        plain2 = [v.cpu().numpy() for _, v in subs2]
        # (tuning, stack budget): with CTF, budgets of 1 MiB (16 slab images of 128^2, about two
        # tilts per chunk) and 1 byte (one tilt per chunk) besides the default
        runs = [(tu, 0) for tu in TUNINGS]
        if ctf:
            runs += [((0, 1, 2, 4), 1 << 20), ((0, 0, 1, 2), 1)]
        for tu, limit in runs:
            print(f"subvol batched ctf={ctf} tuning={tu} stack limit={limit}", flush=True)
            ctx.set_stack_limit(limit)
            subs = device_subs(c)
            tuning = A.SubvolTuning(*tu) if tu is not None else None
            plan = A.SubvolPlan(ctx, subs, c["geoms"], images, os_, active=c["active"], tuning=tuning,
                                **ctf_kwargs(c))
            t = plan.tuning
            print(f"  tuning autotune={t.autotune} policy={t.tile_policy} qpt={t.quads_per_thread} "
                  f"bpc={t.blocks_per_cu}", flush=True)
            # the autotuner must leave the sub-volumes untouched
            torch.cuda.synchronize()
            for (_, v), v0 in zip(subs, c["subs_vol"]):
                assert np.array_equal(v.cpu().numpy(), v0)
            plan.execute()
            torch.cuda.synchronize()
            got1 = [v.cpu().numpy() for _, v in subs]
            plan.execute()
            torch.cuda.synchronize()
            got2 = [v.cpu().numpy() for _, v in subs]
            plan.close()
            for a, b in zip(got1, plain1):
                assert np.array_equal(a, b), (tu, float(np.abs(a - b).max()))
            for a, b in zip(got2, plain2):
                assert np.array_equal(a, b), (tu, "second replay", float(np.abs(a - b).max()))
    finally:
        ctx.close()


def test_subvol_plan_rejects_reference_numerics(gpu):
    c = make_case()
    ctx = A.ArtiaContext(0, numerics="reference")
    try:
        with pytest.raises(A.ArtiaError):
            A.SubvolPlan(ctx, device_subs(c), c["geoms"], torch.from_numpy(c["images"]).cuda(), 1)
    finally:
        ctx.close()
