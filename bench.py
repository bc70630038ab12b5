#!/usr/bin/env python3
"""SART throughput benchmark on one MI355X.

Workload (default): plain SART, fp32 volume 1024x1024x256 of 400 seeded ellipsoids + 2000 gold
beads (r = 3 voxels, density 5); 61 tilts -60..60 step 2 deg, 1024x1024 projections, tilt-axis
rotation psi ~ U[-5, 5] deg, per-tilt shifts ~ U[-5, 5] px, Gaussian noise at SNR 0.5,
oversampling 2, ordered-subset size 1, lambda 0.3.  A step is one SART iteration (61 x
[forward projection, compare, cropBorder, backprojection]) with all data resident in HBM.

Output: one JSON line.  value = voxel updates (V * T per iteration) per second, in G.
roofline: the dominant kernel's algorithmic bytes per launch / its average launch time (HIP events
on the driver's stream, inside the timed region) against MI355X's 8 TB/s HBM peak.
cpu_baseline: the float CPU oracle (reference operation order, OpenMP over the host cores) on a
bounded sample of the same workload, extrapolated to one iteration.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: dims, W, H, tilt step, n ellipsoids, n beads, psi range, shift range, OS, seed
    "sart_1024x1024x256_61tilts": dict(dims=(1024, 1024, 256), W=1024, H=1024, step=2.0, n_ell=400, n_beads=2000,
                                       psi=5.0, shift=5.0, os=2, seed=2, noise_snr=0.5),
    "sart_256x256x64_41tilts": dict(dims=(256, 256, 64), W=256, H=256, step=3.0, n_ell=50, n_beads=0, psi=0.0,
                                    shift=0.0, os=1, seed=1, noise_rel=0.05),
}


# SART with 3D-CTF slab correction: 2048 x 2048 x 512 ellipsoid phantom, 61 tilts of 2048^2 at
# 0.25 nm/px, 300 kV, Cs 2.7 mm, amplitude contrast 0.07, defocus U[3, 5] um per tilt with 0.2 um
# astigmatism at a random angle, slabs of 16 z-slices (4 nm), oversampling 2.  Projections come from
# the forward projector with Poisson noise at 60 e/A^2 over the series (the timing does not depend
# on the data).
CTF_WORKLOADS = {
    "sart_ctf_2048x2048x512_61tilts": dict(dims=(2048, 2048, 512), W=2048, H=2048, step=2.0, n_ell=400, os=2,
                                           seed=3, pixelsize_nm=0.25, slab_nm=4.0, amp_contrast=0.07,
                                           dose_e_per_A2=60.0),
}


# Sub-volume WBP workloads (EmSARTSubVolumes): cubes of 32/64/96/128 voxels in ratio 4:3:2:1 with
# centres uniform in the field, per-(cube, tilt) extra shifts U[-4, 4] px, seeded random WBP-filtered
# projections; the batched plan against the per-call plain schedule on the same inputs.
SUBVOL_WORKLOADS = {
    "subvol_wbp_2000cubes_4096_61tilts": dict(field=(4096, 4096, 1024), W=4096, H=4096, cubes=2000, step=2.0,
                                              os=2, seed=5, ctf=False),
    "subvol_wbp_ctf_200cubes_1024_61tilts": dict(field=(1024, 1024, 256), W=1024, H=1024, cubes=200, step=2.0,
                                                 os=2, seed=6, ctf=True),
    # the full-size case with 3D CTF and exact-weighted projections; the plain schedule (a launch
    # per sub-volume, tilt and slab) is timed on every 10th tilt and scaled by the slab count
    "subvol_wbp_ctf_2000cubes_4096_61tilts": dict(field=(4096, 4096, 1024), W=4096, H=4096, cubes=2000, step=2.0,
                                                  os=2, seed=7, ctf=True, weighting="exact", plain_every=10),
}


# Gradient-refinement workload (EmSARTRefine's per-(sub-volume, tilt) shift search): 64^3
# sub-volumes with zero margins in a 512 x 512 x 128 field, 41 tilts of 512 x 512, true shifts
# U[-3, 3] px per (sub-volume, tilt); 50 fused loss-and-gradient evaluations against the
# integer cross-correlation search over the maxShift disk.
REFINE_WORKLOADS = {
    "refine_500subvols_64_512x512x128_41tilts": dict(name="refine_500subvols_64_512x512x128_41tilts",
                                                     field=(512, 512, 128), W=512, H=512, n=64, cubes=500,
                                                     step=3.0, true_shift=3.0, max_shift=5, margin=8, evals=50,
                                                     seed=9),
}


# Projection preparation workloads (PrepareProjection for every projection of a stack): 16-bit
# counts with a gradient, noise and hot pixels; EmSART's default band-pass (LP 800, LPS 200 spectrum
# pixels), bad-pixel correction at 10x the mean, mean normalisation, dose weighting.
PREP_WORKLOADS = {
    "prep_61x4096x4096_uint16": dict(n=61, W=4096, H=4096, seed=12, pixelsize_nm=0.25),
}


def bench_prep(args) -> int:
    import numpy as np
    import torch

    import artiatomi_amd as A

    w = PREP_WORKLOADS[args.workload]
    n, W, H = w["n"], w["W"], w["H"]
    N = max(W, H)
    gen = torch.Generator(device="cuda").manual_seed(w["seed"])
    yy = torch.arange(H, device="cuda", dtype=torch.float32)[:, None]
    xx = torch.arange(W, device="cuda", dtype=torch.float32)[None, :]
    base = 800.0 + 0.05 * xx - 0.03 * yy
    imgs = torch.empty((n, H, W), device="cuda", dtype=torch.uint16)
    for i in range(n):
        a = base + 25.0 * torch.randn((H, W), generator=gen, device="cuda")
        hot = torch.randint(0, H * W, (50,), generator=gen, device="cuda")
        a.view(-1)[hot] = 60000.0
        imgs[i] = a.clamp(0, 65535).round().to(torch.int32).to(torch.uint16)
    out = torch.empty((n, H, W), device="cuda")
    doses = np.linspace(1.0, 60.0, n).astype(np.float32)
    pp = A.prep_params(dtype="uint16", normalization="mean", dose_weighting=True, pixelsize_nm=w["pixelsize_nm"])
    ctx = A.ArtiaContext(0)
    for _ in range(args.warmup):
        ctx.prepare_projections(pp, imgs, out, dose=doses)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mean, std, bad = ctx.prepare_projections(pp, imgs, out, dose=doses)
    torch.cuda.synchronize()
    elapsed = (time.perf_counter() - t0) / args.steps
    # per projection: 2 B in and 4 B out per pixel, and the square through an R2C/C2R pair
    # (24 N^2 in the one-pass-per-FFT model)
    alg = n * (6.0 * W * H + 24.0 * N * N)
    achieved = alg / elapsed / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle composition (float64 FFT, OpenMP kernels) on one projection, scaled to the stack
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as O
        import prep_compose as P

        cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        O.set_threads(cores)
        one = imgs[n // 2].cpu().numpy()
        t1 = time.perf_counter()
        P.prepare(one, normalization="mean", dose_weighting=True, pixelsize_nm=w["pixelsize_nm"],
                  dose=float(doses[n // 2]))
        t_one = time.perf_counter() - t1
        cpu = {"value": 1.0 / t_one, "unit": "projections/s", "cores": O.load().oracle_get_threads(),
               "kind": "port", "sample": f"one {W}x{H} projection through the oracle composition, scaled to the stack"}
    out_line = {
        "metric": "projections/s prepared (PrepareProjection)",
        "value": n / elapsed,
        "unit": "projections/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (16-bit counts with a gradient, Gaussian noise and 50 hot pixels per projection)",
        "config": {"workload": args.workload, "projections": n, "projection": [W, H], "input": "uint16",
                   "band_pass": [800, 200, 0, 0], "bad_pixel_value": 10, "normalization": "mean",
                   "dose_weighting": True, "parallelism": "single GPU"},
        "roofline": {"kernel": "preparation pipeline (per stack)", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "bad_pixels_mean": float(np.mean(bad)),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out_line), flush=True)
    ctx.close()
    return 0


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary of `workload`
    (FETCH_SIZE and WRITE_SIZE passes, calibrated against a copy kernel), or None when the summary
    has no row for it.  The counters need their own profiler runs, so this is the last measurement,
    not one taken in this process."""
    path = Path(__file__).resolve().parent / "profiles" / f"{workload}.md"
    try:
        text = path.read_text()
    except OSError:
        return None
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        name = cells[0].strip("`") if cells else ""
        # an exact name, or a kernel name without template arguments matching any instance
        if len(cells) == 7 and (name == kernel or ("<" not in kernel and name.startswith(kernel + "<"))):
            try:
                return {"bytes": float(cells[6]) * 1e9, "read_bytes": float(cells[4]) * 1e9,
                        "written_bytes": float(cells[5]) * 1e9, "source": f"profiles/{workload}.md"}
            except ValueError:
                return None
    return None


def measured_valu(workload: str, kernel: str):
    """VALU counters of `kernel` from the committed rocprofv3 summary of `workload` (its VALU table:
    instructions per work unit, wave-level issue rate per SIMD per cycle, share of wave time waiting
    on an operand), or None; like measured_traffic, the last profiler run's figures."""
    path = Path(__file__).resolve().parent / "profiles" / f"{workload}.md"
    try:
        text = path.read_text()
    except OSError:
        return None
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        name = cells[0].strip("`") if cells else ""
        if len(cells) == 11 and (name == kernel or ("<" not in kernel and name.startswith(kernel + "<"))):
            try:
                return {"valu_per_unit": cells[4], "issue_rate": float(cells[6]),
                        "dependency_wait": cells[7], "source": f"profiles/{workload}.md"}
            except ValueError:
                return None
    return None


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_workload(name: str, ctx):
    import numpy as np
    import torch

    import artiatomi_amd as A
    from artiatomi_amd import phantom

    w = WORKLOADS[name]
    nx, ny, nz = w["dims"]
    W, H = w["W"], w["H"]
    tilts = phantom.tilt_series(-60.0, 60.0, w["step"])
    rng = np.random.default_rng(w["seed"])
    psis = rng.uniform(-w["psi"], w["psi"], len(tilts)) if w["psi"] else np.zeros(len(tilts))
    shifts = rng.uniform(-w["shift"], w["shift"], (len(tilts), 2)) if w["shift"] else np.zeros((len(tilts), 2))
    vg = A.volume_geometry(nx, ny, nz)
    acq = A.Acquisition.build(vg, W, H, tilts, psis_deg=psis.tolist(), shifts=shifts.tolist())
    t0 = time.time()
    shapes = phantom.make_shapes(w["dims"], w["n_ell"], w["seed"], n_beads=w["n_beads"])
    # rasterised on the host so the GPU sees only the reconstruction kernels (short profiles)
    truth = phantom.rasterize(w["dims"], shapes, device="cpu").cuda()
    projs = torch.zeros((len(tilts), H, W), device="cuda")
    dist = torch.zeros((H, W), device="cuda")
    for i, g in enumerate(acq.geoms):
        ctx.fp_march(vg, g, truth, projs[i], dist)
    torch.cuda.synchronize()
    gen = torch.Generator(device="cuda").manual_seed(w["seed"])
    if "noise_snr" in w:
        sigma = float(torch.sqrt(projs.var() / w["noise_snr"]))
    else:
        sigma = w["noise_rel"] * float(projs.mean())
    projs += torch.randn(projs.shape, generator=gen, device="cuda") * sigma
    torch.cuda.synchronize()
    log(f"workload {name}: phantom + {len(tilts)} projections in {time.time() - t0:.1f} s (noise sigma {sigma:.3g})")
    del truth
    return vg, acq, projs, w


def fp_samples_per_launch(ctx, vg, acq, vol):
    """Approximate trilinear samples per forward projection (4 sub-rays x (chord / 0.15 + 0.5)),
    averaged over the tilt series, from the mean chord image of each tilt."""
    import torch

    H, W = acq.height, acq.width
    total = 0.0
    dist = torch.zeros((H, W), device="cuda")
    proj = torch.zeros((H, W), device="cuda")
    for g in acq.geoms[:: max(1, len(acq.geoms) // 8)]:
        ctx.vol_traversal(vg, g, dist.zero_())
        hit = dist > 0
        total += float((4 * (dist[hit] / (0.15 * vg.voxel_size[0]) + 0.5)).sum())
    del proj
    return total / len(acq.geoms[:: max(1, len(acq.geoms) // 8)])


def cpu_baseline(vg, acq, vol_dev, projs_dev, w):
    """Float CPU oracle on a bounded sample: one tilt's forward projection over a band of rows and
    its backprojection over a slab of z slices, scaled to the full image / volume and to all
    tilts (compare and cropBorder run on the full image)."""
    import ctypes as C

    import numpy as np

    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O

    lib = O.load()
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    O.set_threads(cores)
    cores = lib.oracle_get_threads()
    nx, ny, nz = w["dims"]
    W, H = acq.width, acq.height
    ti = len(acq.geoms) // 4   # a tilt of about -30 deg
    g = acq.geoms[ti]
    vol = vol_dev.detach().float().cpu().numpy()
    rows = H if W >= 1024 else 64   # one whole projection at 1024^2: about 6 s
    y0 = H // 2 - rows // 2
    t0 = time.perf_counter()
    proj, dist = O.march(vg, g, vol, W, H, dtype=np.float32, roi=(0, y0, W, y0 + rows))
    t_fp = (time.perf_counter() - t0) * (H / rows)
    real = projs_dev[ti].detach().cpu().numpy()
    dist_full = np.ones((H, W), np.float32) * 2.0
    t0 = time.perf_counter()
    res = O.compare(real, np.zeros((H, W), np.float32), dist_full, acq.real_lengths[ti])
    res = O.crop_border(res, acq.hitpoints[ti])
    t_cmp = time.perf_counter() - t0
    zs = 128 if nz >= 128 else 16   # half the slices at 1024^2 x 256: about 5.5 s
    z0 = nz // 2
    slab = type(vg)()
    C.memmove(C.byref(slab), C.byref(vg), C.sizeof(vg))
    slab.bbox_min[2] = vg.bbox_min[2] + z0 * vg.voxel_size[2]
    slab.bbox_max[2] = slab.bbox_min[2] + zs * vg.voxel_size[2]
    slab.dim[2] = float(zs)
    slab.nz = zs
    t0 = time.perf_counter()
    O.back_project(slab, g, vol[z0:z0 + zs], res, 0.3, os_=w["os"], dtype=np.float32)
    t_bp = (time.perf_counter() - t0) * (nz / zs)
    t_iter = (t_fp + t_cmp + t_bp) * len(acq.geoms)
    V = nx * ny * nz
    return {
        "value": V * len(acq.geoms) / t_iter / 1e9,
        "unit": "Gvoxel-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"float oracle, tilt {acq.tilts_deg[ti]:+.0f} deg: forward projection of {rows} of {H} rows "
                   f"({t_fp * rows / H:.2f} s), backprojection of {zs} of {nz} z-slices "
                   f"({t_bp * zs / nz:.2f} s), compare+cropBorder full image ({t_cmp:.3f} s); extrapolated "
                   f"x{H // rows} rows, x{nz // zs} slices, x{len(acq.geoms)} tilts -> {t_iter:.0f} s/iteration"),
    }


def bench_ctf(args) -> int:
    import numpy as np
    import torch

    import artiatomi_amd as A
    from artiatomi_amd import phantom

    w = CTF_WORKLOADS[args.workload]
    nx, ny, nz = w["dims"]
    W, H = w["W"], w["H"]
    tilts = phantom.tilt_series(-60.0, 60.0, w["step"])
    T = len(tilts)
    rng = np.random.default_rng(w["seed"])
    ctx = A.ArtiaContext(0)
    vg = A.volume_geometry(nx, ny, nz)
    acq = A.Acquisition.build(vg, W, H, tilts)
    t0 = time.time()
    shapes = phantom.make_shapes(w["dims"], w["n_ell"], w["seed"])
    truth = phantom.rasterize(w["dims"], shapes, device="cuda")
    projs = torch.zeros((T, H, W), device="cuda")
    dist = torch.zeros((H, W), device="cuda")
    for i, g in enumerate(acq.geoms):
        ctx.fp_march(vg, g, truth, projs[i], dist)
    del truth
    # Poisson counts at the dose per pixel and tilt, attenuated by the projection, back to -log
    dose = w["dose_e_per_A2"] / T * (10.0 * w["pixelsize_nm"]) ** 2
    gen = torch.Generator(device="cuda").manual_seed(w["seed"])
    scale = 1.0 / max(float(projs.mean()), 1e-30)
    counts = torch.poisson(dose * torch.exp(-projs * scale), generator=gen)
    projs = -torch.log(counts.clamp_min(0.5) / dose) / scale
    del counts
    torch.cuda.synchronize()
    dz = rng.uniform(3000.0, 5000.0, T)
    ang = rng.uniform(-90.0, 90.0, T)
    defocus = [A.TiltDefocus(float(d), float(d) + 200.0, float(a), float(t)) for d, a, t in zip(dz, ang, tilts)]
    ctf = A.ctf_params(pixelsize_nm=w["pixelsize_nm"], slice_thickness_nm=w["slab_nm"],
                       amp_contrast=w["amp_contrast"])
    sp = A.sart_params(lambda_=0.3, sirt_count=1, oversampling=w["os"])
    log(f"workload {args.workload}: phantom + {T} noisy projections in {time.time() - t0:.1f} s")
    vol = torch.zeros((nz, ny, nx), device="cuda")
    empty = True
    for i in range(args.warmup):
        _, empty = ctx.sart_ctf_run(vg, acq, defocus, projs, vol, sp, ctf, iterations=1, volume_is_empty=empty)
        torch.cuda.synchronize()
        log(f"warmup iteration {i + 1}/{args.warmup} done")
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, empty = ctx.sart_ctf_run(vg, acq, defocus, projs, vol, sp, ctf, iterations=1, volume_is_empty=empty)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    log(f"timed {args.steps} iterations in {elapsed:.3f} s; per-kernel totals {prof}")
    V = nx * ny * nz
    per_kernel = {name: {"total_ms_per_iteration": ms / args.steps, "launches_per_iteration": n / args.steps,
                         "share": ms / (elapsed * 1e3)} for name, (ms, n) in prof.items() if n}
    # per iteration, every tilt reads the volume once across its slab passes (FP) and reads and
    # writes every voxel once (BP)
    alg = {"fp": (4.0 * V + 8.0 * W * H) * T, "bp": (8.0 * V + 4.0 * W * H) * T}
    dom = max(("fp", "bp"), key=lambda k: prof[k][0])
    achieved = alg[dom] / (prof[dom][0] / args.steps / 1e3) / 1e9
    out = {
        "metric": "Gvoxel-updates/s per CTF-SART iteration",
        "value": V * T * args.steps / elapsed / 1e9,
        "unit": "Gvoxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, forward-projected, Poisson noise at the stated dose)",
        "config": {"workload": args.workload, "volume": list(w["dims"]), "projection": [W, H], "tilts": T,
                   "pixelsize_nm": w["pixelsize_nm"], "ctf_slab_nm": w["slab_nm"], "oversampling": w["os"],
                   "amp_contrast": w["amp_contrast"], "numerics": "exact", "parallelism": "single GPU"},
        "roofline": {"kernel": "march (slicer passes)" if dom == "fp" else "backProjection (slab passes)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "kernels": per_kernel,
    }
    dom_kernel = "slicer_segments_lds_kernel<float>" if dom == "fp" else "bp_stack_kernel<2, float>"
    measured = measured_traffic(args.workload, dom_kernel)
    if measured:
        out["roofline"]["traffic"] = measured.pop("bytes")
        out["roofline"]["traffic_detail"] = measured
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def bench_subvol(args) -> int:
    print(json.dumps(run_subvol(args.workload, args.steps, plain=not args.no_plain)), flush=True)
    return 0


def run_subvol(name: str, steps: int, plain: bool = True) -> dict:
    """The batched sub-volume plan against the per-call plain schedule on one workload; returns the
    benchmark line (value = steady-state batched Gvoxel-updates/s).  plain=False runs the batched
    plan alone (for profiling it)."""
    import numpy as np
    import torch

    import artiatomi_amd as A
    from artiatomi_amd import phantom

    w = SUBVOL_WORKLOADS[name]
    rng = np.random.default_rng(w["seed"])
    FX, FY, FZ = w["field"]
    W, H = w["W"], w["H"]
    tilts = phantom.tilt_series(-60.0, 60.0, w["step"])
    T = len(tilts)
    full = A.volume_geometry(FX, FY, FZ)
    n_classes = [32, 64, 96, 128]
    counts = [w["cubes"] * r // 10 for r in (4, 3, 2, 1)]
    sizes = [n for n, c in zip(n_classes, counts) for _ in range(c)]
    rng.shuffle(sizes)
    t0 = time.time()
    subs_a, subs_b, geoms = [], [], []
    for n in sizes:
        pos = rng.uniform([n / 2, n / 2, n / 2], [FX - n / 2, FY - n / 2, FZ - n / 2]) + 1.0
        g = A.subvolume_geometry(n, 1.0, (1.0, 1.0, 1.0), full, pos.tolist())
        subs_a.append((g, torch.zeros((n, n, n), device="cuda")))
        subs_b.append((g, torch.zeros((n, n, n), device="cuda")))
        extra = rng.uniform(-4.0, 4.0, (T, 2))
        geoms.append([A.subvolume_tilt_geometry(t, W, H, extra_shift=extra[i]) for i, t in enumerate(tilts)])
    gen = torch.Generator(device="cuda").manual_seed(w["seed"])
    images = torch.rand((T, H, W), generator=gen, device="cuda")
    ctx = A.ArtiaContext(0)
    if w.get("weighting"):
        # WBP weighting of every projection (the same input to both schedules)
        weighted = torch.empty_like(images)
        for t in range(T):
            ctx.wbp_filter(images[t], weighted[t], 0.0, method=w["weighting"])
        images = weighted
        del weighted
    kw = {}
    if w["ctf"]:
        ctf = A.ctf_params(pixelsize_nm=0.25, slice_thickness_nm=4.0, amp_contrast=0.07)
        base = [A.tilt_geometry(t, W, H) for t in tilts]
        dz = rng.uniform(2000.0, 5000.0, T)
        kw = dict(ctf=ctf, slab_ranges=[A.defocus_distances(full, g) for g in base],
                  defocus=[A.TiltDefocus(float(d), float(d) + 200.0, float(a), float(t))
                           for d, a, t in zip(dz, rng.uniform(-90, 90, T), tilts)])
    torch.cuda.synchronize()
    vox = sum(n ** 3 for n in sizes)
    log(f"workload {name}: {len(sizes)} cubes ({vox / 1e6:.0f} M voxels), {T} tilts of {W}x{H}, "
        f"inputs in {time.time() - t0:.1f} s")
    every = w.get("plain_every", 1)
    plain_info = {}
    t_plain, equal = None, None
    if plain and every > 1:
        # plain on a tilt sample, one call per tilt; a plan over the same tilts for the bitwise check
        sample = list(range(0, T, every))

        def pick(seq):
            return [seq[t] for t in sample]

        kws = dict(ctf=kw["ctf"], slab_ranges=pick(kw["slab_ranges"]), defocus=pick(kw["defocus"])) if kw else {}
        img_s = images[sample].contiguous()
        plan_s = A.SubvolPlan(ctx, subs_b, [pick(g) for g in geoms], img_s, w["os"], **kws)
        plan_s.execute()
        plan_s.close()
        torch.cuda.synchronize()
        t_sample = 0.0
        for t in sample:
            kwt = dict(ctf=kw["ctf"], slab_ranges=[kw["slab_ranges"][t]], defocus=[kw["defocus"][t]]) if kw else {}
            t0 = time.perf_counter()
            A.subvol_wbp_plain(ctx, subs_a, [[g[t]] for g in geoms], images[t:t + 1], w["os"], **kwt)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            t_sample += dt
            log(f"plain tilt {t}: {dt:.3f} s")
        equal = all(torch.equal(a, b) for (_, a), (_, b) in zip(subs_a, subs_b))
        del img_s, subs_a

        def nslab(t):
            if not kw:
                return 1
            step = np.float32(kw["ctf"].slice_thickness_nm) / np.float32(kw["ctf"].pixelsize_nm)
            a, b = (np.float32(x) for x in kw["slab_ranges"][t])
            n = 0
            while a < b:
                n += 1
                a = np.float32(a + step)
            return n

        counts_all = [nslab(t) for t in range(T)]
        scale = sum(counts_all) / sum(counts_all[t] for t in sample)
        t_plain = t_sample * scale
        plain_info = {"extrapolated": True, "sampled_tilts": sample, "sampled_seconds": t_sample,
                      "scale": scale, "scaled_by": "slab count" if kw else "tilt count"}
        log(f"plain: {t_sample:.3f} s on {len(sample)} tilts, {t_plain:.1f} s extrapolated (x{scale:.2f})")
        subs_b = [(g, torch.zeros_like(v)) for g, v in subs_b]
    t0 = time.time()
    plan = A.SubvolPlan(ctx, subs_b, geoms, images, w["os"], **kw)
    tu = plan.tuning
    log(f"plan + autotune in {time.time() - t0:.1f} s: policy {tu.tile_policy}, quads/thread "
        f"{tu.quads_per_thread}, blocks/CU {tu.blocks_per_cu}")
    torch.cuda.synchronize()
    if plain and every == 1:
        t0 = time.perf_counter()
        A.subvol_wbp_plain(ctx, subs_a, geoms, images, w["os"], **kw)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        log(f"plain: {t_plain:.3f} s")
    t0 = time.perf_counter()
    plan.execute()
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    log(f"batched: {t_batched:.3f} s")
    if plain and every == 1:
        equal = all(torch.equal(a, b) for (_, a), (_, b) in zip(subs_a, subs_b))
    # further replays accumulate on top; timed for the steady-state rate
    reps = max(1, steps)
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.execute()
    torch.cuda.synchronize()
    t_rep = (time.perf_counter() - t0) / reps
    plan.close()
    updates = vox * T
    out = {
        "metric": "Gvoxel-updates/s per sub-volume WBP pass",
        "value": updates / t_rep / 1e9,
        "unit": "Gvoxel-updates/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": 1,
        "ms_per_step": t_rep * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded cube placements and shifts, random projections)",
        "config": {"workload": name, "field": list(w["field"]), "projection": [W, H], "tilts": T,
                   "cubes": len(sizes), "cube_sizes": n_classes, "cube_counts": counts, "oversampling": w["os"],
                   "ctf": w["ctf"], "weighting": w.get("weighting"), "numerics": "exact",
                   "parallelism": "single GPU"},
        "plain": ({"seconds": t_plain, "Gvoxel_updates_per_s": updates / t_plain / 1e9, **plain_info}
                  if t_plain is not None else None),
        "batched_first": {"seconds": t_batched},
        "speedup_vs_plain": t_plain / t_rep if t_plain is not None else None,
        "bitwise_equal_to_plain": bool(equal) if equal is not None else None,
        "tuning": {"tile_policy": tu.tile_policy, "quads_per_thread": tu.quads_per_thread,
                   "blocks_per_cu": tu.blocks_per_cu, "autotuned": bool(tu.autotune)},
    }
    del subs_b, images
    ctx.close()
    return out


def plain_gradient_sample(ctx, ps, vols, subs, targets, shifts, geoms, W, H, npairs):
    """One gradient evaluation as per-pair calls (artia_fp_march + residual + artia_fp_shift_grad
    on the pair's window), the schedule without pair sets, timed over the first npairs pairs."""
    import torch

    proj = torch.zeros((H, W), device="cuda")
    dist = torch.zeros((H, W), device="cuda")
    gimg = torch.zeros((H, W), device="cuda")
    out = torch.zeros((npairs, 2), device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in range(npairs):
        k, x0, y0, w, h, off = ps.windows[p]
        vg, voff = subs[k]
        vol = vols[voff:voff + vg.nx * vg.ny * vg.nz]
        roi = (x0, y0, x0 + w, y0 + h)
        g = geoms[p]
        proj[y0:y0 + h, x0:x0 + w].zero_()
        ctx.fp_march(vg, g, vol, proj, dist, roi=roi)
        gimg[y0:y0 + h, x0:x0 + w] = 2.0 * (proj[y0:y0 + h, x0:x0 + w] - targets[off:off + w * h].view(h, w))
        ctx.fp_shift_grad(vg, g, vol, gimg, out[p], roi=roi)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def plain_cc_sample(ps, models, targets, square, max_shift, npairs):
    """GetDisplacement's per-pair schedule (Reconstructor.cpp:1746-1959: square images, means,
    R2C of both, conj multiply, C2R, maxShift mask, max index) with torch.fft on the GPU, timed over
    the first npairs pairs: the cost of the reference's full-square search on this GPU."""
    import torch

    N = square
    idx = torch.arange(N, device="cuda", dtype=torch.float32)
    m = torch.where(idx > N // 2, N - idx, idx)
    mask = torch.sqrt(m[None, :] ** 2 + m[:, None] ** 2) <= max_shift
    peaks = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in range(npairs):
        _, x0, y0, w, h, off = ps.windows[p]
        a = torch.zeros((N, N), device="cuda")
        b = torch.zeros((N, N), device="cuda")
        a[y0:y0 + h, x0:x0 + w] = targets[off:off + w * h].view(h, w)
        b[y0:y0 + h, x0:x0 + w] = models[off:off + w * h].view(h, w)
        a -= a.mean()
        b -= b.mean()
        den = N * N * torch.sqrt((a * a).sum() * (b * b).sum())
        cc = torch.fft.irfft2(torch.conj(torch.fft.rfft2(a)) * torch.fft.rfft2(b), s=(N, N)) * (N * N) / den
        cc = torch.where(mask, cc, torch.zeros_like(cc))
        peaks.append(int(torch.argmax(cc)))
    torch.cuda.synchronize()
    return time.perf_counter() - t0, peaks


def bench_refine(args) -> int:
    print(json.dumps(run_refine(args.workload, args.steps)), flush=True)
    return 0


def run_refine(name: str, steps: int, lite: bool = False) -> dict:
    """The fused pair-set loss-and-gradient launch against per-pair calls, the dot test of the
    whole-volume projector and adjoint, and (not lite) the gradient refinement, the CC search and
    the adjoint launches; returns the benchmark line."""
    import numpy as np
    import torch

    import artiatomi_amd as A
    from artiatomi_amd import autograd as AG
    from artiatomi_amd import phantom

    w = REFINE_WORKLOADS[name]
    rng = np.random.default_rng(w["seed"])
    FX, FY, FZ = w["field"]
    W, H, n, K = w["W"], w["H"], w["n"], w["cubes"]
    tilts = phantom.tilt_series(-60.0, 60.0, w["step"])
    T = len(tilts)
    full = A.volume_geometry(FX, FY, FZ)
    ctx = A.ArtiaContext(0)
    t0 = time.time()
    vols = phantom.blob_subvolumes(n, K, w["seed"])
    base = [A.tilt_geometry(t, W, H) for t in tilts]
    subs, pairs, geoms = [], [], []
    for k in range(K):
        pos = rng.uniform([n / 2, n / 2, n / 2], [FX - n / 2, FY - n / 2, FZ - n / 2]) + 1.0
        vg = A.subvolume_geometry(n, 1.0, (1.0, 1.0, 1.0), full, pos.tolist())
        subs.append((vg, k * n ** 3))
        for g in base:
            x0, y0, ww, hh = A.footprint_window(vg, g, W, H, w["margin"])
            pairs.append((k, x0, y0, ww, hh))
            geoms.append(g)
    ps = A.PairSet(ctx, subs, pairs, geoms)
    P = ps.npairs
    s_true = torch.from_numpy(rng.uniform(-w["true_shift"], w["true_shift"], (P, 2))).cuda()
    targets = ps.project(vols, s_true)
    torch.cuda.synchronize()
    log(f"workload {name}: {K} sub-volumes of {n}^3, {T} tilts, {P} pairs, {ps.n_pixels / 1e6:.1f} M "
        f"window pixels; inputs in {time.time() - t0:.1f} s")

    # fused loss + shift gradient: kernel time (HIP events) and the 50-evaluation refinement
    zero = torch.zeros((P, 2), device="cuda", dtype=torch.float64)
    ps.loss_grad(vols, zero, targets)
    ms = [ps.loss_grad(vols, zero, targets, timed=True)[2] for _ in range(max(1, steps))]
    t_eval = sum(ms) / len(ms) / 1e3
    if lite:
        return refine_lite(ctx, ps, vols, subs, targets, zero, geoms, base, w, t_eval, K, n, T, P)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s_grad = AG.refine_pair_shifts(vols, targets, ps, steps=w["evals"])
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    e_grad = (s_grad - s_true).norm(dim=1)
    log(f"gradient refinement: {w['evals']} evaluations in {t_grad:.2f} s ({t_eval * 1e3:.1f} ms per fused launch), "
        f"rms error {float(e_grad.pow(2).mean().sqrt()):.2e} px, max {float(e_grad.max()):.2e} px")
    # the error after each evaluation (an untimed second run that records its iterates)
    trace = {}
    AG.refine_pair_shifts(vols, targets, ps, steps=w["evals"],
                          callback=lambda k, s: trace.__setitem__(k, float((s - s_true).norm(dim=1).max())))
    to_1e3 = next((k for k in sorted(trace) if trace[k] <= 1e-3), None)
    log(f"max error by evaluation: {[f'{trace[k]:.1e}' for k in sorted(trace)[:12]]}; <= 1e-3 px after {to_1e3}")

    # integer cross-correlation search: one projection + one search launch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s_cc, cc, failed = AG.cc_search_shifts(vols, targets, ps, W, w["max_shift"])
    torch.cuda.synchronize()
    t_cc = time.perf_counter() - t0
    e_cc = (s_cc - s_true).norm(dim=1)
    log(f"CC search: {t_cc:.3f} s, rms error {float(e_cc.pow(2).mean().sqrt()):.3f} px, "
        f"max {float(e_cc.max()):.3f} px, failures {int(failed.sum())}")

    # per-pair schedules on a sample, extrapolated to all pairs
    sample = min(P, 10 * T)
    t_plain_grad = plain_gradient_sample(ctx, ps, vols, subs, targets, zero, geoms, W, H, sample) * P / sample
    models = ps.project(vols)
    t_plain_cc_s, _ = plain_cc_sample(ps, models, targets, W, w["max_shift"], sample)
    t_plain_cc = t_plain_cc_s * P / sample
    log(f"per-pair calls (extrapolated from {sample} pairs): gradient evaluation {t_plain_grad:.2f} s, "
        f"full-square FFT search {t_plain_cc:.2f} s")

    # dot-product test of the whole-volume projector and its adjoint at this size
    dot_rel = dot_test(ctx, FX, FY, FZ, W, H, base, w["seed"])
    log(f"dot test {FX}x{FY}x{FZ}, {T} tilts: rel {dot_rel:.2e}")
    vg_full = A.volume_geometry(FX, FY, FZ)
    gen = torch.Generator(device="cuda").manual_seed(w["seed"] + 1)
    xv = torch.randn((FZ, FY, FX), generator=gen, device="cuda")
    proj = torch.empty((H, W), device="cuda")
    dist = torch.empty((H, W), device="cuda")
    # adjoint and forward launch times on the whole volume, LDS-staged and direct
    yv = torch.randn((H, W), generator=gen, device="cuda")
    pty = torch.zeros_like(xv)
    adj = {}
    for staged in (True, False):
        ctx.set_fp_staging(staged)
        ctx.fp_adjoint(vg_full, base[0], yv, pty)
        ta = [ctx.fp_adjoint(vg_full, g, yv, pty, timed=True) for g in base]
        tf = [ctx.fp_march(vg_full, g, xv, proj, dist, timed=True) for g in base]
        adj["staged" if staged else "direct"] = {"adjoint_ms": sum(ta) / len(ta), "forward_ms": sum(tf) / len(tf)}
    ctx.set_fp_staging(True)
    ms_pair_adj = ps.adjoint(zero, targets, torch.zeros_like(vols), timed=True)
    ctx.set_fp_staging(False)
    ps_direct = A.PairSet(ctx, subs, pairs, geoms)
    ctx.set_fp_staging(True)
    ms_pair_adj_direct = ps_direct.adjoint(zero, targets, torch.zeros_like(vols), timed=True)
    ms_pair_lg_direct = ps_direct.loss_grad(vols, zero, targets, normal=True, timed=True)[3]
    ps_direct.close()
    log(f"whole-volume adjoint per tilt: {adj['staged']['adjoint_ms']:.2f} ms staged, "
        f"{adj['direct']['adjoint_ms']:.2f} ms direct; forward {adj['staged']['forward_ms']:.2f} / "
        f"{adj['direct']['forward_ms']:.2f} ms; pair-set adjoint launch {ms_pair_adj:.1f} ms staged, "
        f"{ms_pair_adj_direct:.1f} ms direct; pair-set loss+gradient launch direct {ms_pair_lg_direct:.1f} ms")

    vox_tilts = K * n ** 3 * T
    bytes_alg = 4.0 * K * n ** 3 + 4.0 * ps.n_pixels + 24.0 * P
    achieved = bytes_alg / t_eval / 1e9
    out = {
        "metric": "Gvoxel-updates/s per fused loss + shift-gradient evaluation",
        "value": vox_tilts / t_eval / 1e9,
        "unit": "Gvoxel-updates/s",
        "n_gpus": 1,
        "steps": len(ms),
        "warmup": 1,
        "ms_per_step": t_eval * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded zero-margin blob sub-volumes, targets projected at seeded true shifts, no noise)",
        "config": {"workload": name, "field": list(w["field"]), "projection": [W, H], "tilts": T,
                   "subvolumes": K, "subvolume_edge": n, "pairs": P, "true_shift_px": w["true_shift"],
                   "max_shift": w["max_shift"], "numerics": "exact", "parallelism": "single GPU"},
        "roofline": {"kernel": "pair_march_kernel<loss+grad>", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "gradient_refinement": {"method": "Levenberg-Marquardt on 2 J^T r and J^T J of the fused launch",
                                "evaluations": w["evals"], "seconds": t_grad,
                                "rms_error_px": float(e_grad.pow(2).mean().sqrt()),
                                "max_error_px": float(e_grad.max()),
                                "max_error_by_evaluation": [trace[k] for k in sorted(trace)],
                                "evaluations_to_1e-3px": to_1e3,
                                "seconds_to_1e-3px": None if to_1e3 is None else to_1e3 * t_grad / w["evals"]},
        "cc_search": {"seconds": t_cc, "rms_error_px": float(e_cc.pow(2).mean().sqrt()),
                      "max_error_px": float(e_cc.max()), "failures": int(failed.sum())},
        "per_pair_calls_extrapolated": {"sample_pairs": sample, "gradient_evaluation_s": t_plain_grad,
                                        "fft_cc_search_s": t_plain_cc,
                                        "speedup_gradient_evaluation": t_plain_grad / t_eval,
                                        "speedup_cc_search": t_plain_cc / t_cc},
        "dot_test_full_volume_rel": dot_rel,
        "whole_volume_per_tilt_ms": adj,
        "pair_adjoint_launch_ms": {"staged": ms_pair_adj, "direct": ms_pair_adj_direct},
        "pair_loss_grad_launch_direct_ms": ms_pair_lg_direct,
    }
    ps.close()
    ctx.close()
    return out


def dot_test(ctx, FX, FY, FZ, W, H, base, seed):
    """<P x, y> against <x, P^T y> over a tilt series for random x and y, the whole-volume projector
    and its adjoint at FX x FY x FZ; the relative difference."""
    import torch

    import artiatomi_amd as A

    vg_full = A.volume_geometry(FX, FY, FZ)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    xv = torch.randn((FZ, FY, FX), generator=gen, device="cuda")
    lhs = rhs = 0.0
    proj = torch.empty((H, W), device="cuda")
    dist = torch.empty((H, W), device="cuda")
    for g in base:
        yv = torch.randn((H, W), generator=gen, device="cuda")
        proj.zero_()
        ctx.fp_march(vg_full, g, xv, proj, dist)
        pty = torch.zeros_like(xv)
        ctx.fp_adjoint(vg_full, g, yv, pty)
        lhs += float((proj.double() * yv.double()).sum())
        rhs += float((xv.double() * pty.double()).sum())
    return abs(lhs - rhs) / abs(lhs)


def refine_lite(ctx, ps, vols, subs, targets, zero, geoms, base, w, t_eval, K, n, T, P) -> dict:
    """The adjoint/gradient form against the plain per-pair schedule: one fused loss + shift-gradient
    launch over every (sub-volume, tilt) pair against per-pair projection, residual and
    shift-gradient calls (timed on a sample of pairs, scaled to all), plus the dot test."""
    FX, FY, FZ = w["field"]
    W, H = w["W"], w["H"]
    sample = min(P, 10 * T)
    t_plain = plain_gradient_sample(ctx, ps, vols, subs, targets, zero, geoms, W, H, sample) * P / sample
    dot = dot_test(ctx, FX, FY, FZ, W, H, base, w["seed"])
    log(f"adjoint vs plain: fused {t_eval * 1e3:.1f} ms, per-pair {t_plain:.2f} s (from {sample} pairs); "
        f"dot test {dot:.2e}")
    ps.close()
    ctx.close()
    return {"workload": w["name"], "pairs": P, "subvolumes": K, "subvolume_edge": n, "tilts": T,
            "fused_launch_ms": t_eval * 1e3, "Gvoxel_updates_per_s": K * n ** 3 * T / t_eval / 1e9,
            "plain_per_pair_s": t_plain, "plain_extrapolated_from_pairs": sample,
            "speedup_vs_plain": t_plain / t_eval, "dot_test_rel": dot}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sart_1024x1024x256_61tilts",
                    choices=sorted(WORKLOADS) + sorted(CTF_WORKLOADS) + sorted(SUBVOL_WORKLOADS) +
                    sorted(REFINE_WORKLOADS) + sorted(PREP_WORKLOADS))
    ap.add_argument("--numerics", default="exact", choices=["exact", "reference"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--volume-f16", action="store_true",
                    help="store the SART volume as IEEE half (the reference's FP16Volume mode)")
    ap.add_argument("--no-fp-staging", action="store_true",
                    help="forward-project with direct volume reads instead of LDS-staged slabs")
    ap.add_argument("--no-plain", action="store_true",
                    help="sub-volume workloads: run the batched plan alone, without the plain schedule")
    ap.add_argument("--no-extras", action="store_true",
                    help="SART line only: skip the reference-numerics iteration and the batched and adjoint "
                         "comparisons (for profiling runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only an elementwise copy of 1 GiB (three launches) to calibrate HBM byte counters")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch

        n = 1 << 28   # 268M floats = 1 GiB read + 1 GiB written per launch
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "elementwise copy", "bytes_read_per_launch": 4 * n,
                          "bytes_written_per_launch": 4 * n, "launches": 3}), flush=True)
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    if args.workload in SUBVOL_WORKLOADS:
        return bench_subvol(args)
    if args.workload in REFINE_WORKLOADS:
        return bench_refine(args)
    if args.workload in CTF_WORKLOADS:
        return bench_ctf(args)
    if args.workload in PREP_WORKLOADS:
        return bench_prep(args)

    import torch

    import artiatomi_amd as A

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    ctx = A.ArtiaContext(0, numerics=args.numerics)
    ctx.set_fp_staging(not args.no_fp_staging)
    vg, acq, projs, w = build_workload(args.workload, ctx)
    nx, ny, nz = w["dims"]
    V, T = nx * ny * nz, len(acq.geoms)
    sp = A.sart_params(lambda_=0.3, sirt_count=1, oversampling=w["os"])
    vol = torch.zeros((nz, ny, nx), device="cuda", dtype=torch.float16 if args.volume_f16 else torch.float32)
    vbytes = 2.0 if args.volume_f16 else 4.0   # bytes per stored voxel

    empty = True
    for i in range(args.warmup):
        _, empty = ctx.sart_run(vg, acq, projs, vol, sp, iterations=1, volume_is_empty=empty)
        torch.cuda.synchronize()
        log(f"warmup iteration {i + 1}/{args.warmup} done")

    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, empty = ctx.sart_run(vg, acq, projs, vol, sp, iterations=1, volume_is_empty=empty)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    log(f"timed {args.steps} iterations in {elapsed:.3f} s; per-kernel totals {prof}")

    H, W = acq.height, acq.width
    per_kernel = {}
    for name, (ms, n) in prof.items():
        if n:
            per_kernel[name] = {"avg_ms": ms / n, "launches": n, "share": ms / (elapsed * 1e3)}
    bytes_bp = 2 * vbytes * V + 4.0 * W * H
    bytes_fp = vbytes * V + 8.0 * W * H
    bp_avg = prof["bp"][0] / max(1, prof["bp"][1]) / 1e3
    fp_avg = prof["fp"][0] / max(1, prof["fp"][1]) / 1e3
    samples = fp_samples_per_launch(ctx, vg, acq, vol)
    per_kernel["bp"].update({"algorithmic_bytes": bytes_bp, "achieved_GBs": bytes_bp / bp_avg / 1e9,
                             "hbm_frac": bytes_bp / bp_avg / 1e9 / HBM_PEAK_GBS,
                             "Gvoxel_updates_per_s": V / bp_avg / 1e9})
    per_kernel["fp"].update({"algorithmic_bytes": bytes_fp, "achieved_GBs": bytes_fp / fp_avg / 1e9,
                             "hbm_frac": bytes_fp / fp_avg / 1e9 / HBM_PEAK_GBS,
                             "Gsamples_per_s_approx": samples / fp_avg / 1e9})
    dom = "fp" if prof["fp"][0] >= prof["bp"][0] else "bp"
    dom_bytes, dom_avg = (bytes_fp, fp_avg) if dom == "fp" else (bytes_bp, bp_avg)
    achieved = dom_bytes / dom_avg / 1e9
    roofline = {"kernel": "march" if dom == "fp" else "backProjection", "bound": "hbm", "achieved": achieved,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    # the profile covers the default configuration's kernels only
    default_kernels = args.numerics == "exact" and not args.no_fp_staging and not args.volume_f16
    # the driver's forward projection is the residual form of the march (compare and cropBorder in its
    # epilogue), the instance without the y axis for tilts about y; the backprojection runs the
    # window form, one instance per window size
    cands = (["march_lds_kernel<false, true, float, true>", "march_lds_kernel<false, true, float, false>"]
             if dom == "fp" else ["bp_win_kernel"])
    measured = None
    for dom_kernel in cands if default_kernels else []:
        measured = measured_traffic(args.workload, dom_kernel)
        if measured:
            break
    if measured:
        roofline["traffic"] = measured.pop("bytes")
        roofline["traffic_detail"] = measured
    # the kernels are bound by their VALU issue, not by HBM: their counters from the same profile
    if default_kernels:
        for kname, kcands in (("fp", ["march_lds_kernel<false, true, float, true>",
                                      "march_lds_kernel<false, true, float, false>"]),
                              ("bp", ["bp_win_kernel"])):
            for kc in kcands:
                valu = measured_valu(args.workload, kc)
                if valu and kname in per_kernel:
                    per_kernel[kname]["valu_detail"] = valu
                    if kname == dom:
                        roofline["valu_detail"] = valu
                    break

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(vg, acq, vol, projs, w)
        except Exception as exc:   # reported, never silently replaced
            log(f"cpu baseline failed: {exc!r}")
            cpu = {"error": repr(exc)}

    value = V * T * args.steps / elapsed / 1e9
    # the whole iteration against the roofline: per tilt the forward projection reads the volume,
    # the backprojection reads and writes it, and the images move 28 bytes per pixel
    # (FP 8, compare 16, BP 4)
    iter_bytes = T * (3.0 * vbytes * V + 28.0 * W * H)
    iter_gbs = iter_bytes / (elapsed / args.steps) / 1e9
    ref_numerics = None
    if not args.no_extras and args.numerics == "exact":
        ref_numerics = reference_numerics_iteration(ctx, vg, acq, projs, vol, sp)
    del projs, vol
    torch.cuda.empty_cache()
    copy = copy_bandwidth()
    if cpu is not None and "error" not in cpu:
        cpu["cpu_model"] = cpu_model()
    extras = {}
    if not args.no_extras:
        extras = extra_comparisons()
    out = {
        "metric": "Gvoxel-updates/s per SART iteration",
        "value": value,
        "unit": "Gvoxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid + gold-bead phantom, projections from the forward projector + noise)",
        "config": {"workload": args.workload, "volume": list(w["dims"]), "projection": [W, H], "tilts": T,
                   "tilt_range_deg": [-60, 60], "oversampling": w["os"], "sirt_count": 1, "lambda": 0.3,
                   "numerics": args.numerics, "fp_staging": not args.no_fp_staging,
                   "volume_storage": "f16" if args.volume_f16 else "f32", "parallelism": "single GPU"},
        "roofline": roofline,
        "iteration_roofline": {"algorithmic_bytes": iter_bytes, "achieved_GBs": iter_gbs, "peak": HBM_PEAK_GBS,
                               "frac": iter_gbs / HBM_PEAK_GBS, "frac_of_measured_copy": iter_gbs / copy["GBs"]},
        "copy_bandwidth": copy,
        "kernels": per_kernel,
        "reference_numerics": ref_numerics,
        **extras,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def cpu_model() -> str:
    """The host CPU's model name (the first `model name` of /proc/cpuinfo, which lscpu reports)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def copy_bandwidth() -> dict:
    """Device-to-device copy of 1 GiB (1 GiB read + 1 GiB written per copy), timed with events over
    five copies: the achievable HBM bandwidth on this GPU."""
    import torch

    n = 1 << 28
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    del src, dst
    return {"GBs": 8.0 * n / ms / 1e6, "bytes_per_copy": 8 * n, "ms_per_copy": ms,
            "how": "torch copy_ of a 1 GiB float32 tensor, 5 copies between HIP events"}


def reference_numerics_iteration(ctx, vg, acq, projs, vol, sp) -> dict:
    """One SART iteration in the reference's float operation order (ARTIA_NUMERICS_REFERENCE) from a
    copy of the current volume: the price of bit-faithfulness to the reference's kernel bodies."""
    import torch

    v = vol.clone()
    ctx.set_numerics("reference")
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.sart_run(vg, acq, projs, v, sp, iterations=1, volume_is_empty=False)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    finally:
        ctx.set_numerics("exact")
    V = vg.nx * vg.ny * vg.nz
    log(f"reference numerics: one iteration in {dt:.3f} s")
    del v
    return {"seconds_per_iteration": dt, "Gvoxel_updates_per_s": V * len(acq.geoms) / dt / 1e9}


def extra_comparisons() -> dict:
    """The rest of the headline metric, measured in the same run: the batched sub-volume WBP against
    per-call plain WBP (2000 cubes, 61 tilts of 4096^2; without CTF, and with 3D CTF where the plain
    schedule is timed on every 10th tilt and scaled by the slab count), and the fused adjoint/gradient
    form against per-pair calls (500 sub-volumes of 64^3, 41 tilts of 512^2)."""
    import torch

    out = {"batched_vs_plain": {}, "adjoint_vs_plain": None}
    for key, name in (("no_ctf", "subvol_wbp_2000cubes_4096_61tilts"),
                      ("ctf", "subvol_wbp_ctf_2000cubes_4096_61tilts")):
        t0 = time.time()
        r = run_subvol(name, 2)
        torch.cuda.empty_cache()
        out["batched_vs_plain"][key] = {
            "workload": name, "speedup_vs_plain": r["speedup_vs_plain"], "batched_s": r["ms_per_step"] / 1e3,
            "plain_s": r["plain"]["seconds"], "plain_extrapolated": bool(r["plain"].get("extrapolated", False)),
            "batched_Gvoxel_updates_per_s": r["value"], "bitwise_equal_to_plain": r["bitwise_equal_to_plain"],
            "wall_s": time.time() - t0}
    t0 = time.time()
    out["adjoint_vs_plain"] = run_refine("refine_500subvols_64_512x512x128_41tilts", 3, lite=True)
    out["adjoint_vs_plain"]["wall_s"] = time.time() - t0
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    sys.exit(main())
