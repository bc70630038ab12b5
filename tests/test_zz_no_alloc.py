"""The allocation-free SART entry point (artia_sart_run_noalloc) through the C ABI: two calls in a
row, the second on a smaller problem, each equal bit for bit to artia_sart_run on the same inputs;
sizes beyond its fixed workspace are rejected with a clear error before anything runs."""
import numpy as np
import pytest
import torch

import artiatomi_amd as A
from artiatomi_amd import phantom

pytestmark = pytest.mark.gpu


def case(dims, W, H, step, seed, psi=0.0, shift=0.0):
    rng = np.random.default_rng(seed)
    tilts = phantom.tilt_series(-60, 60, step)
    psis = rng.uniform(-psi, psi, len(tilts)) if psi else np.zeros(len(tilts))
    shifts = rng.uniform(-shift, shift, (len(tilts), 2)) if shift else np.zeros((len(tilts), 2))
    vg = A.volume_geometry(*dims)
    acq = A.Acquisition.build(vg, W, H, tilts, psis_deg=psis.tolist(), shifts=shifts.tolist())
    projs = torch.from_numpy(rng.random((len(tilts), H, W), dtype=np.float32)).cuda()
    return vg, acq, projs


def test_sart_run_noalloc_equals_sart_run(gpu):
    ctx = A.ArtiaContext(0)
    try:
        # larger first, then smaller: the second call reuses the same per-GPU array
        for dims, W, H, step, sirt, os_, seed in (((96, 80, 24), 104, 96, 6.0, 3, 2, 31),
                                                  ((40, 32, 12), 48, 40, 10.0, 1, 1, 32)):
            print(f"noalloc SART {dims} {W}x{H} S={sirt} OS={os_}", flush=True)
            vg, acq, projs = case(dims, W, H, step, seed, psi=3.0, shift=2.0)
            sp = A.sart_params(lambda_=0.3, sirt_count=sirt, oversampling=os_, cut_length=(1, 1),
                               dim_length=(2, 2))
            nz, ny, nx = dims[2], dims[1], dims[0]
            ref = torch.zeros((nz, ny, nx), device="cuda")
            got = torch.zeros((nz, ny, nx), device="cuda")
            _, e_ref = ctx.sart_run(vg, acq, projs, ref, sp, iterations=2)
            _, e_got = ctx.sart_run(vg, acq, projs, got, sp, iterations=2, noalloc=True)
            torch.cuda.synchronize()
            assert e_ref == e_got and not e_got
            assert float(ref.abs().max()) > 0
            assert torch.equal(got, ref)
    finally:
        ctx.close()


def test_sart_run_noalloc_rejects_oversized_workspace(gpu):
    ctx = A.ArtiaContext(0)
    try:
        # (1 + 1) x 4096 x 4096 floats = 32 Mi > the 16 Mi floats of the fixed workspace
        vg = A.volume_geometry(32, 32, 8)
        acq = A.Acquisition.build(vg, 4096, 4096, [0.0])
        projs = torch.zeros((1, 4096, 4096), device="cuda")
        vol = torch.zeros((8, 32, 32), device="cuda")
        with pytest.raises(A.ArtiaError, match="allocation-free"):
            ctx.sart_run(vg, acq, projs, vol, A.sart_params(), iterations=1, noalloc=True)
        torch.cuda.synchronize()
        assert float(vol.abs().max()) == 0.0
    finally:
        ctx.close()
