"""The HIP path (exact numerics) against the committed golden vectors of the FP64 oracle
(tests/golden/sart_32x32x8_5tilts.npz): forward projection and mean chord per tilt, OS-2
backprojection at two tilts, one SART iteration; each within 1e-5 relative L2."""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

import artiatomi_amd as A
from conftest import rel_l2

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_golden as MG  # noqa: E402

pytestmark = pytest.mark.gpu
FIXTURE = Path(__file__).resolve().parent / "golden" / "sart_32x32x8_5tilts.npz"


def test_hip_path_matches_golden_vectors(gpu):
    with np.load(FIXTURE, allow_pickle=False) as f:
        gold = {k: f[k] for k in f.files}
    vg, acq, sp = MG.setup()
    ctx = A.ArtiaContext(0)
    try:
        vol = torch.from_numpy(gold["volume"]).cuda()
        resid = torch.from_numpy(gold["residual"]).cuda()
        bpi = 0
        for i, g in enumerate(acq.geoms):
            print(f"golden tilt {MG.TILTS[i]}", flush=True)
            proj = torch.zeros((MG.H, MG.W), device="cuda")
            dist = torch.zeros((MG.H, MG.W), device="cuda")
            ctx.fp_march(vg, g, vol, proj, dist)
            assert rel_l2(proj.cpu().numpy(), gold["march_proj"][i]) <= 1e-5
            assert rel_l2(dist.cpu().numpy(), gold["march_dist"][i]) <= 1e-5
            if i in MG.BP_TILTS:
                out = torch.zeros(MG.DIMS[::-1], device="cuda")
                ctx.bp(vg, g, out, resid, 0.3, oversampling=2)
                assert rel_l2(out.cpu().numpy(), gold["bp_os2"][bpi]) <= 1e-5
                bpi += 1
        out = torch.zeros(MG.DIMS[::-1], device="cuda")
        ctx.sart_run(vg, acq, torch.from_numpy(gold["projections"]).cuda(), out, sp, iterations=1)
        err = rel_l2(out.cpu().numpy(), gold["sart_1iter"])
        print(f"golden sart 1 iteration: rel L2 {err:.2e}", flush=True)
        assert err <= 1e-5
    finally:
        ctx.close()
