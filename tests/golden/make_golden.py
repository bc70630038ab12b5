"""Writes tests/golden/sart_32x32x8_5tilts.npz: small golden vectors of the oracle (FP64) for the
SART path, so that a change of the oracle or of the kernels shows against committed numbers.

Inputs (stored in the file): a seeded 32 x 32 x 8 volume, five tilts with psi and shifts, 40 x 36
projections and one residual image.  Outputs: per tilt the forward projection and mean chord of the
volume, the backprojection of the residual into a zero volume (OS 2) at two tilts, and one SART
iteration from an empty volume (OS 2, lambda 0.3, a 2-pixel border crop).  Run from the repository
root:
    python tests/golden/make_golden.py
The geometry comes from libartia's host geometry, the numbers from the oracle; the reference is
not involved (it ships no golden vectors of its own)."""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))
import artiatomi_amd as A  # noqa: E402
import oracle_lib as O  # noqa: E402

DIMS, W, H = (32, 32, 8), 40, 36
TILTS = [-54.0, -21.0, 0.0, 17.0, 48.0]
PSIS = [1.5, -2.0, 0.0, 2.5, -1.0]
SHIFTS = [(0.5, -1.0), (1.2, 0.3), (0.0, 0.0), (-0.8, 0.6), (-1.5, -0.4)]
BP_TILTS = (0, 3)   # tilts whose backprojection is stored


def inputs():
    rng = np.random.default_rng(2024)
    vol = rng.random(DIMS[::-1]).astype(np.float32)
    projs = (rng.random((len(TILTS), H, W)) * 8.0).astype(np.float32)
    resid = rng.normal(0.0, 1.0, (H, W)).astype(np.float32)
    return vol, projs, resid


def setup():
    vg = A.volume_geometry(*DIMS)
    acq = A.Acquisition.build(vg, W, H, TILTS, psis_deg=PSIS, shifts=SHIFTS)
    sp = A.sart_params(lambda_=0.3, oversampling=2, cut_length=(2, 2), dim_length=(3, 3))
    return vg, acq, sp


def compute():
    vol, projs, resid = inputs()
    vg, acq, sp = setup()
    out = dict(volume=vol, projections=projs, residual=resid, tilts=np.array(TILTS), psis=np.array(PSIS),
               shifts=np.array(SHIFTS), index_list=np.array(acq.index_list, np.int32))
    fp, dist, bp = [], [], []
    for i, g in enumerate(acq.geoms):
        p, d = O.march(vg, g, vol.astype(np.float64), W, H, dtype=np.float64)
        fp.append(p)
        dist.append(d)
        if i in BP_TILTS:
            bp.append(O.back_project(vg, g, np.zeros(DIMS[::-1]), resid, 0.3, os_=2, dtype=np.float64))
    out["march_proj"], out["march_dist"], out["bp_os2"] = np.stack(fp), np.stack(dist), np.stack(bp)
    out["sart_1iter"], _ = O.sart(vg, vg, acq.geoms, acq.index_list, projs, np.zeros(DIMS[::-1]), sp, 1,
                                  dtype=np.float64)
    return out


if __name__ == "__main__":
    data = compute()
    np.savez_compressed(HERE / "sart_32x32x8_5tilts.npz", **data)
    print("wrote", HERE / "sart_32x32x8_5tilts.npz", {k: v.shape for k, v in data.items()})
