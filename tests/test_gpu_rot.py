"""GPU checks of the batched motive rotation (artia_rot3d, RotKernel / rot3d) through the C ABI:
  - zero angles copy the cube bit for bit, a quarter turn about z is the exact index permutation
    with the clamped edge;
  - smooth cubes rotated by 50 angle triples (two launches of the batch) match the oracle's
    float64 rotation (oracle/rot3d.py) to 1e-6 relative L2 and 2e-5 of the peak per voxel, at an
    even and an odd edge length;
  - bad sizes are rejected with a status.
"""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

import artiatomi_amd as A
from conftest import rel_l2

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import rot3d as RO  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu):
    c = A.ArtiaContext(0)
    yield c
    c.close()


def smooth_cube(S, seed):
    rng = np.random.default_rng(seed)
    g = np.arange(S, dtype=np.float64)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    v = np.zeros((S, S, S))
    for _ in range(4):
        cx, cy, cz = rng.uniform(0.25 * S, 0.75 * S, 3)
        v += rng.uniform(0.5, 1.0) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2 * 3.0 ** 2))
    return v.astype(np.float32)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_identity_and_quarter_turn_are_exact(ctx):
    rng = np.random.default_rng(3)
    for S in (16, 21):
        v = rng.standard_normal((S, S, S)).astype(np.float32)
        out = host(ctx.rot3d(torch.from_numpy(v).cuda(), [(0, 0, 0), (90, 0, 0)]))
        assert np.array_equal(out[0], v)
        c = S // 2
        z, y, x = np.meshgrid(np.arange(S), np.arange(S), np.arange(S), indexing="ij")
        assert np.array_equal(out[1], v[z, np.clip(2 * c - x, 0, S - 1), y])


@pytest.mark.parametrize("S", [32, 37])
def test_rotations_match_oracle(ctx, S):
    rng = np.random.default_rng(S)
    angles = [tuple(rng.uniform(-180, 360, 3)) for _ in range(46)] + [(30, 45, 60), (135, 270, 315), (0, 0, 180),
                                                                       (-30, 90, 45)]
    v = smooth_cube(S, S)
    out, ms = ctx.rot3d(torch.from_numpy(v).cuda(), angles, timed=True)
    out = host(out)
    peak = float(np.abs(v).max())
    worst = 0.0
    for r, (phi, psi, theta) in enumerate(angles):
        ref = RO.rot3d(v, phi, psi, theta)
        e = rel_l2(out[r], ref)
        worst = max(worst, e)
        assert e <= 1e-6, (r, angles[r], e)
        assert np.abs(out[r] - ref).max() <= 2e-5 * peak, (r, angles[r])
    gbs = len(angles) * S ** 3 * 4 / (ms * 1e-3) / 1e9
    print(f"rot3d {len(angles)} x {S}^3: worst rel L2 {worst:.2e}, {ms:.3f} ms ({gbs:.1f} GB/s written)", flush=True)


def test_bad_arguments_raise(ctx):
    v = torch.zeros((4, 4, 4), device="cuda")
    with pytest.raises(A.ArtiaError):
        ctx.rot3d(torch.zeros((0, 0, 0), device="cuda"), [(0, 0, 0)])
    out = ctx.rot3d(v, [])
    assert out.shape == (0, 4, 4, 4)
