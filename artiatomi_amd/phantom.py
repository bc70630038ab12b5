"""Seeded synthetic workloads: ellipsoid / gold-bead phantoms and noise.

Every random draw comes from SplitMix64 (Steele, Lea, Flood 2014) so a phantom is identical on
any machine for a given seed.  Ellipsoids: centre uniform in the volume, semi-axes U[4, 40]
voxels, uniformly random orientation (Shoemake's unit quaternion), density U[0.1, 1],
additive; gold beads: spheres of radius 3 voxels and density 5.  A voxel takes a shape's
density when its centre lies inside the shape.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


@dataclass
class Shape:
    centre: tuple      # voxel coordinates (x, y, z), voxel j centred at j + 0.5
    axes: tuple        # semi-axes in voxels
    rot: np.ndarray    # 3x3 rotation, body -> volume frame
    density: float


def _random_rotation(rng: SplitMix64) -> np.ndarray:
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
         math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def make_shapes(dims, n_ellipsoids: int, seed: int, n_beads: int = 0, axes=(4.0, 40.0),
                density=(0.1, 1.0), bead_radius: float = 3.0, bead_density: float = 5.0) -> list:
    nx, ny, nz = dims
    rng = SplitMix64(seed)
    shapes = []
    for _ in range(n_ellipsoids):
        c = (rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz))
        a = (rng.uniform(*axes), rng.uniform(*axes), rng.uniform(*axes))
        shapes.append(Shape(c, a, _random_rotation(rng), rng.uniform(*density)))
    for _ in range(n_beads):
        c = (rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz))
        shapes.append(Shape(c, (bead_radius,) * 3, np.eye(3), bead_density))
    return shapes


def rasterize(dims, shapes, device: str = "cpu"):
    """Returns a float32 torch tensor [nz, ny, nx] with the summed shape densities."""
    import torch

    nx, ny, nz = dims
    vol = torch.zeros((nz, ny, nx), dtype=torch.float32, device=device)
    for s in shapes:
        r = max(s.axes)
        lo = [max(0, int(math.floor(s.centre[i] - r - 1))) for i in range(3)]
        hi = [min(dims[i], int(math.ceil(s.centre[i] + r + 1))) for i in range(3)]
        if any(hi[i] <= lo[i] for i in range(3)):
            continue
        zs = torch.arange(lo[2], hi[2], device=device, dtype=torch.float64) + 0.5 - s.centre[2]
        ys = torch.arange(lo[1], hi[1], device=device, dtype=torch.float64) + 0.5 - s.centre[1]
        xs = torch.arange(lo[0], hi[0], device=device, dtype=torch.float64) + 0.5 - s.centre[0]
        Z, Y, X = torch.meshgrid(zs, ys, xs, indexing="ij")
        rt = torch.as_tensor(s.rot.T, dtype=torch.float64, device=device)
        bx = rt[0, 0] * X + rt[0, 1] * Y + rt[0, 2] * Z
        by = rt[1, 0] * X + rt[1, 1] * Y + rt[1, 2] * Z
        bz = rt[2, 0] * X + rt[2, 1] * Y + rt[2, 2] * Z
        inside = (bx / s.axes[0]) ** 2 + (by / s.axes[1]) ** 2 + (bz / s.axes[2]) ** 2 <= 1.0
        vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += inside.to(torch.float32) * float(s.density)
    return vol


def phantom(dims, n_ellipsoids: int, seed: int, n_beads: int = 0, device: str = "cpu"):
    return rasterize(dims, make_shapes(dims, n_ellipsoids, seed, n_beads=n_beads), device=device)


def gaussian_noise(shape, sigma: float, seed: int) -> np.ndarray:
    return (np.random.default_rng(seed).standard_normal(shape) * sigma).astype(np.float32)


def tilt_series(lo: float, hi: float, step: float) -> list:
    n = int(round((hi - lo) / step)) + 1
    return [lo + i * step for i in range(n)]


def blob_subvolumes(n: int, count: int, seed: int):
    """count seeded n^3 sub-volumes of three Gaussian blobs each (centres U[0.38, 0.62] of the edge,
    amplitudes U[0.5, 1], width 0.09 of the edge), below 1e-6 at the faces, so the projection of a
    sub-volume is smooth in a detector shift (zero margins).  Returned as one flat float32 CUDA
    buffer, sub-volume k at [k * n^3, (k + 1) * n^3)."""
    import torch

    gen = torch.Generator(device="cuda").manual_seed(seed)
    g = (torch.arange(n, device="cuda", dtype=torch.float32) + 0.5) / n
    z, y, x = torch.meshgrid(g, g, g, indexing="ij")
    out = torch.empty((count, n, n, n), device="cuda")
    for c0 in range(0, count, 50):
        c1 = min(count, c0 + 50)
        cen = 0.38 + 0.24 * torch.rand((c1 - c0, 3, 3), generator=gen, device="cuda")
        amp = 0.5 + 0.5 * torch.rand((c1 - c0, 3), generator=gen, device="cuda")
        acc = torch.zeros((c1 - c0, n, n, n), device="cuda")
        for b in range(3):
            d2 = ((x[None] - cen[:, b, 0, None, None, None]) ** 2 + (y[None] - cen[:, b, 1, None, None, None]) ** 2
                  + (z[None] - cen[:, b, 2, None, None, None]) ** 2)
            acc += amp[:, b, None, None, None] * torch.exp(-d2 / 0.09 ** 2)
        out[c0:c1] = acc
    return out.reshape(-1)
