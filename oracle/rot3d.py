"""CPU oracle of the motive rotation (RotKernel::computeRotMat and the rot3d kernel).

TEST INFRASTRUCTURE ONLY.  Nothing in the product (artiatomi_amd/) may import or call this
module; only tests/ use it, as the checker.

Restated from src/EmSART/Kernels.cpp:1227-1327 (computeRotMat) and
src/EmSART/cuda_kernels/wbpWeighting.cu:388-405 (rot3d), paths relative to the uermel/Artiatomi
source tree:
  - rot_matrix: the float matrix from phi, psi, theta in degrees; the sixteen tabulated angles take
    exact cosines and sines (sqrt(3.0f) / 2.0f, sqrt(2.0f) / 2.0f, 0.5f, 0, 1), any other angle
    cos/sin of angle * (float)M_PI / 180.0f in float;
  - rot3d: out[z][y][x] = tex3D(texVol, p + 0.5) with p = c + R^T (v - c), c = (float)(size / 2);
    the texture filters LINEAR with CLAMP addressing, i.e. trilinear interpolation at voxel
    position p with every texel index clamped to [0, size - 1].  The interpolation runs in
    float64 here (exact weights).
"""
import numpy as np

_TABLE = (0, 30, 45, 60, 90, 120, 135, 150, 180, 210, 225, 240, 270, 300, 315, 330)


def _cos_sin(a: float):
    a = np.float32(a)
    r3 = np.float32(np.sqrt(np.float32(3.0)) / np.float32(2.0))
    r2 = np.float32(np.sqrt(np.float32(2.0)) / np.float32(2.0))
    cs = (1.0, r3, r2, 0.5, 0.0, -0.5, -r2, -r3, -1.0, -r3, -r2, -0.5, 0.0, 0.5, r2, r3)
    sn = (0.0, 0.5, r2, r3, 1.0, r3, r2, 0.5, 0.0, -0.5, -r2, -r3, -1.0, -r3, -r2, -0.5)
    hit = None
    for i, d in enumerate(_TABLE):
        if np.float32(d) == a:
            hit = i   # the reference's loop keeps the last match; the table has no duplicates
    if hit is not None:
        return np.float32(cs[hit]), np.float32(sn[hit])
    rad = np.float32(np.float32(a * np.float32(np.pi)) / np.float32(180.0))
    return np.cos(rad, dtype=np.float32), np.sin(rad, dtype=np.float32)


def rot_matrix(phi: float, psi: float, theta: float) -> np.ndarray:
    """computeRotMat: float32 3 x 3, R[i][j] = rotMat[i][j]."""
    cphi, sphi = _cos_sin(phi)
    cpsi, spsi = _cos_sin(psi)
    cth, sth = _cos_sin(theta)
    f = np.float32
    R = np.empty((3, 3), np.float32)
    R[0, 0] = f(cpsi * cphi) - f(f(cth * spsi) * sphi)
    R[1, 0] = f(spsi * cphi) + f(f(cth * cpsi) * sphi)
    R[2, 0] = f(sth * sphi)
    R[0, 1] = f(-cpsi * sphi) - f(f(cth * spsi) * cphi)
    R[1, 1] = f(-spsi * sphi) + f(f(cth * cpsi) * cphi)
    R[2, 1] = f(sth * cphi)
    R[0, 2] = f(sth * spsi)
    R[1, 2] = f(-sth * cpsi)
    R[2, 2] = cth
    return R


def sample_clamped(vol: np.ndarray, px, py, pz) -> np.ndarray:
    """Trilinear interpolation of vol[z][y][x] at voxel positions (px, py, pz), indices clamped."""
    S = vol.shape
    v = vol.astype(np.float64)
    fx, fy, fz = np.floor(px), np.floor(py), np.floor(pz)
    tx, ty, tz = px - fx, py - fy, pz - fz
    ix, iy, iz = fx.astype(np.int64), fy.astype(np.int64), fz.astype(np.int64)

    def at(dx, dy, dz):
        return v[np.clip(iz + dz, 0, S[0] - 1), np.clip(iy + dy, 0, S[1] - 1), np.clip(ix + dx, 0, S[2] - 1)]

    out = np.zeros(np.shape(px))
    for dz, wz in ((0, 1 - tz), (1, tz)):
        for dy, wy in ((0, 1 - ty), (1, ty)):
            for dx, wx in ((0, 1 - tx), (1, tx)):
                out += wz * wy * wx * at(dx, dy, dz)
    return out


def rot3d(vol: np.ndarray, phi: float, psi: float, theta: float) -> np.ndarray:
    """The cube vol (size^3, [z][y][x]) rotated as rot3d does, float64."""
    S = vol.shape[0]
    R = rot_matrix(phi, psi, theta).astype(np.float64)
    c = float(S // 2)
    z, y, x = np.meshgrid(np.arange(S, dtype=np.float64), np.arange(S, dtype=np.float64),
                          np.arange(S, dtype=np.float64), indexing="ij")
    vx, vy, vz = x - c, y - c, z - c
    px = c + R[0, 0] * vx + R[1, 0] * vy + R[2, 0] * vz
    py = c + R[0, 1] * vx + R[1, 1] * vy + R[2, 1] * vz
    pz = c + R[0, 2] * vx + R[1, 2] * vy + R[2, 2] * vz
    return sample_clamped(vol, px, py, pz)
