"""Pins the oracle of the motive rotation (oracle/rot3d.py: computeRotMat, Kernels.cpp:1227-1327,
and rot3d, wbpWeighting.cu:388-405) with checks independent of it, and checks the library's host
matrix (artia_rot_matrix) against it."""
import sys
from pathlib import Path

import numpy as np
import pytest
from scipy import ndimage

import artiatomi_amd as A

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import rot3d as RO  # noqa: E402


def rz(a):
    c, s = np.cos(np.radians(a)), np.sin(np.radians(a))
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def rx(a):
    c, s = np.cos(np.radians(a)), np.sin(np.radians(a))
    return np.array([[1.0, 0, 0], [0, c, -s], [0, s, c]])


ANGLES = [(0, 0, 0), (90, 0, 0), (0, 0, 90), (30, 45, 60), (12.5, -71.0, 33.3), (200.0, 315, 17.0),
          (-45.0, 123.4, 180)]


@pytest.mark.parametrize("phi,psi,theta", ANGLES)
def test_matrix_is_zxz_euler_product(phi, psi, theta):
    """computeRotMat = Rz(psi) Rx(theta) Rz(phi), orthonormal with determinant 1."""
    R = RO.rot_matrix(phi, psi, theta).astype(np.float64)
    assert np.abs(R - rz(psi) @ rx(theta) @ rz(phi)).max() <= 1e-6
    assert np.abs(R @ R.T - np.eye(3)).max() <= 1e-6
    assert abs(np.linalg.det(R) - 1.0) <= 1e-6


def test_tabulated_angles_are_exact():
    assert np.array_equal(RO.rot_matrix(90, 0, 0), np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float32))
    assert np.array_equal(RO.rot_matrix(0, 0, 180), np.array([[1, 0, 0], [0, -1, 0], [0, 0, -1]], np.float32))
    R = RO.rot_matrix(0, 0, 60)
    assert R[2, 2] == np.float32(0.5) and R[1, 2] == -(np.sqrt(np.float32(3.0)) / np.float32(2.0))


@pytest.mark.parametrize("phi,psi,theta", ANGLES)
def test_library_matrix_matches_oracle(phi, psi, theta):
    """artia_rot_matrix (host code of libartia, no GPU) gives the oracle's float matrix."""
    assert np.abs(A.rot_matrix(phi, psi, theta) - RO.rot_matrix(phi, psi, theta)).max() <= 3e-7


def test_identity_and_quarter_turn():
    rng = np.random.default_rng(1)
    for S in (8, 9):
        v = rng.standard_normal((S, S, S)).astype(np.float32)
        assert np.array_equal(RO.rot3d(v, 0, 0, 0), v)
        # phi = 90: R^T (v - c) = (y - c, c - x, z - c), so out[z][y][x] = in[z][clamp(2c - x)][y]
        c = S // 2
        out = RO.rot3d(v, 90, 0, 0)
        z, y, x = np.meshgrid(np.arange(S), np.arange(S), np.arange(S), indexing="ij")
        assert np.array_equal(out, v[z, np.clip(2 * c - x, 0, S - 1), y])


def test_output_is_input_rotated_by_R():
    """A point at c + d in the input lands at c + R d in the output."""
    S = 16
    c = S // 2
    v = np.zeros((S, S, S), np.float32)
    v[c, c, c + 3] = 1.0                       # d = (3, 0, 0)
    for (phi, psi, theta), want in (((90, 0, 0), (0, 3, 0)), ((0, 0, 90), (3, 0, 0)), ((90, 0, 90), (0, 0, 3)), ((0, 90, 90), (0, 3, 0))):
        d = np.rint(RO.rot_matrix(phi, psi, theta).astype(np.float64) @ np.array([3.0, 0, 0])).astype(int)
        assert tuple(d) == want
        out = RO.rot3d(v, phi, psi, theta)
        z, y, x = np.unravel_index(np.argmax(out), out.shape)
        assert (x - c, y - c, z - c) == want and out[z, y, x] == 1.0


@pytest.mark.parametrize("phi,psi,theta", ANGLES[3:])
def test_interpolation_matches_scipy(phi, psi, theta):
    """Trilinear with clamped indices equals scipy's order-1 map_coordinates with mode 'nearest'."""
    rng = np.random.default_rng(2)
    S = 11
    v = rng.standard_normal((S, S, S)).astype(np.float32)
    R = RO.rot_matrix(phi, psi, theta).astype(np.float64)
    c = float(S // 2)
    z, y, x = np.meshgrid(np.arange(S), np.arange(S), np.arange(S), indexing="ij")
    p = c + np.einsum("ji,jk->ik", R, np.stack([x.ravel() - c, y.ravel() - c, z.ravel() - c]))
    ref = ndimage.map_coordinates(v.astype(np.float64), [p[2], p[1], p[0]], order=1, mode="nearest")
    assert np.abs(RO.rot3d(v, phi, psi, theta).ravel() - ref).max() <= 1e-12
