"""Host geometry: the product's C-ABI geometry against the oracle restatement (bit-exact), and
both against closed forms that do not share their code:
  - tilt 0 / psi 0 detector frame (u = (0,1,0), v = (-1,0,0), n = (0,0,-1));
  - orthonormality and the detector-matrix inverse identity;
  - matlab/+artia/+geo/project3D.m + recon2real.m point projection (an independently written
    closed form of the same acquisition geometry);
  - the projection index list against the running C library's srand(0)/rand().
"""
import ctypes
import math

import numpy as np
import pytest

import artiatomi_amd as A
import oracle_lib as O

CASES = [
    dict(tilt=0.0, psi=0.0, shift=(0.0, 0.0)),
    dict(tilt=2.0, psi=0.0, shift=(0.0, 0.0)),
    dict(tilt=-60.0, psi=0.0, shift=(3.5, -1.25)),
    dict(tilt=37.0, psi=4.3, shift=(-4.9, 2.2)),
    dict(tilt=-13.0, psi=-2.7, shift=(0.3, 0.0), mag=1.03),
    dict(tilt=51.0, psi=1.0, shift=(1.0, 1.0), mag=0.5),   # mag below 0.8 is ignored
]


@pytest.mark.parametrize("case", CASES)
def test_tilt_geometry_product_matches_oracle(case):
    W, H = 200, 160
    kw = dict(psi_deg=case["psi"], shift=case["shift"], mag=case.get("mag", 1.0))
    g = O.tilt_geometry(case["tilt"], W, H, **kw)
    gp = A.tilt_geometry(case["tilt"], W, H, **kw)
    assert bytes(g) == bytes(gp)


def test_global_params_and_mag_aniso_match():
    gl = A.GlobalParams(1.5, 0.7, 12.0, 0, 0.0)
    for fixpsi in (0, 1):
        gl.use_fix_psi = fixpsi
        gl.fix_psi_deg = 3.0
        g = O.tilt_geometry(25.0, 128, 96, psi_deg=-2.0, glob=gl, mag_aniso=(1.02, 33.0))
        gp = A.tilt_geometry(25.0, 128, 96, psi_deg=-2.0, glob=gl, mag_aniso=(1.02, 33.0))
        assert bytes(g) == bytes(gp)


def test_volume_subvolume_hitpoints_reallength_match():
    vg = O.volume_geometry(96, 80, 24, vs=(1.0, 1.0, 1.0), shift=(0.5, -2.0, 3.0))
    vgp = A.volume_geometry(96, 80, 24, voxel_size=(1.0, 1.0, 1.0), shift=(0.5, -2.0, 3.0))
    assert bytes(vg) == bytes(vgp)
    sv = O.subvolume_geometry(32, 1.0, (1.0, 1.0, 1.0), vg, (40.0, 17.0, 9.0), (0.25, -0.5, 1.0))
    svp = A.subvolume_geometry(32, 1.0, (1.0, 1.0, 1.0), vgp, (40.0, 17.0, 9.0), (0.25, -0.5, 1.0))
    assert bytes(sv) == bytes(svp)
    for c in CASES:
        g = O.tilt_geometry(c["tilt"], 128, 128, psi_deg=c["psi"], shift=c["shift"])
        assert O.hit_points(vg, g) == A.hitpoints(vgp, g)
        assert O.real_length(vg, g) == A.real_length(vgp, g)


def test_tilt_zero_closed_form():
    W, H = 256, 256
    g = O.tilt_geometry(0.0, W, H)
    assert list(g.u) == [0.0, 1.0, 0.0]
    assert list(g.v) == [-1.0, 0.0, 0.0]
    assert list(g.n) == [0.0, 0.0, -1.0]
    assert list(g.d) == [H / 2, -W / 2, 10000.0]
    # a point on the detector maps to its pixel coordinates
    M = np.array(g.det_matrix, dtype=np.float64).reshape(4, 4)
    p = np.array(g.d, np.float64) + 17.25 * np.array(g.u) + 3.5 * np.array(g.v) + 42.0 * np.array(g.n)
    a = M @ np.append(p, 1.0)
    np.testing.assert_allclose(a[:3], [17.25, 3.5, 42.0], atol=1e-9)


@pytest.mark.parametrize("case", CASES)
def test_frame_orthonormal_and_inverse(case):
    g = O.tilt_geometry(case["tilt"], 180, 140, psi_deg=case["psi"], shift=case["shift"], mag=case.get("mag", 1.0))
    u, v, n = (np.array(x, np.float64) for x in (g.u, g.v, g.n))
    uh, vh = u / np.linalg.norm(u), v / np.linalg.norm(v)
    assert abs(uh @ vh) < 1e-6 and abs(uh @ n) < 1e-6 and abs(vh @ n) < 1e-6
    assert abs(np.linalg.norm(n) - 1) < 1e-6
    M = np.array(g.det_matrix, np.float64).reshape(4, 4)
    rng = np.random.default_rng(1)
    for _ in range(5):
        a, b, c = rng.uniform(-300, 300, 3)
        p = np.array(g.d, np.float64) + a * u + b * v + c * n
        r = M @ np.append(p, 1.0)
        np.testing.assert_allclose(r[:3], [a, b, c], rtol=0, atol=2e-3)


def project3d_matlab(p, theta, psi, shift, W, H, mag=1.0):
    """matlab/+artia/+geo/project3D.m:106-143 with recon2real.m (x_m = y, y_m = -x), phi = 0."""
    x, y, z = p[1], -p[0], p[2]
    c, s = math.cos(math.radians(psi)), math.sin(math.radians(psi))
    ct, st = math.cos(math.radians(theta)), math.sin(math.radians(theta))
    m11, m12, m13 = c / mag, (-s * ct) / mag, (-s * st) / mag
    m21, m22, m23 = s / mag, (c * ct) / mag, (c * st) / mag
    px = m11 * x + m12 * y + m13 * z + shift[0] + W / 2
    py = m21 * x + m22 * y + m23 * z + shift[1] + H / 2
    return px, py


@pytest.mark.parametrize("case", CASES[:5])
def test_detector_mapping_matches_matlab_project3d(case):
    W, H = 200, 160
    g = O.tilt_geometry(case["tilt"], W, H, psi_deg=case["psi"], shift=case["shift"], mag=case.get("mag", 1.0))
    M = np.array(g.det_matrix, np.float64).reshape(4, 4)
    rng = np.random.default_rng(7)
    for _ in range(10):
        p = rng.uniform(-60, 60, 3)
        r = M @ np.append(p, 1.0)
        px, py = project3d_matlab(p, case["tilt"], case["psi"], case["shift"], W, H, case.get("mag", 1.0))
        np.testing.assert_allclose([r[0], r[1]], [px, py], rtol=0, atol=2e-3)


def test_index_list_matches_libc_rand():
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    tilts = [-60.0 + 3 * i for i in range(41)]
    libc.srand(0)
    ref = list(range(len(tilts)))
    for i in range(len(ref) - 1):
        r = i + libc.rand() % (len(ref) - i)
        ref[i], ref[r] = ref[r], ref[i]
    z = min(range(len(tilts)), key=lambda i: (abs(tilts[i]), i))
    j = ref.index(z)
    ref[0], ref[j] = ref[j], ref[0]
    assert O.index_list(tilts) == ref
    assert A.projection_index_list(tilts) == ref
    libc.srand(0)
    assert O.glibc_rand(0, 50) == [libc.rand() for _ in range(50)]


def test_mag_aniso_closed_form():
    amount, ang, W, H = 1.05, 30.0, 128.0, 96.0
    g = O.tilt_geometry(0.0, int(W), int(H), mag_aniso=(amount, ang))
    A_ = np.array(g.mag_aniso, np.float64).reshape(3, 3)
    a = math.radians(ang)
    R = lambda t: np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0], [0, 0, 1]])
    T = lambda x, y: np.array([[1, 0, x], [0, 1, y], [0, 0, 1.0]])
    ref = T(W / 2, H / 2) @ R(-a) @ np.diag([amount, 1, 1]) @ R(a) @ T(-W / 2, -H / 2)
    np.testing.assert_allclose(A_, ref, atol=2e-5)
    Ai = np.array(g.mag_aniso_inv, np.float64).reshape(3, 3)
    np.testing.assert_allclose(Ai @ A_, np.eye(3), atol=5e-5)
    ident = O.tilt_geometry(10.0, 64, 64)
    assert list(ident.mag_aniso) == [1, 0, 0, 0, 1, 0, 0, 0, 1]
