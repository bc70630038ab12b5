"""Oracle kernel bodies against checks that do not reuse their formulas:
  march / slicer / volTraversal
    - FP64 brute force with scipy.ndimage.map_coordinates (independent trilinear) and a
      Liang-Barsky clip about the box centre (independent ray/box intersection);
    - constant-volume closed form proj = c * delta * sum_k count_k / 4;
    - linearity in the volume;
    - float vs double: the documented step bias of the float accumulation t += 0.15 at
      |t| ~ 1e4, bounded and visible;
    - slicer over a partition of slabs equals march minus the detector-point first samples.
  compare / cropBorder / makeSquare: closed-form special values, numpy.pad index maps.
  backProjection
    - constant residual r: every voxel that any sub-ray hits gets exactly lambda * r;
    - FP64 brute force: numpy Liang-Barsky chords over the global sub-ray lattice (the footprint
      is complete for psi = 0 away from the image edge), weighted mean per voxel.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle_lib as O
from conftest import rel_l2


def clip_line(origin, direction, lo, hi, unit=True):
    """Parameter interval of the line origin + t*direction inside [lo, hi], computed about the
    box centre (Liang-Barsky).  unit=True normalises the direction (t is a length, as in the BP
    chord); unit=False keeps it (the marchers step t along the float-normalised n as given)."""
    d = np.asarray(direction, np.float64)
    if unit:
        d = d / np.linalg.norm(d)
    c = (np.asarray(lo, np.float64) + np.asarray(hi, np.float64)) / 2
    h = (np.asarray(hi, np.float64) - np.asarray(lo, np.float64)) / 2
    o = np.asarray(origin, np.float64) - c
    t0, t1 = -np.inf, np.inf
    for i in range(3):
        if d[i] == 0.0:
            if not (-h[i] < o[i] < h[i]):
                return 0.0, None
            continue
        a, b = (-h[i] - o[i]) / d[i], (h[i] - o[i]) / d[i]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    return max(0.0, t1 - t0), (t0, t1)


def g64(g):
    return {k: np.array(getattr(g, k), np.float64) for k in ("d", "u", "v", "n")}


def brute_march(vg, g, vol, W, H):
    G = g64(g)
    lo, hi = np.array(vg.bbox_min, np.float64), np.array(vg.bbox_max, np.float64)
    vs = np.array(vg.voxel_size, np.float64)
    proj = np.zeros((H, W))
    dist = np.zeros((H, W))
    step = vs[0] * 0.15
    for y in range(H):
        for x in range(W):
            acc, hits, length = 0.0, 0, 0.0
            for ox in (0.25, 0.75):
                for oy in (0.25, 0.75):
                    c = G["d"] + (x + ox) * G["u"] + (y + oy) * G["v"]
                    chord, ts = clip_line(c, G["n"], lo, hi, unit=False)
                    if chord <= 0:
                        continue
                    nn = G["n"]
                    t_in = ts[0]
                    hits += 1
                    length += chord
                    k = np.arange(int(math.floor(chord / step)) + 1)
                    pts = c[None, :] + (t_in + k[:, None] * step) * nn[None, :]
                    # voxel coordinate with the uploaded float c_volumeBBoxRcp (Kernels.cpp:996)
                    scale = np.array(vg.bbox_rcp, np.float64) * np.array(vg.dim, np.float64)
                    q = (pts - lo[None, :]) * scale[None, :] - 0.5
                    vals = ndimage.map_coordinates(vol.astype(np.float64), [q[:, 2], q[:, 1], q[:, 0]], order=1,
                                                   mode="nearest")
                    acc += vals.sum() * step
            proj[y, x] = acc * 0.25
            dist[y, x] = length / hits if hits else 0.0
    return proj, dist


GEOMS = [(0.0, 0.0, (0.0, 0.0)), (27.0, 0.0, (1.3, -0.7)), (-58.0, 3.1, (-2.0, 1.5)), (2.0, -4.0, (0.0, 0.0))]


@pytest.mark.parametrize("tilt,psi,shift", GEOMS)
def test_march_fp64_matches_bruteforce(tilt, psi, shift):
    W = H = 12
    vg = O.volume_geometry(8, 8, 6)
    g = O.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
    vol = np.random.default_rng(3).random((6, 8, 8))
    p, d = O.march(vg, g, vol, W, H, dtype=np.float64)
    pb, db = brute_march(vg, g, vol, W, H)
    assert rel_l2(p, pb) < 1e-9
    assert rel_l2(d, db) < 1e-9


def test_march_constant_volume_closed_form():
    W = H = 48
    vg = O.volume_geometry(32, 32, 12)
    vol = np.full((12, 32, 32), 2.5)
    g = O.tilt_geometry(33.0, W, H, psi_deg=1.5, shift=(0.4, -0.3))
    p, d = O.march(vg, g, vol, W, H, dtype=np.float64)
    G = g64(g)
    step = 0.15
    ref = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            cnt = 0
            for ox in (0.25, 0.75):
                for oy in (0.25, 0.75):
                    c = G["d"] + (x + ox) * G["u"] + (y + oy) * G["v"]
                    chord, _ = clip_line(c, G["n"], vg.bbox_min, vg.bbox_max, unit=False)
                    if chord > 0:
                        cnt += int(math.floor(chord / step)) + 1
            ref[y, x] = 2.5 * step * cnt * 0.25
    assert rel_l2(p, ref) < 1e-12


def test_march_linearity_fp64():
    W = H = 24
    vg = O.volume_geometry(16, 16, 8)
    g = O.tilt_geometry(-41.0, W, H, psi_deg=2.0)
    r = np.random.default_rng(5)
    a, b = r.random((8, 16, 16)), r.random((8, 16, 16))
    pa, _ = O.march(vg, g, a, W, H, dtype=np.float64)
    pb, _ = O.march(vg, g, b, W, H, dtype=np.float64)
    pab, _ = O.march(vg, g, 2.0 * a - 3.0 * b, W, H, dtype=np.float64)
    assert rel_l2(pab, 2.0 * pa - 3.0 * pb) < 1e-12


def test_march_float_step_bias_is_bounded_and_visible():
    # 64 voxels thick, tilt 60: ~850 samples per ray; float t accumulation at |t|~1e4 uses a
    # step of 154 ulp = 0.1504 instead of 0.15, so the float march is ~0.26 % low.
    W = H = 32
    vg = O.volume_geometry(32, 32, 64)
    vol = np.ones((64, 32, 32))
    g = O.tilt_geometry(60.0, W, H)
    p32, _ = O.march(vg, g, vol.astype(np.float32), W, H, dtype=np.float32)
    p64, _ = O.march(vg, g, vol, W, H, dtype=np.float64)
    mask = p64 > 1
    ratio = p32[mask].sum() / p64[mask].sum()
    assert 0.995 < ratio < 0.9995


def test_vol_traversal_closed_form():
    W, H = 40, 40
    vg = O.volume_geometry(24, 24, 10)
    for tilt, psi in [(0.0, 0.0), (45.0, 2.0), (-60.0, -1.0)]:
        g = O.tilt_geometry(tilt, W, H, psi_deg=psi)
        d = O.vol_traversal(vg, g, W, H, dtype=np.float64)
        G = g64(g)
        ref = np.zeros((H, W))
        for y in range(H):
            for x in range(W):
                c = G["d"] + (x + 0.5) * G["u"] + (y + 0.5) * G["v"]
                ref[y, x] = clip_line(c, G["n"], vg.bbox_min, vg.bbox_max, unit=False)[0]
        assert rel_l2(d, ref) < 1e-12
    g0 = O.tilt_geometry(0.0, W, H)
    d0 = O.vol_traversal(vg, g0, W, H, dtype=np.float32)
    assert d0[20, 20] == 10.0 and d0[0, 0] == 0.0


def test_slicer_first_sample_quirk():
    """slicer over an all-covering slab = march with its entry sample replaced by a sample at the
    detector point (ForwardProjectionSlicer.cu:129-133); with zero boundary faces both read 0."""
    W = H = 20
    vg = O.volume_geometry(12, 12, 8)
    g = O.tilt_geometry(20.0, W, H)
    vol = np.random.default_rng(9).random((8, 12, 12))
    p, _ = O.march(vg, g, vol, W, H, dtype=np.float64)
    s = O.slicer(vg, g, vol, W, H, -1e9, 1e9, dtype=np.float64)
    G = g64(g)
    lo, hi = np.array(vg.bbox_min, np.float64), np.array(vg.bbox_max, np.float64)
    scale = np.array(vg.bbox_rcp, np.float64) * np.array(vg.dim, np.float64)

    def tex(pt):
        q = (pt - lo) * scale - 0.5
        return ndimage.map_coordinates(vol, [[q[2]], [q[1]], [q[0]]], order=1, mode="nearest")[0]

    expect = p.copy()
    for y in range(H):
        for x in range(W):
            for ox in (0.25, 0.75):
                for oy in (0.25, 0.75):
                    c = G["d"] + (x + ox) * G["u"] + (y + oy) * G["v"]
                    chord, ts = clip_line(c, G["n"], lo, hi, unit=False)
                    if chord > 0:
                        expect[y, x] += (tex(c) - tex(c + ts[0] * G["n"])) * 0.15 * 0.25
    assert rel_l2(s, expect) < 1e-9
    vz = np.zeros_like(vol)
    vz[1:-1, 1:-1, 1:-1] = vol[1:-1, 1:-1, 1:-1]
    pz, _ = O.march(vg, g, vz, W, H, dtype=np.float64)
    sz = O.slicer(vg, g, vz, W, H, -1e9, 1e9, dtype=np.float64)
    assert rel_l2(sz, pz) < 1e-12


def test_compare_special_values():
    H, W = 8, 10
    real = np.random.default_rng(1).random((H, W))
    virt = np.random.default_rng(2).random((H, W))
    L = 5.0
    dist = np.full((H, W), L)
    r = O.compare(real, virt, dist, L, dtype=np.float64)
    np.testing.assert_allclose(r, (real - virt) / L, atol=1e-14)   # dist == L: r = (b - p) / L
    r0 = O.compare(real, virt, np.full((H, W), 0.5), L, dtype=np.float64)
    assert np.all(r0 == 0)
    # border weights: bottom rows ramp 1-exp(-9 w^2), top rows are a zero band
    cut, dim = (0, 0, 1, 2), (0, 0, 2, 2)
    rb = O.compare(real, virt, dist, L, crop=cut, crop_dim=dim, dtype=np.float64)
    w_row1 = 0.0
    w_row2 = 1 - math.exp(-9 * 0.25)
    res = (real - virt) / L
    np.testing.assert_allclose(rb[1], res[1] * w_row1, atol=1e-14)
    np.testing.assert_allclose(rb[2], res[2] * w_row2, atol=1e-14)
    np.testing.assert_allclose(rb[3], res[3], atol=1e-14)
    assert np.all(rb[H - 4:] == 0)   # y > H - dim.w - cut.w - 1 -> weight 0


def test_crop_border_quadrilateral():
    H, W = 20, 20
    img = np.ones((H, W))
    pts = [2, 17, 17, 17, 2, 2, 17, 2]   # p1 (2,17) p2 (17,17) p3 (2,2) p4 (17,2)
    out = O.crop_border(img, pts, dtype=np.float64)
    inside = np.zeros((H, W), bool)
    inside[3:17, 3:17] = True
    assert np.all(out[inside] == 1.0) and np.all(out[~inside] == 0.0)
    # dimming: distance to the nearest edge 1 -> 1 - exp(-9 (1/2)^2) with dim 2
    out2 = O.crop_border(img, pts, cut=(0, 0), dim=(2, 2), dtype=np.float64)
    assert out2[10, 10] == 1.0
    np.testing.assert_allclose(out2[10, 3], 1 - math.exp(-9 * 0.25), atol=1e-14)


def test_make_square_index_maps():
    img = np.arange(6 * 10, dtype=np.float64).reshape(6, 10)
    sq = O.make_square(img, 10, 0, 2, fill_zero=True, dtype=np.float64)
    np.testing.assert_array_equal(sq, np.pad(img, ((2, 2), (0, 0))))
    sqm = O.make_square(img, 10, 0, 2, fill_zero=True, mirror_y=True, dtype=np.float64)
    np.testing.assert_array_equal(sqm, np.pad(img[::-1], ((2, 2), (0, 0))))
    sqr = O.make_square(img, 10, 0, 2, fill_zero=False, dtype=np.float64)
    np.testing.assert_array_equal(sqr, np.pad(img, ((2, 2), (0, 0)), mode="symmetric"))


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_bp_constant_residual_invariant(dtype, tol):
    W = H = 40
    vg = O.volume_geometry(16, 16, 8)
    for tilt, psi, os_ in [(0.0, 0.0, 1), (31.0, 2.5, 2), (-59.0, -4.0, 3)]:
        g = O.tilt_geometry(tilt, W, H, psi_deg=psi)
        vol0 = np.random.default_rng(4).random((8, 16, 16))
        out = O.back_project(vg, g, vol0, np.full((H, W), 1.75), 0.3, os_=os_, dtype=dtype)
        delta = out - vol0.astype(dtype)
        hit = np.abs(delta) > 0
        assert hit.mean() > 0.9
        lam = float(np.float32(0.3))   # lambda crosses the interface as a float (Kernels.cpp:339)
        np.testing.assert_allclose(delta[hit], lam * 1.75, rtol=tol, atol=tol)


def brute_bp(vg, g, vol, img, lam, os_):
    """Weighted mean over the global sub-ray lattice (px + (i+0.5)/OS) with Liang-Barsky chords;
    valid when the quad footprint is complete (psi = 0) and the volume projects inside the image."""
    G = g64(g)
    nz, ny, nx = vol.shape
    out = vol.astype(np.float64).copy()
    H, W = img.shape
    M = np.array(g.det_matrix, np.float64).reshape(4, 4)
    lo0 = np.array(vg.bbox_min, np.float64)
    vs = np.array(vg.voxel_size, np.float64)
    sub = (np.arange(os_) + 0.5) / os_
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                lo = lo0 + np.array([x, y, z]) * vs
                hi = lo + vs
                corners = np.array([[lo[0] if i == 0 else hi[0], lo[1] if j == 0 else hi[1], lo[2] if k == 0 else hi[2], 1]
                                    for i in (0, 1) for j in (0, 1) for k in (0, 1)])
                pc = corners @ M.T
                num = den = 0.0
                for py in range(int(math.floor(pc[:, 1].min())), int(math.ceil(pc[:, 1].max()))):
                    for sy in sub:
                        for px in range(int(math.floor(pc[:, 0].min())), int(math.ceil(pc[:, 0].max()))):
                            for sx in sub:
                                r0 = G["d"] + (px + sx) * G["u"] + (py + sy) * G["v"]
                                c, _ = clip_line(r0, G["n"], lo, hi)
                                if c > 0:
                                    num += img[py, px] * c
                                    den += c
                if den > 0:
                    out[z, y, x] += lam * num / den
    return out


@pytest.mark.parametrize("tilt,os_", [(0.0, 1), (23.0, 2), (-47.0, 2), (58.0, 1)])
def test_bp_fp64_matches_bruteforce(tilt, os_):
    W = H = 32
    vg = O.volume_geometry(8, 8, 4)
    g = O.tilt_geometry(tilt, W, H, shift=(0.37, -0.61))
    vol0 = np.random.default_rng(11).random((4, 8, 8))
    img = np.random.default_rng(12).random((H, W))
    out = O.back_project(vg, g, vol0, img, 0.5, os_=os_, dtype=np.float64)
    ref = brute_bp(vg, g, vol0, img, 0.5, os_)
    assert rel_l2(out - vol0, ref - vol0) < 1e-9


def test_bp_float_vs_double_reference_noise():
    # the float kernel body carries the reference's own rounding (ray origin r0 + 1e5*n in
    # float): its update differs from the double evaluation by ~1e-4..1e-2, never more
    W = H = 48
    vg = O.volume_geometry(24, 24, 8)
    g = O.tilt_geometry(35.0, W, H, psi_deg=1.0)
    img = np.random.default_rng(2).random((H, W)).astype(np.float32)
    v0 = np.zeros((8, 24, 24), np.float32)
    o32 = O.back_project(vg, g, v0, img, 1.0, os_=2, dtype=np.float32)
    o64 = O.back_project(vg, g, v0, img, 1.0, os_=2, dtype=np.float64)
    e = rel_l2(o32, o64)
    assert 1e-6 < e < 3e-2
