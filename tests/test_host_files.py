"""Host readers of the reference's data carriers (no GPU): the MarkerFile item layout
(MarkerFile.cpp:1216-1221, items MarkerFile.h:34-46, CheckIfProjIndexIsGood :143-146) and the
CtfFile rows (CtfFile.cpp:37-56), on arrays laid out as the files' float data."""
import math

import numpy as np
import pytest

import artiatomi_amd as A


def marker_array(markers, projections, items=10, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(-50.0, 50.0, (markers, projections, items)).astype(np.float32)


def test_marker_items_of_reference_marker():
    m = marker_array(3, 7)
    for ref in (0, 2):
        for p in (0, 4, 6):
            m[ref, p, 1] = 12.0   # valid coordinates
            m[ref, p, 2] = 3.0
            tp, good = A.marker_tilt_params(m, p, ref_marker=ref)
            flat = m.reshape(-1)
            dx, dy = m.shape[2], m.shape[1]
            item = lambda i: flat[ref * dx * dy + p * dx + i]
            assert good
            assert tp.tilt_deg == item(0) and tp.shift_x == item(4) and tp.shift_y == item(5)
            assert tp.mag == item(8) and tp.psi_deg == item(9)


@pytest.mark.parametrize("x,y,good", [(-1.0, 5.0, False), (5.0, -1.0, False), (-0.5, 0.0, True), (-2.0, -3.0, False)])
def test_projection_good_flag(x, y, good):
    m = marker_array(1, 4)
    m[0, 2, 1], m[0, 2, 2] = x, y
    assert A.marker_tilt_params(m, 2)[1] is good


def test_marker_file_with_more_items():
    # files may carry more than the ten items per projection; the stride is dim_x
    m = marker_array(2, 5, items=12)
    tp, _ = A.marker_tilt_params(m, 3, ref_marker=1)
    assert tp.psi_deg == m[1, 3, 9] and tp.tilt_deg == m[1, 3, 0]


def test_ctf_file_rows():
    rows, n = 6, 9
    d = np.arange(rows * n, dtype=np.float32).reshape(rows, n) * 10.0
    d[4] = np.linspace(-1.2, 1.5, n).astype(np.float32)   # astigmatism angle in radians
    for i in (0, 5, 8):
        r = A.ctf_file_defocus(d, i, tilt_deg=-42.0)
        assert r.defocus_min_nm == d[1, i] and r.defocus_max_nm == d[2, i]
        assert r.astig_angle_deg == pytest.approx(math.degrees(float(d[4, i])), rel=1e-6)
        assert r.tilt_deg == -42.0


def test_readers_reject_bad_shapes():
    with pytest.raises(A.ArtiaError):
        A.marker_tilt_params(np.zeros((1, 4, 9), np.float32), 0)   # fewer than ten items
    with pytest.raises(A.ArtiaError):
        A.marker_tilt_params(np.zeros((1, 4, 10), np.float32), 4)  # projection out of range
    with pytest.raises(A.ArtiaError):
        A.ctf_file_defocus(np.zeros((4, 3), np.float32), 0, 0.0)   # needs the angle row
