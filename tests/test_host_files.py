"""Host readers of the reference's data carriers (no GPU): the MarkerFile item layout
(MarkerFile.cpp:1216-1221, items MarkerFile.h:34-46, CheckIfProjIndexIsGood :143-146) and the
CtfFile rows (CtfFile.cpp:37-56), on arrays laid out as the files' float data; the EM and MRC
headers (EmHeader.h, mrcHeader.h) against layouts restated with struct, and the volume headers
EmSART writes (EmSart.cpp:717-759)."""
import math
import struct

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
        A.marker_tilt_params(np.zeros((2, 4, 10), np.float32), 0, ref_marker=2)  # marker out of range
    with pytest.raises(A.ArtiaError):
        A.ctf_file_defocus(np.zeros((4, 3), np.float32), 0, 0.0)   # needs the angle row


# ---------------------------------------------------------------- EM and MRC headers
# The headers are restated here from the formats' published layouts with struct, independently of
# the library: EM (EmHeader.h: 4 coding bytes, DimX/Y/Z, 80-byte comment, 40 ints, 256 bytes of
# user data) and MRC (mrcHeader.h / MRC2014: NX NY NZ MODE at 0-15, MX MY MZ at 28, Xlen at 40,
# MAPC/R/S at 64, NEXT at 92, 1024 bytes in all).


def em_header(nx, ny, nz, dtype_code, coding=6, object_pixel_size=0):
    h = bytearray(512)
    h[0], h[3] = coding, dtype_code
    struct.pack_into("<3i", h, 4, nx, ny, nz)
    struct.pack_into("<i", h, 96 + 6 * 4, object_pixel_size)   # Voltage, Cs, Aperture, Mag, PostMag, Exp
    return bytes(h)


def mrc_header(nx, ny, nz, mode, xlen=0.0, next_bytes=0, order="<"):
    h = bytearray(1024)
    struct.pack_into(order + "4I", h, 0, nx, ny, nz, mode)
    struct.pack_into(order + "3I", h, 28, nx, ny, nz)
    struct.pack_into(order + "3f", h, 40, xlen, xlen, xlen)
    struct.pack_into(order + "3I", h, 64, 1, 2, 3)
    struct.pack_into(order + "I", h, 92, next_bytes)
    return bytes(h)


@pytest.mark.parametrize("code,dtype,nbytes", [(1, -1, 1), (2, 0, 2), (4, 2, 4), (5, 4, 4), (8, -1, 8),
                                                (9, -1, 8), (10, -1, 2)])
def test_em_header_types(code, dtype, nbytes):
    info = A.parse_file_header("em", em_header(37, 21, 5, code))
    assert (info.nx, info.ny, info.nz) == (37, 21, 5)
    assert (info.dtype, info.raw_type, info.elem_bytes, info.data_offset) == (dtype, code, nbytes, 512)
    assert info.pixel_size_nm == 1.0 and info.byte_swapped == 0


def test_em_header_rejects_non_pc_coding_and_unknown_types():
    with pytest.raises(A.ArtiaError):   # EmFile.cpp:75-78
        A.parse_file_header("em", em_header(8, 8, 1, 5, coding=3))
    with pytest.raises(A.ArtiaError):
        A.parse_file_header("em", em_header(8, 8, 1, 3))
    with pytest.raises(A.ArtiaError):   # too short
        A.parse_file_header("em", em_header(8, 8, 1, 5)[:400])


@pytest.mark.parametrize("mode,dtype,nbytes", [(0, -1, 1), (1, 0, 2), (2, 4, 4), (4, -1, 8), (6, 1, 2),
                                                (12, -1, 2), (64, -1, 2)])
@pytest.mark.parametrize("order", ["<", ">"])
def test_mrc_header_types_offsets_and_byte_order(mode, dtype, nbytes, order):
    info = A.parse_file_header("mrc", mrc_header(4096, 3000, 61, mode, xlen=4096 * 2.62, next_bytes=61 * 128,
                                                 order=order))
    assert (info.nx, info.ny, info.nz) == (4096, 3000, 61)
    assert (info.dtype, info.raw_type, info.elem_bytes) == (dtype, mode, nbytes)
    assert info.data_offset == 1024 + 61 * 128
    assert info.byte_swapped == (1 if order == ">" else 0)
    # MRCFile::GetPixelsize: Xlen / NX / 10 in float
    want = np.float32(np.float32(np.float32(4096 * 2.62) / np.float32(4096)) / np.float32(10))
    assert info.pixel_size_nm == pytest.approx(float(want), rel=0, abs=0)


def test_mrc_header_without_xlen_has_no_pixel_size():
    assert A.parse_file_header("mrc", mrc_header(64, 64, 3, 2)).pixel_size_nm == 0.0


@pytest.mark.parametrize("half", [False, True])
def test_volume_headers_as_emsart_writes_them(half):
    nx, ny, nz, ps, vs = 1024, 900, 256, 0.262, (2.0, 2.0, 2.0)
    f32 = np.float32
    m = A.volume_file_header("mrc", nx, ny, nz, ps, vs, half=half)
    assert len(m) == 1024
    assert struct.unpack_from("<4I", m, 0) == (nx, ny, nz, 64 if half else 2)
    assert struct.unpack_from("<3I", m, 28) == (nx, ny, nz)
    xl = struct.unpack_from("<3f", m, 40)
    for i, d in enumerate((nx, ny, nz)):
        assert xl[i] == f32(f32(f32(f32(ps) * f32(d)) * f32(vs[i])) * f32(10))
    assert struct.unpack_from("<3I", m, 64) == (1, 2, 3)
    assert not any(m[76:]) and not any(m[16:28])
    info = A.parse_file_header("mrc", m)
    assert (info.nx, info.ny, info.nz, info.data_offset) == (nx, ny, nz, 1024)
    e = A.volume_file_header("em", nx, ny, nz, ps, vs, half=half)
    assert len(e) == 512 and e[0] == 6 and e[3] == (10 if half else 5)
    assert struct.unpack_from("<3i", e, 4) == (nx, ny, nz)
    assert struct.unpack_from("<i", e, 120)[0] == int(f32(f32(ps) * f32(vs[0])) * f32(1000))
    assert sum(e) == 6 + e[3] + sum(e[4:16]) + sum(e[120:124])


def test_read_file_em_and_mrc(tmp_path):
    rng = np.random.default_rng(3)
    a = rng.integers(-3000, 3000, (4, 6, 10)).astype("<i2")
    p = tmp_path / "stack.mrc"
    p.write_bytes(mrc_header(10, 6, 4, 1, xlen=25.0, next_bytes=256) + bytes(256) + a.tobytes())
    info, d = A.read_file(p)
    assert info.dtype == 0 and info.pixel_size_nm == pytest.approx(0.25)
    assert np.array_equal(d, a)
    b = rng.random((3, 5, 7)).astype("<f4")
    q = tmp_path / "vol.em"
    q.write_bytes(em_header(7, 5, 3, 5) + b.tobytes())
    info, d = A.read_file(q)
    assert info.dtype == 4 and np.array_equal(d, b)
    # a big-endian MRC stack reads with its byte order
    c = rng.integers(0, 60000, (2, 3, 4)).astype(">u2")
    r = tmp_path / "be.mrc"
    r.write_bytes(mrc_header(4, 3, 2, 6, order=">") + c.tobytes())
    info, d = A.read_file(r)
    assert info.byte_swapped == 1 and np.array_equal(d.astype(np.uint16), c.astype(np.uint16))


# ---------------------------------------------------------------- ShiftFile and MotiveList
# ShiftFile (ShiftFile.cpp:68-84): EM float data of DimX = motives, DimY = projections, DimZ = 2;
# operator() negates.  MotiveList (MotiveListe.h:81-106, MotiveList.cpp): rows of 20 floats, GetAt
# scales the 2D and 3D coordinates by the click binning and the shifts by the shift binning.

MOTIVE_FIELDS = ["cc_coeff", "x_coord_2d", "y_coord_2d", "part_nr", "tomo_nr", "part_nr_in_tomo", "wedge_idx",
                 "x", "y", "z", "x_shift", "y_shift", "z_shift", "x_shift_before", "y_shift_before",
                 "z_shift_before", "phi", "psi", "theta", "class_no"]


def test_shift_file_layout_and_sign(tmp_path):
    P, M = 7, 5
    rng = np.random.default_rng(11)
    sx, sy = rng.uniform(-4, 4, (P, M)).astype("<f4"), rng.uniform(-4, 4, (P, M)).astype("<f4")
    p = tmp_path / "shifts.em"
    p.write_bytes(em_header(M, P, 2, 5) + sx.tobytes() + sy.tobytes())   # DimX motives, DimY projections
    info, d = A.read_file(p)
    assert (info.nx, info.ny, info.nz) == (M, P, 2)
    for proj, mot in ((0, 0), (3, 4), (6, 2)):
        assert A.shift_file_get(d, proj, mot) == (-float(sx[proj, mot]), -float(sy[proj, mot]))
    out = np.zeros((2, P, M), np.float32)
    A.shift_file_set(out, 5, 1, (1.25, -2.5))
    assert out[0, 5, 1] == 1.25 and out[1, 5, 1] == -2.5 and np.count_nonzero(out) == 2
    assert A.shift_file_get(out, 5, 1) == (-1.25, 2.5)
    for bad in ((P, 0), (0, M), (-1, 0)):
        with pytest.raises(A.ArtiaError):
            A.shift_file_get(d, *bad)
    with pytest.raises(A.ArtiaError):
        A.shift_file_get(np.zeros((1, P, M), np.float32), 0, 0)   # one component only


def motive_rows(n, seed, classes=None):
    rng = np.random.default_rng(seed)
    rows = rng.uniform(-20, 20, (n, 20)).astype(np.float32)
    rows[:, 7:10] = rng.uniform(1, 100, (n, 3)).astype(np.float32)
    rows[:, 19] = classes if classes is not None else rng.integers(0, 3, n)
    return rows


def test_motive_list_row_layout_and_binning(tmp_path):
    rows = motive_rows(9, 4)
    p = tmp_path / "motl.em"
    p.write_bytes(em_header(20, 9, 1, 5) + rows.astype("<f4").tobytes())
    ml = A.MotiveList.from_file(p, binning_click=2.0, binning_shift=0.5)
    assert len(ml) == 9
    f32 = np.float32
    for i in (0, 4, 8):
        m = ml[i]
        for k, name in enumerate(MOTIVE_FIELDS):
            scale = 2.0 if name in ("x_coord_2d", "y_coord_2d", "x", "y", "z") else \
                0.5 if name in ("x_shift", "y_shift", "z_shift") else 1.0
            assert getattr(m, name) == float(f32(rows[i, k]) * f32(scale)), name
    with pytest.raises(A.ArtiaError):
        ml[9]
    with pytest.raises(A.ArtiaError):
        A.MotiveList(np.zeros((3, 19), np.float32))


def float_dist(a, b):
    d = (a - b).astype(np.float32)
    return np.float32(np.sqrt(np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])))


def test_motive_groups_by_class():
    # class numbers in file order 2, 0, 2.5 (truncated to 2 but compared as float: a class of its own
    # whose members are those equal to 2.0), -1 (no members)
    cls = np.array([2, 0, 2, 2.5, -1, 0, 2], np.float32)
    ml = A.MotiveList(motive_rows(len(cls), 5, cls))
    assert ml.group_count("by_class") == 4
    assert ml.group("by_class", 0) == [0, 2, 6]
    assert ml.group("by_class", 1) == [1, 5]
    assert ml.group("by_class", 2) == [0, 2, 6]
    assert ml.group("by_class", 3) == []
    assert ml.group("by_class", 7) == []


def test_motive_groups_by_distance_and_count():
    rows = motive_rows(40, 6)
    rows[17, 7:10] = rows[3, 7:10]   # a particle at the same place as particle 3
    ml = A.MotiveList(rows, binning_click=1.5)
    pos = rows[:, 7:10] * np.float32(1.5)
    assert ml.group_count("max_dist") == 40 and ml.group_count("max_count") == 40
    for g in (0, 3, 17, 39):
        d = np.array([float_dist(pos[i], pos[g]) for i in range(40)], np.float32)
        want = [g] + [i for i in range(40) if i != g and d[i] <= np.float32(30.0)]
        assert ml.group("max_dist", g, max_distance=30.0) == want
        order = sorted(range(40), key=lambda i: (d[i], i))
        assert ml.group("max_count", g, group_size=6) == order[:6]
    # particle 3 and 17 coincide: the (distance, index) order puts 3 first in both groups
    assert ml.group("max_count", 17, group_size=2) == [3, 17]
    assert ml.group("max_count", 5, group_size=100) == sorted(range(40), key=lambda i: (
        float_dist(pos[i], pos[5]), i))


def test_subvolume_job_from_motive_list_and_shift_file():
    full = A.volume_geometry(256, 256, 64)
    acq = A.Acquisition.build(full, 128, 128, [-30.0, 0.0, 30.0], shifts=[(1.0, -2.0), (0, 0), (0.5, 0.25)])
    rows = motive_rows(4, 7)
    rows[:, 10:13] = np.float32([[1, 0, 0], [0, -1, 0], [0, 0, 2], [0.5, 0.5, 0.5]])
    ml = A.MotiveList(rows)
    shifts = np.zeros((2, 3, 4), np.float32)
    shifts[0, 1, 2], shifts[1, 1, 2] = 3.0, 4.0    # |shift| = 5 for motive 2 at projection 1
    vgs, geoms, active = A.subvolume_job(full, acq, ml, shifts, 32, 1.0, max_shift=5.4)
    assert len(vgs) == 4 and len(geoms) == 12
    assert active == [True] * 7 + [False] + [True] * 4       # 5 < 5.4 - 0.5 fails for pair (2, 1)
    m = ml[1]
    ref = A.subvolume_geometry(32, 1.0, (1.0, 1.0, 1.0), full, (m.x, m.y, m.z), (m.x_shift, m.y_shift, m.z_shift))
    assert bytes(vgs[1]) == bytes(ref)
    g = A.subvolume_tilt_geometry(0.0, 128, 128, shift=(0.0, 0.0), extra_shift=(-3.0, -4.0))
    assert bytes(geoms[2 * 3 + 1]) == bytes(g)
