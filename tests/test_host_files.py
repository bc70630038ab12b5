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
