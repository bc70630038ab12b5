"""Pins the oracle of EmSARTRefine's shift search (oracle/cc_search.py, a restatement of
Reconstructor::GetDisplacement, Reconstructor.cpp:1746-1959) with checks independent of it."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import cc_search  # noqa: E402


def brute_cc(A, B):
    """Direct circular correlation of the mean-free squares, normalised: sum_x A'(x) B'(x + r)."""
    N = A.shape[0]
    a = A - A.mean()
    b = B - B.mean()
    out = np.zeros((N, N))
    for ry in range(N):
        for rx in range(N):
            out[ry, rx] = (a * np.roll(np.roll(b, -ry, axis=0), -rx, axis=1)).sum()
    return out / np.sqrt((a * a).sum() * (b * b).sum())


def blob(N, cx, cy, s=2.0):
    y, x = np.mgrid[0:N, 0:N]
    return np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))


def test_matches_brute_force_correlation():
    rng = np.random.default_rng(3)
    N, M = 12, 4
    A = rng.standard_normal((N, N))
    B = rng.standard_normal((N, N))
    rx, ry, val, cc = cc_search.get_displacement(A, B, M)
    ref = brute_cc(A, B)
    idx = np.arange(N)
    m = np.minimum(idx, N - idx).astype(np.float32)
    inside = np.sqrt(m[None, :] ** 2 + m[:, None] ** 2, dtype=np.float32) <= M
    assert np.allclose(cc[inside], ref[inside], atol=1e-12)
    assert np.all(cc[~inside] == 0)
    # the maximum over the disk, wrapped to a signed shift
    masked = np.where(inside, ref, -np.inf)
    y, x = np.unravel_index(np.argmax(masked), masked.shape)
    if masked[y, x] > 0:
        assert (rx, ry) == ((x - N if x > N // 2 else x), (y - N if y > N // 2 else y))
        assert abs(val - masked[y, x]) < 1e-12
    else:
        assert (rx, ry) == (-1000, -1000)


def test_recovers_circular_shift_exactly():
    rng = np.random.default_rng(4)
    N, M = 32, 5
    B = rng.standard_normal((N, N))
    for dx, dy in [(0, 0), (3, -2), (-4, 1), (2, 4), (-3, -3)]:
        A = np.roll(np.roll(B, dy, axis=0), dx, axis=1)   # A(x) = B(x - d)
        rx, ry, val, _ = cc_search.get_displacement(A, B, M)
        assert (rx, ry) == (-dx, -dy)
        assert abs(val - 1.0) < 1e-12


def test_failure_value_for_anticorrelated_target():
    N = 32
    B = blob(N, 16, 16, 2.0)
    rx, ry, val, _ = cc_search.get_displacement(-B, B, 4)
    assert (rx, ry) == (-1000, -1000)
    assert val <= 0


def test_disk_excludes_corners():
    N, M = 32, 3
    B = blob(N, 16, 16, 1.0)
    A = blob(N, 16 - 3, 16 - 3, 1.0)   # true lag (3, 3): sqrt(18) > 3, outside the disk
    rx, ry, _, _ = cc_search.get_displacement(A, B, M)
    assert np.hypot(rx, ry) <= M
    assert (rx, ry) in [(2, 2), (3, 1), (1, 3), (3, 0), (0, 3)]


def test_ties_go_to_first_in_row_major_order():
    N = 16
    B = np.zeros((N, N))
    B[8, 8] = 1.0
    A = np.zeros((N, N))
    A[8, 6] = 1.0    # lag r = (2, 0): row 0 of the square, column 2
    A[6, 8] = 1.0    # lag r = (0, 2): row 2 of the square, column 0
    rx, ry, _, cc = cc_search.get_displacement(A, B, 3)
    assert cc[0, 2] == cc[2, 0]
    assert (rx, ry) == (2, 0)
    A = np.zeros((N, N))
    A[8, 10] = 1.0   # lag (-2, 0): row 0, column N - 2
    A[10, 8] = 1.0   # lag (0, -2): row N - 2, column 0
    rx, ry, _, _ = cc_search.get_displacement(A, B, 3)
    assert (rx, ry) == (-2, 0)


def test_window_sum_equals_square_correlation():
    """The windowed form the device search evaluates,
    CC(r) = (sum_y B(y) A(y - r) - sum A sum B / N^2) / sqrt(SumA SumB), with both images zero
    outside one window, equals the square's circular correlation when w, h <= N - 2 M."""
    rng = np.random.default_rng(5)
    N, M, w, h = 40, 4, 20, 17
    Aw = rng.standard_normal((h, w))
    Bw = rng.standard_normal((h, w))
    _, _, _, cc = cc_search.get_displacement(cc_search.to_square(Aw, N, 9, 13), cc_search.to_square(Bw, N, 9, 13), M)
    n2 = float(N * N)
    sA, sB = Aw.sum(), Bw.sum()
    sumA = (Aw * Aw).sum() - sA * sA / n2
    sumB = (Bw * Bw).sum() - sB * sB / n2
    win = cc_search.cc_map_of(cc, M)
    for ry in range(-M, M + 1):
        for rx in range(-M, M + 1):
            if np.sqrt(np.float32(rx * rx + ry * ry), dtype=np.float32) > M:
                continue
            c = 0.0
            for y in range(h):
                for x in range(w):
                    if 0 <= y - ry < h and 0 <= x - rx < w:
                        c += Bw[y, x] * Aw[y - ry, x - rx]
            assert abs((c - sA * sB / n2) / np.sqrt(sumA * sumB) - win[ry + M, rx + M]) < 1e-12


def _map(N, entries):
    cc = np.zeros((N, N), np.float32)
    for x, y, v in entries:
        cc[y % N, x % N] = v
    return cc


def test_multi_peak_takes_nearest_peak_above_nine_tenths():
    N, M = 16, 6
    assert cc_search.multi_peak_index(_map(N, [(5, 0, 1.0), (1, 1, 0.95)]), 1.0, M) == (1, 1)
    assert cc_search.multi_peak_index(_map(N, [(5, 0, 1.0), (1, 1, 0.85)]), 1.0, M) == (5, 0)
    assert cc_search.multi_peak_index(_map(N, [(5, 0, 1.0), (-1, -1, 0.95)]), 1.0, M) == (N - 1, N - 1)
    # outside the disk a peak weighs nothing: sqrt(50) > 6
    assert cc_search.multi_peak_index(_map(N, [(5, 0, 1.0), (5, 5, 0.99)]), 1.0, M) == (5, 0)
    # the threshold is maxVal * 0.9f, reached with equality
    t = np.float32(np.float32(1.0) * np.float32(0.9))
    assert cc_search.multi_peak_index(_map(N, [(5, 0, 1.0), (1, 0, t)]), 1.0, M) == (1, 0)
    below = np.nextafter(t, np.float32(0))
    assert cc_search.multi_peak_index(_map(N, [(5, 0, 1.0), (1, 0, below)]), 1.0, M) == (5, 0)


def test_multi_peak_neighbour_rule_is_asymmetric():
    """findPeak compares the two pixels of row y - 1 at x -/+ 1 with <=, all other neighbours with <."""
    N, M = 16, 6
    # equal pair on the main diagonal: only the lower one, whose equal neighbour is upper-left, is a peak
    assert cc_search.multi_peak_index(_map(N, [(3, 3, 1.0), (4, 4, 1.0)]), 1.0, M) == (4, 4)
    # equal pair on the anti-diagonal: only (3, 4), whose equal neighbour is upper-right
    assert cc_search.multi_peak_index(_map(N, [(4, 3, 1.0), (3, 4, 1.0)]), 1.0, M) == (3, 4)
    # equal pair in a row or a column: neither is a peak, the farther single peak wins
    assert cc_search.multi_peak_index(_map(N, [(1, 1, 1.0), (2, 1, 1.0), (0, 5, 0.95)]), 1.0, M) == (0, 5)
    assert cc_search.multi_peak_index(_map(N, [(1, 1, 1.0), (1, 2, 1.0), (0, 5, 0.95)]), 1.0, M) == (0, 5)
    # neighbours wrap around the square: (0, 2) sits right of (-1, 2)
    assert cc_search.multi_peak_index(_map(N, [(5, 0, 1.0), (0, 2, 0.95), (-1, 2, 0.97)]), 1.0, M) == (N - 1, 2)
    # equal distance: the first in row-major order
    assert cc_search.multi_peak_index(_map(N, [(5, 0, 1.0), (0, 2, 0.95), (2, 0, 0.95)]), 1.0, M) == (2, 0)


def test_get_displacement_multi_peak_on_two_copies():
    """A target holding the model twice, moved by d1 (full strength) and by d2 (0.95): the plain
    search finds -d1, MultiPeakDetection the nearer -d2; at 0.8 strength the second copy's peak
    drops below 0.9 of the maximum."""
    N, M = 48, 8
    B = blob(N, 24, 24, 1.0)
    d1, d2 = (5, -3), (1, 2)
    for amp, want in ((0.95, d2), (0.8, d1)):
        A = np.roll(np.roll(B, d1[1], axis=0), d1[0], axis=1) + amp * np.roll(np.roll(B, d2[1], axis=0), d2[0], axis=1)
        rx, ry, val, _ = cc_search.get_displacement(A, B, M)
        assert (rx, ry) == (-d1[0], -d1[1])
        mx, my, mval, _ = cc_search.get_displacement(A, B, M, multi_peak=True)
        assert (mx, my) == (-want[0], -want[1])
        assert mval == val
    rx, ry, _, _ = cc_search.get_displacement(-B, B, 4, multi_peak=True)
    assert (rx, ry) == (-1000, -1000)
