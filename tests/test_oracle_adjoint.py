"""Pins the oracle's transpose of the forward projector and its shifted forward projection:
the dot-product identity <P x, y> = <x, P^T y> against march itself in FP64, and the shift
identities (zero shift is march; an integer pixel shift at tilt 0 translates the image)."""
import numpy as np
import pytest

import oracle_lib as O

GEOMS = [(0.0, 0.0, (0.0, 0.0)), (33.0, 2.5, (1.2, -0.7)), (-57.0, -4.0, (-2.0, 3.0))]


@pytest.mark.parametrize("tilt,psi,shift", GEOMS)
def test_adjoint_dot_product_fp64(tilt, psi, shift):
    vg = O.volume_geometry(14, 12, 6)
    W, H = 20, 18
    g = O.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((6, 12, 14))
    y = rng.standard_normal((H, W))
    px, _ = O.march(vg, g, x, W, H, dtype=np.float64)
    pty = O.march_adjoint(vg, g, y, dtype=np.float64)
    lhs, rhs = float((px * y).sum()), float((x * pty).sum())
    assert abs(px).max() > 0.1
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    # the scale argument is linear
    scaled = O.march_adjoint(vg, g, y, dtype=np.float64, scale=-2.5)
    assert np.abs(scaled + 2.5 * pty).max() <= 1e-12 * np.abs(pty).max()


def test_shifted_march_identities():
    vg = O.volume_geometry(16, 16, 8)
    W, H = 24, 24
    rng = np.random.default_rng(6)
    x = rng.random((8, 16, 16))
    g = O.tilt_geometry(21.0, W, H, psi_deg=1.5)
    base, _ = O.march(vg, g, x, W, H, dtype=np.float64)
    np.testing.assert_array_equal(O.march_shifted(vg, g, x, W, H, 0.0, 0.0), base)
    g0 = O.tilt_geometry(0.0, W, H)
    p0, _ = O.march(vg, g0, x, W, H, dtype=np.float64)
    # shifting by one pixel along u samples, at pixel x, the rays of pixel x - 1
    ps = O.march_shifted(vg, g0, x, W, H, 1.0, 0.0)
    np.testing.assert_allclose(ps[:, 1:], p0[:, :-1], rtol=1e-12, atol=1e-12)
    ps = O.march_shifted(vg, g0, x, W, H, 0.0, -2.0)
    np.testing.assert_allclose(ps[:-2, :], p0[2:, :], rtol=1e-12, atol=1e-12)
