"""The identities the forward projector's cell-wise summation and its adjoint rest on, checked in
FP64 on random cells: the m samples a_j = a0 + j da (likewise b, c) of a trilinear cell interpolant
sum to  m f(centre) + S2c (AB da db + AC da dc + BC db dc + ABC (a db dc + b da dc + c da db))
with S2c = m (m^2 - 1) / 12 about the centre sample, and the same sum is sum_c v_c W_c with the
corner weights W_c = m w_c(centre) + S2c kappa_c that the adjoint scatters."""
import numpy as np


def direct(v, a, b, c):
    v000, v100, v010, v110, v001, v101, v011, v111 = v
    return (v000 * (1 - a) * (1 - b) * (1 - c) + v100 * a * (1 - b) * (1 - c) + v010 * (1 - a) * b * (1 - c)
            + v110 * a * b * (1 - c) + v001 * (1 - a) * (1 - b) * c + v101 * a * (1 - b) * c
            + v011 * (1 - a) * b * c + v111 * a * b * c)


def test_cell_sum_closed_form_and_transpose():
    rng = np.random.default_rng(1)
    worst = 0.0
    for _ in range(3000):
        v = rng.standard_normal(8)
        v000, v100, v010, v110, v001, v101, v011, v111 = v
        m = int(rng.integers(1, 13))
        da, db, dc = rng.uniform(-0.15, 0.15, 3)
        a0, b0, c0 = rng.uniform(0, 1, 3)
        j = np.arange(m)
        ref = direct(v, a0 + j * da, b0 + j * db, c0 + j * dc).sum()
        kc = 0.5 * (m - 1)
        a, b, c = a0 + kc * da, b0 + kc * db, c0 + kc * dc
        s2c = m * (m * m - 1) / 12
        assert s2c == float((m * (m * m - 1)) // 6) * 0.5          # the kernel's exact integer form
        AB = (v110 - v100) - (v010 - v000)
        AC = (v101 - v100) - (v001 - v000)
        BC = (v011 - v010) - (v001 - v000)
        ABC = ((v111 - v110) - (v101 - v100)) - ((v011 - v010) - (v001 - v000))
        pab, pac, pbc = da * db, da * dc, db * dc
        q = a * pbc + b * pac + c * pab
        closed = m * direct(v, a, b, c) + s2c * (AB * pab + AC * pac + BC * pbc + ABC * q)
        a1, b1, c1 = 1 - a, 1 - b, 1 - c
        W = [m * a1 * b1 * c1 + s2c * (pab + pac + pbc - q), m * a * b1 * c1 + s2c * (-pab - pac + q),
             m * a1 * b * c1 + s2c * (-pab - pbc + q), m * a * b * c1 + s2c * (pab - q),
             m * a1 * b1 * c + s2c * (-pac - pbc + q), m * a * b1 * c + s2c * (pac - q),
             m * a1 * b * c + s2c * (pbc - q), m * a * b * c + s2c * q]
        worst = max(worst, abs(closed - ref), abs(float(np.dot(W, v)) - ref))
    assert worst < 1e-12
