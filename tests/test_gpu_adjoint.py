"""GPU checks of the differentiable projector (adjoint and shift gradient of the exact forward
projector) through the C ABI:
  - dot-product test <P x, y> = <x, P^T y> with the GPU forward projector and its GPU adjoint,
    <= 1e-6 relative;
  - the GPU adjoint against the oracle's FP64 transpose of its forward projector, <= 1e-5;
  - the shift gradient against FP64 central finite differences of the oracle's shifted forward
    projection: <= 1e-5 relative at h = 1e-6 px (converged to a few 1e-6), <= 1e-3 at h = 1e-3 px
    (whose own truncation across the trilinear kinks is a few 1e-4);
  - the autograd Function's gradients equal the direct kernel calls, the volume gradient points
    along SART's backprojection of the same residual, and 50 gradient steps recover per-tilt
    shifts drawn from U[-3, 3] px.
"""
import numpy as np
import pytest
import torch

import artiatomi_amd as A
import oracle_lib as O
from artiatomi_amd import autograd as AG
from conftest import rel_l2

pytestmark = pytest.mark.gpu

GEOMS = [(0.0, 0.0, (0.0, 0.0)), (28.0, 2.0, (1.5, -0.5)), (-55.0, -3.5, (-2.0, 2.5)), (47.0, 0.0, (0.0, 0.0))]


@pytest.fixture(scope="module")
def ctx(gpu):
    c = A.ArtiaContext(0)
    yield c
    c.close()


def blob(nx, ny, nz, centre=(0.5, 0.5, 0.5), sigma=0.18, zfac=1.6):
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx,
                          indexing="ij")
    r2 = ((x - centre[0]) ** 2 + (y - centre[1]) ** 2) / sigma ** 2 + (z - centre[2]) ** 2 / (zfac * sigma) ** 2
    return np.exp(-r2).astype(np.float32)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("tilt,psi,shift", GEOMS)
def test_dot_product(ctx, tilt, psi, shift):
    print(f"dot test tilt={tilt} psi={psi}", flush=True)
    vg = A.volume_geometry(48, 40, 16)
    W, H = 64, 56
    g = A.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((16, 40, 48)).astype(np.float32)
    y = rng.standard_normal((H, W)).astype(np.float32)
    px, dist = dev(np.zeros((H, W))), dev(np.zeros((H, W)))
    ctx.fp_march(vg, g, dev(x), px, dist)
    pty = dev(np.zeros((16, 40, 48)))
    ctx.fp_adjoint(vg, g, dev(y), pty)
    lhs = float((host(px).astype(np.float64) * y).sum())
    rhs = float((x.astype(np.float64) * host(pty)).sum())
    err = abs(lhs - rhs) / abs(lhs)
    print(f"  <Px,y> = {lhs:.9e}, <x,P^T y> = {rhs:.9e}, rel = {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("tilt,psi,shift", GEOMS[1:3])
def test_adjoint_matches_oracle(ctx, tilt, psi, shift):
    vg = A.volume_geometry(32, 28, 12)
    W, H = 40, 36
    g = A.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
    y = np.random.default_rng(2).standard_normal((H, W)).astype(np.float32)
    got = dev(np.zeros((12, 28, 32)))
    ctx.fp_adjoint(vg, g, dev(y), got, scale=0.5)
    ref = O.march_adjoint(vg, g, y.astype(np.float64), dtype=np.float64, scale=0.5)
    err = rel_l2(host(got), ref)
    print(f"adjoint vs FP64 oracle tilt={tilt}: {err:.2e}", flush=True)
    assert err <= 1e-5


@pytest.mark.parametrize("tilt,psi,shift", GEOMS)
def test_shift_gradient_vs_fd(ctx, tilt, psi, shift):
    vg = A.volume_geometry(40, 40, 16)
    W, H = 56, 56
    g = A.tilt_geometry(tilt, W, H, psi_deg=psi, shift=shift)
    # the phantom vanishes at the volume faces (< 1e-7): a sample entering or leaving a sub-ray under
    # a shift is a discontinuity of the discretised projector that finite differences would pick up
    vol = blob(40, 40, 16, centre=(0.47, 0.53, 0.5), sigma=0.12, zfac=1.0)
    gout = np.random.default_rng(3).standard_normal((H, W)).astype(np.float32)
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    ctx.fp_shift_grad(vg, g, dev(vol), dev(gout), out)
    got = host(out)
    v64, g64 = vol.astype(np.float64), gout.astype(np.float64)
    # P is piecewise linear in the shift (trilinear interpolation), so a central difference over
    # +-1e-3 px straddles kinks and carries a few 1e-4 of truncation at tilted geometries; over
    # +-1e-6 px it agrees with +-1e-7 px to a few 1e-6 (its remaining kink crossings)
    for h, tol in ((1e-6, 1e-5), (1e-3, 1e-3)):
        fd = []
        for ex, ey in ((1, 0), (0, 1)):
            lp = (O.march_shifted(vg, g, v64, W, H, h * ex, h * ey) * g64).sum()
            lm = (O.march_shifted(vg, g, v64, W, H, -h * ex, -h * ey) * g64).sum()
            fd.append((lp - lm) / (2 * h))
        fd = np.array(fd)
        err = np.abs(got - fd).max() / np.abs(fd).max()
        print(f"shift gradient tilt={tilt} h={h}: gpu {got}, fd {fd}, rel {err:.2e}", flush=True)
        assert err <= tol, (h, err)


def series_setup(ctx, n=40, nz=16, W=56, tilts=(-54, -36, -18, 0, 18, 36, 54)):
    vg = A.volume_geometry(n, n, nz)
    return AG.ProjectionSetup(ctx, vg, W, W, list(map(float, tilts)))


def test_autograd_matches_kernels_and_sart_direction(ctx):
    setup = series_setup(ctx)
    vol = dev(blob(40, 40, 16))
    T = len(setup.tilts_deg)
    shifts = torch.zeros((T, 2), dtype=torch.float64, device="cuda", requires_grad=True)
    x = vol.clone().requires_grad_(True)
    target = AG.project(dev(blob(40, 40, 16, centre=(0.55, 0.45, 0.5))), torch.zeros((T, 2), dtype=torch.float64,
                        device="cuda"), setup)
    p = AG.project(x, shifts, setup)
    resid = p - target
    (resid ** 2).sum().backward()
    geoms = setup.geometries(torch.zeros((T, 2), dtype=torch.float64))
    gv = torch.zeros_like(vol)
    gs = torch.zeros((T, 2), dtype=torch.float64, device="cuda")
    g2 = (2 * resid).detach().contiguous()
    for t, geo in enumerate(geoms):
        ctx.fp_adjoint(setup.vg, geo, g2[t], gv)
        ctx.fp_shift_grad(setup.vg, geo, vol, g2[t], gs[t])
    assert rel_l2(host(x.grad), host(gv)) <= 1e-6
    assert rel_l2(host(shifts.grad), host(gs)) <= 1e-12
    # SART's update for the same residual: the lambda-1 backprojection of -resid / chord summed over
    # the tilts points along the negative volume gradient
    bp = dev(np.zeros((16, 40, 40)))
    for t, geo in enumerate(geoms):
        ctx.bp(setup.vg, geo, bp, (-resid[t]).detach().contiguous(), 1.0, oversampling=2)
    a, b = host(bp).ravel().astype(np.float64), -host(x.grad).ravel().astype(np.float64)
    cos = float(a @ b / np.linalg.norm(a) / np.linalg.norm(b))
    print(f"cos(SART update, -gradient) = {cos:.4f}", flush=True)
    assert cos > 0.9


def test_gradient_shift_refinement(ctx):
    setup = series_setup(ctx)
    vol = dev(blob(40, 40, 16, sigma=0.14))
    T = len(setup.tilts_deg)
    truth = torch.from_numpy(np.random.default_rng(7).uniform(-3, 3, (T, 2))).cuda()
    with torch.no_grad():
        target = AG.project(vol, truth, setup)
    est = AG.refine_shifts(vol, target, setup, steps=50)
    err = (est - truth).abs().max().item()
    print(f"max shift error after 50 steps: {err:.4f} px", flush=True)
    assert err < 0.05
