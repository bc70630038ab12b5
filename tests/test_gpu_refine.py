"""GPU checks of the batched refinement pair sets (artia_pairs_*) through the C ABI:
  - the windowed projection equals artia_fp_march bit for bit at zero shift, and the FP64 oracle's
    shifted march (<= 1e-6 relative L2) at per-pair shifts;
  - the fused loss-and-gradient launch: residuals equal projection minus target exactly, the loss
    is their sum of squares, and the shift gradient matches FP64 central differences of the
    oracle's loss (h = 1e-6 px, <= 1e-5 relative) and the per-tilt shift-gradient kernel;
  - the batched adjoint passes the dot-product test with shifts (<= 1e-6) and matches the oracle's
    FP64 transpose per sub-volume (<= 1e-5);
  - the cross-correlation search returns the oracle's (oracle/cc_search.py, GetDisplacement)
    peaks exactly, with its correlation values and maps to 1e-5, with and without
    MultiPeakDetection (on projections, and on two-spot targets where the two searches differ);
  - the Gauss-Newton matrix J^T J of the same launch matches the oracle's FP64 Jacobian;
  - 12 fused evaluations (Levenberg-Marquardt) or 50 (Barzilai-Borwein) recover per-pair shifts
    from U[-3, 3] px, where the integer search is off by its quantisation.
"""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

import artiatomi_amd as A
import oracle_lib as O
from artiatomi_amd import autograd as AG
from conftest import rel_l2

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import cc_search  # noqa: E402

pytestmark = pytest.mark.gpu

W = H = 64
TILTS = (-48.0, -21.0, 0.0, 17.0, 44.0)
# (edge, 1-based centre position in the 64 x 64 x 24 volume)
SUBS = ((16, (22, 30, 12)), (12, (40, 24, 10)), (20, (30, 42, 13)))
MARGIN = 6


@pytest.fixture(scope="module")
def ctx(gpu):
    c = A.ArtiaContext(0)
    yield c
    c.close()


def blob(n, centre, sigma):
    g = (np.arange(n) + 0.5) / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    r2 = ((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2) / sigma ** 2
    return np.exp(-r2)


def phantom(n, seed):
    """Two Gaussian blobs, < 1e-7 at the faces (no sample-count jumps under small shifts)."""
    rng = np.random.default_rng(seed)
    v = np.zeros((n, n, n))
    for _ in range(2):
        v += rng.uniform(0.5, 1.0) * blob(n, rng.uniform(0.42, 0.58, 3), 0.11)
    return v.astype(np.float32)


def dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def setup(ctx):
    full = A.volume_geometry(64, 64, 24)
    vols, data, off = [], [], 0
    for k, (n, pos) in enumerate(SUBS):
        vg = A.subvolume_geometry(n, 1.0, (1.0, 1.0, 1.0), full, pos)
        vols.append((vg, off))
        data.append(phantom(n, 10 + k))
        off += n ** 3
    flat = np.concatenate([d.ravel() for d in data])
    pairs, geoms, meta = [], [], []
    for k, (vg, _) in enumerate(vols):
        for t in TILTS:
            g = A.tilt_geometry(t, W, H, shift=(0.7, -0.4))
            x0, y0, w, h = A.footprint_window(vg, g, W, H, MARGIN)
            assert w > 0 and h > 0
            pairs.append((k, x0, y0, w, h))
            geoms.append(g)
            meta.append((k, g))
    ps = A.PairSet(ctx, vols, pairs, geoms)
    yield dict(ps=ps, vols=vols, data=data, flat=flat, dflat=dev(flat), meta=meta, pairs=pairs)
    ps.close()


def window_of(img, pair):
    _, x0, y0, w, h = pair
    return img[y0:y0 + h, x0:x0 + w]


def test_project_equals_fp_march_at_zero_shift(ctx, setup):
    ps = setup["ps"]
    got = host(ps.project(setup["dflat"]))
    for p, ((k, g), pr) in enumerate(zip(setup["meta"], setup["pairs"])):
        vg = setup["vols"][k][0]
        proj, dist = dev(np.zeros((H, W))), dev(np.zeros((H, W)))
        _, x0, y0, w, h = pr
        ctx.fp_march(vg, g, dev(setup["data"][k]), proj, dist, roi=(x0, y0, x0 + w, y0 + h))
        ref = window_of(host(proj), pr)
        assert np.array_equal(got[pr_slice(ps, p)].reshape(h, w), ref), p


def test_project_shifted_vs_fp64_oracle(ctx, setup):
    ps = setup["ps"]
    rng = np.random.default_rng(7)
    s = rng.uniform(-3, 3, (ps.npairs, 2))
    got = host(ps.project(setup["dflat"], dev(s, torch.float64)))
    worst = 0.0
    for p, ((k, g), pr) in enumerate(zip(setup["meta"], setup["pairs"])):
        vg = setup["vols"][k][0]
        _, x0, y0, w, h = pr
        ref = O.march_shifted(vg, g, setup["data"][k].astype(np.float64), W, H, s[p, 0], s[p, 1],
                              roi=(x0, y0, x0 + w, y0 + h))
        e = rel_l2(got[pr_slice(ps, p)].reshape(h, w), window_of(ref, pr))
        worst = max(worst, e)
    print(f"pairs project vs FP64 shifted march: worst rel L2 {worst:.2e}", flush=True)
    assert worst <= 1e-6


def pr_slice(ps, p):
    _, _, _, w, h, off = ps.windows[p]
    return slice(off, off + w * h)


def test_loss_grad_vs_fd_and_per_tilt_kernel(ctx, setup):
    ps = setup["ps"]
    rng = np.random.default_rng(8)
    s_true = rng.uniform(-2, 2, (ps.npairs, 2))
    target = ps.project(setup["dflat"], dev(s_true, torch.float64))
    target = target + 0.01 * torch.from_numpy(rng.standard_normal(ps.n_pixels).astype(np.float32)).cuda()
    s = dev(s_true + rng.uniform(-0.8, 0.8, (ps.npairs, 2)), torch.float64)
    resid = torch.empty_like(target)
    loss, grad, jtj = ps.loss_grad(setup["dflat"], s, target, resid=resid, normal=True)
    loss2, grad2 = ps.loss_grad(setup["dflat"], s, target)
    assert torch.equal(loss, loss2) and torch.equal(grad, grad2)
    model = ps.project(setup["dflat"], s)
    assert torch.equal(resid, model - target)
    r64 = host(resid).astype(np.float64)
    tgt = host(target).astype(np.float64)
    for p in range(ps.npairs):
        assert abs(host(loss)[p] - (r64[pr_slice(ps, p)] ** 2).sum()) <= 1e-6 * host(loss)[p]
    # FP64 central differences of the oracle's loss, a few pairs
    sh, gh, nh = host(s), host(grad), host(jtj)
    hstep = 1e-6
    worst = worst_n = 0.0
    for p in (0, 3, 7, 11, 14):
        k, g = setup["meta"][p]
        vg = setup["vols"][k][0]
        _, x0, y0, w, h = setup["pairs"][p]
        v64 = setup["data"][k].astype(np.float64)
        b = tgt[pr_slice(ps, p)].reshape(h, w)
        fd, jac = [], []
        for ex, ey in ((1, 0), (0, 1)):
            ls, ps_ = [], []
            for sg in (1, -1):
                P = O.march_shifted(vg, g, v64, W, H, sh[p, 0] + sg * hstep * ex, sh[p, 1] + sg * hstep * ey,
                                    roi=(x0, y0, x0 + w, y0 + h))
                ps_.append(window_of(P, setup["pairs"][p]))
                ls.append(((ps_[-1] - b) ** 2).sum())
            fd.append((ls[0] - ls[1]) / (2 * hstep))
            jac.append(((ps_[0] - ps_[1]) / (2 * hstep)).ravel())
        fd = np.array(fd)
        e = np.abs(gh[p] - fd).max() / np.abs(fd).max()
        jx, jy = jac
        nfd = np.array([(jx * jx).sum(), (jx * jy).sum(), (jy * jy).sum()])
        en = np.abs(nh[p] - nfd).max() / np.abs(nfd).max()
        print(f"pair {p}: loss grad {gh[p]}, FP64 FD {fd}, rel {e:.2e}; J^T J {nh[p]} vs FD {nfd}, rel {en:.2e}",
              flush=True)
        worst = max(worst, e)
        worst_n = max(worst_n, en)
    assert worst <= 1e-5
    assert worst_n <= 1e-5
    # per-sample formula of the per-tilt kernel at zero shift, weights on the windows
    wts = torch.from_numpy(rng.standard_normal(ps.n_pixels).astype(np.float32)).cuda()
    gp = host(ps.shift_grad(setup["dflat"], None, wts))
    wh = host(wts)
    worst = 0.0
    for p, ((k, g), pr) in enumerate(zip(setup["meta"], setup["pairs"])):
        vg = setup["vols"][k][0]
        _, x0, y0, w, h = pr
        img = np.zeros((H, W), np.float32)
        img[y0:y0 + h, x0:x0 + w] = wh[pr_slice(ps, p)].reshape(h, w)
        out = torch.zeros(2, dtype=torch.float64, device="cuda")
        ctx.fp_shift_grad(vg, g, dev(setup["data"][k]), dev(img), out)
        ref = host(out)
        worst = max(worst, np.abs(gp[p] - ref).max() / np.abs(ref).max())
    print(f"pairs shift grad vs per-tilt kernel: worst rel {worst:.2e}", flush=True)
    assert worst <= 1e-5


def test_adjoint_dot_product_and_oracle(ctx, setup):
    ps = setup["ps"]
    rng = np.random.default_rng(9)
    s = dev(rng.uniform(-3, 3, (ps.npairs, 2)), torch.float64)
    x = torch.from_numpy(rng.standard_normal(setup["flat"].size).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.standard_normal(ps.n_pixels).astype(np.float32)).cuda()
    px = ps.project(x, s)
    pty = torch.zeros_like(x)
    ps.adjoint(s, y, pty)
    lhs = float((host(px).astype(np.float64) * host(y)).sum())
    rhs = float((host(x).astype(np.float64) * host(pty)).sum())
    err = abs(lhs - rhs) / abs(lhs)
    print(f"pairs dot test: <Px,y> = {lhs:.9e}, <x,P^T y> = {rhs:.9e}, rel {err:.2e}", flush=True)
    assert err <= 1e-6
    # zero shift: each sub-volume's gradient against the FP64 transpose summed over its tilts
    got = torch.zeros_like(x)
    ps.adjoint(None, y, got, scale=0.5)
    got = host(got)
    yh = host(y).astype(np.float64)
    for k, (vg, off) in enumerate(setup["vols"]):
        n = vg.nx
        ref = np.zeros((n, n, n))
        for p, ((kk, g), pr) in enumerate(zip(setup["meta"], setup["pairs"])):
            if kk != k:
                continue
            _, x0, y0, w, h = pr
            img = np.zeros((H, W))
            img[y0:y0 + h, x0:x0 + w] = yh[pr_slice(ps, p)].reshape(h, w)
            ref += O.march_adjoint(vg, g, img, scale=0.5, roi=(x0, y0, x0 + w, y0 + h))
        e = rel_l2(got[off:off + n ** 3].reshape(n, n, n), ref)
        print(f"pairs adjoint vs FP64 transpose, sub-volume {k}: {e:.2e}", flush=True)
        assert e <= 1e-5


def check_cc_search(ps, pairs, models, targets, max_shift, multi_peak):
    """The device search against the oracle's GetDisplacement on each pair's square; returns the peaks."""
    peak, cc, ccmap = ps.cc_search(models, targets, W, max_shift, want_map=True, multi_peak=multi_peak)
    peak, cc, ccmap = host(peak), host(cc), host(ccmap)
    th, mh = host(targets).astype(np.float64), host(models).astype(np.float64)
    for p, pr in enumerate(pairs):
        _, x0, y0, w, h = pr
        a = cc_search.to_square(th[pr_slice(ps, p)].reshape(h, w), W, x0, y0)
        b = cc_search.to_square(mh[pr_slice(ps, p)].reshape(h, w), W, x0, y0)
        rx, ry, val, full = cc_search.get_displacement(a, b, max_shift, multi_peak=multi_peak)
        assert (peak[p, 0], peak[p, 1]) == (rx, ry), (p, peak[p], (rx, ry))
        assert abs(cc[p] - val) <= 1e-5
        assert np.abs(ccmap[p] - cc_search.cc_map_of(full, max_shift)).max() <= 1e-5
    return peak


@pytest.mark.parametrize("multi_peak", [False, True])
@pytest.mark.parametrize("max_shift", [0, 4, 7])
def test_cc_search_matches_oracle(ctx, setup, max_shift, multi_peak):
    ps = setup["ps"]
    rng = np.random.default_rng(11 + max_shift)
    s_true = rng.uniform(-3, 3, (ps.npairs, 2))
    targets = ps.project(setup["dflat"], dev(s_true, torch.float64))
    targets = targets + 0.02 * torch.from_numpy(rng.standard_normal(ps.n_pixels).astype(np.float32)).cuda()
    models = ps.project(setup["dflat"])
    check_cc_search(ps, setup["pairs"], models, targets, max_shift, multi_peak)


def spot(h, w, cx, cy, s=0.8):
    y, x = np.mgrid[0:h, 0:w]
    return np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))


@pytest.mark.parametrize("amp", [0.95, 0.8])
def test_cc_search_multi_peak_two_copies(ctx, setup, amp):
    """Each target holds its model spot twice, moved by d1 (full strength) and by the nearer d2
    (strength amp): MultiPeakDetection picks -d2 when amp >= 0.9, else -d1 like the plain search."""
    ps = setup["ps"]
    d1s = ((5, -3), (-4, 4), (6, 1), (-2, -6))
    d2s = ((1, 2), (-1, 0), (0, -2), (2, 1))
    models = np.zeros(ps.n_pixels, np.float32)
    targets = np.zeros(ps.n_pixels, np.float32)
    for p, pr in enumerate(setup["pairs"]):
        _, x0, y0, w, h = pr
        cx, cy = w // 2, h // 2
        (ax, ay), (bx, by) = d1s[p % 4], d2s[p % 4]
        models[pr_slice(ps, p)] = spot(h, w, cx, cy).ravel()
        targets[pr_slice(ps, p)] = (spot(h, w, cx + ax, cy + ay) + amp * spot(h, w, cx + bx, cy + by)).ravel()
    m, t = dev(models), dev(targets)
    plain = check_cc_search(ps, setup["pairs"], m, t, 7, False)
    multi = check_cc_search(ps, setup["pairs"], m, t, 7, True)
    for p in range(ps.npairs):
        d = d2s[p % 4] if amp >= 0.9 else d1s[p % 4]
        assert tuple(plain[p]) == (-d1s[p % 4][0], -d1s[p % 4][1])
        assert tuple(multi[p]) == (-d[0], -d[1]), (p, multi[p])


def test_refinement_recovers_shifts(ctx, setup):
    ps = setup["ps"]
    rng = np.random.default_rng(12)
    s_true = rng.uniform(-3, 3, (ps.npairs, 2))
    targets = ps.project(setup["dflat"], dev(s_true, torch.float64))
    for method, steps, tol in (("gn", 12, 1e-4), ("bb", 50, 1e-2)):
        s = AG.refine_pair_shifts(setup["dflat"], targets, ps, steps=steps, method=method)
        err = np.abs(host(s) - s_true)
        print(f"gradient refinement ({method}, {steps} evaluations): max error {err.max():.2e} px", flush=True)
        assert err.max() <= tol
    sc, cc, failed = AG.cc_search_shifts(setup["dflat"], targets, ps, W, 5)
    assert not bool(host(failed).any())
    ec = np.abs(host(sc) - s_true)
    print(f"integer CC search: max error {ec.max():.3f} px, rms {np.sqrt((ec ** 2).mean()):.3f} px", flush=True)
    assert ec.max() <= 1.0


def test_autograd_pairs_matches_kernels(ctx, setup):
    ps = setup["ps"]
    rng = np.random.default_rng(13)
    s_true = rng.uniform(-2, 2, (ps.npairs, 2))
    targets = ps.project(setup["dflat"], dev(s_true, torch.float64))
    vol = setup["dflat"].clone().requires_grad_(True)
    s = dev(s_true + 0.5, torch.float64).requires_grad_(True)
    loss = ((AG.project_pairs(vol, s, ps) - targets) ** 2).sum()
    loss.backward()
    l_k, g_k = ps.loss_grad(setup["dflat"], s.detach().contiguous(), targets)
    assert abs(float(loss.detach()) - float(l_k.sum())) <= 1e-5 * float(l_k.sum())
    e = float((s.grad - g_k).abs().max() / g_k.abs().max())
    assert e <= 1e-6, e
    resid = ps.project(setup["dflat"], s.detach().contiguous()) - targets
    gv = torch.zeros_like(vol)
    ps.adjoint(s.detach().contiguous(), 2.0 * resid, gv)
    assert float((vol.grad - gv).abs().max()) <= 1e-6 * float(gv.abs().max())


def test_bad_arguments_raise(ctx, setup):
    vg = setup["vols"][0][0]
    g = setup["meta"][0][1]
    with pytest.raises(A.ArtiaError):
        A.PairSet(ctx, [(vg, 0)], [(0, 0, 0, 0, 5)], [g])
    with pytest.raises(A.ArtiaError):
        A.PairSet(ctx, [(vg, 0)], [(1, 0, 0, 4, 4)], [g])
    ps = setup["ps"]
    with pytest.raises(A.ArtiaError):
        ps.cc_search(ps.project(setup["dflat"]), ps.project(setup["dflat"]), W, 33)
