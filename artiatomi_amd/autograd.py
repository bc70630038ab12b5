"""Differentiable forward projection for refinement (volumes and per-tilt 2D shifts).

`project(vol, shifts, setup)` returns the projections P(vol, shifts) of one volume (a whole
tomogram or a motive-list sub-volume) through a tilt series, each tilt shifted by its row of
`shifts` (pixels, added to the tilt's marker shift as Projection::GetPosition adds extra shifts,
Projection.cpp:183-195).  The backward pass runs the library's hand-written kernels over the C
ABI: the exact transpose of the forward projector for the volume gradient (artia_fp_adjoint) and
the shift derivative (artia_fp_shift_grad), both recomputing the ray samples instead of storing
them.  The forward projector is the exact-numerics one of artia_fp_march; the HIP library is
required (there is no other implementation).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import torch

from . import core
from ._abi import GlobalParams, VolGeom


@dataclass
class ProjectionSetup:
    """Fixed part of the projection: volume placement, tilt series and detector size."""

    ctx: "core.ArtiaContext"
    vg: VolGeom
    width: int
    height: int
    tilts_deg: Sequence[float]
    psis_deg: Optional[Sequence[float]] = None
    base_shifts: Optional[Sequence[Sequence[float]]] = None
    glob: Optional[GlobalParams] = None
    roi: Optional[Sequence[int]] = None          # {x0, y0, x1, y1}; None: the whole image
    _dist: Optional[torch.Tensor] = field(default=None, repr=False)

    def geometries(self, shifts: torch.Tensor) -> list:
        s = shifts.detach().to("cpu", torch.float64).numpy()
        T = len(self.tilts_deg)
        if s.shape != (T, 2):
            raise core.ArtiaError(f"shifts must have shape ({T}, 2), got {tuple(s.shape)}")
        psis = self.psis_deg if self.psis_deg is not None else [0.0] * T
        base = self.base_shifts if self.base_shifts is not None else [(0.0, 0.0)] * T
        return [core.subvolume_tilt_geometry(float(t), self.width, self.height, psi_deg=float(p),
                                             shift=(float(b[0]), float(b[1])), extra_shift=(float(e[0]), float(e[1])),
                                             glob=self.glob)
                for t, p, b, e in zip(self.tilts_deg, psis, base, s)]

    def scratch(self) -> torch.Tensor:
        if self._dist is None:
            self._dist = torch.empty((self.height, self.width), device="cuda")
        return self._dist


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol: torch.Tensor, shifts: torch.Tensor, setup: ProjectionSetup):
        if setup.ctx.numerics != "exact":
            raise core.ArtiaError("the differentiable projector needs an ARTIA_NUMERICS_EXACT context")
        if vol.dtype != torch.float32:
            # the adjoint and shift-gradient kernels read fp32 volumes only
            raise core.ArtiaError(f"the differentiable projector takes a float32 volume, got {vol.dtype}")
        vol = vol.contiguous()
        geoms = setup.geometries(shifts)
        T = len(geoms)
        out = torch.zeros((T, setup.height, setup.width), device=vol.device, dtype=torch.float32)
        dist = setup.scratch()
        for t, g in enumerate(geoms):
            setup.ctx.fp_march(setup.vg, g, vol, out[t], dist, roi=setup.roi)
        ctx.save_for_backward(vol, shifts)
        ctx.geoms = geoms
        ctx.setup = setup
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        vol, shifts = ctx.saved_tensors
        setup: ProjectionSetup = ctx.setup
        g = grad_out.contiguous().to(torch.float32)
        gvol = torch.zeros_like(vol) if ctx.needs_input_grad[0] else None
        gshift = None
        if ctx.needs_input_grad[1]:
            gshift = torch.zeros((len(ctx.geoms), 2), device=vol.device, dtype=torch.float64)
        for t, geo in enumerate(ctx.geoms):
            if gvol is not None:
                setup.ctx.fp_adjoint(setup.vg, geo, g[t], gvol, roi=setup.roi)
            if gshift is not None:
                setup.ctx.fp_shift_grad(setup.vg, geo, vol, g[t], gshift[t], roi=setup.roi)
        if gshift is not None:
            gshift = gshift.to(shifts.dtype)
        return gvol, gshift, None


def project(vol: torch.Tensor, shifts: torch.Tensor, setup: ProjectionSetup) -> torch.Tensor:
    """Projections (T, H, W) of `vol` ([z][y][x] float32, CUDA) with per-tilt shifts (T, 2)."""
    return _Project.apply(vol, shifts, setup)


def _bb_descent(grad_at, s0: torch.Tensor, steps: int, first_step_px: float, max_step_px: float,
                callback=None) -> torch.Tensor:
    """Per-row Barzilai-Borwein descent on a separable objective: `steps` evaluations of
    grad_at(s) -> (n, 2); the first step is `first_step_px` along each row's gradient, every step
    is capped at `max_step_px`.  callback(k, s) sees the iterate after evaluation k."""
    s = s0.detach().clone()
    g = grad_at(s)
    if callback:
        callback(1, s)
    gn = g.norm(dim=1, keepdim=True)
    alpha = torch.where(gn > 0, first_step_px / gn.clamp_min(1e-300), torch.zeros_like(gn))
    for k in range(steps - 1):
        d = alpha * g
        dn = d.norm(dim=1, keepdim=True)
        d = torch.where(dn > max_step_px, d * (max_step_px / dn.clamp_min(1e-300)), d)
        s_new = s - d
        g_new = grad_at(s_new)
        ds, dg = s_new - s, g_new - g
        den = (ds * dg).sum(1, keepdim=True)
        alpha = torch.where(den > 0, (ds * ds).sum(1, keepdim=True) / den.clamp_min(1e-300), alpha)
        s, g = s_new, g_new
        if callback:
            callback(k + 2, s)
    return s


def refine_shifts(vol: torch.Tensor, target: torch.Tensor, setup: ProjectionSetup, steps: int = 50,
                  init: Optional[torch.Tensor] = None, first_step_px: float = 0.1,
                  max_step_px: float = 2.0) -> torch.Tensor:
    """Gradient-based replacement of EmSARTRefine's cross-correlation shift search: minimises
    ||P(vol, s) - target||^2 over the per-tilt shifts s with `steps` gradient evaluations.  A tilt's
    shift changes only its own projection, so the problem separates per tilt; each tilt takes
    Barzilai-Borwein steps (first step `first_step_px` along its gradient, every step capped at
    `max_step_px`)."""
    T = len(setup.tilts_deg)
    s0 = init.clone() if init is not None else torch.zeros((T, 2), dtype=torch.float64, device="cuda")

    def grad_at(x):
        x = x.clone().requires_grad_(True)
        ((project(vol, x, setup) - target) ** 2).sum().backward()
        return x.grad.detach()

    return _bb_descent(grad_at, s0, steps, first_step_px, max_step_px)


class _ProjectPairs(torch.autograd.Function):
    @staticmethod
    def forward(ctx, volumes: torch.Tensor, shifts: torch.Tensor, pairs: "core.PairSet"):
        volumes = volumes.contiguous()
        s = shifts.detach().to(torch.float64).contiguous()
        out = pairs.project(volumes, s)
        ctx.save_for_backward(volumes, s)
        ctx.pairs = pairs
        ctx.shift_dtype = shifts.dtype
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        volumes, s = ctx.saved_tensors
        pairs: core.PairSet = ctx.pairs
        g = grad_out.contiguous().to(torch.float32)
        gvol = gshift = None
        if ctx.needs_input_grad[0]:
            gvol = torch.zeros_like(volumes)
            pairs.adjoint(s, g, gvol)
        if ctx.needs_input_grad[1]:
            gshift = pairs.shift_grad(volumes, s, g).to(ctx.shift_dtype)
        return gvol, gshift, None


def project_pairs(volumes: torch.Tensor, shifts: torch.Tensor, pairs: "core.PairSet") -> torch.Tensor:
    """Packed window images of every (sub-volume, tilt) pair of `pairs` (a core.PairSet): sub-volumes
    in the flat float32 buffer `volumes`, per-pair 2D shifts (npairs, 2).  Differentiable in both."""
    return _ProjectPairs.apply(volumes, shifts, pairs)


def _lm_descent(eval_at, s0: torch.Tensor, steps: int, mu0: float = 1e-3, callback=None) -> torch.Tensor:
    """Per-row Levenberg-Marquardt on a separable least-squares objective: `steps` evaluations of
    eval_at(s) -> (loss (n,), grad (n, 2) = 2 J^T r, J^T J (n, 3) as xx, xy, yy).  A row's step
    solves (J^T J + mu diag(J^T J)) d = -J^T r; it is kept when the loss drops (mu / 3), otherwise
    discarded (mu * 4).  callback(k, s) sees the best iterate after evaluation k."""
    s = s0.detach().clone()
    loss, g, n = eval_at(s)
    mu = torch.full_like(loss, mu0)
    if callback:
        callback(1, s)
    for k in range(steps - 1):
        a = n[:, 0] * (1.0 + mu)
        b = n[:, 1]
        c = n[:, 2] * (1.0 + mu)
        det = a * c - b * b
        ok = det > 0
        det = torch.where(ok, det, torch.ones_like(det))
        hx, hy = 0.5 * g[:, 0], 0.5 * g[:, 1]
        d = torch.stack([(c * hx - b * hy) / det, (a * hy - b * hx) / det], dim=1)
        d = torch.where(ok[:, None], d, torch.zeros_like(d))
        s_new = s - d
        l2, g2, n2 = eval_at(s_new)
        acc = l2 < loss
        s = torch.where(acc[:, None], s_new, s)
        loss = torch.where(acc, l2, loss)
        g = torch.where(acc[:, None], g2, g)
        n = torch.where(acc[:, None], n2, n)
        mu = torch.where(acc, mu / 3.0, mu * 4.0)
        if callback:
            callback(k + 2, s)
    return s


def refine_pair_shifts(volumes: torch.Tensor, targets: torch.Tensor, pairs: "core.PairSet", steps: int = 50,
                       init: Optional[torch.Tensor] = None, method: str = "gn", first_step_px: float = 0.1,
                       max_step_px: float = 2.0, callback=None) -> torch.Tensor:
    """Per-(sub-volume, tilt) shifts minimising ||P_s - target||^2 of every pair, with `steps`
    evaluations of the fused loss-and-gradient launch (artia_pairs_loss_grad).  method "gn":
    per-pair Levenberg-Marquardt on the launch's gradient 2 J^T r and Gauss-Newton matrix J^T J
    (J = dP_s/ds, first derivatives only); "bb": per-pair Barzilai-Borwein steps as
    refine_shifts.  callback(k, s) sees the iterate after evaluation k."""
    s0 = init.to(torch.float64) if init is not None else torch.zeros((pairs.npairs, 2), dtype=torch.float64,
                                                                     device=volumes.device)
    if method == "gn":
        return _lm_descent(lambda x: pairs.loss_grad(volumes, x.contiguous(), targets, normal=True), s0, steps,
                           callback=callback)
    if method == "bb":
        return _bb_descent(lambda x: pairs.loss_grad(volumes, x.contiguous(), targets)[1], s0, steps, first_step_px,
                           max_step_px, callback=callback)
    raise core.ArtiaError(f"unknown refinement method {method!r} (gn or bb)")


def cc_search_shifts(volumes: torch.Tensor, targets: torch.Tensor, pairs: "core.PairSet", square: int,
                     max_shift: int, init: Optional[torch.Tensor] = None, multi_peak: bool = False):
    """EmSARTRefine's search (GetDisplacement, Reconstructor.cpp:1746-1959) for every pair at once:
    the models are the pairs projected at `init` (zero by default); a pair whose target matches its
    model moved by d gets the shift init + d (integer d within the max_shift disk).  Pairs whose
    search fails keep `init`.  multi_peak selects MultiPeakDetection (Reconstructor.cpp:1887-1935).  Returns (shifts (npairs, 2) float64, peak CC (npairs,), failed mask)."""
    s0 = init.to(torch.float64).contiguous() if init is not None else torch.zeros(
        (pairs.npairs, 2), dtype=torch.float64, device=volumes.device)
    models = pairs.project(volumes, s0)
    peak, cc, _ = pairs.cc_search(models, targets, square, max_shift, multi_peak=multi_peak)
    failed = peak[:, 0] == -1000
    s = torch.where(failed[:, None], s0, s0 - peak.to(torch.float64))
    return s, cc, failed


__all__ = ["ProjectionSetup", "cc_search_shifts", "project", "project_pairs", "refine_pair_shifts",
           "refine_shifts"]
