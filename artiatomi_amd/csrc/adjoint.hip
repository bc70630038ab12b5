// Adjoint and shift gradient of the forward projector (exact numerics), for gradient-based
// refinement.  No reference counterpart: the reference's backprojector (BackProjectionSquareOS.cu)
// is SART's normalised update, not the transpose of its ray marcher.  Here
//   P x [pixel] = 0.25 * step * sum over the pixel's 4 sub-rays and their samples q_k of trilinear(x, q_k)
// is the forward projector of fp.hip (march_exact_kernel), and
//   fp_adjoint:    vol_grad += scale * P^T r, the transpose of the forward projector's per-cell
//                  closed form along the same sub-rays and cells (recomputed, never stored);
//   fp_shift_grad: out += sum_pixels g[pixel] * dP[pixel]/d(shift_x, shift_y), where the tilt's 2D
//                  shift moves the detector origin by -shift_x*u - shift_y*v (Projection.cpp:183-195)
//                  and with it every sample: dq_k/ds = scale * (-u + (dt_in/ds) n) per sub-ray.
#include "artia_internal.hpp"
#include "fp_common.hpp"
#include "tube_walk.hpp"

namespace artia {
namespace {

struct Roi {
    int x0, y0, x1, y1;
};

// Cell and weights of a sample at coordinates already shifted by -0.5, with the forward projector's
// clamp (coordinate in [0, n-1], lower texel at most n-2).
struct Cell {
    int ib, jb, kb;
    float a, b, c;
};

__device__ __forceinline__ Cell cell_of(const VolG& v, float xb, float yb, float zb) {
    Cell q;
    const float xc = __builtin_amdgcn_fmed3f(xb, 0.0f, (float)(v.nx - 1));
    const float yc = __builtin_amdgcn_fmed3f(yb, 0.0f, (float)(v.ny - 1));
    const float zc = __builtin_amdgcn_fmed3f(zb, 0.0f, (float)(v.nz - 1));
    q.ib = min((int)xc, v.nx - 2);
    q.jb = min((int)yc, v.ny - 2);
    q.kb = min((int)zc, v.nz - 2);
    q.a = xc - (float)q.ib;
    q.b = yc - (float)q.jb;
    q.c = zc - (float)q.kb;
    return q;
}

// One lane per sub-ray, the forward projector's mapping (8x8 pixels per 256-thread block).  The
// lane walks the same cells as the forward projector's march_sum and scatters the transpose of its
// per-cell closed form (cell_weights) with one set of eight atomics per cell.
template <bool kBig>
__global__ void __launch_bounds__(256) fp_adjoint_kernel(VolG v, TiltG g, const float* __restrict__ resid,
                                                         size_t pitch, Roi roi, float scale,
                                                         float* __restrict__ grad) {
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int sub = lane & 3, pix = lane >> 2;
    const int x = roi.x0 + (int)blockIdx.x * 8 + (wave & 1) * 4 + (pix & 3);
    const int y = roi.y0 + (int)blockIdx.y * 8 + (wave >> 1) * 4 + (pix >> 2);
    if (x >= roi.x1 || y >= roi.y1) return;
    const double step = (double)v.vs[0] * 0.15;
    // proj = (float)(0.25 * step * sum): the same factor per sample
    const double r = (double)resid[(size_t)y * pitch + (size_t)x] * (double)scale * (step * 0.25);
    if (r == 0.0) return;
    SubraySamples s;
    const double px = (double)x + ((sub & 2) ? 0.75 : 0.25);
    const double py = (double)y + ((sub & 1) ? 0.75 : 0.25);
    if (!subray_samples(v, g, px, py, step, s)) return;
    for_each_cell(s.q0, s.dq, s.count, [&](const int cell[3], int k, int mi) {
        size_t o[8];
        cell_corners<kBig>(v, cell, o);
        double W[8];
        cell_weights(cell_moments(s.q0, s.dq, cell, k, mi), W);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float add = (float)(r * W[i]);
            if (add != 0.0f) unsafeAtomicAdd(grad + o[i], add);
        }
    });
}

// The same transpose with each tile's rays walked through LDS-staged slabs (tube_walk.hpp): the
// weights are summed per slab in a box of shared memory and added to grad with one atomic per box
// voxel, not eight per cell.  Tiles are dealt to the XCDs in groups, as in the forward projector.
__global__ void __launch_bounds__(256) fp_adjoint_lds_kernel(VolG v, TiltG g, const float* __restrict__ resid,
                                                             size_t pitch, Roi roi, float scale,
                                                             float* __restrict__ grad, TubePlan tp) {
    __shared__ float chunk[kChunkFloats];
    __shared__ int zr[2];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int sub = lane & 3, pix = lane >> 2;
    const int xcd = (int)(blockIdx.x & 7u), slot = (int)(blockIdx.x >> 3);
    const int grp = (slot / (kXcdGroup * kXcdGroup)) * 8 + xcd, mem = slot % (kXcdGroup * kXcdGroup);
    const int gtx = (grp % tp.groups_x) * kXcdGroup + mem % kXcdGroup;
    const int gty = (grp / tp.groups_x) * kXcdGroup + mem / kXcdGroup;
    if (gtx >= tp.tiles_x || gty >= tp.tiles_y) return;   // whole workgroup
    const int tx0 = roi.x0 + gtx * 8, ty0 = roi.y0 + gty * 8;
    const int x = tx0 + (wave & 1) * 4 + (pix & 3);
    const int y = ty0 + (wave >> 1) * 4 + (pix >> 2);
    const double step = (double)v.vs[0] * 0.15;
    double r = 0.0;
    SubraySamples s;
    s.count = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) s.q0[c] = s.dq[c] = 0.0f;
    if (x < roi.x1 && y < roi.y1) {
        // proj = (float)(0.25 * step * sum): the same factor per sample
        r = (double)resid[(size_t)y * pitch + (size_t)x] * (double)scale * (step * 0.25);
        const double px = (double)x + ((sub & 2) ? 0.75 : 0.25);
        const double py = (double)y + ((sub & 1) ? 0.75 : 0.25);
        if (r == 0.0 || !subray_samples(v, g, px, py, step, s)) s.count = 0;
    }
    const TubeDir td = tube_dir(v, g, step);
    const TubeBox tb = tube_box(v, g, tp, (double)tx0 + 0.25, (double)tx0 + 7.75, (double)ty0 + 0.25,
                                (double)ty0 + 7.75, 0.0, 0.0);
    tube_scatter(v, grad, chunk, zr, tp, tb, td, s.q0, s.count, [&](float wgt[8], const int cell[3], int k, int mi) {
        double W[8];
        cell_weights(cell_moments(s.q0, td.dq, cell, k, mi), W);
#pragma unroll
        for (int i = 0; i < 8; i++) wgt[i] = (float)(r * W[i]);
    });
}

// Gradient of the trilinear read with respect to the (shifted) voxel coordinates; zero along an
// axis where the coordinate is clamped.
__device__ __forceinline__ void tri_grad(const float* __restrict__ vol, const VolG& v, float xb, float yb, float zb,
                                         float gq[3]) {
    const Cell q = cell_of(v, xb, yb, zb);
    const size_t sy = (size_t)v.nx, sz = (size_t)v.nx * (size_t)v.ny;
    const float* p = vol + (size_t)q.kb * sz + (size_t)q.jb * sy + (size_t)q.ib;
    const float v000 = p[0], v100 = p[1], v010 = p[sy], v110 = p[sy + 1];
    const float v001 = p[sz], v101 = p[sz + 1], v011 = p[sz + sy], v111 = p[sz + sy + 1];
    const float c00 = lerpf(v000, v100, q.a), c10 = lerpf(v010, v110, q.a);
    const float c01 = lerpf(v001, v101, q.a), c11 = lerpf(v011, v111, q.a);
    const float e0 = lerpf(c00, c10, q.b), e1 = lerpf(c01, c11, q.b);
    const float dx = lerpf(lerpf(v100 - v000, v110 - v010, q.b), lerpf(v101 - v001, v111 - v011, q.b), q.c);
    const float dy = lerpf(c10 - c00, c11 - c01, q.c);
    const float dz = e1 - e0;
    gq[0] = (xb >= 0.0f && xb <= (float)(v.nx - 1)) ? dx : 0.0f;
    gq[1] = (yb >= 0.0f && yb <= (float)(v.ny - 1)) ? dy : 0.0f;
    gq[2] = (zb >= 0.0f && zb <= (float)(v.nz - 1)) ? dz : 0.0f;
}

__global__ void __launch_bounds__(256) fp_shift_grad_kernel(VolG v, TiltG g, const float* __restrict__ vol,
                                                            const float* __restrict__ gout, size_t pitch, Roi roi,
                                                            double* __restrict__ out) {
    __shared__ double part[2][4];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int sub = lane & 3, pix = lane >> 2;
    const int x = roi.x0 + (int)blockIdx.x * 8 + (wave & 1) * 4 + (pix & 3);
    const int y = roi.y0 + (int)blockIdx.y * 8 + (wave >> 1) * 4 + (pix >> 2);
    const double step = (double)v.vs[0] * 0.15;
    double gsx = 0.0, gsy = 0.0;
    if (x < roi.x1 && y < roi.y1) {
        const float go = gout[(size_t)y * pitch + (size_t)x];
        SubraySamples s;
        const double px = (double)x + ((sub & 2) ? 0.75 : 0.25);
        const double py = (double)y + ((sub & 1) ? 0.75 : 0.25);
        if (go != 0.0f && subray_samples(v, g, px, py, step, s)) {
            double G[3] = {0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < s.count; k0 += 256) {
                const int k1 = min(s.count, k0 + 256);
                float f[3] = {0.0f, 0.0f, 0.0f};
                for (int k = k0; k < k1; k++) {
                    const float fk = (float)k;
                    float gq[3];
                    tri_grad(vol, v, __builtin_fmaf(fk, s.dq[0], s.q0[0]), __builtin_fmaf(fk, s.dq[1], s.q0[1]),
                             __builtin_fmaf(fk, s.dq[2], s.q0[2]), gq);
                    f[0] += gq[0];
                    f[1] += gq[1];
                    f[2] += gq[2];
                }
                for (int i = 0; i < 3; i++) G[i] += (double)f[i];
            }
            // d p_k / d shift = -u (-v) + (d t_in / d shift) n, d t_in / d shift = u_a / n_a (v_a / n_a)
            const int a = s.entry_axis;
            const double tu = (double)g.u[a] / (double)g.n[a], tv = (double)g.v[a] / (double)g.n[a];
            double dx = 0.0, dy = 0.0;
            for (int i = 0; i < 3; i++) {
                const double scale = (double)v.rcp[i] * (double)v.dim[i];
                dx += G[i] * scale * (-(double)g.u[i] + tu * (double)g.n[i]);
                dy += G[i] * scale * (-(double)g.v[i] + tv * (double)g.n[i]);
            }
            const double w = (double)go * step * 0.25;
            gsx = w * dx;
            gsy = w * dy;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        gsx += __shfl_down(gsx, o);
        gsy += __shfl_down(gsy, o);
    }
    if (lane == 0) {
        part[0][wave] = gsx;
        part[1][wave] = gsy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsafeAtomicAdd(out, ((part[0][0] + part[0][1]) + part[0][2]) + part[0][3]);
        unsafeAtomicAdd(out + 1, ((part[1][0] + part[1][1]) + part[1][2]) + part[1][3]);
    }
}

Roi clip_roi(const int* roi, int W, int H) {
    Roi r{0, 0, W, H};
    if (roi) {
        r.x0 = roi[0] < 0 ? 0 : roi[0];
        r.y0 = roi[1] < 0 ? 0 : roi[1];
        r.x1 = roi[2] > W ? W : roi[2];
        r.y1 = roi[3] > H ? H : roi[3];
    }
    return r;
}

dim3 grid_of(const Roi& r) {
    return dim3((unsigned)((r.x1 - r.x0 + 7) / 8), (unsigned)((r.y1 - r.y0 + 7) / 8));
}

artia_status check_args(artia_ctx* ctx, const char* fn, const artia_vol_geom* vg, const artia_tilt_geom* tg, int W,
                        int H, size_t pitch) {
    if (!valid_geometry(vg, tg))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: null pointer or invalid geometry", fn);
    if (vg->nx < 2 || vg->ny < 2 || vg->nz < 2)
        return fail(ctx, ARTIA_ERR_UNSUPPORTED, "%s: volumes need at least 2 voxels per axis", fn);
    if (W <= 0 || H <= 0 || pitch < (size_t)W)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: bad image size %d x %d (pitch %zu)", fn, W, H, pitch);
    return ARTIA_OK;
}

}  // namespace
}  // namespace artia

using namespace artia;

extern "C" {

artia_status artia_fp_adjoint(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                              const float* resid, int width, int height, size_t pitch, float* vol_grad,
                              const int* roi, float scale, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    artia_status st = check_args(ctx, "artia_fp_adjoint", vg, tg, width, height, pitch);
    if (st != ARTIA_OK) return st;
    if (!resid || !vol_grad) return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_fp_adjoint: null pointer");
    const Roi r = clip_roi(roi, width, height);
    Timing t(ctx, ms);
    const VolG vd = to_device(*vg);
    const TiltG gd = flat_beam(to_device(*tg));
    TubePlan tp;
    if (r.x1 > r.x0 && r.y1 > r.y0 && ctx->march_lds && tube_plan(vd, gd, tp)) {
        const dim3 grid = grid_of(r);
        tp.tiles_x = (int)grid.x;
        tp.tiles_y = (int)grid.y;
        tp.groups_x = (tp.tiles_x + kXcdGroup - 1) / kXcdGroup;
        const int groups = tp.groups_x * ((tp.tiles_y + kXcdGroup - 1) / kXcdGroup);
        const dim3 grid1((unsigned)(((groups + 7) / 8) * 8 * kXcdGroup * kXcdGroup));
        hipLaunchKernelGGL(fp_adjoint_lds_kernel, grid1, dim3(256), 0, ctx->stream, vd, gd, resid, pitch, r, scale,
                           vol_grad, tp);
    } else if (r.x1 > r.x0 && r.y1 > r.y0)
    {
        const bool big = (double)vg->nx * (double)vg->ny * (double)vg->nz >= 1073741824.0;
        if (big)
            hipLaunchKernelGGL(fp_adjoint_kernel<true>, grid_of(r), dim3(256), 0, ctx->stream, vd,
                               gd, resid, pitch, r, scale, vol_grad);
        else
            hipLaunchKernelGGL(fp_adjoint_kernel<false>, grid_of(r), dim3(256), 0, ctx->stream, vd,
                               gd, resid, pitch, r, scale, vol_grad);
    }
    st = check_hip(ctx, hipGetLastError(), "artia_fp_adjoint launch");
    if (st != ARTIA_OK) return st;
    return t.finish();
}

artia_status artia_fp_shift_grad(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                                 const float* vol, const float* grad_proj, int width, int height, size_t pitch,
                                 const int* roi, double* out, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    artia_status st = check_args(ctx, "artia_fp_shift_grad", vg, tg, width, height, pitch);
    if (st != ARTIA_OK) return st;
    if (!vol || !grad_proj || !out) return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_fp_shift_grad: null pointer");
    const Roi r = clip_roi(roi, width, height);
    Timing t(ctx, ms);
    if (r.x1 > r.x0 && r.y1 > r.y0)
        hipLaunchKernelGGL(fp_shift_grad_kernel, grid_of(r), dim3(256), 0, ctx->stream, to_device(*vg),
                           flat_beam(to_device(*tg)), vol, grad_proj, pitch, r, out);
    st = check_hip(ctx, hipGetLastError(), "artia_fp_shift_grad launch");
    if (st != ARTIA_OK) return st;
    return t.finish();
}

}  // extern "C"
