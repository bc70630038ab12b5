// Batched refinement over (sub-volume, tilt) pairs: the forward projector, its shift derivative
// and its transpose restricted to each pair's detector window, and EmSARTRefine's
// cross-correlation shift search (GetDisplacement, Reconstructor.cpp:1746-1959), each over all
// pairs of a refinement step in one launch.
//
// One 256-thread workgroup per pair walks its window in 8x8-pixel tiles with one lane per sub-ray
// (the forward projector's mapping), so per-pair sums need no atomics and come out in a fixed
// order.  Workgroup b is dispatched to XCD b mod 8; pairs are ordered by sub-volume and dealt to
// XCDs in contiguous runs, so the tilts of one sub-volume share one XCD's L2.
//
// Per sub-ray the samples are summed cell by cell in closed form (fp_common.hpp); the shift
// derivative of the sub-ray is (sum of the interpolant's gradients over its samples) . dq0/ds,
// because a detector shift translates all samples of a sub-ray alike: dq0/ds = scale * (-u +
// (u_a / n_a) n) with a the axis whose face sets the entry point (likewise v).
#include <algorithm>
#include <numeric>
#include <vector>

#include "artia_internal.hpp"
#include "fp_common.hpp"
#include "tube_walk.hpp"

namespace artia {
namespace {

struct PairDev {
    TiltG g;
    VolG v;
    TubePlan tp;            // LDS-staged walk of the window's tiles (staged != 0)
    int staged;
    long long vol_offset;   // floats into the volume buffer
    long long img_offset;   // floats into image buffers
    int x0, y0, w, h;
};

enum { kProject = 0, kLossGrad = 1, kShiftGrad = 2 };

struct PairArgs {
    const PairDev* pairs;
    const int* order;        // slot -> pair, grouped by sub-volume
    int npairs, per_xcd;
    const float* vols;
    const double* shifts;    // 2 per pair or null
    const float* in;         // targets (kLossGrad), weights (kShiftGrad), residuals (adjoint)
    float* out;              // images (kProject), residuals (kLossGrad, may be null)
    double* loss;
    double* grad;
    double* normal;          // J^T J per pair (kLossGrad, may be null)
};

// Slot of workgroup b: XCD b % 8 takes slots [xcd * per_xcd, (xcd + 1) * per_xcd).
__device__ __forceinline__ int slot_of_block(int per_xcd) {
    return (int)(blockIdx.x & 7u) * per_xcd + (int)(blockIdx.x >> 3);
}

__device__ __forceinline__ void load_cell(const __amdgpu_buffer_rsrc_t rs, const VolG& v, const int cell[3],
                                          float val[8], uint32_t o[8]) {
    cell_corners_bytes(v, cell, o);
#pragma unroll
    for (int i = 0; i < 8; i++)   // the b32 load returns the word's bits
        val[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)o[i], 0, 0));
}

// Block sum of three doubles in a fixed order; the result is valid in thread 0.  Ends with a
// barrier, so it can be called again right away.
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c) {
    __shared__ double part[3][4];
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o);
        b += __shfl_down(b, o);
        c += __shfl_down(c, o);
    }
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    if (lane == 0) {
        part[0][wave] = a;
        part[1][wave] = b;
        part[2][wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = ((part[0][0] + part[0][1]) + part[0][2]) + part[0][3];
        b = ((part[1][0] + part[1][1]) + part[1][2]) + part[1][3];
        c = ((part[2][0] + part[2][1]) + part[2][2]) + part[2][3];
    }
    __syncthreads();
}

// Five waves per SIMD (96 registers, a few spilled): measured faster than the unconstrained
// allocation (3-4 waves) for the LDS-staged walk.
template <int kMode>
__global__ void __launch_bounds__(256, 5) pair_march_kernel(PairArgs a) {
    __shared__ float chunk[kChunkFloats];
    __shared__ int zr[2];
    const int slot = slot_of_block(a.per_xcd);
    if (slot >= a.npairs) return;   // whole workgroup
    const int p = a.order[slot];
    const PairDev& P = a.pairs[p];
    const int x0 = P.x0, y0 = P.y0, w = P.w, h = P.h;
    const long long io = P.img_offset;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int sub = lane & 3, pix = lane >> 2;
    const int tx = (wave & 1) * 4 + (pix & 3), ty = (wave >> 1) * 4 + (pix >> 2);
    const int ntx = (w + 7) >> 3, ntiles = ntx * ((h + 7) >> 3);
    double lsum = 0.0, gx = 0.0, gy = 0.0, jxx = 0.0, jxy = 0.0, jyy = 0.0;
    for (int t = 0; t < ntiles; t++) {
        // The geometry is re-read (scalar loads) per tile: hoisted out of the loop, its double
        // conversions would stay live in vector registers and halve the occupancy.
        __asm__ volatile("" ::: "memory");
        const VolG v = P.v;
        const TiltG g = P.g;
        const __amdgpu_buffer_rsrc_t rs = volume_rsrc(a.vols + P.vol_offset, v);
        const double sx = a.shifts ? a.shifts[2 * p] : 0.0, sy = a.shifts ? a.shifts[2 * p + 1] : 0.0;
        const double step = (double)v.vs[0] * 0.15;
        const int i = (t % ntx) * 8 + tx, j = (t / ntx) * 8 + ty;
        const bool active = i < w && j < h;
        const long long pofs = io + (long long)j * w + i;
        float win = 0.0f;
        if (kMode != kProject && active) win = a.in[pofs];
        double acc = 0.0, dxs = 0.0, dys = 0.0;
        SubraySamples s;
        const double px = (double)(x0 + i) + ((sub & 2) ? 0.75 : 0.25);
        const double py = (double)(y0 + j) + ((sub & 1) ? 0.75 : 0.25);
        const bool ok = active && (kMode != kShiftGrad || win != 0.0f) && subray_samples(v, g, px, py, step, s, sx, sy);
        if (!ok) {
            s.count = 0;
#pragma unroll
            for (int c = 0; c < 3; c++) s.q0[c] = s.dq[c] = 0.0f;
        }
        double G[3] = {0.0, 0.0, 0.0};
        auto cell_fn = [&](const float val[8], const auto* cell, auto k, auto mi) {
            const CellMoments cm = cell_moments(s.q0, s.dq, cell, k, mi);
            if (kMode != kShiftGrad) acc += cell_sum_values(val, cm, mi);
            if (kMode != kProject) cell_grad_sum(val, cm, mi, G);
        };
        if (P.staged) {   // uniform: the whole block walks the tile's tube through LDS
            const TubeDir td = tube_dir(v, g, step);
            const double pi0 = (double)(x0 + (t % ntx) * 8), pj0 = (double)(y0 + (t / ntx) * 8);
            const TubeBox tb = tube_box(v, g, P.tp, pi0 + 0.25, pi0 + 7.75, pj0 + 0.25, pj0 + 7.75, sx, sy);
            // rays without a y component (n_y == 0 after flat_beam): the walk without the y axis,
            // the same cells and sums (block-uniform: one pair per block)
            if (g.n[1] == 0.0f)
                tube_walk<true>(v, a.vols + P.vol_offset, chunk, zr, P.tp, tb, td, s.q0, s.count, cell_fn);
            else
                tube_walk<false>(v, a.vols + P.vol_offset, chunk, zr, P.tp, tb, td, s.q0, s.count, cell_fn);
        } else if (ok) {
            for_each_cell(s.q0, s.dq, s.count, [&](const int cell[3], int k, int mi) {
                float val[8];
                uint32_t o[8];
                load_cell(rs, v, cell, val, o);
                cell_fn(val, cell, k, mi);
            });
        }
        if (ok) {
            if (kMode != kProject) {
                // entry-axis components by select (a dynamic index would put the geometry in scratch)
                const int ax = s.entry_axis;
                const float ua = ax == 0 ? g.u[0] : (ax == 1 ? g.u[1] : g.u[2]);
                const float va = ax == 0 ? g.v[0] : (ax == 1 ? g.v[1] : g.v[2]);
                const float na = ax == 0 ? g.n[0] : (ax == 1 ? g.n[1] : g.n[2]);
                const double tu = (double)ua / (double)na, tv = (double)va / (double)na;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double scale = (double)v.rcp[c] * (double)v.dim[c];
                    dxs += G[c] * scale * (-(double)g.u[c] + tu * (double)g.n[c]);
                    dys += G[c] * scale * (-(double)g.v[c] + tv * (double)g.n[c]);
                }
            }
        }
        // the four sub-rays of a pixel are lanes 4q .. 4q+3
        if (kMode != kShiftGrad) {
            acc += __shfl_xor(acc, 1);
            acc += __shfl_xor(acc, 2);
        }
        if (kMode != kProject) {
            dxs += __shfl_xor(dxs, 1);
            dxs += __shfl_xor(dxs, 2);
            dys += __shfl_xor(dys, 1);
            dys += __shfl_xor(dys, 2);
        }
        if (active && sub == 0) {
            const double dpx = dxs * step * 0.25, dpy = dys * step * 0.25;
            if (kMode == kProject) {
                a.out[pofs] = (float)(acc * step * 0.25);
            } else if (kMode == kLossGrad) {
                const float pv = (float)(acc * step * 0.25);
                const double r = (double)pv - (double)win;
                if (a.out) a.out[pofs] = (float)r;
                lsum += r * r;
                gx += 2.0 * r * dpx;
                gy += 2.0 * r * dpy;
                jxx += dpx * dpx;
                jxy += dpx * dpy;
                jyy += dpy * dpy;
            } else {
                gx += (double)win * dpx;
                gy += (double)win * dpy;
            }
        }
    }
    if (kMode == kProject) return;
    block_sum3(lsum, gx, gy);
    if (threadIdx.x == 0) {
        if (kMode == kLossGrad) a.loss[p] = lsum;
        a.grad[2 * p] = gx;
        a.grad[2 * p + 1] = gy;
    }
    if (kMode == kLossGrad && a.normal) {   // uniform branch
        block_sum3(jxx, jxy, jyy);
        if (threadIdx.x == 0) {
            a.normal[3 * p] = jxx;
            a.normal[3 * p + 1] = jxy;
            a.normal[3 * p + 2] = jyy;
        }
    }
}

// The transpose summed per sub-volume slab in LDS.  A workgroup owns cell layers [zs, zs + L) of
// one sub-volume (voxel layers max(zs, 0) .. min(zs + L, nz - 1) in LDS) and walks every sub-ray
// of every tilt of that sub-volume through them only: the samples between the sub-ray's crossings
// of the slab's two z faces (the walker's own face-crossing formula, so neighbouring slabs split
// each sub-ray's samples without overlap or gap).  The tilts' weights meet in LDS (shared-memory
// atomics) and reach the gradient with one atomic per voxel and slab, instead of one per box voxel
// and tube slab of every pair.  A corner outside the layers (a sample rounded across a slab face)
// goes to global memory directly.  The layers accumulate in double: a voxel receives the weights of
// every sub-ray of every tilt near it (hundreds of terms of both signs for a signed image), whose
// float sum would carry about 1e-6 of rounding and depend on the atomics' order.
constexpr int kAdjSlabDoubles = 8192;   // 64 KiB of LDS: the x-y layers of a slab

// Cell layers per slab of a sub-volume (0: its x-y layer pair does not fit).
__host__ __device__ __forceinline__ int adj_slab_layers(int nx, int ny) {
    const long long nxy = (long long)nx * (long long)ny;
    const long long layers = (long long)kAdjSlabDoubles / (nxy > 0 ? nxy : 1);
    return layers >= 2 ? (int)(layers - 1) : 0;
}

struct AdjSlabArgs {
    const int2* items;      // (sub-volume, first cell layer zs), per XCD in runs of per_xcd; x < 0: none
    int per_xcd;
    const int* vfirst;      // per sub-volume: first slot of its pairs in `order`, and their count
    const int* vcount;
};

__global__ void __launch_bounds__(256) pair_adjoint_slab_kernel(PairArgs a, AdjSlabArgs sa, float scale,
                                                                float* __restrict__ grads) {
    __shared__ double box[kAdjSlabDoubles];
    const int pos = (int)(blockIdx.x >> 3);
    if (pos >= sa.per_xcd) return;   // whole workgroup
    const int2 it = sa.items[(int)(blockIdx.x & 7u) * sa.per_xcd + pos];
    if (it.x < 0) return;   // whole workgroup
    const int first = sa.vfirst[it.x], npair = sa.vcount[it.x];
    const VolG v0 = a.pairs[a.order[first]].v;
    const int L = adj_slab_layers(v0.nx, v0.ny);
    const int zs = it.y, ze = min(zs + L, v0.nz);   // cell layers [zs, ze); layer c spans q_z in [c, c + 1)
    const int bz0 = max(zs, 0), bz1 = min(ze, v0.nz - 1);
    const int nxy = v0.nx * v0.ny, nbox = nxy * (bz1 - bz0 + 1);
    for (int e = (int)threadIdx.x; e < nbox; e += 256) box[e] = 0.0;
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int sub = lane & 3, pix = lane >> 2;
    const int tx = (wave & 1) * 4 + (pix & 3), ty = (wave >> 1) * 4 + (pix >> 2);
    for (int q = 0; q < npair; q++) {
        const int p = a.order[first + q];
        const PairDev& P = a.pairs[p];
        const int x0 = P.x0, y0 = P.y0, w = P.w, h = P.h;
        const long long io = P.img_offset;
        const int ntx = (w + 7) >> 3, ntiles = ntx * ((h + 7) >> 3);
        for (int t = 0; t < ntiles; t++) {
            __asm__ volatile("" ::: "memory");   // geometry re-read per tile, as in pair_march_kernel
            const VolG v = P.v;
            const TiltG g = P.g;
            float* __restrict__ gv = grads + P.vol_offset;
            const double sx = a.shifts ? a.shifts[2 * p] : 0.0, sy = a.shifts ? a.shifts[2 * p + 1] : 0.0;
            const double step = (double)v.vs[0] * 0.15;
            const int i = (t % ntx) * 8 + tx, j = (t / ntx) * 8 + ty;
            if (!(i < w && j < h)) continue;
            // proj = (float)(0.25 * step * sum): the same factor per sample
            const double r = (double)a.in[io + (long long)j * w + i] * (double)scale * (step * 0.25);
            if (r == 0.0) continue;
            SubraySamples s;
            const double px = (double)(x0 + i) + ((sub & 2) ? 0.75 : 0.25);
            const double py = (double)(y0 + j) + ((sub & 1) ? 0.75 : 0.25);
            if (!subray_samples(v, g, px, py, step, s, sx, sy)) continue;
            int sgn[3];
            float inv[3];
#pragma unroll
            for (int c = 0; c < 3; c++) walk_dir(s.dq[c], sgn[c], inv[c]);
            // samples of this slab: from the crossing of its entry face to that of its exit face, as
            // the walker computes face crossings (CellWalker::step), clamped to the sub-ray
            const float offz = -s.q0[2] * inv[2];
            const float fin = (float)(sgn[2] > 0 ? zs : ze), fout = (float)(sgn[2] > 0 ? ze : zs);
            const int kA = max(0, min(s.count, (int)ceilf(__builtin_fmaf(fin, inv[2], offz))));
            const int kB = max(0, min(s.count, (int)ceilf(__builtin_fmaf(fout, inv[2], offz))));
            if (kB <= kA) continue;
            CellWalker wk;
            wk.init_from(s.q0, s.dq, sgn, inv, kA, kB);
            while (wk.active()) {
                wk.step([&](const int cell[3], int k, int mi) {
                    double W[8];
                    cell_weights(cell_moments(s.q0, s.dq, cell, k, mi), W);
                    // clamp-addressed corners (cell_corners): 0 step where an index is clamped
                    const int xa = clampi(cell[0], 0, v.nx - 1), ya = clampi(cell[1], 0, v.ny - 1);
                    const int za = clampi(cell[2], 0, v.nz - 1);
                    const int dx = (cell[0] + 1 > xa && xa + 1 < v.nx) ? 1 : 0;
                    const int dy = (cell[1] + 1 > ya && ya + 1 < v.ny) ? v.nx : 0;
                    const int zb = (cell[2] + 1 > za && za + 1 < v.nz) ? za + 1 : za;
                    const int o0 = ya * v.nx + xa;
#pragma unroll
                    for (int c = 0; c < 8; c++) {
                        const double wc = r * W[c];
                        if (wc == 0.0) continue;
                        const int zc = (c & 4) ? zb : za;
                        const int o = o0 + ((c & 1) ? dx : 0) + ((c & 2) ? dy : 0);
                        if (zc >= bz0 && zc <= bz1)
                            unsafeAtomicAdd(box + (zc - bz0) * nxy + o, wc);
                        else
                            unsafeAtomicAdd(gv + (size_t)zc * (size_t)nxy + (size_t)o, (float)wc);
                    }
                });
            }
        }
    }
    __syncthreads();
    float* __restrict__ gv = grads + a.pairs[a.order[first]].vol_offset + (size_t)bz0 * (size_t)nxy;
    for (int e = (int)threadIdx.x; e < nbox; e += 256)
        if (box[e] != 0.0) unsafeAtomicAdd(gv + e, (float)box[e]);
}

// vol_grads += scale * P_s^T in: the transpose of the per-cell closed form, one set of eight
// atomics per cell a sub-ray crosses (as fp_adjoint_kernel).
__global__ void __launch_bounds__(256) pair_adjoint_kernel(PairArgs a, float scale, float* __restrict__ grads) {
    __shared__ float chunk[kChunkFloats];
    __shared__ int zr[2];
    const int slot = slot_of_block(a.per_xcd);
    if (slot >= a.npairs) return;
    const int p = a.order[slot];
    const PairDev& P = a.pairs[p];
    const int x0 = P.x0, y0 = P.y0, w = P.w, h = P.h;
    const long long io = P.img_offset;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int sub = lane & 3, pix = lane >> 2;
    const int tx = (wave & 1) * 4 + (pix & 3), ty = (wave >> 1) * 4 + (pix >> 2);
    const int ntx = (w + 7) >> 3, ntiles = ntx * ((h + 7) >> 3);
    for (int t = 0; t < ntiles; t++) {
        __asm__ volatile("" ::: "memory");   // geometry re-read per tile, as in pair_march_kernel
        const VolG v = P.v;
        const TiltG g = P.g;
        float* __restrict__ gv = grads + P.vol_offset;
        const double sx = a.shifts ? a.shifts[2 * p] : 0.0, sy = a.shifts ? a.shifts[2 * p + 1] : 0.0;
        const double step = (double)v.vs[0] * 0.15;
        const int i = (t % ntx) * 8 + tx, j = (t / ntx) * 8 + ty;
        double r = 0.0;
        SubraySamples s;
        s.count = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) s.q0[c] = s.dq[c] = 0.0f;
        if (i < w && j < h) {
            // proj = (float)(0.25 * step * sum): the same factor per sample
            r = (double)a.in[io + (long long)j * w + i] * (double)scale * (step * 0.25);
            const double px = (double)(x0 + i) + ((sub & 2) ? 0.75 : 0.25);
            const double py = (double)(y0 + j) + ((sub & 1) ? 0.75 : 0.25);
            if (r == 0.0 || !subray_samples(v, g, px, py, step, s, sx, sy)) s.count = 0;
        }
        auto weights = [&](float wgt[8], const int cell[3], int k, int mi) {
            double W[8];
            cell_weights(cell_moments(s.q0, s.dq, cell, k, mi), W);
#pragma unroll
            for (int c = 0; c < 8; c++) wgt[c] = (float)(r * W[c]);
        };
        if (P.staged) {   // uniform: per-slab sums in LDS, one atomic per box voxel
            const TubeDir td = tube_dir(v, g, step);
            const double pi0 = (double)(x0 + (t % ntx) * 8), pj0 = (double)(y0 + (t / ntx) * 8);
            const TubeBox tb = tube_box(v, g, P.tp, pi0 + 0.25, pi0 + 7.75, pj0 + 0.25, pj0 + 7.75, sx, sy);
            tube_scatter(v, gv, chunk, zr, P.tp, tb, td, s.q0, s.count, weights);
        } else if (s.count > 0) {
            for_each_cell(s.q0, s.dq, s.count, [&](const int cell[3], int k, int mi) {
                uint32_t o[8];
                cell_corners_bytes(v, cell, o);
                float wgt[8];
                weights(wgt, cell, k, mi);
#pragma unroll
                for (int c = 0; c < 8; c++)
                    if (wgt[c] != 0.0f) unsafeAtomicAdd(gv + (o[c] >> 2), wgt[c]);
            });
        }
    }
}

constexpr int kMaxCcShift = 32;

// GetDisplacement per pair (Reconstructor.cpp:1746-1959) as a direct sum over the window: the
// raw correlation C(r) = sum_y B(y) A(y - r) for every integer r in the disk, then the square's
// mean and energy terms, the maxShift disk (wbpWeighting.cu:265-294) and nppiMaxIndx's first
// maximum in the square's row-major order.  Wave w takes candidates w, w + 4, ...; lanes run
// along window rows.
// multi != 0: MultiPeakDetection (Reconstructor.cpp:1887-1935): among the peaks of the float map
// (findPeak, wbpWeighting.cu:329-386: above the 8 neighbours, the two upper diagonal ones or equal,
// and at least 0.9f of the maximum) the one nearest to zero shift (maxShiftWeighted,
// wbpWeighting.cu:295-327: weight 1 / (dist + 0.0001f)), first in row-major order; the CC value and
// the failure test keep the first search's maximum.  Candidates outside the disk are 0 in the map.
__global__ void __launch_bounds__(256) pair_cc_kernel(const PairDev* __restrict__ pairs, int npairs,
                                                      const float* __restrict__ models,
                                                      const float* __restrict__ targets, int N, int M,
                                                      int multi, int* __restrict__ peak, float* __restrict__ ccv,
                                                      float* __restrict__ ccmap) {
    __shared__ double craw[(2 * kMaxCcShift + 1) * (2 * kMaxCcShift + 1)];
    const int p = (int)blockIdx.x;
    if (p >= npairs) return;
    const PairDev& P = pairs[p];
    const int w = P.w, h = P.h;
    const float* __restrict__ A = targets + P.img_offset;
    const float* __restrict__ B = models + P.img_offset;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    // window sums of A, A^2, B, B^2 (the square is zero outside the window)
    double sa = 0.0, saa = 0.0, sb = 0.0;
    double sbb = 0.0;
    for (long long q = threadIdx.x; q < (long long)w * h; q += 256) {
        const double av = (double)A[q], bv = (double)B[q];
        sa += av;
        saa += av * av;
        sb += bv;
        sbb += bv * bv;
    }
    block_sum3(sa, saa, sb);
    __shared__ double tot[4];
    if (threadIdx.x == 0) {
        tot[0] = sa;
        tot[1] = saa;
        tot[2] = sb;
    }
    double d = 0.0, e = 0.0;
    block_sum3(sbb, d, e);
    if (threadIdx.x == 0) tot[3] = sbb;
    const int side = 2 * M + 1, ncand = side * side;
    for (int c = wave; c < ncand; c += 4) {
        const int rx = c % side - M, ry = c / side - M;
        double acc = 0.0;
        if (sqrtf((float)(rx * rx + ry * ry)) <= (float)M) {
            const int ylo = max(0, ry), yhi = min(h, h + ry);   // rows with 0 <= y - ry < h
            const int xlo = max(0, rx), xhi = min(w, w + rx);
            for (int y = ylo; y < yhi; y++) {
                const float* __restrict__ brow = B + (long long)y * w;
                const float* __restrict__ arow = A + (long long)(y - ry) * w - rx;
                for (int x = xlo + lane; x < xhi; x += 64) acc += (double)brow[x] * (double)arow[x];
            }
            for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
        }
        if (lane == 0) craw[c] = acc;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double SA = tot[0], SB = tot[2], n2 = (double)N * (double)N;
    const double sumA = tot[1] - SA * SA / n2, sumB = tot[3] - SB * SB / n2;
    const double den = sqrt(sumA * sumB), off = SA * SB / n2;
    const bool ok = sumA * sumB > 0.0;
    // candidates in the square's row-major order: rows ry = 0..M, then -M..-1; likewise columns
    double best = 0.0;
    int bx = 0, by = 0;
    bool found = false;
    for (int yy = 0; yy < side; yy++) {
        const int ry = yy <= M ? yy : yy - side;
        for (int xx = 0; xx < side; xx++) {
            const int rx = xx <= M ? xx : xx - side;
            if (!(sqrtf((float)(rx * rx + ry * ry)) <= (float)M)) continue;
            const int c = (ry + M) * side + (rx + M);
            const double cc = ok ? (craw[c] - off) / den : 0.0;
            if (ccmap) ccmap[(long long)p * ncand + c] = (float)cc;
            if (!found || cc > best) {
                best = cc;
                bx = rx;
                by = ry;
                found = true;
            }
        }
    }
    if (ccmap) {
        for (int c = 0; c < ncand; c++) {
            const int rx = c % side - M, ry = c / side - M;
            if (!(sqrtf((float)(rx * rx + ry * ry)) <= (float)M)) ccmap[(long long)p * ncand + c] = 0.0f;
        }
    }
    if (multi) {
        // the float map value at shift (rx, ry), wrapped around the N x N square as findPeak's
        // neighbours are: 0 outside the disk (maxShift)
        auto map = [&](int rx, int ry) -> float {
            rx = ((rx % N) + N) % N;
            ry = ((ry % N) + N) % N;
            rx = rx > N / 2 ? rx - N : rx;
            ry = ry > N / 2 ? ry - N : ry;
            if (rx < -M || rx > M || ry < -M || ry > M || !(sqrtf((float)(rx * rx + ry * ry)) <= (float)M))
                return 0.0f;
            return ok ? (float)((craw[(ry + M) * side + (rx + M)] - off) / den) : 0.0f;
        };
        const float thr = (float)best * 0.9f;
        float bestw = 0.0f;
        int mx = 0, my = 0;
        bool seen = false;
        for (int yy = 0; yy < side; yy++) {
            const int ry = yy <= M ? yy : yy - side;
            for (int xx = 0; xx < side; xx++) {
                const int rx = xx <= M ? xx : xx - side;
                const float d = sqrtf((float)(rx * rx + ry * ry));
                float wv = 0.0f;
                if (d <= (float)M) {
                    const float val = map(rx, ry);
                    const bool pk = map(rx - 1, ry) < val && map(rx + 1, ry) < val && map(rx, ry - 1) < val &&
                                    map(rx, ry + 1) < val && map(rx - 1, ry - 1) <= val && map(rx - 1, ry + 1) < val &&
                                    map(rx + 1, ry - 1) <= val && map(rx + 1, ry + 1) < val && val >= thr;
                    wv = pk ? 1.0f / (d + 0.0001f) : 0.0f;
                }
                if (!seen || wv > bestw) {
                    bestw = wv;
                    mx = rx;
                    my = ry;
                    seen = true;
                }
            }
        }
        bx = mx;
        by = my;
    }
    const bool fail = !ok || !(best > 0.0);
    peak[2 * p] = fail ? -1000 : bx;
    peak[2 * p + 1] = fail ? -1000 : by;
    ccv[p] = ok ? (float)best : 0.0f;
}

}  // namespace
}  // namespace artia

using namespace artia;

struct artia_pairs {
    artia_ctx* ctx = nullptr;
    PairDev* pairs = nullptr;
    int* order = nullptr;
    int npairs = 0, per_xcd = 0;
    // the slab-summed adjoint (pair_adjoint_slab_kernel): work items and per-volume slot ranges,
    // or none when a sub-volume's layers do not fit or a tilt runs parallel to its z layers
    int2* slab_items = nullptr;
    int* vfirst = nullptr;
    int* vcount = nullptr;
    int slab_per_xcd = 0;
};

namespace {

artia_status pairs_fail(artia_pairs* ps, artia_status code, const char* what) {
    return fail(ps->ctx, code, "%s", what);
}

PairArgs base_args(const artia_pairs* ps) {
    PairArgs a{};
    a.pairs = ps->pairs;
    a.order = ps->order;
    a.npairs = ps->npairs;
    a.per_xcd = ps->per_xcd;
    return a;
}

dim3 pair_grid(const artia_pairs* ps) { return dim3((unsigned)(8 * ps->per_xcd)); }

}  // namespace

extern "C" {

artia_status artia_pairs_create(artia_ctx* ctx, const artia_pair_volume* volumes, int nvol, const artia_pair* pairs,
                                const artia_tilt_geom* geoms, int npairs, artia_pairs** out) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!out || !volumes || !pairs || !geoms || nvol <= 0 || npairs <= 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_pairs_create: null pointer or empty set");
    *out = nullptr;
    for (int i = 0; i < nvol; i++) {
        const artia_vol_geom& g = volumes[i].geom;
        if (!valid_geometry(&g, &geoms[0]) || volumes[i].offset < 0)
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_pairs_create: invalid sub-volume %d", i);
        if (g.nx < 2 || g.ny < 2 || g.nz < 2 || g.nx >= (1 << 24) || (double)g.ny * g.nz >= 16777216.0 ||
            (double)g.nx * g.ny * g.nz >= 1073741824.0)
            return fail(ctx, ARTIA_ERR_UNSUPPORTED,
                        "artia_pairs_create: sub-volume %d is %d x %d x %d (needs 2 .. 2^24-1 voxels per axis, "
                        "ny*nz < 2^24, fewer than 2^30 voxels)",
                        i, g.nx, g.ny, g.nz);
    }
    std::vector<PairDev> host((size_t)npairs);
    for (int p = 0; p < npairs; p++) {
        const artia_pair& q = pairs[p];
        if (q.volume < 0 || q.volume >= nvol || q.w <= 0 || q.h <= 0 || q.offset < 0)
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_pairs_create: pair %d has a bad volume or window", p);
        if (!valid_geometry(&volumes[q.volume].geom, &geoms[p]))
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_pairs_create: invalid geometry of pair %d", p);
        PairDev& d = host[(size_t)p];
        d.g = flat_beam(to_device(geoms[p]));   // as the exact forward projector takes it
        d.v = to_device(volumes[q.volume].geom);
        d.staged = ctx->march_lds && tube_plan(d.v, d.g, d.tp) ? 1 : 0;
        d.vol_offset = volumes[q.volume].offset;
        d.img_offset = q.offset;
        d.x0 = q.x0;
        d.y0 = q.y0;
        d.w = q.w;
        d.h = q.h;
    }
    // slots grouped by sub-volume (stable), dealt to the 8 XCDs in contiguous runs
    std::vector<int> order((size_t)npairs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pairs[a].volume < pairs[b].volume; });
    artia_pairs* ps = new artia_pairs();
    ps->ctx = ctx;
    ps->npairs = npairs;
    ps->per_xcd = (npairs + 7) / 8;
    // slab work items of the summed adjoint: every sub-volume's cell layers -1 .. nz - 1 in runs of
    // its slab depth, all on the XCD of the sub-volume (volume index mod 8), so the pairs' windows
    // are read through one L2
    std::vector<int> vfirst((size_t)nvol, 0), vcount((size_t)nvol, 0);
    for (int s = 0; s < npairs; s++) {
        const int vi = pairs[order[(size_t)s]].volume;
        if (vcount[(size_t)vi]++ == 0) vfirst[(size_t)vi] = s;
    }
    // (the staging switch off keeps the per-pair scatter, as it keeps the per-pair walks direct)
    bool slab_ok = ctx->march_lds;
    for (int p = 0; p < npairs && slab_ok; p++)
        slab_ok = std::fabs(host[(size_t)p].g.n[2]) > 1e-3f;   // rays cross the z layers
    std::vector<std::vector<int2>> per((size_t)8);
    for (int vi = 0; vi < nvol && slab_ok; vi++) {
        if (vcount[(size_t)vi] == 0) continue;
        const artia_vol_geom& g = volumes[vi].geom;
        const int L = adj_slab_layers(g.nx, g.ny);
        if (L <= 0) {
            slab_ok = false;
            break;
        }
        for (int zs = -1; zs < g.nz; zs += L) per[(size_t)(vi % 8)].push_back(make_int2(vi, zs));
    }
    std::vector<int2> items;
    if (slab_ok) {
        size_t longest = 0;
        for (const auto& l : per) longest = std::max(longest, l.size());
        ps->slab_per_xcd = (int)longest;
        items.assign(8 * longest, make_int2(-1, 0));
        for (size_t x = 0; x < 8; x++)
            for (size_t k = 0; k < per[x].size(); k++) items[x * longest + k] = per[x][k];
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMalloc(&ps->pairs, sizeof(PairDev) * (size_t)npairs);
    if (e == hipSuccess) e = hipMalloc(&ps->order, sizeof(int) * (size_t)npairs);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ps->pairs, host.data(), sizeof(PairDev) * (size_t)npairs, hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ps->order, order.data(), sizeof(int) * (size_t)npairs, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && !items.empty()) {
        e = hipMalloc(&ps->slab_items, sizeof(int2) * items.size());
        if (e == hipSuccess) e = hipMalloc(&ps->vfirst, sizeof(int) * (size_t)nvol);
        if (e == hipSuccess) e = hipMalloc(&ps->vcount, sizeof(int) * (size_t)nvol);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ps->slab_items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice,
                               ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ps->vfirst, vfirst.data(), sizeof(int) * (size_t)nvol, hipMemcpyHostToDevice,
                               ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ps->vcount, vcount.data(), sizeof(int) * (size_t)nvol, hipMemcpyHostToDevice,
                               ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        artia_pairs_destroy(ps);
        return check_hip(ctx, e, "artia_pairs_create");
    }
    *out = ps;
    return ARTIA_OK;
}

artia_status artia_pairs_destroy(artia_pairs* ps) {
    if (!ps) return ARTIA_OK;
    if (ps->pairs) (void)hipFree(ps->pairs);
    if (ps->order) (void)hipFree(ps->order);
    if (ps->slab_items) (void)hipFree(ps->slab_items);
    if (ps->vfirst) (void)hipFree(ps->vfirst);
    if (ps->vcount) (void)hipFree(ps->vcount);
    delete ps;
    return ARTIA_OK;
}

artia_status artia_pairs_project(artia_pairs* ps, const float* volumes, const double* shifts, float* images,
                                 float* ms) {
    if (!ps) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!volumes || !images) return pairs_fail(ps, ARTIA_ERR_INVALID_ARGUMENT, "artia_pairs_project: null pointer");
    artia_ctx* ctx = ps->ctx;
    PairArgs a = base_args(ps);
    a.vols = volumes;
    a.shifts = shifts;
    a.out = images;
    Timing t(ctx, ms);
    hipLaunchKernelGGL(pair_march_kernel<kProject>, pair_grid(ps), dim3(256), 0, ctx->stream, a);
    artia_status st = check_hip(ctx, hipGetLastError(), "artia_pairs_project launch");
    if (st != ARTIA_OK) return st;
    return t.finish();
}

artia_status artia_pairs_loss_grad(artia_pairs* ps, const float* volumes, const double* shifts, const float* targets,
                                   double* loss, double* grad, double* normal, float* resid, float* ms) {
    if (!ps) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!volumes || !targets || !loss || !grad)
        return pairs_fail(ps, ARTIA_ERR_INVALID_ARGUMENT, "artia_pairs_loss_grad: null pointer");
    artia_ctx* ctx = ps->ctx;
    PairArgs a = base_args(ps);
    a.vols = volumes;
    a.shifts = shifts;
    a.in = targets;
    a.out = resid;
    a.loss = loss;
    a.grad = grad;
    a.normal = normal;
    Timing t(ctx, ms);
    hipLaunchKernelGGL(pair_march_kernel<kLossGrad>, pair_grid(ps), dim3(256), 0, ctx->stream, a);
    artia_status st = check_hip(ctx, hipGetLastError(), "artia_pairs_loss_grad launch");
    if (st != ARTIA_OK) return st;
    return t.finish();
}

artia_status artia_pairs_shift_grad(artia_pairs* ps, const float* volumes, const double* shifts,
                                    const float* weights, double* grad, float* ms) {
    if (!ps) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!volumes || !weights || !grad)
        return pairs_fail(ps, ARTIA_ERR_INVALID_ARGUMENT, "artia_pairs_shift_grad: null pointer");
    artia_ctx* ctx = ps->ctx;
    PairArgs a = base_args(ps);
    a.vols = volumes;
    a.shifts = shifts;
    a.in = weights;
    a.grad = grad;
    Timing t(ctx, ms);
    hipLaunchKernelGGL(pair_march_kernel<kShiftGrad>, pair_grid(ps), dim3(256), 0, ctx->stream, a);
    artia_status st = check_hip(ctx, hipGetLastError(), "artia_pairs_shift_grad launch");
    if (st != ARTIA_OK) return st;
    return t.finish();
}

artia_status artia_pairs_adjoint(artia_pairs* ps, const double* shifts, const float* images, float scale,
                                 float* vol_grads, float* ms) {
    if (!ps) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!images || !vol_grads) return pairs_fail(ps, ARTIA_ERR_INVALID_ARGUMENT, "artia_pairs_adjoint: null pointer");
    artia_ctx* ctx = ps->ctx;
    PairArgs a = base_args(ps);
    a.shifts = shifts;
    a.in = images;
    Timing t(ctx, ms);
    if (ps->slab_items) {
        const AdjSlabArgs sa{ps->slab_items, ps->slab_per_xcd, ps->vfirst, ps->vcount};
        hipLaunchKernelGGL(pair_adjoint_slab_kernel, dim3((unsigned)(8 * ps->slab_per_xcd)), dim3(256), 0, ctx->stream,
                           a, sa, scale, vol_grads);
    } else {
        hipLaunchKernelGGL(pair_adjoint_kernel, pair_grid(ps), dim3(256), 0, ctx->stream, a, scale, vol_grads);
    }
    artia_status st = check_hip(ctx, hipGetLastError(), "artia_pairs_adjoint launch");
    if (st != ARTIA_OK) return st;
    return t.finish();
}

artia_status artia_pairs_cc_search(artia_pairs* ps, const float* models, const float* targets, int square,
                                   int max_shift, int multi_peak, int* peak, float* cc_value, float* cc_map,
                                   float* ms) {
    if (!ps) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!models || !targets || !peak || !cc_value)
        return pairs_fail(ps, ARTIA_ERR_INVALID_ARGUMENT, "artia_pairs_cc_search: null pointer");
    if (max_shift < 0 || max_shift > kMaxCcShift || square < 2 * max_shift + 1)
        return fail(ps->ctx, ARTIA_ERR_INVALID_ARGUMENT,
                    "artia_pairs_cc_search: max_shift %d must be in 0..%d and square %d at least 2 max_shift + 1",
                    max_shift, kMaxCcShift, square);
    artia_ctx* ctx = ps->ctx;
    Timing t(ctx, ms);
    hipLaunchKernelGGL(pair_cc_kernel, dim3((unsigned)ps->npairs), dim3(256), 0, ctx->stream, ps->pairs, ps->npairs,
                       models, targets, square, max_shift, multi_peak ? 1 : 0, peak, cc_value,
                       cc_map);
    artia_status st = check_hip(ctx, hipGetLastError(), "artia_pairs_cc_search launch");
    if (st != ARTIA_OK) return st;
    return t.finish();
}

}  // extern "C"
