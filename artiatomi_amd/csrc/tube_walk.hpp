// LDS-staged cell walk of a tile's parallel sub-rays (the exact forward projector and the
// refinement pair kernels).
//
// The 256 sub-rays of an 8x8-pixel tile are parallel, so the lattice cells they cross between two
// z planes lie in a box that grows linearly with the slab: the block walks the volume in slabs of
// kSlabCells cell layers along the rays' z direction, stages each slab's box of voxels in LDS with
// coalesced row loads, and reads the eight corners of every cell from LDS instead of through the
// texture addresser.  The box is padded with the clamp-addressed border voxels (indices -1 and n),
// so cells need no clamping.  A lane whose cells in a slab provably lie in the box walks with an
// LDS offset updated per step; otherwise each cell is tested, and a cell outside the box (not
// expected) is read from global memory.  Every lane runs the same CellWalker steps in the same
// order as an unstaged walk, so sums are bit-identical to it.
#pragma once

#include <climits>
#include <cmath>

#include "fp_common.hpp"

namespace artia {

constexpr int kSlabCells = 8;        // cell layers per slab
constexpr int kChunkFloats = 5120;   // LDS staging capacity per block (20 KB)
constexpr int kXcdGroup = 4;         // tiles per side of the tile groups dealt to one XCD
constexpr int kSlicerXcdGroup = 6;   // the same for the one-pass CTF slicer (deep volumes: 6 measured
                                     // 4 % faster than 4 at 2048x2048x512)

struct TubePlan {
    float bx, by;   // dq_x / dq_z and dq_y / dq_z (texel units; equal for all sub-rays)
    int zdir;       // sign of dq_z (nonzero)
    int tiles_x, tiles_y, groups_x;   // launch layout of the forward projector (8x8-pixel tiles)
};

// The plan of one tilt geometry, or false when the staged walk does not apply: volumes below 2
// voxels on an axis, rays parallel to the z layers, or a tile's slab box larger than the LDS chunk
// (steep rays).  The box of an 8x8-pixel tile spans the tile's extent in a_x = Q_x - bx Q_z
// (affine in the pixel) plus the rays' drift over a slab.
// span: the tile's extent in pixels between its outermost sub-rays (7.5 for 8 pixels at offsets
// 0.25 .. 7.75); slab: cell layers per slab.
inline bool tube_plan(const VolG& v, const TiltG& g, TubePlan& tp, double span = 7.5, int slab = kSlabCells) {
    // (nz bounded so that the walk's float box offsets stay exact: |cell z * CXY| < 2^24, CXY being
    // at most kChunkFloats / 2)
    if (v.nx < 2 || v.ny < 2 || v.nz < 2 || (long long)(v.nz + 2) * (kChunkFloats / 2) >= (1LL << 24)) return false;
    double sc[3], N[3], dpx[3], dpy[3];
    for (int i = 0; i < 3; i++) {
        sc[i] = (double)v.rcp[i] * (double)v.dim[i];
        N[i] = (double)g.n[i] * sc[i];
        dpx[i] = ((double)g.A[0] * g.u[i] + (double)g.A[3] * g.v[i]) * sc[i];
        dpy[i] = ((double)g.A[1] * g.u[i] + (double)g.A[4] * g.v[i]) * sc[i];
    }
    if (!(std::fabs(N[2]) > 1e-3 * (std::fabs(N[0]) + std::fabs(N[1]) + std::fabs(N[2])))) return false;
    const double bx = N[0] / N[2], by = N[1] / N[2];
    const double wx = span * (std::fabs(dpx[0] - bx * dpx[2]) + std::fabs(dpy[0] - bx * dpy[2]));
    const double wy = span * (std::fabs(dpx[1] - by * dpx[2]) + std::fabs(dpy[1] - by * dpy[2]));
    const double cx = wx + std::fabs(bx) * (slab + 1) + 5.0, cy = wy + std::fabs(by) * (slab + 1) + 5.0;
    if (cx * cy * (slab + 1) > (double)kChunkFloats) return false;
    tp.bx = (float)bx;
    tp.by = (float)by;
    tp.zdir = N[2] > 0.0 ? 1 : -1;
    tp.tiles_x = tp.tiles_y = tp.groups_x = 0;
    return true;
}

// A value equal in every lane, moved to a scalar register.
__device__ __forceinline__ float wave_uniform(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Sample step shared by all sub-rays (subray_samples' dq) and the walk's direction part, in scalar
// registers.
struct TubeDir {
    float dq[3], inv[3];
    int sgn[3];
};

__device__ __forceinline__ TubeDir tube_dir(const VolG& v, const TiltG& g, double step) {
    TubeDir d;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        d.dq[i] = wave_uniform((float)(step * (double)g.n[i] * ((double)v.rcp[i] * (double)v.dim[i])));
        float inv;
        walk_dir(d.dq[i], d.sgn[i], inv);
        d.inv[i] = wave_uniform(inv);
    }
    return d;
}

// Every sub-ray of the tile lies on a line Q(z) = (a_x + bx z, a_y + by z, z) in texel
// coordinates; the tile's corner sub-rays (pixel coordinates px_lo..px_hi, py_lo..py_hi, detector
// shift (sx, sy) as in subray_samples) bound a_x and a_y.
struct TubeBox {
    float axl, axh, ayl, ayh;
};

__device__ __forceinline__ TubeBox tube_box(const VolG& v, const TiltG& g, const TubePlan& tp, double px_lo,
                                            double px_hi, double py_lo, double py_hi, double sx, double sy) {
    float axl = INFINITY, axh = -INFINITY, ayl = INFINITY, ayh = -INFINITY;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const double px = (c & 1) ? px_hi : px_lo, py = (c & 2) ? py_hi : py_lo;
        const double xa = (((double)g.A[0] * px + (double)g.A[1] * py) + (double)g.A[2]) - sx;
        const double ya = (((double)g.A[3] * px + (double)g.A[4] * py) + (double)g.A[5]) - sy;
        double q[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double ci = ((double)g.d[i] + xa * (double)g.u[i]) + ya * (double)g.v[i];
            q[i] = (ci - (double)v.bbMin[i]) * ((double)v.rcp[i] * (double)v.dim[i]) - 0.5 -
                   (i == 2 ? (double)v.zshift : 0.0);
        }
        const float ax = (float)(q[0] - q[2] * (double)tp.bx), ay = (float)(q[1] - q[2] * (double)tp.by);
        axl = fminf(axl, ax);
        axh = fmaxf(axh, ax);
        ayl = fminf(ayl, ay);
        ayh = fmaxf(ayh, ay);
    }
    return TubeBox{wave_uniform(axl), wave_uniform(axh), wave_uniform(ayl), wave_uniform(ayh)};
}

// Corners of a cell from the staged box when it lies inside, otherwise from global memory with
// clamped indices.  The LDS reads are issued unconditionally (at offset 0 when outside, which stays
// inside the box) and overwritten on the rare global path: branching between the two sources would
// let the compiler merge them into generic loads that wait on both memory counters.
template <typename T>
__device__ __forceinline__ void chunk_corners(const VolG& v, const T* __restrict__ vol,
                                              const float* __restrict__ chunk, bool staged, int X0, int Y0, int Z0,
                                              int CX, int CY, int CZ, const int cell[3], float val[8]) {
    const int cx = cell[0] - X0, cy = cell[1] - Y0, cz = cell[2] - Z0;
    const bool inside = staged && (unsigned)cx < (unsigned)(CX - 1) && (unsigned)cy < (unsigned)(CY - 1) &&
                        (unsigned)cz < (unsigned)(CZ - 1);
    // outside: every read at chunk[0] or chunk[1]
    const int o = inside ? (cz * CY + cy) * CX + cx : 0;
    const int dy = inside ? CX : 0, dz = inside ? CX * CY : 0;
    val[0] = chunk[o];
    val[1] = chunk[o + 1];
    val[2] = chunk[o + dy];
    val[3] = chunk[o + dy + 1];
    val[4] = chunk[o + dz];
    val[5] = chunk[o + dz + 1];
    val[6] = chunk[o + dz + dy];
    val[7] = chunk[o + dz + dy + 1];
    if (!inside) {
        size_t g[8];
        cell_corners<true>(v, cell, g);
#pragma unroll
        for (int i = 0; i < 8; i++) val[i] = vox(vol, g[i]);
    }
}

// Box entry e = (zz * CY + yy) * CX + xx of the thread's strided pass over a padded box (entries
// threadIdx.x, threadIdx.x + 256, ...), stepped incrementally instead of dividing per entry.
struct BoxCursor {
    int xx, yy, zz;
    int sx, sy, sz;   // the step of 256 entries, as (x, y, z) increments before carries
    int CX, CY;
    __device__ __forceinline__ BoxCursor(int e, int cx, int cy) : CX(cx), CY(cy) {
        const int row = e / cx;
        xx = e - row * cx;
        zz = row / cy;
        yy = row - zz * cy;
        const int rs = 256 / cx;
        sx = 256 - rs * cx;
        sz = rs / cy;
        sy = rs - sz * cy;
    }
    __device__ __forceinline__ void next() {
        xx += sx;
        yy += sy;
        zz += sz;
        if (xx >= CX) {
            xx -= CX;
            yy++;
        }
        if (yy >= CY) {
            yy -= CY;
            zz++;
        }
    }
    // global index of the entry (clamp-addressed: -1 and n map to the edge voxels)
    // (rows z * ny + y of the volume number below 2^31)
    __device__ __forceinline__ size_t voxel(const VolG& v, int X0, int Y0, int Z0) const {
        const int gx = clampi(X0 + xx, 0, v.nx - 1), gy = clampi(Y0 + yy, 0, v.ny - 1);
        const int gz = clampi(Z0 + zz, 0, v.nz - 1);
        return (size_t)(unsigned)(gz * v.ny + gy) * (size_t)(unsigned)v.nx + (size_t)(unsigned)gx;
    }
};

// Stages the padded box [X0, X0 + CX) x [Y0, ..) x [Z0, ..) (n = CX * CY * CZ entries) into chunk:
// four loads in flight per thread before their LDS writes.
template <typename T>
__device__ __forceinline__ void stage_box(const VolG& v, const T* __restrict__ vol, float* __restrict__ chunk, int X0,
                                          int Y0, int Z0, int CX, int CY, int n) {
    int e = (int)threadIdx.x;
    if (e >= n) return;
    BoxCursor c(e, CX, CY);
    for (; e + 3 * 256 < n; e += 4 * 256) {
        float val[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            val[k] = vox(vol, c.voxel(v, X0, Y0, Z0));
            c.next();
        }
#pragma unroll
        for (int k = 0; k < 4; k++) chunk[e + k * 256] = val[k];
    }
    for (; e < n; e += 256) {
        chunk[e] = vox(vol, c.voxel(v, X0, Y0, Z0));
        c.next();
    }
}


// stage_box row by row, for boxes up to 256 entries wide (wider ones go to stage_box): thread t keeps column t mod CX and takes every (256 / CX)-th row from row t / CX on, so its
// x index and clamp are fixed and a row step is two adds and a carry (the rows z * ny + y of the
// volume number below 2^31, z and ny each below 2^23: 24-bit multiplies).
template <typename T>
__device__ __forceinline__ void stage_box_rows(const VolG& v, const T* __restrict__ vol, float* __restrict__ chunk,
                                               int X0, int Y0, int Z0, int CX, int CY, int n) {
    const int P = 256 / CX;   // rows per pass
    if (P > 0) {
        const int rows = n / CX;
        const int t = (int)threadIdx.x, r0 = t / CX, x = t - r0 * CX;
        if (r0 >= P || r0 >= rows) return;
        const int dz = P / CY, dy = P - dz * CY;
        int zz = r0 / CY, yy = r0 - zz * CY;
        if (X0 >= 0 && X0 + CX <= v.nx && Y0 >= 0 && Y0 + CY <= v.ny && Z0 >= 0 && Z0 + rows / CY <= v.nz) {
            // the box inside the volume (no clamps, block-uniform): a voxel pointer and an LDS
            // pointer stepped per row, the voxel step one of two (with or without the y carry)
            const long long nx = v.nx, nxy = nx * (long long)v.ny;
            const T* __restrict__ pv = vol + ((long long)(Z0 + zz) * nxy + (long long)(Y0 + yy) * nx + (X0 + x));
            const long long stepN = (long long)dy * nx + (long long)dz * nxy, stepC = stepN - (long long)CY * nx + nxy;
            float* __restrict__ q = chunk + (r0 * CX + x);
            const int qs = P * CX;
            auto next = [&]() {
                yy += dy;
                const bool c = yy >= CY;
                yy = c ? yy - CY : yy;
                pv += c ? stepC : stepN;
            };
            int r = r0;
            for (; r + 3 * P < rows; r += 4 * P) {
                float val[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    val[k] = vox(pv, 0);
                    next();
                }
#pragma unroll
                for (int k = 0; k < 4; k++) q[k * qs] = val[k];
                q += 4 * qs;
            }
            for (; r < rows; r += P) {
                *q = vox(pv, 0);
                next();
                q += qs;
            }
            return;
        }
        const size_t gx = (size_t)(unsigned)clampi(X0 + x, 0, v.nx - 1);
        auto load = [&]() {
            const int gy = clampi(Y0 + yy, 0, v.ny - 1), gz = clampi(Z0 + zz, 0, v.nz - 1);
            const float val = vox(vol, (size_t)(unsigned)(__mul24(gz, v.ny) + gy) * (size_t)(unsigned)v.nx + gx);
            yy += dy;
            zz += dz;
            if (yy >= CY) {
                yy -= CY;
                zz++;
            }
            return val;
        };
        int r = r0;
        for (; r + 3 * P < rows; r += 4 * P) {
            float val[4];
#pragma unroll
            for (int k = 0; k < 4; k++) val[k] = load();
#pragma unroll
            for (int k = 0; k < 4; k++) chunk[__mul24(r + k * P, CX) + x] = val[k];
        }
        for (; r < rows; r += P) chunk[__mul24(r, CX) + x] = load();
        return;
    }
    stage_box(v, vol, chunk, X0, Y0, Z0, CX, CY, n);
}

// Walks every lane's samples q0 + k*dq (k < count) cell by cell, slab by slab through LDS.  Must be
// called by all 256 threads of the block (it synchronises); chunk holds kChunkFloats floats and
// zr two ints of shared memory.
//   gather (kScatter false): f(const float val[8], cell, k, m) for every cell holding samples
//     k..k+m-1, with the cell's eight corner values read from vol (staged in LDS); f may also see
//     cells with m = 0, to which it must add nothing but +-0;
//   scatter (kScatter true): f(float wgt[8], cell, k, m) returns eight corner weights, which are
//     added to the slab's box in LDS (shared-memory atomics) and then to grad with one global
//     atomic per box voxel, instead of eight global atomics per cell.
// Box offset cell[2] * CXY + cell[0] + K of a cell of the walk without the y axis: float cells
// (integers; tube_plan keeps |cell[2] * CXY| below 2^24, so every term and sum is exact) in one fma,
// int cells in one 24-bit multiply-add.
__device__ __forceinline__ int flat_box_offset(const float* cell, int CXY, int K) {
    return (int)__builtin_fmaf(cell[2], (float)CXY, cell[0] + (float)K);
}
__device__ __forceinline__ int flat_box_offset(const int* cell, int CXY, int K) {
    return __mul24(cell[2], CXY) + (cell[0] + K);
}

template <bool kScatter, bool kFlat, bool kRows, typename T, class F>
__device__ __forceinline__ void tube_slabs(const VolG& v, const T* __restrict__ vol, float* __restrict__ grad,
                                           float* __restrict__ chunk, int* zr, const TubePlan& tp, const TubeBox& tb,
                                           const TubeDir& td, const float q0[3], int count, F&& f) {
    // cell layers the block's sub-rays cross
    __syncthreads();   // a previous walk's readers are done with zr and chunk
    if (threadIdx.x == 0) {
        zr[0] = INT_MAX;
        zr[1] = INT_MIN;
    }
    __syncthreads();
    if (count > 0) {
        const int za = (int)floorf(q0[2]);
        const int zb = (int)floorf(__builtin_fmaf((float)(count - 1), td.dq[2], q0[2]));
        atomicMin(&zr[0], min(za, zb) - 1);
        atomicMax(&zr[1], max(za, zb) + 1);
    }
    __syncthreads();
    const int zlo = max(zr[0], -1), zhi = min(zr[1], v.nz - 1);

    // corner weights into the box (LDS) or, outside it, straight to grad
    auto scatter_global = [&](const int cell[3], const float wgt[8]) {
        size_t o[8];
        cell_corners<true>(v, cell, o);
#pragma unroll
        for (int i = 0; i < 8; i++)
            if (wgt[i] != 0.0f) unsafeAtomicAdd(grad + o[i], wgt[i]);
    };
    CellWalkerT<kFlat> w;
    w.init_dir(q0, td.sgn, td.inv, count);
    const int nslab = zlo <= zhi ? (zhi - zlo) / kSlabCells + 1 : 0;
    for (int j = 0; j < nslab; j++) {
        // cell layers [zs, ze] of this slab, in the rays' z direction
        const int zs = tp.zdir > 0 ? zlo + j * kSlabCells : zhi - (j + 1) * kSlabCells + 1;
        const int ze = zs + kSlabCells - 1;
        // box of lattice indices the slab's cells touch, in index space padded by the clamp border
        // (-1 .. n); the margins cover the float rounding of the walk
        const float z0 = (float)zs, z1 = (float)(ze + 1);
        const int X0 = wave_uniform(max((int)floorf(tb.axl + fminf(tp.bx * z0, tp.bx * z1)) - 1, -1));
        const int X1 = min((int)floorf(tb.axh + fmaxf(tp.bx * z0, tp.bx * z1)) + 2, v.nx);
        const int Y0 = wave_uniform(max((int)floorf(tb.ayl + fminf(tp.by * z0, tp.by * z1)) - 1, -1));
        const int Y1 = min((int)floorf(tb.ayh + fmaxf(tp.by * z0, tp.by * z1)) + 2, v.ny);
        const int Z0 = max(zs, -1), Z1 = min(ze + 1, v.nz);
        const int CX = wave_uniform(X1 - X0 + 1), CY = wave_uniform(Y1 - Y0 + 1), CZ = Z1 - Z0 + 1;
        const bool staged = CX > 1 && CY > 1 && CZ > 1 && CX * CY * CZ <= kChunkFloats;
        const int n = CX * CY * CZ;
        __syncthreads();   // the previous slab's reads (or flush) are done
        if (staged) {
            if (kScatter)
                for (int e = (int)threadIdx.x; e < n; e += 256) chunk[e] = 0.0f;
            else if (kRows)
                stage_box_rows(v, vol, chunk, X0, Y0, Z0, CX, CY, n);
            else
                stage_box(v, vol, chunk, X0, Y0, Z0, CX, CY, n);
        }
        __syncthreads();
        // A lane whose cells in this slab provably lie in the box (its samples here run from the
        // current cell to the slab's exit plane; one cell of margin covers the walk's rounding)
        // walks with an LDS offset updated per step and no per-cell test.
        bool fast = false;
        if (staged && w.active()) {
            const float zp = tp.zdir > 0 ? (float)(ze + 1) : (float)zs;
            const float ke = fminf(fmaxf((zp - q0[2]) * td.inv[2], (float)w.k), (float)(w.count - 1));
            const int xb = (int)floorf(__builtin_fmaf(ke, td.dq[0], q0[0]));
            const int yb = (int)floorf(__builtin_fmaf(ke, td.dq[1], q0[1]));
            const int zl = tp.zdir > 0 ? w.icell(2) : max(zs, -1);
            const int zh = tp.zdir > 0 ? min(ze + 1, v.nz) : w.icell(2) + 1;
            fast = min(w.icell(0), xb) - 1 >= X0 && max(w.icell(0), xb) + 2 <= X1 && min(w.icell(1), yb) - 1 >= Y0 &&
                   max(w.icell(1), yb) + 2 <= Y1 && zl >= Z0 && zh <= Z1;
        }
        if (fast) {
            const int CXY = CX * CY;
            const int st0 = td.sgn[0], st1 = td.sgn[1] * CX, st2 = td.sgn[2] * CXY;
            const int off[8] = {0, 1, CX, CX + 1, CXY, CXY + 1, CXY + CX, CXY + CX + 1};
            int o = ((w.icell(2) - Z0) * CY + (w.icell(1) - Y0)) * CX + (w.icell(0) - X0);
            // a fast lane enters with its cell in [zs, ze] and moves monotonically in z, so staying
            // in the slab is one range test
            if (kScatter) {
                while (w.active() && w.z_in(zs, kSlabCells)) {
                    w.step(
                        [&](const auto* cell, auto k, auto mi) {
                            float val[8];
                            f(val, cell, k, mi);
#pragma unroll
                            for (int i = 0; i < 8; i++)
                                if (val[i] != 0.0f) atomicAdd(chunk + o + off[i], val[i]);
                        },
                        [&](int i) { o += i == 0 ? st0 : (i == 1 ? st1 : st2); });
                }
            } else {
                // every step reads its cell's corners and calls f, also for a cell without samples
                // (f adds +-0 then: the gather callbacks add cell and gradient sums), so the loop
                // body has no branch
                if (kFlat) {
                    // y is fixed: the box offset is one 24-bit multiply-add of the x and z indices
                    // per step instead of a select and add per axis
                    const int K = (w.icell(1) - Y0) * CX - X0 - Z0 * CXY;
                    while (w.active() && w.z_in(zs, kSlabCells)) {
                        w.step_all(
                            [&](const auto* cell, auto k, auto mi) {
                                const int oc = flat_box_offset(cell, CXY, K);
                                float val[8];
#pragma unroll
                                for (int i = 0; i < 8; i++) val[i] = chunk[oc + off[i]];
                                f(val, cell, k, mi);
                            },
                            [](int) {});
                    }
                } else {
                    while (w.active() && w.z_in(zs, kSlabCells)) {
                        w.step_all(
                            [&](const auto* cell, auto k, auto mi) {
                                float val[8];
#pragma unroll
                                for (int i = 0; i < 8; i++) val[i] = chunk[o + off[i]];
                                f(val, cell, k, mi);
                            },
                            [&](int i) { o += i == 0 ? st0 : (i == 1 ? st1 : st2); });
                    }
                }
            }
        } else {
            while (w.active() && (tp.zdir > 0 ? w.icell(2) <= ze : w.icell(2) >= zs)) {
                w.step([&](const auto* cell, auto k, auto mi) {
                    float val[8];
                    if (kScatter) {
                        f(val, cell, k, mi);
                        const int ic[3] = {(int)cell[0], (int)cell[1], (int)cell[2]};
                        const int cx = ic[0] - X0, cy = ic[1] - Y0, cz = ic[2] - Z0;
                        if (staged && (unsigned)cx < (unsigned)(CX - 1) && (unsigned)cy < (unsigned)(CY - 1) &&
                            (unsigned)cz < (unsigned)(CZ - 1)) {
                            const int o = (cz * CY + cy) * CX + cx, CXY = CX * CY;
                            const int off[8] = {0, 1, CX, CX + 1, CXY, CXY + 1, CXY + CX, CXY + CX + 1};
#pragma unroll
                            for (int i = 0; i < 8; i++)
                                if (val[i] != 0.0f) atomicAdd(chunk + o + off[i], val[i]);
                        } else {
                            scatter_global(ic, val);
                        }
                    } else {
                        const int ic[3] = {(int)cell[0], (int)cell[1], (int)cell[2]};
                        chunk_corners(v, vol, chunk, staged, X0, Y0, Z0, CX, CY, CZ, ic, val);
                        f(val, cell, k, mi);
                    }
                });
            }
        }
        if (kScatter && staged) {
            __syncthreads();   // the box is complete
            if ((int)threadIdx.x < n) {
                BoxCursor c((int)threadIdx.x, CX, CY);
                for (int e = (int)threadIdx.x; e < n; e += 256) {
                    const float a = chunk[e];
                    if (a != 0.0f) unsafeAtomicAdd(grad + c.voxel(v, X0, Y0, Z0), a);
                    c.next();
                }
            }
        }
    }
    // cells beyond the slabs (not expected): global memory
    while (w.active()) {
        w.step([&](const auto* cell, auto k, auto mi) {
            float val[8];
            const int ic[3] = {(int)cell[0], (int)cell[1], (int)cell[2]};
            if (kScatter) {
                f(val, cell, k, mi);
                scatter_global(ic, val);
            } else {
                size_t o[8];
                cell_corners<true>(v, ic, o);
#pragma unroll
                for (int i = 0; i < 8; i++) val[i] = vox(vol, o[i]);
                f(val, cell, k, mi);
            }
        });
    }
}

// The gather walk over a run of segments per lane: consecutive sample sets along the lane's ray,
// each with its own start q0 and count (a slicer sample lattice per CTF slab).  seg.next() finishes
// the lane's current segment, if any, and loads the next into seg.q0 and seg.count, returning false
// when none is left; f(val, cell, k, m) sees every cell of the current segment.  The segments follow
// each other along the rays, so one pass over the slabs of cell layers serves all of them; za..zb
// are cell layers bounding every sample of the lane (za > zb: none).  Each segment is walked with
// tube_slabs' CellWalker steps, so its cells, their order and their values are those of a
// tube_walk of that segment alone (kFlat: the walk without the y axis, as tube_walk<true>).  Must
// be called by all 256 threads of the block.
template <bool kFlat, bool kRows, typename T, class Seg, class F>
__device__ __forceinline__ void tube_walk_segments(const VolG& v, const T* __restrict__ vol,
                                                   float* __restrict__ chunk, int* zr, const TubePlan& tp,
                                                   const TubeBox& tb, const TubeDir& td, int za, int zb, Seg& seg,
                                                   F&& f) {
    __syncthreads();
    if (threadIdx.x == 0) {
        zr[0] = INT_MAX;
        zr[1] = INT_MIN;
    }
    __syncthreads();
    if (za <= zb) {
        atomicMin(&zr[0], za - 1);
        atomicMax(&zr[1], zb + 1);
    }
    __syncthreads();
    const int zlo = max(zr[0], -1), zhi = min(zr[1], v.nz - 1);

    CellWalkerT<kFlat> w;
    w.init_dir(seg.q0, td.sgn, td.inv, 0);   // no segment yet: the first pass below loads one
    bool more = true;
    const int nslab = zlo <= zhi ? (zhi - zlo) / kSlabCells + 1 : 0;
    // slabs of cell layers, then (j == nslab) the cells beyond them (not expected) from global memory
    for (int j = 0; j <= nslab; j++) {
        const bool tail = j == nslab;
        const int zs = tp.zdir > 0 ? zlo + j * kSlabCells : zhi - (j + 1) * kSlabCells + 1;
        const int ze = zs + kSlabCells - 1;
        const float z0 = (float)zs, z1 = (float)(ze + 1);
        const int X0 = wave_uniform(max((int)floorf(tb.axl + fminf(tp.bx * z0, tp.bx * z1)) - 1, -1));
        const int X1 = min((int)floorf(tb.axh + fmaxf(tp.bx * z0, tp.bx * z1)) + 2, v.nx);
        const int Y0 = wave_uniform(max((int)floorf(tb.ayl + fminf(tp.by * z0, tp.by * z1)) - 1, -1));
        const int Y1 = min((int)floorf(tb.ayh + fmaxf(tp.by * z0, tp.by * z1)) + 2, v.ny);
        const int Z0 = max(zs, -1), Z1 = min(ze + 1, v.nz);
        const int CX = wave_uniform(X1 - X0 + 1), CY = wave_uniform(Y1 - Y0 + 1), CZ = Z1 - Z0 + 1;
        const bool staged = !tail && CX > 1 && CY > 1 && CZ > 1 && CX * CY * CZ <= kChunkFloats;
        const int n = CX * CY * CZ;
        if (!tail) {
            __syncthreads();
            if (staged && kRows)
                stage_box_rows(v, vol, chunk, X0, Y0, Z0, CX, CY, n);
            else if (staged)
                stage_box(v, vol, chunk, X0, Y0, Z0, CX, CY, n);
            __syncthreads();
        }
        for (;;) {
            if (!w.active()) {
                if (!more) break;
                more = seg.next();
                w.init_dir(seg.q0, td.sgn, td.inv, more ? seg.count : 0);
                continue;
            }
            if (!tail && (tp.zdir > 0 ? w.icell(2) > ze : w.icell(2) < zs)) break;   // resumes in a later slab
            bool fast = false;
            if (staged) {
                const float zp = tp.zdir > 0 ? (float)(ze + 1) : (float)zs;
                const float ke = fminf(fmaxf((zp - seg.q0[2]) * td.inv[2], (float)w.k), (float)(w.count - 1));
                const int xb = (int)floorf(__builtin_fmaf(ke, td.dq[0], seg.q0[0]));
                const int yb = (int)floorf(__builtin_fmaf(ke, td.dq[1], seg.q0[1]));
                const int zl = tp.zdir > 0 ? w.icell(2) : max(zs, -1);
                const int zh = tp.zdir > 0 ? min(ze + 1, v.nz) : w.icell(2) + 1;
                fast = min(w.icell(0), xb) - 1 >= X0 && max(w.icell(0), xb) + 2 <= X1 && min(w.icell(1), yb) - 1 >= Y0 &&
                       max(w.icell(1), yb) + 2 <= Y1 && zl >= Z0 && zh <= Z1;
            }
            if (fast) {
                const int CXY = CX * CY;
                const int st0 = td.sgn[0], st1 = td.sgn[1] * CX, st2 = td.sgn[2] * CXY;
                const int off[8] = {0, 1, CX, CX + 1, CXY, CXY + 1, CXY + CX, CXY + CX + 1};
                int o = ((w.icell(2) - Z0) * CY + (w.icell(1) - Y0)) * CX + (w.icell(0) - X0);
                // every step reads its cell's corners and calls f, also for a cell without
                // samples (f adds +-0 then), so the loop body has no branch
                if (kFlat) {
                    // y is fixed: the box offset is one 24-bit multiply-add of the x and z indices
                    const int K = (w.icell(1) - Y0) * CX - X0 - Z0 * CXY;
                    while (w.active() && w.z_in(zs, kSlabCells)) {
                        w.step_all(
                            [&](const auto* cell, auto k, auto mi) {
                                const int oc = flat_box_offset(cell, CXY, K);
                                float val[8];
#pragma unroll
                                for (int i = 0; i < 8; i++) val[i] = chunk[oc + off[i]];
                                f(val, cell, k, mi);
                            },
                            [](int) {});
                    }
                } else {
                    while (w.active() && w.z_in(zs, kSlabCells)) {
                        w.step_all(
                            [&](const auto* cell, auto k, auto mi) {
                                float val[8];
#pragma unroll
                                for (int i = 0; i < 8; i++) val[i] = chunk[o + off[i]];
                                f(val, cell, k, mi);
                            },
                            [&](int i) { o += i == 0 ? st0 : (i == 1 ? st1 : st2); });
                    }
                }
            } else {
                while (w.active() && (tail || (tp.zdir > 0 ? w.icell(2) <= ze : w.icell(2) >= zs))) {
                    w.step([&](const auto* cell, auto k, auto mi) {
                        float val[8];
                        const int ic[3] = {(int)cell[0], (int)cell[1], (int)cell[2]};
                        chunk_corners(v, vol, chunk, staged, X0, Y0, Z0, CX, CY, CZ, ic, val);
                        f(val, cell, k, mi);
                    });
                }
            }
        }
    }
}

// The gather walk (see tube_slabs); kFlat: the rays have no y component (td.dq[1] == 0); kRows:
// the boxes staged by stage_box_rows.
template <bool kFlat = false, bool kRows = false, typename T, class F>
__device__ __forceinline__ void tube_walk(const VolG& v, const T* __restrict__ vol, float* __restrict__ chunk,
                                          int* zr, const TubePlan& tp, const TubeBox& tb, const TubeDir& td,
                                          const float q0[3], int count, F&& f) {
    tube_slabs<false, kFlat, kRows>(v, vol, nullptr, chunk, zr, tp, tb, td, q0, count, f);
}

// The scatter walk (see tube_slabs): grad += the weights f returns.
template <class F>
__device__ __forceinline__ void tube_scatter(const VolG& v, float* __restrict__ grad, float* __restrict__ chunk,
                                             int* zr, const TubePlan& tp, const TubeBox& tb, const TubeDir& td,
                                             const float q0[3], int count, F&& f) {
    tube_slabs<true, false, false>(v, static_cast<const float*>(nullptr), grad, chunk, zr, tp, tb, td, q0, count, f);
}

}  // namespace artia
