// Exact-numerics ray sampling shared by the forward projector (fp.hip) and its adjoint and
// shift-gradient kernels (adjoint.hip): double-precision ray setup and slab clip, the affine sample
// coordinates of a sub-ray, and the clamp-addressed trilinear read
// (ForwardProjectionRayMarcher_TL.cu:57-160 evaluated stably).
#pragma once

#include <type_traits>

#include <cmath>
#include <cstdint>

#include "device_common.hpp"

namespace artia {

// Exact numerics take a beam whose y component is a float residue of the rotation (|n_y| below
// 1e-6 of |n|; every tilt about the y axis whose rotation leaves one) as parallel to the x-z plane:
// n_y = 0, so the walk runs without the y axis (CellWalkerT<true>).  A ray then moves by less than
// 1e-6 voxels over a 1000-voxel path.  Every exact forward-family launcher (march, slicer, traversal
// length, adjoint, shift gradient, pair sets) takes the geometry through here, so they stay exact
// transposes and bit-identical to one another.
inline TiltG flat_beam(TiltG g) {
    const float s = std::fabs(g.n[0]) + std::fabs(g.n[1]) + std::fabs(g.n[2]);
    if (g.n[1] != 0.0f && std::fabs(g.n[1]) < 1e-6f * s) g.n[1] = 0.0f;
    return g;
}

struct RayD {
    double c[3];
    double t_in, t_out;
};

__device__ __forceinline__ RayD ray_setup_d(const VolG& v, const TiltG& g, double px, double py) {
    RayD r;
    const double xa = ((double)g.A[0] * px + (double)g.A[1] * py) + (double)g.A[2];
    const double ya = ((double)g.A[3] * px + (double)g.A[4] * py) + (double)g.A[5];
    double lo[3], hi[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        r.c[i] = ((double)g.d[i] + xa * (double)g.u[i]) + ya * (double)g.v[i];
        const double tE = ((double)v.bbMin[i] - r.c[i]) / (double)g.n[i];
        const double tX = ((double)v.bbMax[i] - r.c[i]) / (double)g.n[i];
        lo[i] = fmin(tE, tX);
        hi[i] = fmax(tE, tX);
    }
    r.t_in = fmax(fmax(lo[0], lo[1]), lo[2]);
    r.t_out = fmin(fmin(hi[0], hi[1]), hi[2]);
    return r;
}

// Pair of x-neighbours vol[o], vol[o+1] (4-byte aligned; the compiler picks the widest legal load).
template <bool kBig>
__device__ __forceinline__ float2 load_pair(const float* __restrict__ vol, uint64_t o64, uint32_t o32) {
    float2 p;
    const float* a = kBig ? vol + o64 : (const float*)((const char*)vol + (o32 << 2));
    __builtin_memcpy(&p, a, sizeof(float2));
    return p;
}

// Trilinear read at voxel coordinates already shifted by -0.5 (texel j centred at j), clamp
// addressing, for volumes of at least 2 voxels per axis.  The coordinate is clamped to
// [0, n-1] and the lower texel to n-2, so each x-neighbour pair is a single load and no
// per-lane selects are needed; beyond the last texel centre the weight is exactly 1 and the
// result is the edge texel up to the rounding of one lerp.
template <bool kBig>
__device__ __forceinline__ float tri_pairs(const float* __restrict__ vol, int nx, int ny, int nz, float xb, float yb,
                                           float zb) {
    const float xc = __builtin_amdgcn_fmed3f(xb, 0.0f, (float)(nx - 1));
    const float yc = __builtin_amdgcn_fmed3f(yb, 0.0f, (float)(ny - 1));
    const float zc = __builtin_amdgcn_fmed3f(zb, 0.0f, (float)(nz - 1));
    const int ib = min((int)xc, nx - 2), jb = min((int)yc, ny - 2), kb = min((int)zc, nz - 2);
    const float a = xc - (float)ib, b = yc - (float)jb, c = zc - (float)kb;
    const uint32_t sy = (uint32_t)nx, sz = (uint32_t)nx * (uint32_t)ny;
    uint32_t o32 = 0;
    uint64_t o64 = 0;
    if (kBig)
        o64 = (uint64_t)kb * sz + (uint64_t)jb * sy + (uint64_t)ib;
    else
        o32 = __umul24((uint32_t)kb, sz) + __umul24((uint32_t)jb, sy) + (uint32_t)ib;
    const float2 p00 = load_pair<kBig>(vol, o64, o32);
    const float2 p10 = load_pair<kBig>(vol, o64 + sy, o32 + sy);
    const float2 p01 = load_pair<kBig>(vol, o64 + sz, o32 + sz);
    const float2 p11 = load_pair<kBig>(vol, o64 + sz + sy, o32 + sz + sy);
    const float c00 = lerpf(p00.x, p00.y, a), c10 = lerpf(p10.x, p10.y, a);
    const float c01 = lerpf(p01.x, p01.y, a), c11 = lerpf(p11.x, p11.y, a);
    return lerpf(lerpf(c00, c10, b), lerpf(c01, c11, b), c);
}

// Sample grid of one sub-ray of the forward projector: samples at t_in + k*step for
// k = 0..count-1 (t <= t_out), as voxel coordinates shifted by -0.5 (texel j centred at j):
// q_k = q0 + k*dq.  Returns false when the sub-ray misses the volume.  (sx, sy) is an extra 2D
// detector shift in pixels, applied in double to the sub-ray origin (c -= sx*u + sy*v, as the
// marker shifts enter the detector origin, Projection.cpp:183-195); zero leaves the sub-ray as is.
struct SubraySamples {
    float q0[3], dq[3];
    int count;
    int entry_axis;   // axis whose slab sets t_in (the derivative of t_in follows it)
};

__device__ __forceinline__ bool subray_samples(const VolG& v, const TiltG& g, double px, double py, double step,
                                               SubraySamples& s, double sx = 0.0, double sy = 0.0) {
    const double xa = (((double)g.A[0] * px + (double)g.A[1] * py) + (double)g.A[2]) - sx;
    const double ya = (((double)g.A[3] * px + (double)g.A[4] * py) + (double)g.A[5]) - sy;
    double c[3], lo[3], hi[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        c[i] = ((double)g.d[i] + xa * (double)g.u[i]) + ya * (double)g.v[i];
        const double tE = ((double)v.bbMin[i] - c[i]) / (double)g.n[i];
        const double tX = ((double)v.bbMax[i] - c[i]) / (double)g.n[i];
        lo[i] = fmin(tE, tX);
        hi[i] = fmax(tE, tX);
    }
    double t_in = lo[0];
    s.entry_axis = 0;
    if (lo[1] > t_in) {
        t_in = lo[1];
        s.entry_axis = 1;
    }
    if (lo[2] > t_in) {
        t_in = lo[2];
        s.entry_axis = 2;
    }
    const double t_out = fmin(fmin(hi[0], hi[1]), hi[2]);
    if (!((t_out - t_in) > 0.0)) return false;
    s.count = (int)fmin(floor((t_out - t_in) / step) + 1.0, (double)kMaxSamplesPerRay);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double scale = (double)v.rcp[i] * (double)v.dim[i];
        const double p = (c[i] + t_in * (double)g.n[i]) - (double)v.bbMin[i];
        s.q0[i] = (float)(p * scale - 0.5 - (i == 2 ? (double)v.zshift : 0.0));
        s.dq[i] = (float)(step * (double)g.n[i] * scale);
    }
    return true;
}

// Walk over the cells of the texel lattice crossed by the samples q0 + k*dq, k = 0..count-1, in
// order.  The sample index of the next crossing on axis i is recomputed from the cell index,
// (face - q0_i) / dq_i, as one fma (its float error, a few 1e-4 samples for rays of thousands of
// samples, only moves a sample lying within that distance of a face to the neighbouring cell,
// where the continuous trilinear interpolant has the same value).  Every step either consumes
// samples or advances an axis; the walk is bounded by 4 * count + 64 steps.  init() sets it up,
// step() handles one cell (calling f(cell, k, m) when it holds samples k..k+m-1) and advances to
// the next, active() says whether samples remain, so the walk can be resumed (the LDS-staged
// projector runs it slab by slab).
//
// kFlat: the rays have no y component (dq[1] == 0 exactly, every tilt about the y axis whose beam
// direction rounds to n_y = 0), so the walk never crosses a y face and the y axis drops out of the
// step.  With dq[1] == 0 both forms visit the same cells with the same sample ranges.
// Direction of one axis: the sign of dq and 1 / dq, or no crossings (sign 0) when dq is 0 or so
// small that 1 / dq overflows (such an axis would cross no face within the walk's samples).
__device__ __forceinline__ void walk_dir(float dq, int& sgn, float& inv) {
    const float r = 1.0f / dq;
    const bool ok = dq != 0.0f && __builtin_isfinite(r);
    sgn = ok ? (dq > 0.0f ? 1 : -1) : 0;
    inv = ok ? r : 0.0f;
}

// The walk without the y axis keeps its cell indices and sample counts in float (cell, k, and the m
// passed to f): integers below 2^24, exact in float, which the cell moments and the crossing
// indices take without conversions.
template <bool kFlat>
struct CellWalkerT {
    using Index = typename std::conditional<kFlat, float, int>::type;
    Index cell[3];
    int sgn[3];
    float inv[3], off[3], kb[3];
    float fcount;
    Index k;
    int count, guard;

    __device__ __forceinline__ int icell(int i) const { return (int)cell[i]; }
    // whether the cell's z index lies in [zs, zs + n)
    __device__ __forceinline__ bool z_in(int zs, int n) const {
        if (kFlat) return cell[2] >= (Index)zs && cell[2] <= (Index)(zs + n - 1);
        return (unsigned)((int)cell[2] - zs) <= (unsigned)(n - 1);
    }

    __device__ __forceinline__ void init(const float q0[3], const float dq[3], int n) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            cell[i] = (Index)floorf(q0[i]);
            walk_dir(dq[i], sgn[i], inv[i]);
            off[i] = -q0[i] * inv[i];
            kb[i] = sgn[i] != 0 ? __builtin_fmaf((float)(cell[i] + (Index)(sgn[i] > 0 ? 1 : 0)), inv[i], off[i])
                                : INFINITY;
        }
        count = n;
        fcount = (float)n;
        k = (Index)0;
        guard = 0;
    }
    // The same walk with the direction part (sgn, inv) precomputed: when every lane walks parallel
    // samples the caller passes them as wave-uniform values, which then stay in scalar registers.
    __device__ __forceinline__ void init_dir(const float q0[3], const int sgn_u[3], const float inv_u[3], int n) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            cell[i] = (Index)floorf(q0[i]);
            sgn[i] = sgn_u[i];
            inv[i] = inv_u[i];
            off[i] = -q0[i] * inv[i];
            kb[i] = sgn[i] != 0 ? __builtin_fmaf((float)(cell[i] + (Index)(sgn[i] > 0 ? 1 : 0)), inv[i], off[i])
                                : INFINITY;
        }
        count = n;
        fcount = (float)n;
        k = (Index)0;
        guard = 0;
    }
    // The walk of samples k0 .. n-1 of the same sub-ray, starting in the cell of sample k0 (crossing
    // indices stay relative to q0, so moments use the sub-ray's own sample indices).
    __device__ __forceinline__ void init_from(const float q0[3], const float dq[3], const int sgn_u[3],
                                              const float inv_u[3], int k0, int n) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            cell[i] = (Index)floorf(__builtin_fmaf((float)k0, dq[i], q0[i]));
            sgn[i] = sgn_u[i];
            inv[i] = inv_u[i];
            off[i] = -q0[i] * inv[i];
            kb[i] = sgn[i] != 0 ? __builtin_fmaf((float)(cell[i] + (Index)(sgn[i] > 0 ? 1 : 0)), inv[i], off[i])
                                : INFINITY;
        }
        count = n;
        fcount = (float)n;
        k = (Index)k0;
        guard = 0;
    }
    __device__ __forceinline__ bool active() const {
        return (kFlat ? (float)k < fcount : (int)k < count) && guard < 4 * count + 64;
    }
    // One step; moved(i) is called for every axis i the walk advances along.  The next crossing kn
    // is the least of the axes' crossing indices (never NaN: finite, or INFINITY on an axis without
    // crossings), found with compares, and every axis at kn advances.  The samples in the cell run
    // up to kend = clamp(ceil(kn), 0, count).
    template <class F, class M>
    __device__ __forceinline__ void step(F&& f, M&& moved) {
        step_impl<false>(f, moved);
    }
    // The same step with f called for every cell, also one holding no samples (m = 0): for callers
    // whose callbacks add terms that are +-0 when m = 0 (cell sums and gradient sums: m = 0 and
    // S2c = -0), which then run without a branch per step.
    template <class F, class M>
    __device__ __forceinline__ void step_all(F&& f, M&& moved) {
        step_impl<true>(f, moved);
    }
    template <bool kAll, class F, class M>
    __device__ __forceinline__ void step_impl(F&& f, M&& moved) {
        bool adv[3];
        float kn;
        if (kFlat) {
            adv[0] = kb[0] <= kb[2];
            adv[1] = false;
            adv[2] = kb[2] <= kb[0];
            kn = adv[0] ? kb[0] : kb[2];
        } else {
            const float kxy = kb[0] <= kb[1] ? kb[0] : kb[1];
            kn = kxy <= kb[2] ? kxy : kb[2];
#pragma unroll
            for (int i = 0; i < 3; i++) adv[i] = kb[i] == kn;
        }
        const Index kend = (Index)__builtin_amdgcn_fmed3f(ceilf(kn), 0.0f, fcount);
        if (kAll) {
            const Index mi = kend - k > (Index)0 ? kend - k : (Index)0;
            f(cell, k, mi);
            k += mi;
        } else if (kend > k) {
            f(cell, k, kend - k);
            k = kend;
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (kFlat && i == 1) continue;
            if (adv[i]) {
                cell[i] += (Index)sgn[i];
                kb[i] = __builtin_fmaf((float)(cell[i] + (Index)(sgn[i] > 0 ? 1 : 0)), inv[i], off[i]);
                moved(i);
            }
        }
        guard++;
    }
    template <class F>
    __device__ __forceinline__ void step(F&& f) {
        step(f, [](int) {});
    }
};

using CellWalker = CellWalkerT<false>;

// The whole walk: f(cell, k, m) for every cell holding samples k..k+m-1, in order.
template <bool kFlat = false, class F>
__device__ __forceinline__ void for_each_cell(const float q0[3], const float dq[3], int count, F&& f) {
    CellWalkerT<kFlat> w;
    w.init(q0, dq, count);
    while (w.active()) w.step(f);
}

// The m samples of one cell: centre coordinates relative to the cell and the second moment of the
// sample offsets about the centre, S2c = sum_j (j - (m-1)/2)^2 = m (m^2 - 1) / 12 (exact).  Centred
// on the middle sample the odd moments vanish, so with the cell interpolant
//   f = v000 + A a + B b + C c + AB ab + AC ac + BC bc + ABC abc
// the samples sum to  m f(centre) + S2c (AB da db + AC da dc + BC db dc + ABC (a db dc + b da dc + c da db)).
struct CellMoments {
    float a, b, c, S2c, m;
    float pab, pac, pbc, q;   // da db, da dc, db dc, a db dc + b da dc + c da db
    bool flat;                // db == 0: b is the same for every sample
};

// kFlat (dq[1] == 0): b is the same for every sample, da db = db dc = 0 and q = b da dc; the
// values equal the general form's (fma(kc, 0, q0) = q0, fma(x, 0, y) = y for finite operands).
// cell, k and mi are ints, or (the walk without the y axis) floats holding the same integers.
template <bool kFlat = false, typename C, typename K>
__device__ __forceinline__ CellMoments cell_moments(const float q0[3], const float dq[3], const C* cell, K k, K mi) {
    CellMoments cm;
    // k + (m - 1) / 2, exact: integer sums below 2^24 halved; in float as m / 2 + (k - 1/2), whose
    // half-integer terms and sum are exact too
    const float kc = std::is_same<K, float>::value ? __builtin_fmaf((float)mi, 0.5f, (float)k - 0.5f)
                                                   : (float)(k + (k + mi) - 1) * 0.5f;
    cm.a = __builtin_fmaf(kc, dq[0], q0[0]) - (float)cell[0];
    cm.b = (kFlat ? q0[1] : __builtin_fmaf(kc, dq[1], q0[1])) - (float)cell[1];
    cm.c = __builtin_fmaf(kc, dq[2], q0[2]) - (float)cell[2];
    cm.m = (float)mi;
    cm.S2c = cm.m * __builtin_fmaf(cm.m, cm.m, -1.0f) * (1.0f / 12.0f);   // m, m^2 - 1 and their product exact
    cm.pac = dq[0] * dq[2];
    cm.flat = kFlat || dq[1] == 0.0f;
    if (kFlat) {
        cm.pab = cm.pbc = 0.0f;
        cm.q = cm.b * cm.pac;
    } else {
        cm.pab = dq[0] * dq[1];
        cm.pbc = dq[1] * dq[2];
        cm.q = __builtin_fmaf(cm.a, cm.pbc, __builtin_fmaf(cm.b, cm.pac, cm.c * cm.pab));
    }
    return cm;
}

// Corner offsets of a cell (clamped indices: clamp addressing), order 000 100 010 110 001 101 011 111:
// the 000 corner plus per-axis steps, each 0 where the index is clamped.
template <bool kBig>
__device__ __forceinline__ void cell_corners(const VolG& v, const int cell[3], size_t off[8]) {
    const int x0 = clampi(cell[0], 0, v.nx - 1), y0 = clampi(cell[1], 0, v.ny - 1), z0 = clampi(cell[2], 0, v.nz - 1);
    const size_t sy = (size_t)v.nx, sz = (size_t)v.nx * (size_t)v.ny;
    const size_t dx = (cell[0] + 1 > x0 && x0 + 1 < v.nx) ? 1 : 0;
    const size_t dy = (cell[1] + 1 > y0 && y0 + 1 < v.ny) ? sy : 0;
    const size_t dz = (cell[2] + 1 > z0 && z0 + 1 < v.nz) ? sz : 0;
    off[0] = (size_t)z0 * sz + (size_t)y0 * sy + (size_t)x0;
    off[1] = off[0] + dx;
    off[2] = off[0] + dy;
    off[3] = off[2] + dx;
    off[4] = off[0] + dz;
    off[5] = off[4] + dx;
    off[6] = off[4] + dy;
    off[7] = off[6] + dx;
}

// The same offsets in 32 bits, in bytes, for volumes below 2^30 voxels with ny * nz < 2^24 (buffer
// loads; the row index z * ny + y then fits the 24-bit multiplier).
template <uint32_t kElem = 4>
__device__ __forceinline__ void cell_corners_bytes(const VolG& v, const int cell[3], uint32_t off[8]) {
    const int x0 = clampi(cell[0], 0, v.nx - 1), y0 = clampi(cell[1], 0, v.ny - 1), z0 = clampi(cell[2], 0, v.nz - 1);
    const uint32_t sy = (uint32_t)v.nx * kElem, sz = (uint32_t)v.nx * (uint32_t)v.ny * kElem;
    const uint32_t dx = (cell[0] + 1 > x0 && x0 + 1 < v.nx) ? kElem : 0u;
    const uint32_t dy = (cell[1] + 1 > y0 && y0 + 1 < v.ny) ? sy : 0u;
    const uint32_t dz = (cell[2] + 1 > z0 && z0 + 1 < v.nz) ? sz : 0u;
    const uint32_t row = __umul24((uint32_t)z0, (uint32_t)v.ny) + (uint32_t)y0;
    off[0] = (__umul24(row, (uint32_t)v.nx) + (uint32_t)x0) * kElem;
    off[1] = off[0] + dx;
    off[2] = off[0] + dy;
    off[3] = off[2] + dx;
    off[4] = off[0] + dz;
    off[5] = off[4] + dx;
    off[6] = off[4] + dy;
    off[7] = off[6] + dx;
}

// Buffer resource of a volume for 32-bit byte offsets (raw buffer, gfx9 descriptor word 3); loads
// beyond the volume's bytes return 0.
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t volume_rsrc(const T* vol, const VolG& v) {
    const uint32_t bytes = (uint32_t)v.nx * (uint32_t)v.ny * (uint32_t)v.nz * (uint32_t)sizeof(T);
    return __builtin_amdgcn_make_buffer_rsrc((void*)vol, (short)0, (int)bytes, 0x00020000);
}

// One voxel through a buffer resource at a byte offset: the b32 load returns the word's bits, the
// b16 load a half's bits.
__device__ __forceinline__ float buffer_vox(const __amdgpu_buffer_rsrc_t rs, uint32_t off, const float*) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0));
}
__device__ __forceinline__ float buffer_vox(const __amdgpu_buffer_rsrc_t rs, uint32_t off, const vhalf*) {
    return (float)__builtin_bit_cast(vhalf, __builtin_amdgcn_raw_buffer_load_b16(rs, (int)off, 0, 0));
}

// Sum of one cell's m samples, m f(centre) + S2c corr, rounded once to float (relative error
// 6e-8 per cell); the ray adds the cells in double, so per-cell roundings do not accumulate into a
// float running sum.
// db = 0 (cm.flat; always under kFlat): b is the same for every sample and the cell is a
// bilinear interpolant in (a, c) of its four y-blended edges w_xz = lerp(v_x0z, v_x1z, b): the
// samples sum to m g(centre) + S2c (w11 - w01 - w10 + w00) da dc.  Every kernel takes this form
// for such rays, so their sums agree bit for bit whatever walk they run.
template <bool kFlat = false>
__device__ __forceinline__ double cell_sum_values(const float val[8], const CellMoments& cm, int mi) {
    const float v000 = val[0], v100 = val[1], v010 = val[2], v110 = val[3];
    const float v001 = val[4], v101 = val[5], v011 = val[6], v111 = val[7];
    (void)mi;
    if (kFlat || cm.flat) {
        const float w00 = lerpf(v000, v010, cm.b), w10 = lerpf(v100, v110, cm.b);
        const float w01 = lerpf(v001, v011, cm.b), w11 = lerpf(v101, v111, cm.b);
        const float d0 = w10 - w00, d1 = w11 - w01;
        const float g = lerpf(__builtin_fmaf(cm.a, d0, w00), __builtin_fmaf(cm.a, d1, w01), cm.c);
        // S2c da dc = m (m^2 - 1) (da dc / 12): the integer product exact, the ray's da dc / 12 hoisted
        const float spac = (cm.m * __builtin_fmaf(cm.m, cm.m, -1.0f)) * (cm.pac * (1.0f / 12.0f));
        return (double)__builtin_fmaf(cm.m, g, (d1 - d0) * spac);
    }
    // x differences, shared by the x lerps and the mixed differences
    const float x00 = v100 - v000, x10 = v110 - v010, x01 = v101 - v001, x11 = v111 - v011;
    const float c00 = __builtin_fmaf(cm.a, x00, v000), c10 = __builtin_fmaf(cm.a, x10, v010);
    const float c01 = __builtin_fmaf(cm.a, x01, v001), c11 = __builtin_fmaf(cm.a, x11, v011);
    const float f = lerpf(lerpf(c00, c10, cm.b), lerpf(c01, c11, cm.b), cm.c);
    const float AB = x10 - x00, AC = x01 - x00, ABC = (x11 - x01) - AB;
    const float BC = (v011 - v001) - (v010 - v000);
    const float corr = __builtin_fmaf(ABC, cm.q, __builtin_fmaf(BC, cm.pbc, __builtin_fmaf(AC, cm.pac, AB * cm.pab)));
    return (double)__builtin_fmaf(cm.m, f, cm.S2c * corr);
}

// The same sum as weights on the eight corners (cell_sum = sum_c v_c W_c): the transpose of
// cell_sum, W_c = m w_c(centre) + S2c kappa_c with w_c the trilinear weights and kappa_c the
// corner coefficients of the correction.
__device__ __forceinline__ void cell_weights(const CellMoments& cm, double W[8]) {
    const double a = cm.a, b = cm.b, c = cm.c, m = cm.m, s = cm.S2c;
    const double a1 = 1.0 - a, b1 = 1.0 - b, c1 = 1.0 - c;
    const double pab = cm.pab, pac = cm.pac, pbc = cm.pbc, q = cm.q;
    W[0] = m * (a1 * b1 * c1) + s * (pab + pac + pbc - q);
    W[1] = m * (a * b1 * c1) + s * (-pab - pac + q);
    W[2] = m * (a1 * b * c1) + s * (-pab - pbc + q);
    W[3] = m * (a * b * c1) + s * (pab - q);
    W[4] = m * (a1 * b1 * c) + s * (-pac - pbc + q);
    W[5] = m * (a * b1 * c) + s * (pac - q);
    W[6] = m * (a1 * b * c) + s * (pbc - q);
    W[7] = m * (a * b * c) + s * q;
}

// Sum over the cell's m samples of the interpolant's gradient (d/da, d/db, d/dc in texel units).
// With samples a_j = a + (j - (m-1)/2) da about the centre (likewise b, c), d f/da = A + AB b + AC c
// + ABC b c sums to m (d f/da)(centre) + S2c ABC db dc, and likewise for b and c.  Where a corner
// index is clamped its step vanishes, and with it the derivative along that axis, as for the
// clamp-addressed read.  Accumulated into G.
// db = 0 (cm.flat): the x and z derivatives are those of the bilinear interpolant in (a, c) of the
// y-blended edges w_xz that the cell sum forms (shared with it), the y derivative interpolates the
// y differences in (a, c), and with da db = db dc = 0 only the y derivative keeps a correction.
__device__ __forceinline__ void cell_grad_sum(const float val[8], const CellMoments& cm, int mi, double G[3]) {
    const float v000 = val[0], v100 = val[1], v010 = val[2], v110 = val[3];
    const float v001 = val[4], v101 = val[5], v011 = val[6], v111 = val[7];
    if (cm.flat) {
        const float w00 = lerpf(v000, v010, cm.b), w10 = lerpf(v100, v110, cm.b);
        const float w01 = lerpf(v001, v011, cm.b), w11 = lerpf(v101, v111, cm.b);
        const float ga = lerpf(w10 - w00, w11 - w01, cm.c);
        const float gc = lerpf(w01 - w00, w11 - w10, cm.a);
        const float y00 = v010 - v000, y10 = v110 - v100, y01 = v011 - v001, y11 = v111 - v101;
        const float gb = lerpf(lerpf(y00, y10, cm.a), lerpf(y01, y11, cm.a), cm.c);
        const float ABCf = (y11 - y01) - (y10 - y00);
        G[0] += (double)(cm.m * ga);
        G[1] += (double)__builtin_fmaf(cm.m, gb, (cm.S2c * ABCf) * cm.pac);
        G[2] += (double)(cm.m * gc);
        return;
    }
    const float ABC = ((v111 - v110) - (v101 - v100)) - ((v011 - v010) - (v001 - v000));
    // d f / d a at the centre: the x-differences interpolated in b, then c (likewise for b and c)
    const float ga = lerpf(lerpf(v100 - v000, v110 - v010, cm.b), lerpf(v101 - v001, v111 - v011, cm.b), cm.c);
    const float gb = lerpf(lerpf(v010 - v000, v110 - v100, cm.a), lerpf(v011 - v001, v111 - v101, cm.a), cm.c);
    const float gc = lerpf(lerpf(v001 - v000, v101 - v100, cm.a), lerpf(v011 - v010, v111 - v110, cm.a), cm.b);
    const float s = cm.S2c * ABC;
    G[0] += (double)__builtin_fmaf(cm.m, ga, s * cm.pbc);
    G[1] += (double)__builtin_fmaf(cm.m, gb, s * cm.pac);
    G[2] += (double)__builtin_fmaf(cm.m, gc, s * cm.pab);
}

}  // namespace artia
