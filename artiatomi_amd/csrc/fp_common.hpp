// Exact-numerics ray sampling shared by the forward projector (fp.hip) and its adjoint and
// shift-gradient kernels (adjoint.hip): double-precision ray setup and slab clip, the affine sample
// coordinates of a sub-ray, and the clamp-addressed trilinear read
// (ForwardProjectionRayMarcher_TL.cu:57-160 evaluated stably).
#pragma once

#include <cstdint>

#include "device_common.hpp"

namespace artia {

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
// q_k = q0 + k*dq.  Returns false when the sub-ray misses the volume.
struct SubraySamples {
    float q0[3], dq[3];
    int count;
    int entry_axis;   // axis whose slab sets t_in (the derivative of t_in follows it)
};

__device__ __forceinline__ bool subray_samples(const VolG& v, const TiltG& g, double px, double py, double step,
                                               SubraySamples& s) {
    const double xa = ((double)g.A[0] * px + (double)g.A[1] * py) + (double)g.A[2];
    const double ya = ((double)g.A[3] * px + (double)g.A[4] * py) + (double)g.A[5];
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

}  // namespace artia
