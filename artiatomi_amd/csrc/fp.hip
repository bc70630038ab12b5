// Forward projection on gfx950: ray marcher (march), slab-limited marcher (slicer) and the
// centre-ray volume traversal length (volTraversalLength).
//
// Thread = detector pixel, 64x4 blocks (one wave per 64-pixel row segment).  Every pixel
// shoots four sub-rays at offsets {0.25, 0.75}^2 (ForwardProjectionRayMarcher_TL.cu:80-82)
// and samples the volume every 0.15 voxels with trilinear interpolation.
//
// Two numerics:
//   reference  the kernel bodies in the reference's float operation order, including the
//              float accumulation t += 0.15*s at |t| ~ 1e4 that biases the step
//              (ForwardProjectionRayMarcher_TL.cu:130-151);
//   exact      the ray/box intersection in double, samples at t_in + k*0.15*s as float
//              coordinates affine in k relative to the entry point, and their trilinear values
//              summed cell by cell in closed form (march_sum); this tracks the double-precision
//              evaluation of the same algorithm.
#include <algorithm>
#include <climits>
#include <cmath>

#include "device_common.hpp"
#include "fp_common.hpp"
#include "image_ops.hpp"
#include "tube_walk.hpp"

namespace artia {
namespace {

struct Roi {
    int x0, y0, x1, y1;
};

__device__ __forceinline__ void sub_ray_origin(const TiltG& g, float px, float py, float c[3]) {
    float xa, ya;
    mag_apply(g.A, px, py, xa, ya);
    for (int i = 0; i < 3; i++) c[i] = (g.d[i] + xa * g.u[i]) + ya * g.v[i];
}

__device__ __forceinline__ void slab_test(const VolG& v, const TiltG& g, const float c[3], float& t_in,
                                          float& t_out) {
    float lo[3], hi[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float tE = (v.bbMin[i] - c[i]) / g.n[i];
        const float tX = (v.bbMax[i] - c[i]) / g.n[i];
        lo[i] = fminf(tE, tX);
        hi[i] = fmaxf(tE, tX);
    }
    t_in = fmaxf(fmaxf(lo[0], lo[1]), lo[2]);
    t_out = fminf(fminf(hi[0], hi[1]), hi[2]);
}

// ---------------------------------------------------------------- reference numerics

template <bool kSlicer, typename T>
__global__ void __launch_bounds__(256) march_ref_kernel(VolG v, TiltG g, const T* __restrict__ vol, Roi roi,
                                                        size_t pitch, float* __restrict__ proj,
                                                        float* __restrict__ dist, float tminDef, float tmaxDef) {
    const int x = roi.x0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = roi.y0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= roi.x1 || y >= roi.y1) return;
    const float step = v.vs[0] * 0.15f;
    float temp = 0.0f, hits = 0.0f, len = 0.0f;
    for (float osx = 0.25f; osx < 0.8f; osx += 0.5f) {
        for (float osy = 0.25f; osy < 0.8f; osy += 0.5f) {
            float c[3];
            sub_ray_origin(g, (float)x + osx, (float)y + osy, c);
            float t_in, t_out;
            slab_test(v, g, c, t_in, t_out);
            if ((t_out - t_in) > 0.0f) {
                float f[3];
                if (kSlicer) {
                    t_in = fmaxf(t_in, tminDef);
                    t_out = fminf(t_out, tmaxDef);
                    // the slicer's first sample sits at the detector point itself
                    f[0] = c[0]; f[1] = c[1]; f[2] = c[2];
                } else {
                    hits += 1.0f;
                    len += (t_out - t_in);
                    for (int i = 0; i < 3; i++) f[i] = c[i] + t_in * g.n[i];
                }
                int guard = 0;
                while (t_in <= t_out && guard++ < kMaxSamplesPerRay) {
                    const float qx = ((f[0] - v.bbMin[0]) * v.rcp[0]) * v.dim[0];
                    const float qy = ((f[1] - v.bbMin[1]) * v.rcp[1]) * v.dim[1];
                    const float qz = ((f[2] - v.bbMin[2]) * v.rcp[2]) * v.dim[2] - v.zshift;
                    const float val = tex3d_linear(vol, v.nx, v.ny, v.nz, qx, qy, qz);
                    temp += (val * v.vs[0]) * 0.15f;
                    t_in += step;
                    for (int i = 0; i < 3; i++) f[i] = c[i] + t_in * g.n[i];
                }
            }
        }
    }
    const size_t i = (size_t)y * pitch + (size_t)x;
    proj[i] += temp * 0.25f;
    if (!kSlicer) dist[i] = fmaxf(0.0f, len / hits);
}

__global__ void __launch_bounds__(256) vol_traversal_ref_kernel(VolG v, TiltG g, Roi roi, size_t pitch,
                                                                float* __restrict__ dist) {
    const int x = roi.x0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = roi.y0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= roi.x1 || y >= roi.y1) return;
    float c[3];
    sub_ray_origin(g, (float)x + 0.5f, (float)y + 0.5f, c);
    float t_in, t_out;
    slab_test(v, g, c, t_in, t_out);
    float val = 0.0f;
    if (t_out > t_in) val = t_out - t_in;
    dist[(size_t)y * pitch + (size_t)x] += val;
}

// march_ref_kernel's forward projection (not the slicer) with the volume read from LDS-staged slab
// boxes: the same samples in the same order, the same float arithmetic per sample (tex3d_linear's
// clamps, weights and lerps) and the same voxel values, so the same bits.  A block is a 16x16-pixel
// tile, a thread one pixel; its four sub-rays run one after another as in the reference (the float
// running sum chains them), so the block walks the slabs of cell layers once per sub-ray offset,
// staging for each slab the box the tile's sub-rays of that offset cross.  A sample whose corners
// fall outside the box reads global memory.  Five waves per SIMD (96 registers, a few spilled)
// measured 6 % faster than the unbounded four.
constexpr int kRefTile = 16;   // pixels per side of a block's tile
constexpr int kRefSlab = 4;    // cell layers per slab

// tex3d_linear with the eight voxels taken from the staged box [X0, X0 + CX) x [Y0, ..) x [Z0, ..)
// (padded indices, clamp-addressed voxels) when every clamped corner lies in it, else from vol.
template <typename T>
__device__ __forceinline__ float tex3d_linear_box(const T* __restrict__ vol, int nx, int ny, int nz,
                                                  const float* __restrict__ box, bool staged, int X0, int Y0,
                                                  int Z0, int CX, int CY, int CZ, float x, float y, float z) {
    x = fminf(fmaxf(x, -1.0f), (float)(nx + 1));
    y = fminf(fmaxf(y, -1.0f), (float)(ny + 1));
    z = fminf(fmaxf(z, -1.0f), (float)(nz + 1));
    const float xb = x - 0.5f, yb = y - 0.5f, zb = z - 0.5f;
    const float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    const float a = xb - fx, b = yb - fy, c = zb - fz;
    int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
    const int i1 = clampi(i0 + 1, 0, nx - 1), j1 = clampi(j0 + 1, 0, ny - 1), k1 = clampi(k0 + 1, 0, nz - 1);
    i0 = clampi(i0, 0, nx - 1);
    j0 = clampi(j0, 0, ny - 1);
    k0 = clampi(k0, 0, nz - 1);
    float v000, v100, v010, v110, v001, v101, v011, v111;
    if (staged && i0 >= X0 && i1 < X0 + CX && j0 >= Y0 && j1 < Y0 + CY && k0 >= Z0 && k1 < Z0 + CZ) {
        const int r00 = ((k0 - Z0) * CY + (j0 - Y0)) * CX - X0, r10 = ((k0 - Z0) * CY + (j1 - Y0)) * CX - X0;
        const int r01 = ((k1 - Z0) * CY + (j0 - Y0)) * CX - X0, r11 = ((k1 - Z0) * CY + (j1 - Y0)) * CX - X0;
        v000 = box[r00 + i0];
        v100 = box[r00 + i1];
        v010 = box[r10 + i0];
        v110 = box[r10 + i1];
        v001 = box[r01 + i0];
        v101 = box[r01 + i1];
        v011 = box[r11 + i0];
        v111 = box[r11 + i1];
    } else {
        const size_t sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
        const T* p00 = vol + (size_t)k0 * sz + (size_t)j0 * sy;
        const T* p10 = vol + (size_t)k0 * sz + (size_t)j1 * sy;
        const T* p01 = vol + (size_t)k1 * sz + (size_t)j0 * sy;
        const T* p11 = vol + (size_t)k1 * sz + (size_t)j1 * sy;
        v000 = vox(p00, i0);
        v100 = vox(p00, i1);
        v010 = vox(p10, i0);
        v110 = vox(p10, i1);
        v001 = vox(p01, i0);
        v101 = vox(p01, i1);
        v011 = vox(p11, i0);
        v111 = vox(p11, i1);
    }
    const float c00 = lerpf(v000, v100, a);
    const float c10 = lerpf(v010, v110, a);
    const float c01 = lerpf(v001, v101, a);
    const float c11 = lerpf(v011, v111, a);
    return lerpf(lerpf(c00, c10, b), lerpf(c01, c11, b), c);
}

// The sample's z cell as tex3d_linear finds it (floor of the clamped coordinate minus 1/2).
__device__ __forceinline__ int ref_cell_z(const VolG& v, float qz) {
    return (int)floorf(fminf(fmaxf(qz, -1.0f), (float)(v.nz + 1)) - 0.5f);
}

template <typename T>
__global__ void __launch_bounds__(256, 5) march_ref_lds_kernel(VolG v, TiltG g, const T* __restrict__ vol, Roi roi,
                                                            size_t pitch, float* __restrict__ proj,
                                                            float* __restrict__ dist, TubePlan tp) {
    __shared__ float chunk[kChunkFloats];
    __shared__ int zr[2];
    const int tx0 = roi.x0 + (int)blockIdx.x * kRefTile, ty0 = roi.y0 + (int)blockIdx.y * kRefTile;
    const int x = tx0 + (int)(threadIdx.x & (kRefTile - 1)), y = ty0 + (int)(threadIdx.x / kRefTile);
    const bool active = x < roi.x1 && y < roi.y1;
    const float step = v.vs[0] * 0.15f;
    float temp = 0.0f, hits = 0.0f, len = 0.0f;
    // the sample's texel coordinates, as march_ref_kernel computes them
    auto texel = [&](const float f[3], float q[3]) {
        q[0] = ((f[0] - v.bbMin[0]) * v.rcp[0]) * v.dim[0];
        q[1] = ((f[1] - v.bbMin[1]) * v.rcp[1]) * v.dim[1];
        q[2] = ((f[2] - v.bbMin[2]) * v.rcp[2]) * v.dim[2] - v.zshift;
    };
    for (int s = 0; s < 4; s++) {
        // the reference's sub-ray order: osx outer, osy inner, each 0.25 then 0.75
        const float osx = (s & 2) ? 0.75f : 0.25f, osy = (s & 1) ? 0.75f : 0.25f;
        float c[3] = {0.0f, 0.0f, 0.0f}, f[3] = {0.0f, 0.0f, 0.0f};
        float t_in = 1.0f, t_out = 0.0f;
        bool live = false;
        if (active) {
            sub_ray_origin(g, (float)x + osx, (float)y + osy, c);
            slab_test(v, g, c, t_in, t_out);
            live = (t_out - t_in) > 0.0f;
            if (live) {
                hits += 1.0f;
                len += (t_out - t_in);
                for (int i = 0; i < 3; i++) f[i] = c[i] + t_in * g.n[i];
            }
        }
        int guard = 0;
        // cell layers of the block's samples of this offset (first and last sample, one of margin)
        __syncthreads();   // the previous offset's readers are done with zr and chunk
        if (threadIdx.x == 0) {
            zr[0] = INT_MAX;
            zr[1] = INT_MIN;
        }
        __syncthreads();
        if (live) {
            float q[3], e[3], qe[3];
            texel(f, q);
            for (int i = 0; i < 3; i++) e[i] = c[i] + t_out * g.n[i];
            texel(e, qe);
            const int za = ref_cell_z(v, q[2]), zb = ref_cell_z(v, qe[2]);
            atomicMin(&zr[0], min(za, zb) - 1);
            atomicMax(&zr[1], max(za, zb) + 1);
        }
        __syncthreads();
        const int zlo = max(zr[0], -1), zhi = min(zr[1], v.nz - 1);
        const int nslab = zlo <= zhi ? (zhi - zlo) / kRefSlab + 1 : 0;
        const TubeBox tb = tube_box(v, g, tp, (double)tx0 + osx, (double)tx0 + (kRefTile - 1) + osx,
                                    (double)ty0 + osy, (double)ty0 + (kRefTile - 1) + osy, 0.0, 0.0);
        for (int j = 0; j < nslab; j++) {
            const int zs = tp.zdir > 0 ? zlo + j * kRefSlab : zhi - (j + 1) * kRefSlab + 1;
            const int ze = zs + kRefSlab - 1;
            const float z0 = (float)zs, z1 = (float)(ze + 1);
            const int X0 = wave_uniform(max((int)floorf(tb.axl + fminf(tp.bx * z0, tp.bx * z1)) - 2, -1));
            const int X1 = min((int)floorf(tb.axh + fmaxf(tp.bx * z0, tp.bx * z1)) + 3, v.nx);
            const int Y0 = wave_uniform(max((int)floorf(tb.ayl + fminf(tp.by * z0, tp.by * z1)) - 2, -1));
            const int Y1 = min((int)floorf(tb.ayh + fmaxf(tp.by * z0, tp.by * z1)) + 3, v.ny);
            const int Z0 = max(zs, -1), Z1 = min(ze + 1, v.nz);
            const int CX = wave_uniform(X1 - X0 + 1), CY = wave_uniform(Y1 - Y0 + 1), CZ = Z1 - Z0 + 1;
            const bool staged = CX > 1 && CY > 1 && CZ > 1 && CX * CY * CZ <= kChunkFloats;
            __syncthreads();   // the previous slab's readers are done
            if (staged) stage_box_rows(v, vol, chunk, X0, Y0, Z0, CX, CY, CX * CY * CZ);
            __syncthreads();
            while (live && t_in <= t_out && guard < kMaxSamplesPerRay) {
                float q[3];
                texel(f, q);
                const int kz = ref_cell_z(v, q[2]);
                if (tp.zdir > 0 ? kz > ze : kz < zs) break;   // a later slab's sample
                const float val = tex3d_linear_box(vol, v.nx, v.ny, v.nz, chunk, staged, X0, Y0, Z0, CX, CY, CZ,
                                                   q[0], q[1], q[2]);
                temp += (val * v.vs[0]) * 0.15f;
                guard++;
                t_in += step;
                for (int i = 0; i < 3; i++) f[i] = c[i] + t_in * g.n[i];
            }
        }
        // samples beyond the slabs (not expected): global memory
        while (live && t_in <= t_out && guard < kMaxSamplesPerRay) {
            float q[3];
            texel(f, q);
            temp += (tex3d_linear(vol, v.nx, v.ny, v.nz, q[0], q[1], q[2]) * v.vs[0]) * 0.15f;
            guard++;
            t_in += step;
            for (int i = 0; i < 3; i++) f[i] = c[i] + t_in * g.n[i];
        }
    }
    if (!active) return;
    const size_t i = (size_t)y * pitch + (size_t)x;
    proj[i] += temp * 0.25f;
    dist[i] = fmaxf(0.0f, len / hits);
}

// ---------------------------------------------------------------- exact numerics

// Sum of the trilinear samples q0 + k*dq, k = 0..count-1 (coordinates shifted by -0.5, texel j
// centred at j), cell by cell in closed form (cell_sum, fp_common.hpp).
// kFlat: dq[1] == 0 (the walk, moments and cell sums without the y terms; the same values).
template <bool kBig, bool kFlat, typename T>
__device__ __forceinline__ double march_sum(const VolG& v, const T* __restrict__ vol, const float q0[3],
                                            const float dq[3], int count) {
    double total = 0.0;
    if (kBig) {
        for_each_cell<kFlat>(q0, dq, count, [&](const auto* cell, auto k, auto mi) {
            const int ic[3] = {(int)cell[0], (int)cell[1], (int)cell[2]};
            size_t o[8];
            cell_corners<true>(v, ic, o);
            float val[8];
#pragma unroll
            for (int i = 0; i < 8; i++) val[i] = vox(vol, o[i]);
            total += cell_sum_values<kFlat>(val, cell_moments<kFlat>(q0, dq, cell, k, mi), mi);
        });
    } else {
        const __amdgpu_buffer_rsrc_t rs = volume_rsrc(vol, v);
        for_each_cell<kFlat>(q0, dq, count, [&](const auto* cell, auto k, auto mi) {
            const int ic[3] = {(int)cell[0], (int)cell[1], (int)cell[2]};
            uint32_t o[8];
            cell_corners_bytes<(uint32_t)sizeof(T)>(v, ic, o);
            float val[8];
#pragma unroll
            for (int i = 0; i < 8; i++) val[i] = buffer_vox(rs, o[i], vol);
            total += cell_sum_values<kFlat>(val, cell_moments<kFlat>(q0, dq, cell, k, mi), mi);
        });
    }
    return total;
}

// One sub-ray of the exact marcher: the slicer's first sample at the detector point (acc), the
// chord length and hit count for dist, and the sample grid q0 + k*dq, k < count (count 0: nothing
// to walk).
struct MarchLane {
    float q0[3], dq[3];
    int count;
    double acc, len, hits;
};

// The slicer's first sample, at the detector point c of the ray (ForwardProjectionSlicer.cu:129-133)
template <typename T>
__device__ __forceinline__ float detector_sample(const VolG& v, const T* __restrict__ vol, const RayD& r) {
    return tex3d_linear(vol, v.nx, v.ny, v.nz,
                        (float)((r.c[0] - (double)v.bbMin[0]) * ((double)v.rcp[0] * (double)v.dim[0])),
                        (float)((r.c[1] - (double)v.bbMin[1]) * ((double)v.rcp[1] * (double)v.dim[1])),
                        (float)((r.c[2] - (double)v.bbMin[2]) * ((double)v.rcp[2] * (double)v.dim[2]) -
                                (double)v.zshift));
}

// Number of samples t_first, t_first + step, .. up to t_last (t_first <= t_last)
__device__ __forceinline__ int sample_count(double t_first, double t_last, double step) {
    return (int)fmin(floor((t_last - t_first) / step) + 1.0, (double)kMaxSamplesPerRay);
}

// Texel coordinates of the sample at t_first: (p - bbMin) * rcp * dim (the float-rounded rcp the
// reference uploads), shifted by -0.5 so that texel j is centred at integer j
__device__ __forceinline__ void sample_start(const VolG& v, const TiltG& g, const RayD& r, double t_first,
                                             float q0[3]) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double scale = (double)v.rcp[i] * (double)v.dim[i];
        const double p = (r.c[i] + t_first * (double)g.n[i]) - (double)v.bbMin[i];
        q0[i] = (float)(p * scale - 0.5 - (i == 2 ? (double)v.zshift : 0.0));
    }
}

template <bool kSlicer, typename T>
__device__ __forceinline__ MarchLane march_lane(const VolG& v, const TiltG& g, const T* __restrict__ vol,
                                                bool active, int x, int y, int sub, double step, float tminDef,
                                                float tmaxDef) {
    MarchLane L;
    L.count = 0;
    L.acc = L.len = L.hits = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) L.q0[i] = L.dq[i] = 0.0f;
    if (!active) return L;
    const double px = (double)x + ((sub & 2) ? 0.75 : 0.25);
    const double py = (double)y + ((sub & 1) ? 0.75 : 0.25);
    const RayD r = ray_setup_d(v, g, px, py);
    if (!((r.t_out - r.t_in) > 0.0)) return L;
    double t_first = r.t_in, t_last = r.t_out;
    bool run = true;
    if (kSlicer) {
        const double t_in = fmax(r.t_in, (double)tminDef);
        t_last = fmin(r.t_out, (double)tmaxDef);
        run = t_in <= t_last;
        if (run) {
            L.acc = (double)detector_sample(v, vol, r);
            t_first = t_in + step;
        }
    } else {
        L.hits = 1.0;
        L.len = r.t_out - r.t_in;
    }
    if (run && t_first <= t_last) {
        L.count = sample_count(t_first, t_last, step);
        sample_start(v, g, r, t_first, L.q0);
#pragma unroll
        for (int i = 0; i < 3; i++) L.dq[i] = (float)(step * (double)g.n[i] * ((double)v.rcp[i] * (double)v.dim[i]));
    }
    return L;
}

// The pixel's mean chord over its four sub-rays (lanes 4p .. 4p+3), as dist holds it.
__device__ __forceinline__ float pixel_dist(double len, double hits) {
    len += __shfl_xor(len, 1);
    len += __shfl_xor(len, 2);
    hits += __shfl_xor(hits, 1);
    hits += __shfl_xor(hits, 2);
    return hits > 0.0 ? (float)(len / hits) : 0.0f;
}

__device__ __forceinline__ void march_store_dist(double len, double hits, bool active, int sub, int x, int y,
                                                 size_t pitch, float* __restrict__ dist) {
    const float d = pixel_dist(len, hits);
    if (active && sub == 0) dist[(size_t)y * pitch + (size_t)x] = d;
}

// Sum of a pixel's four sub-rays (lanes 4p .. 4p+3) and the pixel's outputs.
template <bool kSlicer>
__device__ __forceinline__ void march_store(double acc, double len, double hits, bool active, int sub, int x, int y,
                                            size_t pitch, double step, float* __restrict__ proj,
                                            float* __restrict__ dist) {
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    const float d = kSlicer ? 0.0f : pixel_dist(len, hits);
    if (active && sub == 0) {
        const size_t i = (size_t)y * pitch + (size_t)x;
        proj[i] += (float)(acc * step * 0.25);
        if (!kSlicer) dist[i] = d;
    }
}

// The residual epilogue: the pixel's projection (from zero) and mean chord go straight into
// compare and cropBorder (residual_value); out receives the residual instead of the projection.
__device__ __forceinline__ void march_store_resid(double acc, float dist, bool active, int sub, int x, int y,
                                                  size_t pitch, double step, const ResidEpi& e, int W, int H,
                                                  float* __restrict__ out) {
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    if (active && sub == 0)
        out[(size_t)y * pitch + (size_t)x] = residual_value(e, x, y, W, H, pitch, (float)(acc * step * 0.25), dist);
}

// One lane per sub-ray: a wave covers 4x4 pixels x 4 sub-rays, a 256-thread block 8x8 pixels,
// so the rays of a wave stay within a few voxels of each other through the whole volume.
template <bool kSlicer, bool kBig, bool kResid, typename T, bool kFlat = false>
__global__ void __launch_bounds__(256) march_exact_kernel(VolG v, TiltG g, const T* __restrict__ vol, Roi roi,
                                                          size_t pitch, float* __restrict__ proj,
                                                          float* __restrict__ dist, float tminDef, float tmaxDef,
                                                          ResidEpi epi, int W, int H) {
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int sub = lane & 3, pix = lane >> 2;
    const int x = roi.x0 + (int)blockIdx.x * 8 + (wave & 1) * 4 + (pix & 3);
    const int y = roi.y0 + (int)blockIdx.y * 8 + (wave >> 1) * 4 + (pix >> 2);
    const bool active = x < roi.x1 && y < roi.y1;
    const double step = (double)v.vs[0] * 0.15;
    const MarchLane L = march_lane<kSlicer>(v, g, vol, active, x, y, sub, step, tminDef, tmaxDef);
    double acc = L.acc;
    if (L.count > 0) {
        if (v.nx >= 2 && v.ny >= 2 && v.nz >= 2) {
            acc += march_sum<kBig, kFlat>(v, vol, L.q0, L.dq, L.count);
        } else {
            double tot = 0.0;
            for (int k = 0; k < L.count; k++) {
                const float fk = (float)k;
                tot += (double)tex3d_linear(vol, v.nx, v.ny, v.nz, __builtin_fmaf(fk, L.dq[0], L.q0[0]) + 0.5f,
                                            __builtin_fmaf(fk, L.dq[1], L.q0[1]) + 0.5f,
                                            __builtin_fmaf(fk, L.dq[2], L.q0[2]) + 0.5f);
            }
            acc += tot;
        }
    }
    if (kResid)
        march_store_resid(acc, pixel_dist(L.len, L.hits), active, sub, x, y, pitch, step, epi, W, H, proj);
    else
        march_store<kSlicer>(acc, L.len, L.hits, active, sub, x, y, pitch, step, proj, dist);
}

// LDS-staged form of march_exact_kernel (tube_walk.hpp) for volumes of at least 2 voxels per
// axis.  Each lane runs the same cell walk in the same order as march_exact_kernel (the walk is only
// paused at slab boundaries), so the projection is bit-identical.  Six waves per SIMD (80
// registers): the walk is VALU bound, and the occupancy hides the LDS latency better than the few
// registers it spills.
template <bool kSlicer, bool kResid, typename T, bool kFlat = false>
__global__ void __launch_bounds__(256, 6) march_lds_kernel(VolG v, TiltG g, const T* __restrict__ vol, Roi roi,
                                                        size_t pitch, float* __restrict__ proj,
                                                        float* __restrict__ dist, float tminDef, float tmaxDef,
                                                        TubePlan tp, ResidEpi epi, int W, int H) {
    __shared__ float chunk[kChunkFloats];
    __shared__ int zr[2];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int sub = lane & 3, pix = lane >> 2;
    // Workgroup b runs on XCD b mod 8.  Tiles come in groups of kXcdGroup x kXcdGroup, dealt to the
    // XCDs in turn: a group's overlapping ray tubes are staged through one L2, and every XCD gets
    // groups from all over the detector (rows outside the volume's shadow are cheap).
    const int xcd = (int)(blockIdx.x & 7u), slot = (int)(blockIdx.x >> 3);
    const int grp = (slot / (kXcdGroup * kXcdGroup)) * 8 + xcd, mem = slot % (kXcdGroup * kXcdGroup);
    const int gtx = (grp % tp.groups_x) * kXcdGroup + mem % kXcdGroup;
    const int gty = (grp / tp.groups_x) * kXcdGroup + mem / kXcdGroup;
    if (gtx >= tp.tiles_x || gty >= tp.tiles_y) return;   // whole workgroup
    const int tx0 = roi.x0 + gtx * 8, ty0 = roi.y0 + gty * 8;
    const int x = tx0 + (wave & 1) * 4 + (pix & 3);
    const int y = ty0 + (wave >> 1) * 4 + (pix >> 2);
    const bool active = x < roi.x1 && y < roi.y1;
    const double step = (double)v.vs[0] * 0.15;
    const MarchLane L = march_lane<kSlicer>(v, g, vol, active, x, y, sub, step, tminDef, tmaxDef);
    // the mean chord before the walk: stored, or kept for the residual epilogue
    float dval = 0.0f;
    if (kResid) dval = pixel_dist(L.len, L.hits);
    else if (!kSlicer) march_store_dist(L.len, L.hits, active, sub, x, y, pitch, dist);
    const TubeDir td = tube_dir(v, g, step);
    const TubeBox tb = tube_box(v, g, tp, (double)tx0 + 0.25, (double)tx0 + 7.75, (double)ty0 + 0.25,
                                (double)ty0 + 7.75, 0.0, 0.0);
    double total = 0.0;   // summed as march_sum does, then added to the slicer's first sample
    tube_walk<kFlat, true>(v, vol, chunk, zr, tp, tb, td, L.q0, L.count,
                     [&](const float val[8], const auto* cell, auto k, auto mi) {
                         total += cell_sum_values<kFlat>(val, cell_moments<kFlat>(L.q0, td.dq, cell, k, mi), mi);
                     });
    double acc = L.acc;
    if (L.count > 0) acc += total;
    if (kResid)
        march_store_resid(acc, dval, active, sub, x, y, pitch, step, epi, W, H, proj);
    else
        march_store<true>(acc, 0.0, 0.0, active, sub, x, y, pitch, step, proj, nullptr);   // proj only
}

// Every CTF slab of a tilt in one pass: slab k (kbeg <= k < kend) covers bounds[k] .. bounds[k] +
// step_f along the rays.  Each sub-ray walks the volume once, its slabs as consecutive segments
// (tube_walk_segments), and leaves per slab the double sum march_lane / march_lds_kernel<true>
// would hold before the sub-rays are added: part[(k - kbeg) * pstride + 4 * pixel + sub], for the
// slabs range[4 * pixel + sub] = (klo, khi) it may cross (a superset found by bisection on the
// sorted bounds, where bounds[k] + step_f == bounds[k + 1]).  slicer_combine_kernel adds the four
// sub-rays as march_store does.
struct SlicerSegs {
    const VolG& v;
    const TiltG& g;
    RayD r;   // the sub-ray (ray_setup_d)
    const float* __restrict__ bounds;
    double* __restrict__ part;
    size_t pstride;
    double step, total;
    float step_f, px, py, tex;   // px, py: the sub-ray's pixel coordinates (exact in float)
    float tnext;                 // bounds[k + 1], loaded one segment ahead of its use
    float q0[3];
    unsigned slot;
    int count;   // samples of the current slab; -1 when the ray misses it
    int k, klo, khi, kbeg;

    // finish slab k (its sum into part), then set up slab k + 1
    __device__ __forceinline__ bool next() {
        if (k >= klo) {
            // march_lane: the detector sample when the ray meets the slab, plus the walk's sum
            const double acc = count < 0 ? 0.0 : (count > 0 ? (double)tex + total : (double)tex);
            part[(size_t)(k - kbeg) * pstride + slot] = acc;
        }
        if (++k > khi) return false;
        // the slab's start was loaded when the previous segment began; the next one's load is
        // issued now, so that the walk of this segment covers its latency
        const float tmin = tnext, tmax = tmin + step_f;
        if (k < khi) tnext = bounds[k + 1];
        const double t_in = fmax(r.t_in, (double)tmin), t_last = fmin(r.t_out, (double)tmax);
        count = t_in <= t_last ? 0 : -1;
        total = 0.0;
        const double t_first = t_in + step;
        if (count == 0 && t_first <= t_last) {
            count = sample_count(t_first, t_last, step);
            // sample_start with the geometry widened here: hoisted out of the walk, the widened
            // doubles are held in vector registers and spilled (scalar units have no float math)
            float gn[3], bb[3], rc[3], dm[3], zs = v.zshift;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                gn[i] = g.n[i];
                bb[i] = v.bbMin[i];
                rc[i] = v.rcp[i];
                dm[i] = v.dim[i];
                __asm__ volatile("" : "+s"(gn[i]), "+s"(bb[i]), "+s"(rc[i]), "+s"(dm[i]));
            }
            __asm__ volatile("" : "+s"(zs));
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const double scale = (double)rc[i] * (double)dm[i];
                const double p = (r.c[i] + t_first * (double)gn[i]) - (double)bb[i];
                q0[i] = (float)(p * scale - 0.5 - (i == 2 ? (double)zs : 0.0));
            }
        }
        return true;
    }
};

template <typename T, bool kFlat = false>
__global__ void __launch_bounds__(256, 6) slicer_segments_lds_kernel(VolG v, TiltG g, const T* __restrict__ vol,
                                                                  int W, int H, const float* __restrict__ bounds,
                                                                  int kbeg, int kend, float step_f,
                                                                  double* __restrict__ part, size_t pstride,
                                                                  int2* __restrict__ range, TubePlan tp) {
    __shared__ float chunk[kChunkFloats];
    __shared__ int zr[2];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int sub = lane & 3, pix = lane >> 2;
    const int xcd = (int)(blockIdx.x & 7u), slot = (int)(blockIdx.x >> 3);
    constexpr int kG = kSlicerXcdGroup;
    const int grp = (slot / (kG * kG)) * 8 + xcd, mem = slot % (kG * kG);
    const int gtx = (grp % tp.groups_x) * kG + mem % kG;
    const int gty = (grp / tp.groups_x) * kG + mem / kG;
    if (gtx >= tp.tiles_x || gty >= tp.tiles_y) return;   // whole workgroup
    const int tx0 = gtx * 8, ty0 = gty * 8;
    const int x = tx0 + (wave & 1) * 4 + (pix & 3);
    const int y = ty0 + (wave >> 1) * 4 + (pix >> 2);
    const bool active = x < W && y < H;
    const double step = (double)v.vs[0] * 0.15;
    SlicerSegs seg{v, g, RayD{}, bounds, part};
    seg.pstride = pstride;
    seg.step = step;
    seg.total = 0.0;
    seg.step_f = step_f;
    seg.px = (float)x + ((sub & 2) ? 0.75f : 0.25f);
    seg.py = (float)y + ((sub & 1) ? 0.75f : 0.25f);
    seg.tex = 0.0f;
    seg.tnext = 0.0f;
    seg.q0[0] = seg.q0[1] = seg.q0[2] = 0.0f;
    seg.slot = 4u * ((unsigned)y * (unsigned)W + (unsigned)x) + (unsigned)sub;
    seg.count = 0;
    seg.kbeg = kbeg;
    seg.klo = 0;   // no slabs: next() writes nothing and returns false
    seg.k = seg.khi = -1;
    int za = 1, zb = 0;
    if (active) {
        seg.r = ray_setup_d(v, g, (double)seg.px, (double)seg.py);
        const RayD& r = seg.r;
        int klo = kend, khi = kbeg - 1;
        if ((r.t_out - r.t_in) > 0.0) {
            // a slab is crossed only if its end reaches t_in and its start t_out
            int lo = kbeg, hi = kend;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if ((double)(bounds[mid] + step_f) >= r.t_in) hi = mid;
                else lo = mid + 1;
            }
            klo = lo;
            hi = kend;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if ((double)bounds[mid] > r.t_out) hi = mid;
                else lo = mid + 1;
            }
            khi = lo - 1;
        }
        range[seg.slot] = make_int2(klo, khi);
        if (klo <= khi) {
            seg.tex = detector_sample(v, vol, r);
            seg.klo = klo;
            seg.k = klo - 1;
            seg.khi = khi;
            seg.tnext = bounds[klo];
            // cell layers of the ray between the slabs' ends (the walk adds a layer of margin)
            const double scale = (double)v.rcp[2] * (double)v.dim[2];
            const double ta = fmax(r.t_in, (double)bounds[klo]), tb = fmin(r.t_out, (double)(bounds[khi] + step_f));
            const double qa = ((r.c[2] + ta * (double)g.n[2]) - (double)v.bbMin[2]) * scale - 0.5 - (double)v.zshift;
            const double qb = ((r.c[2] + tb * (double)g.n[2]) - (double)v.bbMin[2]) * scale - 0.5 - (double)v.zshift;
            za = (int)floor(fmin(qa, qb));
            zb = (int)floor(fmax(qa, qb));
        }
    }
    const TubeDir td = tube_dir(v, g, step);
    const TubeBox tb = tube_box(v, g, tp, (double)tx0 + 0.25, (double)tx0 + 7.75, (double)ty0 + 0.25,
                                (double)ty0 + 7.75, 0.0, 0.0);
    tube_walk_segments<kFlat, true>(v, vol, chunk, zr, tp, tb, td, za, zb, seg,
                              [&](const float val[8], const auto* cell, auto kk, auto mi) {
                                  seg.total += cell_sum_values<kFlat>(
                                      val, cell_moments<kFlat>(seg.q0, td.dq, cell, kk, mi), mi);
                              });
}

// Slab images from the sub-ray sums: stack[k * stride + y * pitch + x] = 0 + the four sub-rays'
// sums added as march_store does ((s0 + s1) + (s2 + s3)), times step / 4, for kbeg <= k < kend; a
// sub-ray's slabs outside its range hold 0.
__global__ void __launch_bounds__(256) slicer_combine_kernel(const double* __restrict__ part, size_t pstride,
                                                             const int2* __restrict__ range, int W, int H,
                                                             size_t pitch, float* __restrict__ stack, size_t stride,
                                                             int kbeg, int kend, double step) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)blockIdx.y;
    if (x >= W) return;
    // one thread per pixel over the pass's slabs: its sub-rays' slab ranges are read once
    const size_t p4 = 4 * ((size_t)y * (size_t)W + (size_t)x);
    int2 r[4];
#pragma unroll
    for (int s = 0; s < 4; s++) r[s] = range[p4 + s];
    for (int k = kbeg; k < kend; k++) {
        const double* pk = part + (size_t)(k - kbeg) * pstride + p4;
        double a[4];
#pragma unroll
        for (int s = 0; s < 4; s++) a[s] = (k >= r[s].x && k <= r[s].y) ? pk[s] : 0.0;
        const double acc = (a[0] + a[1]) + (a[2] + a[3]);
        stack[(size_t)k * stride + (size_t)y * pitch + (size_t)x] = 0.0f + (float)(acc * step * 0.25);
    }
}

__global__ void __launch_bounds__(256) vol_traversal_exact_kernel(VolG v, TiltG g, Roi roi, size_t pitch,
                                                                  float* __restrict__ dist) {
    const int x = roi.x0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = roi.y0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= roi.x1 || y >= roi.y1) return;
    const RayD r = ray_setup_d(v, g, (double)x + 0.5, (double)y + 0.5);
    double val = 0.0;
    if (r.t_out > r.t_in) val = r.t_out - r.t_in;
    const size_t i = (size_t)y * pitch + (size_t)x;
    dist[i] = (float)((double)dist[i] + val);
}

Roi make_roi(const int* roi, int W, int H) {
    Roi r{0, 0, W, H};
    if (roi) {
        r.x0 = roi[0] < 0 ? 0 : roi[0];
        r.y0 = roi[1] < 0 ? 0 : roi[1];
        r.x1 = roi[2] > W ? W : roi[2];
        r.y1 = roi[3] > H ? H : roi[3];
    }
    return r;
}

dim3 grid_for(const Roi& r) {
    const int w = r.x1 - r.x0, h = r.y1 - r.y0;
    return dim3((unsigned)((w + 63) / 64), (unsigned)((h + 3) / 4));
}

template <bool kResid, typename T>
void launch_exact_t(artia_ctx* ctx, const VolG& v, const TiltG& g_in, const T* vol, const Roi& r, size_t pitch,
                    float* proj, float* dist, bool slicer, float tmin, float tmax, const ResidEpi& epi, int W, int H) {
    const TiltG g = flat_beam(g_in);
    const dim3 grid((unsigned)((r.x1 - r.x0 + 7) / 8), (unsigned)((r.y1 - r.y0 + 7) / 8)), block(256);
    // 32-bit buffer offsets need fewer than 2^30 voxels and a row index z * ny + y below 2^24
    const bool big = (uint64_t)v.nx * (uint64_t)v.ny * (uint64_t)v.nz >= (1ull << 30) ||
                     (uint64_t)v.ny * (uint64_t)v.nz >= (1ull << 24) || v.nx >= (1 << 24);
    // rays without a y component (every tilt about the y axis whose beam direction has n_y == 0):
    // the walk without the y axis, the same projections
    const bool flat = g.n[1] == 0.0f;
    TubePlan tp;
    if (ctx->march_lds && tube_plan(v, g, tp)) {
        tp.tiles_x = (int)grid.x;
        tp.tiles_y = (int)grid.y;
        tp.groups_x = (tp.tiles_x + kXcdGroup - 1) / kXcdGroup;
        const int groups = tp.groups_x * ((tp.tiles_y + kXcdGroup - 1) / kXcdGroup);
        const dim3 grid1((unsigned)(((groups + 7) / 8) * 8 * kXcdGroup * kXcdGroup));
        if (slicer)
            hipLaunchKernelGGL((march_lds_kernel<true, false, T>), grid1, block, 0, ctx->stream, v, g, vol, r, pitch,
                               proj, dist, tmin, tmax, tp, epi, W, H);
        else if (flat)
            hipLaunchKernelGGL((march_lds_kernel<false, kResid, T, true>), grid1, block, 0, ctx->stream, v, g, vol, r,
                               pitch, proj, dist, tmin, tmax, tp, epi, W, H);
        else
            hipLaunchKernelGGL((march_lds_kernel<false, kResid, T>), grid1, block, 0, ctx->stream, v, g, vol, r,
                               pitch, proj, dist, tmin, tmax, tp, epi, W, H);
        return;
    }
    if (slicer) {
        if (big)
            hipLaunchKernelGGL((march_exact_kernel<true, true, false, T>), grid, block, 0, ctx->stream, v, g, vol, r,
                               pitch, proj, dist, tmin, tmax, epi, W, H);
        else
            hipLaunchKernelGGL((march_exact_kernel<true, false, false, T>), grid, block, 0, ctx->stream, v, g, vol, r,
                               pitch, proj, dist, tmin, tmax, epi, W, H);
    } else if (flat) {
        if (big)
            hipLaunchKernelGGL((march_exact_kernel<false, true, kResid, T, true>), grid, block, 0, ctx->stream, v, g,
                               vol, r, pitch, proj, dist, tmin, tmax, epi, W, H);
        else
            hipLaunchKernelGGL((march_exact_kernel<false, false, kResid, T, true>), grid, block, 0, ctx->stream, v, g,
                               vol, r, pitch, proj, dist, tmin, tmax, epi, W, H);
    } else {
        if (big)
            hipLaunchKernelGGL((march_exact_kernel<false, true, kResid, T>), grid, block, 0, ctx->stream, v, g, vol, r,
                               pitch, proj, dist, tmin, tmax, epi, W, H);
        else
            hipLaunchKernelGGL((march_exact_kernel<false, false, kResid, T>), grid, block, 0, ctx->stream, v, g, vol,
                               r, pitch, proj, dist, tmin, tmax, epi, W, H);
    }
}

template <typename T>
void launch_march_t(artia_ctx* ctx, const VolG& v, const TiltG& g, const T* vol, const Roi& r, size_t pitch,
                    float* proj, float* dist, bool slicer, float tmin, float tmax) {
    if (r.x1 <= r.x0 || r.y1 <= r.y0) return;
    if (ctx->numerics == ARTIA_NUMERICS_REFERENCE) {
        TubePlan tp;
        if (!slicer && ctx->march_lds && tube_plan(v, g, tp, (double)(kRefTile - 1), kRefSlab)) {
            const dim3 grid((unsigned)((r.x1 - r.x0 + kRefTile - 1) / kRefTile),
                            (unsigned)((r.y1 - r.y0 + kRefTile - 1) / kRefTile));
            hipLaunchKernelGGL((march_ref_lds_kernel<T>), grid, dim3(256), 0, ctx->stream, v, g, vol, r, pitch, proj,
                               dist, tp);
            return;
        }
        const dim3 grid = grid_for(r), block(64, 4);
        if (slicer)
            hipLaunchKernelGGL((march_ref_kernel<true, T>), grid, block, 0, ctx->stream, v, g, vol, r, pitch, proj,
                               dist, tmin, tmax);
        else
            hipLaunchKernelGGL((march_ref_kernel<false, T>), grid, block, 0, ctx->stream, v, g, vol, r, pitch, proj,
                               dist, tmin, tmax);
        return;
    }
    launch_exact_t<false>(ctx, v, g, vol, r, pitch, proj, dist, slicer, tmin, tmax, ResidEpi{}, 0, 0);
}

void launch_march_impl(artia_ctx* ctx, const VolG& v, const TiltG& g, VolData vol, const Roi& r, size_t pitch,
                       float* proj, float* dist, bool slicer, float tmin, float tmax) {
    if (vol.half)
        launch_march_t(ctx, v, g, static_cast<const vhalf*>(vol.p), r, pitch, proj, dist, slicer, tmin, tmax);
    else
        launch_march_t(ctx, v, g, static_cast<const float*>(vol.p), r, pitch, proj, dist, slicer, tmin, tmax);
}

}  // namespace

void launch_march(artia_ctx* ctx, const VolG& v, const TiltG& g, VolData vol, int W, int H, size_t pitch,
                  float* proj, float* dist, int x0, int y0, int x1, int y1) {
    const int roi[4] = {x0, y0, x1, y1};
    launch_march_impl(ctx, v, g, vol, make_roi(roi, W, H), pitch, proj, dist, false, 0.0f, 0.0f);
}

void launch_march_residual(artia_ctx* ctx, const VolG& v, const TiltG& g, VolData vol, int W, int H, size_t pitch,
                           const float* real, float realLength, const float crop[4], const float cropDim[4],
                           float scale, const float cut[2], const float dim[2], const int* corners, float* resid,
                           float* dist) {
    if (ctx->numerics == ARTIA_NUMERICS_REFERENCE) {
        // the reference's three launches (the kernel bodies keep their own float order)
        launch_fill(ctx, resid, (size_t)H * pitch, 0.0f);
        launch_march(ctx, v, g, vol, W, H, pitch, resid, dist, 0, 0, W, H);
        launch_compare(ctx, W, H, pitch, real, resid, dist, realLength, crop, cropDim, scale);
        if (corners) launch_crop_border(ctx, W, H, pitch, resid, cut, dim, corners);
        return;
    }
    ResidEpi e{};
    e.real = real;
    e.L = realLength;
    e.cut = F4{crop[0], crop[1], crop[2], crop[3]};
    e.dim = F4{cropDim[0], cropDim[1], cropDim[2], cropDim[3]};
    e.S = scale;
    e.crop_on = corners ? 1 : 0;
    if (corners) e.crop = make_crop_params(cut, dim, corners);
    const Roi r{0, 0, W, H};
    if (vol.half)
        launch_exact_t<true>(ctx, v, g, static_cast<const vhalf*>(vol.p), r, pitch, resid, nullptr, false, 0.0f, 0.0f,
                             e, W, H);
    else
        launch_exact_t<true>(ctx, v, g, static_cast<const float*>(vol.p), r, pitch, resid, nullptr, false, 0.0f, 0.0f,
                             e, W, H);
}

void launch_slicer(artia_ctx* ctx, const VolG& v, const TiltG& g, VolData vol, int W, int H, size_t pitch,
                   float* proj, float tmin, float tmax) {
    const int roi[4] = {0, 0, W, H};
    launch_march_impl(ctx, v, g, vol, make_roi(roi, W, H), pitch, proj, nullptr, true, tmin, tmax);
}

size_t slicer_stack_scratch_bytes(int W, int H, int nslab) {
    const size_t npix = (size_t)W * (size_t)H;
    return npix * 4 * sizeof(int2) + (size_t)nslab * npix * 4 * sizeof(double);
}

void launch_slicer_stack(artia_ctx* ctx, const VolG& v, const TiltG& g_in, VolData vol, int W, int H, size_t pitch,
                         float* stack, size_t stride, const float* bounds, const float* d_bounds, int nslab,
                         float step, void* scratch, size_t scratch_bytes) {
    if (nslab <= 0) return;
    // the exact one-pass form takes the flat-beam geometry, as the per-slab launches below do
    const TiltG g = ctx->numerics == ARTIA_NUMERICS_EXACT ? flat_beam(g_in) : g_in;
    const size_t npix = (size_t)W * (size_t)H;
    const size_t range_bytes = npix * 4 * sizeof(int2), slab_bytes = npix * 4 * sizeof(double);
    const int per_pass = scratch_bytes > range_bytes ? (int)std::min<size_t>((scratch_bytes - range_bytes) / slab_bytes,
                                                                             (size_t)nslab)
                                                     : 0;
    TubePlan tp;
    // sub-ray pixel coordinates exact in float, 32-bit sub-ray slots and one grid row per image row
    const bool fits = W < (1 << 22) && H <= 65535 && npix * 4 < ((size_t)1 << 32);
    if (ctx->numerics == ARTIA_NUMERICS_EXACT && ctx->march_lds && per_pass > 0 && fits && tube_plan(v, g, tp)) {
        tp.tiles_x = (W + 7) / 8;
        tp.tiles_y = (H + 7) / 8;
        tp.groups_x = (tp.tiles_x + kSlicerXcdGroup - 1) / kSlicerXcdGroup;
        const int groups = tp.groups_x * ((tp.tiles_y + kSlicerXcdGroup - 1) / kSlicerXcdGroup);
        const dim3 grid1((unsigned)(((groups + 7) / 8) * 8 * kSlicerXcdGroup * kSlicerXcdGroup));
        int2* range = static_cast<int2*>(scratch);
        double* part = reinterpret_cast<double*>(static_cast<char*>(scratch) + range_bytes);
        const double dstep = (double)v.vs[0] * 0.15;
        // rays without a y component: the walk without the y axis (the same cells and sums)
        const bool flat = g.n[1] == 0.0f;
        for (int kb = 0; kb < nslab; kb += per_pass) {
            const int ke = std::min(nslab, kb + per_pass);
            if (vol.half && flat)
                hipLaunchKernelGGL((slicer_segments_lds_kernel<vhalf, true>), grid1, dim3(256), 0, ctx->stream, v, g,
                                   static_cast<const vhalf*>(vol.p), W, H, d_bounds, kb, ke, step, part, npix * 4,
                                   range, tp);
            else if (vol.half)
                hipLaunchKernelGGL((slicer_segments_lds_kernel<vhalf, false>), grid1, dim3(256), 0, ctx->stream, v,
                                   g, static_cast<const vhalf*>(vol.p), W, H, d_bounds, kb, ke, step, part, npix * 4,
                                   range, tp);
            else if (flat)
                hipLaunchKernelGGL((slicer_segments_lds_kernel<float, true>), grid1, dim3(256), 0, ctx->stream, v, g,
                                   static_cast<const float*>(vol.p), W, H, d_bounds, kb, ke, step, part, npix * 4,
                                   range, tp);
            else
                hipLaunchKernelGGL((slicer_segments_lds_kernel<float, false>), grid1, dim3(256), 0, ctx->stream, v,
                                   g, static_cast<const float*>(vol.p), W, H, d_bounds, kb, ke, step, part, npix * 4,
                                   range, tp);
            hipLaunchKernelGGL(slicer_combine_kernel, dim3((unsigned)((W + 255) / 256), (unsigned)H), dim3(256), 0,
                               ctx->stream, part, npix * 4, range, W, H, pitch, stack, stride, kb, ke, dstep);
        }
        return;
    }
    launch_fill(ctx, stack, (size_t)nslab * stride, 0.0f);
    const int roiA[4] = {0, 0, W, H};
    const Roi r = make_roi(roiA, W, H);
    for (int k = 0; k < nslab; k++)
        launch_march_impl(ctx, v, g, vol, r, pitch, stack + (size_t)k * stride, nullptr, true, bounds[k],
                          bounds[k] + step);
}

void launch_vol_traversal(artia_ctx* ctx, const VolG& v, const TiltG& g, int W, int H, size_t pitch, float* dist,
                          int x0, int y0, int x1, int y1) {
    const int roiA[4] = {x0, y0, x1, y1};
    const Roi r = make_roi(roiA, W, H);
    if (r.x1 <= r.x0 || r.y1 <= r.y0) return;
    if (ctx->numerics == ARTIA_NUMERICS_REFERENCE)
        hipLaunchKernelGGL(vol_traversal_ref_kernel, grid_for(r), dim3(64, 4), 0, ctx->stream, v, g, r, pitch, dist);
    else
        hipLaunchKernelGGL(vol_traversal_exact_kernel, grid_for(r), dim3(64, 4), 0, ctx->stream, v, flat_beam(g), r,
                           pitch, dist);
}

}  // namespace artia

using namespace artia;

namespace {
artia_status check_image(artia_ctx* ctx, const char* fn, int W, int H, size_t pitch) {
    if (W <= 0 || H <= 0 || pitch < (size_t)W)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: bad image size %dx%d pitch %zu", fn, W, H, pitch);
    return ARTIA_OK;
}

// artia_fp_march (slicer false: proj and dist) and artia_fp_slicer (proj only)
artia_status fp_entry(artia_ctx* ctx, const char* fn, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                      VolData vol, int width, int height, size_t pitch, float* proj, float* dist, const int* roi,
                      bool slicer, float tmin, float tmax, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!valid_geometry(vg, tg) || !vol.p || !proj || (!slicer && !dist))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: null pointer or invalid geometry", fn);
    artia_status st = check_image(ctx, fn, width, height, pitch);
    if (st != ARTIA_OK) return st;
    Timing t(ctx, ms);
    launch_march_impl(ctx, to_device(*vg), to_device(*tg), vol, make_roi(roi, width, height), pitch, proj, dist,
                      slicer, tmin, tmax);
    return t.finish();
}
}  // namespace

extern "C" {

artia_status artia_fp_march(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, const float* vol,
                            int width, int height, size_t pitch, float* proj, float* dist, const int* roi, float* ms) {
    return fp_entry(ctx, "artia_fp_march", vg, tg, VolData(vol), width, height, pitch, proj, dist, roi, false, 0.0f,
                    0.0f, ms);
}

artia_status artia_fp_march_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                                const uint16_t* vol, int width, int height, size_t pitch, float* proj, float* dist,
                                const int* roi, float* ms) {
    return fp_entry(ctx, "artia_fp_march_f16", vg, tg, VolData::f16(vol), width, height, pitch, proj, dist, roi,
                    false, 0.0f, 0.0f, ms);
}

artia_status artia_fp_slicer(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, const float* vol,
                             int width, int height, size_t pitch, float* proj, float tmin, float tmax, const int* roi,
                             float* ms) {
    return fp_entry(ctx, "artia_fp_slicer", vg, tg, VolData(vol), width, height, pitch, proj, nullptr, roi, true,
                    tmin, tmax, ms);
}

artia_status artia_fp_slicer_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                                 const uint16_t* vol, int width, int height, size_t pitch, float* proj, float tmin,
                                 float tmax, const int* roi, float* ms) {
    return fp_entry(ctx, "artia_fp_slicer_f16", vg, tg, VolData::f16(vol), width, height, pitch, proj, nullptr, roi,
                    true, tmin, tmax, ms);
}

namespace {
artia_status fp_residual_entry(artia_ctx* ctx, const char* fn, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                               VolData vol, int volume_is_empty, int width, int height, size_t pitch, const float* real,
                               float real_length, const float crop[4], const float crop_dim[4], float proj_scale,
                               const float cut_length[2], const float dim_length[2], const int* hitpoints,
                               float* residual, float* dist, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!valid_geometry(vg, tg) || (!vol.p && !volume_is_empty) || !real || !crop || !crop_dim || !residual ||
        (hitpoints && (!cut_length || !dim_length)))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: null pointer or invalid geometry", fn);
    if (!dist && (volume_is_empty || ctx->numerics == ARTIA_NUMERICS_REFERENCE))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: an empty volume or reference numerics need the dist image",
                    fn);
    artia_status st = check_image(ctx, fn, width, height, pitch);
    if (st != ARTIA_OK) return st;
    Timing t(ctx, ms);
    const VolG v = to_device(*vg);
    const TiltG g = to_device(*tg);
    if (volume_is_empty) {
        // EmSart.cpp:598-601: no projection of an empty volume, the centre-ray traversal length only
        launch_fill(ctx, residual, (size_t)height * pitch, 0.0f);
        launch_fill(ctx, dist, (size_t)height * pitch, 0.0f);
        launch_vol_traversal(ctx, v, g, width, height, pitch, dist, 0, 0, width, height);
        launch_compare(ctx, width, height, pitch, real, residual, dist, real_length, crop, crop_dim, proj_scale);
        if (hitpoints) launch_crop_border(ctx, width, height, pitch, residual, cut_length, dim_length, hitpoints);
    } else {
        launch_march_residual(ctx, v, g, vol, width, height, pitch, real, real_length, crop, crop_dim, proj_scale,
                              cut_length, dim_length, hitpoints, residual, dist);
    }
    return t.finish();
}
}  // namespace

artia_status artia_fp_residual(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, const float* vol,
                               int volume_is_empty, int width, int height, size_t pitch, const float* real,
                               float real_length, const float crop[4], const float crop_dim[4], float proj_scale,
                               const float cut_length[2], const float dim_length[2], const int hitpoints[8],
                               float* residual, float* dist, float* ms) {
    return fp_residual_entry(ctx, "artia_fp_residual", vg, tg, VolData(vol), volume_is_empty, width, height, pitch,
                             real, real_length, crop, crop_dim, proj_scale, cut_length, dim_length, hitpoints,
                             residual, dist, ms);
}

artia_status artia_fp_residual_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                                   const uint16_t* vol, int volume_is_empty, int width, int height, size_t pitch,
                                   const float* real, float real_length, const float crop[4], const float crop_dim[4],
                                   float proj_scale, const float cut_length[2], const float dim_length[2],
                                   const int hitpoints[8], float* residual, float* dist, float* ms) {
    return fp_residual_entry(ctx, "artia_fp_residual_f16", vg, tg, VolData::f16(vol), volume_is_empty, width, height,
                             pitch, real, real_length, crop, crop_dim, proj_scale, cut_length, dim_length, hitpoints,
                             residual, dist, ms);
}

artia_status artia_vol_traversal(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, int width,
                                 int height, size_t pitch, float* dist, const int* roi, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!valid_geometry(vg, tg) || !dist)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_vol_traversal: null pointer or invalid geometry");
    artia_status st = check_image(ctx, "artia_vol_traversal", width, height, pitch);
    if (st != ARTIA_OK) return st;
    Timing t(ctx, ms);
    const Roi r = make_roi(roi, width, height);
    launch_vol_traversal(ctx, to_device(*vg), to_device(*tg), width, height, pitch, dist, r.x0, r.y0, r.x1, r.y1);
    return t.finish();
}

}  // extern "C"
