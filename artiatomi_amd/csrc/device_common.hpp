// Device helpers shared by the gfx950 kernels: exact-float texture emulation and small math.
// The filtering formulas are the library's definition of the reference's texture reads
// (clamp addressing, unnormalised coordinates, texel j centred at j + 0.5; EmSart.cpp:421,
// Reconstructor.cpp:318-327, CudaTextures.cpp:792-798): weights are exact float, evaluated
// as fma lerps, rather than the sampler's low-precision fixed point.
#pragma once

#include <hip/hip_runtime.h>

#include "artia_internal.hpp"

namespace artia {

// clamp to [lo, hi] (lo <= hi): one v_med3_i32
__device__ __forceinline__ int clampi(int i, int lo, int hi) { return min(max(i, lo), hi); }
__device__ __forceinline__ float lerpf(float p, float q, float w) { return __builtin_fmaf(w, q - p, p); }

// Volume element types: fp32, or IEEE binary16, the reference's FP16Volume storage (a half-float
// CUDA array behind Volume<unsigned short>, EmSart.cpp:354-374, BackProjectionSquareOS.cu:785-1290).
// A voxel is read as float, exactly; a half voxel is written back rounded to nearest even
// (__float2half_rn).
typedef _Float16 vhalf;
__device__ __forceinline__ float vox(const float* p, size_t i) { return p[i]; }
__device__ __forceinline__ float vox(const vhalf* p, size_t i) { return (float)p[i]; }

// Linear-filtered, clamp-addressed read of a dense [z][y][x] volume of fp32 or half voxels.
template <typename T>
__device__ __forceinline__ float tex3d_linear(const T* __restrict__ vol, int nx, int ny, int nz, float x,
                                              float y, float z) {
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
    const size_t sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
    const T* p00 = vol + (size_t)k0 * sz + (size_t)j0 * sy;
    const T* p10 = vol + (size_t)k0 * sz + (size_t)j1 * sy;
    const T* p01 = vol + (size_t)k1 * sz + (size_t)j0 * sy;
    const T* p11 = vol + (size_t)k1 * sz + (size_t)j1 * sy;
    const float c00 = lerpf(vox(p00, i0), vox(p00, i1), a);
    const float c10 = lerpf(vox(p10, i0), vox(p10, i1), a);
    const float c01 = lerpf(vox(p01, i0), vox(p01, i1), a);
    const float c11 = lerpf(vox(p11, i0), vox(p11, i1), a);
    return lerpf(lerpf(c00, c10, b), lerpf(c01, c11, b), c);
}

// Point- or linear-filtered, clamp-addressed read of a pitched 2D fp32 image.
__device__ __forceinline__ float tex2d(const float* __restrict__ img, int W, int H, size_t pitch, int filter,
                                       float x, float y) {
    x = fminf(fmaxf(x, -1.0f), (float)(W + 1));
    y = fminf(fmaxf(y, -1.0f), (float)(H + 1));
    if (filter == ARTIA_FILTER_POINT) {
        const int i = clampi((int)floorf(x), 0, W - 1), j = clampi((int)floorf(y), 0, H - 1);
        return img[(size_t)j * pitch + (size_t)i];
    }
    const float xb = x - 0.5f, yb = y - 0.5f;
    const float fx = floorf(xb), fy = floorf(yb);
    const float a = xb - fx, b = yb - fy;
    int i0 = (int)fx, j0 = (int)fy;
    const int i1 = clampi(i0 + 1, 0, W - 1), j1 = clampi(j0 + 1, 0, H - 1);
    i0 = clampi(i0, 0, W - 1);
    j0 = clampi(j0, 0, H - 1);
    const float r0 = lerpf(img[(size_t)j0 * pitch + i0], img[(size_t)j0 * pitch + i1], a);
    const float r1 = lerpf(img[(size_t)j1 * pitch + i0], img[(size_t)j1 * pitch + i1], a);
    return lerpf(r0, r1, b);
}

// Mag-anisotropy 3x3 applied to (x, y, 1) (DeviceVariables.cuh:41-48)
__device__ __forceinline__ void mag_apply(const float A[9], float x, float y, float& xo, float& yo) {
    xo = (A[0] * x + A[1] * y) + A[2] * 1.0f;
    yo = (A[3] * x + A[4] * y) + A[5] * 1.0f;
}

// Upper bound on samples per ray; never reached for a positive voxel size, it only keeps a
// degenerate geometry from spinning a wave forever.
constexpr int kMaxSamplesPerRay = 1 << 22;

}  // namespace artia
