// Rotation of a reference cube by motive-list Euler angles (EmSartRefine.cpp:805-865 rotates the
// reference once per motive with RotKernel, Kernels.cpp:1227-1386, kernel rot3d,
// wbpWeighting.cu:388-405), batched: one cube in, one rotated cube out per angle triple.
//
// out[z][y][x] = trilinear(in, c + R^T (v - c)), v = (x, y, z), c = (float)(size / 2) (integer
// division), trilinear with the texel index clamped to [0, size - 1] per axis (the texture's CLAMP
// addressing at texel centres: tex3D(p + 0.5)), weights in exact float (the reference's texture unit
// filters with reduced-precision weights).  R is computeRotMat's matrix, computed on the host.
#include <cmath>

#include "artia_internal.hpp"

namespace artia {
namespace {

// Matrices per launch, passed by value (9 floats each, inside the kernel-argument limit).
constexpr int kRotBatch = 48;
constexpr int kRotZ = 8;   // output slices per thread

struct RotSet {
    float m[kRotBatch][9];   // row-major R[i][j]
};

__device__ __forceinline__ float clamp_fetch(const float* __restrict__ in, int S, int x, int y, int z) {
    x = min(max(x, 0), S - 1);
    y = min(max(y, 0), S - 1);
    z = min(max(z, 0), S - 1);
    return in[((long long)z * S + y) * S + x];
}

// One thread: output column (x, y) over kRotZ slices of rotation blockIdx.z / zt.
__global__ void __launch_bounds__(256) rot3d_kernel(const float* __restrict__ in, int S, RotSet set, int zt,
                                                    float* __restrict__ out) {
    const int x = (int)(blockIdx.x * 64 + threadIdx.x);
    const int y = (int)(blockIdx.y * 4 + threadIdx.y);
    const int r = (int)blockIdx.z / zt;
    const int z0 = ((int)blockIdx.z % zt) * kRotZ;
    if (x >= S || y >= S) return;
    const float* R = set.m[r];
    const float c = (float)(S / 2);
    const float vx = (float)x - c, vy = (float)y - c;
    float* __restrict__ o = out + (long long)r * S * S * S;
    for (int k = 0; k < kRotZ; k++) {
        const int z = z0 + k;
        if (z >= S) break;
        const float vz = (float)z - c;
        // rot3d: position = centre + R^T vox, summed in the kernel's order
        const float px = c + R[0] * vx + R[3] * vy + R[6] * vz;
        const float py = c + R[1] * vx + R[4] * vy + R[7] * vz;
        const float pz = c + R[2] * vx + R[5] * vy + R[8] * vz;
        const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
        const float tx = px - fx, ty = py - fy, tz = pz - fz;
        const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
        const float c000 = clamp_fetch(in, S, ix, iy, iz), c100 = clamp_fetch(in, S, ix + 1, iy, iz);
        const float c010 = clamp_fetch(in, S, ix, iy + 1, iz), c110 = clamp_fetch(in, S, ix + 1, iy + 1, iz);
        const float c001 = clamp_fetch(in, S, ix, iy, iz + 1), c101 = clamp_fetch(in, S, ix + 1, iy, iz + 1);
        const float c011 = clamp_fetch(in, S, ix, iy + 1, iz + 1), c111 = clamp_fetch(in, S, ix + 1, iy + 1, iz + 1);
        const float sx = 1.0f - tx, sy = 1.0f - ty, sz = 1.0f - tz;
        const float a0 = sx * c000 + tx * c100, a1 = sx * c010 + tx * c110;
        const float a2 = sx * c001 + tx * c101, a3 = sx * c011 + tx * c111;
        const float b0 = sy * a0 + ty * a1, b1 = sy * a2 + ty * a3;
        o[((long long)z * S + y) * S + x] = sz * b0 + tz * b1;
    }
}

// computeRotMat's exact cosine and sine for the sixteen multiples of 30 and 45 degrees it tabulates.
bool table_angle(float a, float& c, float& s) {
    static const float deg[16] = {0, 30, 45, 60, 90, 120, 135, 150, 180, 210, 225, 240, 270, 300, 315, 330};
    const float r3 = std::sqrt(3.0f) / 2.0f, r2 = std::sqrt(2.0f) / 2.0f;
    const float cs[16] = {1.0f, r3, r2, 0.5f, 0.0f, -0.5f, -r2, -r3, -1.0f, -r3, -r2, -0.5f, 0.0f, 0.5f, r2, r3};
    const float sn[16] = {0.0f, 0.5f, r2, r3, 1.0f, r3, r2, 0.5f, 0.0f, -0.5f, -r2, -r3, -1.0f, -r3, -r2, -0.5f};
    for (int i = 0; i < 16; i++)
        if (deg[i] == a) {
            c = cs[i];
            s = sn[i];
            return true;
        }
    return false;
}

void cos_sin(float a, float& c, float& s) {
    if (table_angle(a, c, s)) return;
    const float rad = a * (float)M_PI / 180.0f;
    c = std::cos(rad);
    s = std::sin(rad);
}

}  // namespace
}  // namespace artia

using namespace artia;

extern "C" {

void artia_rot_matrix(float phi, float psi, float theta, float R[9]) {
    float cphi, sphi, cpsi, spsi, cth, sth;
    cos_sin(phi, cphi, sphi);
    cos_sin(psi, cpsi, spsi);
    cos_sin(theta, cth, sth);
    R[0] = cpsi * cphi - cth * spsi * sphi;
    R[1] = -cpsi * sphi - cth * spsi * cphi;
    R[2] = sth * spsi;
    R[3] = spsi * cphi + cth * cpsi * sphi;
    R[4] = -spsi * sphi + cth * cpsi * cphi;
    R[5] = -sth * cpsi;
    R[6] = sth * sphi;
    R[7] = sth * cphi;
    R[8] = cth;
}

artia_status artia_rot3d(artia_ctx* ctx, const float* vol, int size, const float* angles, int n, float* out,
                         float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!vol || (n > 0 && (!angles || !out)))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_rot3d: null pointer");
    if (size <= 0 || size > 1024 || n < 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_rot3d: size %d must be in 1..1024 and n %d >= 0", size,
                    n);
    Timing t(ctx, ms);
    const int zt = (size + kRotZ - 1) / kRotZ;
    const dim3 block(64, 4, 1);
    for (int r0 = 0; r0 < n; r0 += kRotBatch) {
        const int nb = n - r0 < kRotBatch ? n - r0 : kRotBatch;
        RotSet set{};
        for (int r = 0; r < nb; r++)
            artia_rot_matrix(angles[3 * (r0 + r)], angles[3 * (r0 + r) + 1], angles[3 * (r0 + r) + 2], set.m[r]);
        const dim3 grid((unsigned)((size + 63) / 64), (unsigned)((size + 3) / 4), (unsigned)(zt * nb));
        hipLaunchKernelGGL(rot3d_kernel, grid, block, 0, ctx->stream, vol, size, set, zt,
                           out + (size_t)r0 * size * size * size);
        artia_status st = check_hip(ctx, hipGetLastError(), "artia_rot3d launch");
        if (st != ARTIA_OK) return st;
    }
    return t.finish();
}

}  // extern "C"
