// Voxel-driven SART backprojection on gfx950 (BackProjectionSquareOS.cu:263-782).
//
// Work unit: a "quad" of four consecutive voxels along x (the reference's thread).  Each quad
// owns the detector footprint of its macro-cell [x, x+5) x [y, y+1) x [z, z+1), clamped to
// [0, W-1] x [0, H-1]; the sub-rays of a voxel are the lattice points
// (px0 + (i+0.5)/OS, py0 + (j+0.5)/OS) of that footprint.  Each voxel receives
// lambda * sum(img * chord) / sum(chord) over the sub-rays whose chord through it is positive.
// One thread per quad; blocks of 64 x 4 quads, so a wave reads and writes 1 KiB of one
// volume row with float4 accesses.
//
// reference numerics: the kernel body in the reference's float operation order, including
//   the sub-ray origin r0 + 1e5*n rounded to float and 6 IEEE divisions per voxel chord.
// exact numerics: the footprint and the gate in double with the reference's operation order
//   (so the sub-ray set is the double-precision one), chords from the ray's offset relative to
//   each voxel (u/v/n frame, double-precision offsets, float slab arithmetic), and only the
//   lattice points inside the voxel's own projected box are visited; the accumulation order
//   per voxel (py outer, px inner) is the reference's.
#include <cmath>

#include "device_common.hpp"

namespace artia {
namespace {

// ---------------------------------------------------------------- reference numerics

__device__ __forceinline__ float bp_src(float r, float n) { return (float)((double)r + 100000.0 * (double)n); }

__global__ void __launch_bounds__(256) bp_ref_kernel(VolG v, TiltG g, float* __restrict__ vol,
                                                     const float* __restrict__ img, int W, int H, size_t pitch,
                                                     float lambda, int os, float osInv, float distMin, float distMax,
                                                     int filter) {
    const int xq = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int z = (int)blockIdx.z;
    if (xq >= v.nx / 4 || y >= v.ny || z >= v.nz) return;
    const unsigned x = (unsigned)xq * 4u;
    const float* n = g.n;
    const float* d = g.d;
    const float mcMin[3] = {v.bbMin[0] + (float)x * v.vs[0], v.bbMin[1] + (float)(unsigned)y * v.vs[1],
                            v.bbMin[2] + (float)(unsigned)z * v.vs[2]};
    const float mcMax[3] = {v.bbMin[0] + (float)(x + 5u) * v.vs[0], v.bbMin[1] + (float)(unsigned)(y + 1) * v.vs[1],
                            v.bbMin[2] + (float)(unsigned)(z + 1) * v.vs[2]};
    const float nd = ((-n[0]) * d[0] - n[1] * d[1]) - n[2] * d[2];
    float bMin[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
    float bMax[3] = {-200000.f, -200000.f, -200000.f};
#pragma unroll
    for (int corner = 0; corner < 8; corner++) {
        const int cx = corner >> 2, cy = (corner >> 1) & 1, cz = corner & 1;
        const float c[3] = {cx ? mcMin[0] + (mcMax[0] - mcMin[0]) : mcMin[0],
                            cy ? mcMin[1] + (mcMax[1] - mcMin[1]) : mcMin[1],
                            cz ? mcMin[2] + (mcMax[2] - mcMin[2]) : mcMin[2]};
        float t = (n[0] * c[0] + n[1] * c[1]) + n[2] * c[2];
        t += nd;
        t = fabsf(t);
        if (corner == 0 && !(t >= distMin && t < distMax)) return;
        for (int i = 0; i < 3; i++) {
            const float h = t * (-n[i]) + c[i];
            bMin[i] = fminf(h, bMin[i]);
            bMax[i] = fmaxf(h, bMax[i]);
        }
    }
    float lo[3], hi[3];
    for (int i = 0; i < 3; i++) {
        lo[i] = fminf(bMin[i], bMax[i]);
        hi[i] = fmaxf(bMin[i], bMax[i]);
    }
    if (n[0] <= 0.0f) {
        const float tmp = lo[0];
        lo[0] = hi[0];
        hi[0] = tmp;
    }
    float plo[2], phi[2];
    for (int r = 0; r < 2; r++) {
        const float* m = g.M + 4 * r;
        plo[r] = ((m[0] * lo[0] + m[1] * lo[1]) + m[2] * lo[2]) + 1.0f * m[3];
        phi[r] = ((m[0] * hi[0] + m[1] * hi[1]) + m[2] * hi[2]) + 1.0f * m[3];
    }
    int px0 = (int)floorf(fminf(plo[0], phi[0]));
    int py0 = (int)floorf(fminf(plo[1], phi[1]));
    int px1 = (int)ceilf(fmaxf(plo[0], phi[0]));
    int py1 = (int)ceilf(fmaxf(plo[1], phi[1]));
    px0 = (int)fminf(fmaxf((float)px0, 0.0f), (float)(W - 1));
    px1 = (int)fminf(fmaxf((float)px1, 0.0f), (float)(W - 1));
    py0 = (int)fminf(fmaxf((float)py0, 0.0f), (float)(H - 1));
    py1 = (int)fminf(fmaxf((float)py1, 0.0f), (float)(H - 1));

    float* vp = vol + ((size_t)z * (size_t)v.ny + (size_t)y) * (size_t)v.nx + x;
    float4 vb = *reinterpret_cast<float4*>(vp);
    float num[4] = {0, 0, 0, 0}, den[4] = {0, 0, 0, 0};
    const float yLo = v.bbMin[1] + (float)(unsigned)y * v.vs[1], yHi = v.bbMin[1] + (float)(unsigned)(y + 1) * v.vs[1];
    const float zLo = v.bbMin[2] + (float)(unsigned)z * v.vs[2], zHi = v.bbMin[2] + (float)(unsigned)(z + 1) * v.vs[2];
    float xPl[5];
    for (int k = 0; k < 5; k++) xPl[k] = v.bbMin[0] + (float)(x + (unsigned)k) * v.vs[0];
    for (float pyv = (float)py0 + osInv * 0.5f; pyv < (float)py1; pyv += osInv) {
        for (float pxv = (float)px0 + osInv * 0.5f; pxv < (float)px1; pxv += osInv) {
            float xa, ya;
            mag_apply(g.Ainv, pxv, pyv, xa, ya);
            float ray[3], src[3];
            for (int i = 0; i < 3; i++) ray[i] = d[i] + pxv * g.u[i];
            for (int i = 0; i < 3; i++) ray[i] = ray[i] + pyv * g.v[i];
            for (int i = 0; i < 3; i++) {
                src[i] = bp_src(ray[i], n[i]);
                ray[i] = ray[i] - src[i];
            }
            const float inv = 1.0f / sqrtf((ray[0] * ray[0] + ray[1] * ray[1]) + ray[2] * ray[2]);
            for (int i = 0; i < 3; i++) ray[i] = ray[i] * inv;
            const float val = tex2d(img, W, H, pitch, filter, xa, ya);
            const float tyE = (yLo - src[1]) / ray[1], tyX = (yHi - src[1]) / ray[1];
            const float tzE = (zLo - src[2]) / ray[2], tzX = (zHi - src[2]) / ray[2];
            float txE = (xPl[0] - src[0]) / ray[0];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float txX = (xPl[k + 1] - src[0]) / ray[0];
                const float t_in = fmaxf(fmaxf(fminf(txE, txX), fminf(tyE, tyX)), fminf(tzE, tzX));
                const float t_out = fminf(fminf(fmaxf(txE, txX), fmaxf(tyE, tyX)), fmaxf(tzE, tzX));
                if ((t_out - t_in) > 0.0f) {
                    num[k] += val * (t_out - t_in);
                    den[k] += (t_out - t_in);
                }
                txE = txX;
            }
        }
    }
    float out[4] = {vb.x, vb.y, vb.z, vb.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        den[k] = fmaxf(0.0f, den[k]);
        if (den[k] != 0.0f) out[k] += (lambda * num[k]) / den[k];
    }
    *reinterpret_cast<float4*>(vp) = make_float4(out[0], out[1], out[2], out[3]);
}

// ---------------------------------------------------------------- exact numerics

// Per-tilt constants for the exact kernel, prepared on the host in double.
struct BPExact {
    // chord through a voxel for a sub-ray offset (a, b) in detector pitch units from the
    // projection of the voxel's min corner: t_c = a*U[c] + b*V[c]; interval
    // [t_c + Smin[c], t_c + Smax[c]] along the unit ray direction.
    float U[3], V[3], Smin[3], Smax[3];
    float uc[3], vc[3], sc[3];   // axes parallel to the rays: inside test on a*u + b*v
    int parallel;                // bit c set when n_c == 0
    // projection of the voxel (i,j,k) min corner: P = p0 + i*pi + j*pj + k*pk
    double pu0, pui, puj, puk, pv0, pvi, pvj, pvk;
    // projected box of one voxel relative to its corner projection (widened by a guard)
    double uMin, uMax, vMin, vMax;
    int identityAniso;
};

BPExact prepare_exact(const VolG& v, const TiltG& g) {
    BPExact e{};
    const double u[3] = {g.u[0], g.u[1], g.u[2]}, w[3] = {g.v[0], g.v[1], g.v[2]};
    const double nl = std::sqrt((double)g.n[0] * g.n[0] + (double)g.n[1] * g.n[1] + (double)g.n[2] * g.n[2]);
    const double n[3] = {g.n[0] / nl, g.n[1] / nl, g.n[2] / nl};
    e.parallel = 0;
    for (int c = 0; c < 3; c++) {
        const double s = (double)v.vs[c];
        e.uc[c] = (float)u[c];
        e.vc[c] = (float)w[c];
        e.sc[c] = (float)s;
        if (n[c] == 0.0) {
            e.parallel |= 1 << c;
            e.U[c] = 0.0f;
            e.V[c] = 0.0f;
            e.Smin[c] = -INFINITY;
            e.Smax[c] = INFINITY;
        } else {
            e.U[c] = (float)(-u[c] / n[c]);
            e.V[c] = (float)(-w[c] / n[c]);
            const double S = s / n[c];
            e.Smin[c] = (float)std::fmin(0.0, S);
            e.Smax[c] = (float)std::fmax(0.0, S);
        }
    }
    // inverse of the column basis [u v n]: p - d = alpha*u + beta*v + gamma*n
    const double B[9] = {u[0], w[0], (double)g.n[0], u[1], w[1], (double)g.n[1], u[2], w[2], (double)g.n[2]};
    const double det = B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                       B[2] * (B[3] * B[7] - B[4] * B[6]);
    const double r0[3] = {(B[4] * B[8] - B[5] * B[7]) / det, -(B[1] * B[8] - B[2] * B[7]) / det,
                          (B[1] * B[5] - B[2] * B[4]) / det};
    const double r1[3] = {-(B[3] * B[8] - B[5] * B[6]) / det, (B[0] * B[8] - B[2] * B[6]) / det,
                          -(B[0] * B[5] - B[2] * B[3]) / det};
    const double o[3] = {(double)v.bbMin[0] - g.d[0], (double)v.bbMin[1] - g.d[1], (double)v.bbMin[2] - g.d[2]};
    e.pu0 = r0[0] * o[0] + r0[1] * o[1] + r0[2] * o[2];
    e.pv0 = r1[0] * o[0] + r1[1] * o[1] + r1[2] * o[2];
    e.pui = r0[0] * v.vs[0];
    e.puj = r0[1] * v.vs[1];
    e.puk = r0[2] * v.vs[2];
    e.pvi = r1[0] * v.vs[0];
    e.pvj = r1[1] * v.vs[1];
    e.pvk = r1[2] * v.vs[2];
    const double guard = 1e-6;
    e.uMin = std::fmin(0.0, e.pui) + std::fmin(0.0, e.puj) + std::fmin(0.0, e.puk) - guard;
    e.uMax = std::fmax(0.0, e.pui) + std::fmax(0.0, e.puj) + std::fmax(0.0, e.puk) + guard;
    e.vMin = std::fmin(0.0, e.pvi) + std::fmin(0.0, e.pvj) + std::fmin(0.0, e.pvk) - guard;
    e.vMax = std::fmax(0.0, e.pvi) + std::fmax(0.0, e.pvj) + std::fmax(0.0, e.pvk) + guard;
    const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    e.identityAniso = 1;
    for (int i = 0; i < 9; i++)
        if (g.Ainv[i] != I[i]) e.identityAniso = 0;
    return e;
}

// Footprint of a quad's macro-cell in double with the reference's operation order; returns
// false when the gate [distMin, distMax) rejects the quad.
__device__ __forceinline__ bool footprint_d(const VolG& v, const TiltG& g, unsigned x, unsigned y, unsigned z, int W,
                                            int H, double distMin, double distMax, int& px0, int& px1, int& py0,
                                            int& py1) {
    const double n[3] = {g.n[0], g.n[1], g.n[2]};
    const double d[3] = {g.d[0], g.d[1], g.d[2]};
    const double vs[3] = {v.vs[0], v.vs[1], v.vs[2]};
    const double bb[3] = {v.bbMin[0], v.bbMin[1], v.bbMin[2]};
    const double mcMin[3] = {bb[0] + (double)x * vs[0], bb[1] + (double)y * vs[1], bb[2] + (double)z * vs[2]};
    const double mcMax[3] = {bb[0] + (double)(x + 5u) * vs[0], bb[1] + (double)(y + 1u) * vs[1],
                             bb[2] + (double)(z + 1u) * vs[2]};
    const double nd = ((-n[0]) * d[0] - n[1] * d[1]) - n[2] * d[2];
    double bMin[3] = {3.4028234663852886e+38, 3.4028234663852886e+38, 3.4028234663852886e+38};
    double bMax[3] = {-200000.0, -200000.0, -200000.0};
#pragma unroll
    for (int corner = 0; corner < 8; corner++) {
        const int cx = corner >> 2, cy = (corner >> 1) & 1, cz = corner & 1;
        const double c[3] = {cx ? mcMin[0] + (mcMax[0] - mcMin[0]) : mcMin[0],
                             cy ? mcMin[1] + (mcMax[1] - mcMin[1]) : mcMin[1],
                             cz ? mcMin[2] + (mcMax[2] - mcMin[2]) : mcMin[2]};
        double t = (n[0] * c[0] + n[1] * c[1]) + n[2] * c[2];
        t += nd;
        t = fabs(t);
        if (corner == 0 && !(t >= distMin && t < distMax)) return false;
        for (int i = 0; i < 3; i++) {
            const double h = t * (-n[i]) + c[i];
            bMin[i] = fmin(h, bMin[i]);
            bMax[i] = fmax(h, bMax[i]);
        }
    }
    double lo[3], hi[3];
    for (int i = 0; i < 3; i++) {
        lo[i] = fmin(bMin[i], bMax[i]);
        hi[i] = fmax(bMin[i], bMax[i]);
    }
    if (n[0] <= 0.0) {
        const double tmp = lo[0];
        lo[0] = hi[0];
        hi[0] = tmp;
    }
    double plo[2], phi[2];
    for (int r = 0; r < 2; r++) {
        const float* m = g.M + 4 * r;
        plo[r] = (((double)m[0] * lo[0] + (double)m[1] * lo[1]) + (double)m[2] * lo[2]) + 1.0 * (double)m[3];
        phi[r] = (((double)m[0] * hi[0] + (double)m[1] * hi[1]) + (double)m[2] * hi[2]) + 1.0 * (double)m[3];
    }
    px0 = (int)fmin(fmax(floor(fmin(plo[0], phi[0])), 0.0), (double)(W - 1));
    py0 = (int)fmin(fmax(floor(fmin(plo[1], phi[1])), 0.0), (double)(H - 1));
    px1 = (int)fmin(fmax(ceil(fmax(plo[0], phi[0])), 0.0), (double)(W - 1));
    py1 = (int)fmin(fmax(ceil(fmax(plo[1], phi[1])), 0.0), (double)(H - 1));
    return true;
}

__global__ void __launch_bounds__(256) bp_exact_kernel(VolG v, TiltG g, BPExact e, float* __restrict__ vol,
                                                       const float* __restrict__ img, int W, int H, size_t pitch,
                                                       float lambda, int os, double distMin, double distMax,
                                                       int filter) {
    const int xq = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int z = (int)blockIdx.z;
    if (xq >= v.nx / 4 || y >= v.ny || z >= v.nz) return;
    const unsigned x = (unsigned)xq * 4u;
    int px0, px1, py0, py1;
    if (!footprint_d(v, g, x, (unsigned)y, (unsigned)z, W, H, distMin, distMax, px0, px1, py0, py1)) return;
    const int nI = (px1 - px0) * os, nJ = (py1 - py0) * os;
    float* vp = vol + ((size_t)z * (size_t)v.ny + (size_t)y) * (size_t)v.nx + x;
    float4 vb = *reinterpret_cast<float4*>(vp);
    float out[4] = {vb.x, vb.y, vb.z, vb.w};
    if (nI <= 0 || nJ <= 0) return;
    const double delta = 1.0 / (double)os;
    const float deltaF = (float)delta;
    const double puBase = e.pu0 + (double)y * e.puj + (double)z * e.puk;
    const double pvBase = e.pv0 + (double)y * e.pvj + (double)z * e.pvk;
#pragma unroll 1
    for (int k = 0; k < 4; k++) {
        const double pu = puBase + (double)(x + (unsigned)k) * e.pui;
        const double pv = pvBase + (double)(x + (unsigned)k) * e.pvi;
        // lattice range inside the voxel's projected box, within the quad footprint
        const int i0 = max(0, (int)ceil((pu + e.uMin - (double)px0) * (double)os - 0.5));
        const int i1 = min(nI - 1, (int)floor((pu + e.uMax - (double)px0) * (double)os - 0.5));
        const int j0 = max(0, (int)ceil((pv + e.vMin - (double)py0) * (double)os - 0.5));
        const int j1 = min(nJ - 1, (int)floor((pv + e.vMax - (double)py0) * (double)os - 0.5));
        float num = 0.0f, den = 0.0f;
        for (int j = j0; j <= j1; j++) {
            const double pyd = (double)py0 + ((double)j + 0.5) * delta;
            const float b = (float)(pyd - pv);
            const float bU[3] = {b * e.V[0], b * e.V[1], b * e.V[2]};
            const float a0 = (float)((double)px0 + ((double)i0 + 0.5) * delta - pu);
            for (int i = i0; i <= i1; i++) {
                const float a = a0 + (float)(i - i0) * deltaF;
                const float t0 = __builtin_fmaf(a, e.U[0], bU[0]);
                const float t1 = __builtin_fmaf(a, e.U[1], bU[1]);
                const float t2 = __builtin_fmaf(a, e.U[2], bU[2]);
                const float tin = fmaxf(fmaxf(t0 + e.Smin[0], t1 + e.Smin[1]), t2 + e.Smin[2]);
                const float tout = fminf(fminf(t0 + e.Smax[0], t1 + e.Smax[1]), t2 + e.Smax[2]);
                float chord = tout - tin;
                if (e.parallel) {
                    for (int c = 0; c < 3; c++) {
                        if (!((e.parallel >> c) & 1)) continue;
                        const float wc = __builtin_fmaf(a, e.uc[c], b * e.vc[c]);
                        if (!(wc > 0.0f && wc < e.sc[c])) chord = 0.0f;
                    }
                }
                if (chord > 0.0f) {
                    const float pxf = (float)((double)px0 + ((double)i + 0.5) * delta);
                    const float pyf = (float)pyd;
                    float val;
                    if (filter == ARTIA_FILTER_POINT && e.identityAniso) {
                        val = img[(size_t)(int)floorf(pyf) * pitch + (size_t)(int)floorf(pxf)];
                    } else {
                        float xa, ya;
                        mag_apply(g.Ainv, pxf, pyf, xa, ya);
                        val = tex2d(img, W, H, pitch, filter, xa, ya);
                    }
                    num = __builtin_fmaf(val, chord, num);
                    den += chord;
                }
            }
        }
        den = fmaxf(0.0f, den);
        if (den != 0.0f) out[k] += (lambda * num) / den;
    }
    *reinterpret_cast<float4*>(vp) = make_float4(out[0], out[1], out[2], out[3]);
}

}  // namespace

void launch_bp(artia_ctx* ctx, const VolG& v, const TiltG& g, float* vol, const float* img, int W, int H, size_t pitch,
               float lambda, int os, float distMin, float distMax, int filter) {
    const int nq = v.nx / 4;
    const dim3 block(64, 4), grid((unsigned)((nq + 63) / 64), (unsigned)((v.ny + 3) / 4), (unsigned)v.nz);
    if (ctx->numerics == ARTIA_NUMERICS_REFERENCE) {
        hipLaunchKernelGGL(bp_ref_kernel, grid, block, 0, ctx->stream, v, g, vol, img, W, H, pitch, lambda, os,
                           1.0f / (float)os, distMin, distMax, filter);
    } else {
        const BPExact e = prepare_exact(v, g);
        hipLaunchKernelGGL(bp_exact_kernel, grid, block, 0, ctx->stream, v, g, e, vol, img, W, H, pitch, lambda, os,
                           (double)distMin, (double)distMax, filter);
    }
}

}  // namespace artia

using namespace artia;

extern "C" {

artia_status artia_bp(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, float* vol,
                      const float* img, int width, int height, size_t pitch, float lambda, int oversampling,
                      float dist_min, float dist_max, int filter, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!valid_geometry(vg, tg) || !vol || !img)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_bp: null pointer or invalid geometry");
    if (vg->nx % 4 != 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_bp: volume x dimension %d is not a multiple of 4", vg->nx);
    if ((reinterpret_cast<uintptr_t>(vol) & 15) != 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_bp: volume pointer must be 16-byte aligned");
    if (width <= 0 || height <= 0 || pitch < (size_t)width || oversampling <= 0 || oversampling > 64)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_bp: bad image size or oversampling");
    if (filter != ARTIA_FILTER_POINT && filter != ARTIA_FILTER_LINEAR)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_bp: unknown filter %d", filter);
    Timing t(ctx, ms);
    launch_bp(ctx, to_device(*vg), to_device(*tg), vol, img, width, height, pitch, lambda, oversampling, dist_min,
              dist_max, filter);
    return t.finish();
}

}  // extern "C"
