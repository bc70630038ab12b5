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
#include <climits>
#include <cmath>

#include "bp_common.hpp"
#include "bp_window.hpp"
#include "device_common.hpp"

namespace artia {
namespace {

// A quad's four voxels: one 16-byte access of fp32 voxels, one 8-byte access of half voxels (x is
// a multiple of 4).  Half voxels are read exactly and written rounded to nearest even, as the
// reference's backProjectionFP16 does (BackProjectionSquareOS.cu:785-1290: the FP32 kernel's
// arithmetic between __half2float reads and __float2half_rn writes).
typedef vhalf vhalf4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_quad(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ float4 load_quad(const vhalf* p) {
    const vhalf4 h = *reinterpret_cast<const vhalf4*>(p);
    return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
}
__device__ __forceinline__ void store_quad(float* p, const float o[4]) {
    *reinterpret_cast<float4*>(p) = make_float4(o[0], o[1], o[2], o[3]);
}
__device__ __forceinline__ void store_quad(vhalf* p, const float o[4]) {
    vhalf4 h;
    h.x = (vhalf)o[0];
    h.y = (vhalf)o[1];
    h.z = (vhalf)o[2];
    h.w = (vhalf)o[3];
    *reinterpret_cast<vhalf4*>(p) = h;
}

// ---------------------------------------------------------------- reference numerics

__device__ __forceinline__ float bp_src(float r, float n) { return (float)((double)r + 100000.0 * (double)n); }

template <typename T>
__global__ void __launch_bounds__(256) bp_ref_kernel(VolG v, TiltG g, T* __restrict__ vol,
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

    T* vp = vol + ((size_t)z * (size_t)v.ny + (size_t)y) * (size_t)v.nx + x;
    float4 vb = load_quad(vp);
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
    store_quad(vp, out);
}

// ---------------------------------------------------------------- exact numerics

// One thread per quad; a 256-thread block covers 16 quads (64 voxels) along x, 4 rows in y and
// 4 slices in z.  kOS >= 0 selects the LDS-tile path (POINT sampling, no magnification
// anisotropy, affine footprint): the block first stages the projection pixels its quads can
// reach.  kOS < 0 is the global-memory path.
template <int kOS, int kPM, typename T>
__global__ void __launch_bounds__(256) bp_exact_kernel(VolG v, TiltG g, BPExact e, T* __restrict__ vol,
                                                       const float* __restrict__ img, int W, int H, size_t pitch,
                                                       float lambda, int os, double distMin, double distMax,
                                                       int filter) {
    constexpr bool kTile = kOS >= 0;
    __shared__ float tile[kTile ? kTileFloats : 1];
    const int tid = (int)threadIdx.x;
    const int xq = (int)blockIdx.x * 16 + (tid & 15);
    const int y = (int)blockIdx.y * 4 + ((tid >> 4) & 3);
    const int z = (int)blockIdx.z * 4 + (tid >> 6);
    const bool inVol = xq < v.nx / 4 && y < v.ny && z < v.nz;
    const unsigned x = (unsigned)xq * 4u;

    int tx0 = 0, ty0 = 0, tw = 0;
    if (kTile) {
        // block pixel range: the union of its quads' footprints (affine, so corners suffice)
        const double bx = (double)blockIdx.x * 64.0, by = (double)blockIdx.y * 4.0, bz = (double)blockIdx.z * 4.0;
        const double base0 = e.G[0][0] * bx + e.G[0][1] * by + e.G[0][2] * bz;
        const double base1 = e.G[1][0] * bx + e.G[1][1] * by + e.G[1][2] * bz;
        tx0 = max(0, (int)floor(base0 + e.tlo[0]) - 1);
        ty0 = max(0, (int)floor(base1 + e.tlo[1]) - 1);
        const int tx1 = min(W - 1, (int)ceil(base0 + e.thi[0]) + 1);
        const int ty1 = min(H - 1, (int)ceil(base1 + e.thi[1]) + 1);
        tw = max(0, tx1 - tx0 + 1);
        const int th = max(0, ty1 - ty0 + 1);
        for (int idx = tid; idx < tw * th; idx += 256) {
            const int r = idx / tw, c = idx - r * tw;
            tile[idx] = img[(size_t)(ty0 + r) * pitch + (size_t)(tx0 + c)];
        }
        __syncthreads();
    }
    if (!inVol) return;
    int px0, px1, py0, py1;
    if (e.affine) {
        if (!footprint_affine(e, x, (unsigned)y, (unsigned)z, W, H, distMin, distMax, px0, px1, py0, py1)) return;
    } else {
        if (!footprint_literal(v, g, x, (unsigned)y, (unsigned)z, W, H, distMin, distMax, px0, px1, py0, py1)) return;
    }
    const int O = kOS > 0 ? kOS : os;
    const int nI = (px1 - px0) * O, nJ = (py1 - py0) * O;
    if (nI <= 0 || nJ <= 0) return;
    T* vp = vol + ((size_t)z * (size_t)v.ny + (size_t)y) * (size_t)v.nx + x;
    const float4 vb = load_quad(vp);
    float out[4] = {vb.x, vb.y, vb.z, vb.w};
    const float osf = (float)O, delta = 1.0f / osf;
    // voxel corner projections relative to the quad footprint origin (small: exact in float to ~1e-6)
    const float du0 = (float)(e.pu0 + (double)x * e.pui + (double)y * e.puj + (double)z * e.puk - (double)px0);
    const float dv0 = (float)(e.pv0 + (double)x * e.pvi + (double)y * e.pvj + (double)z * e.pvk - (double)py0);
    const float dui = (float)e.pui, dvi = (float)e.pvi;
    const int tileOff = kTile ? (py0 - ty0) * tw + (px0 - tx0) : 0;
    bool any = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float du = __builtin_fmaf((float)k, dui, du0);
        const float dv = __builtin_fmaf((float)k, dvi, dv0);
        const int i0 = max(0, (int)ceilf((du + e.uMin) * osf - 0.5f));
        const int i1 = min(nI - 1, (int)floorf((du + e.uMax) * osf - 0.5f));
        const int j0 = max(0, (int)ceilf((dv + e.vMin) * osf - 0.5f));
        const int j1 = min(nJ - 1, (int)floorf((dv + e.vMax) * osf - 0.5f));
        float num = 0.0f, den = 0.0f;
        if (kTile)
            voxel_sums_tile<(kOS > 0 ? kOS : 0), kPM>(e, tile, tw, tileOff, O, delta, du, dv, i0, i1, j0, j1, num,
                                                      den);
        else
            voxel_sums_global(e, g, img, W, H, pitch, filter, px0, py0, O, delta, du, dv, i0, i1, j0, j1, num, den);
        den = fmaxf(0.0f, den);
        if (den != 0.0f) {
            out[k] += (lambda * num) / den;
            any = true;
        }
    }
    if (any) store_quad(vp, out);
}

// ---------------------------------------------------------------- exact numerics, window form

// Up to kWinTilts tilts of an ordered subset in one launch (bp_window.hpp): every quad adds the
// tilts' updates to its voxels in tilt order in registers and writes them once, which is what the
// same number of single-tilt launches write (each launch reads the previous one's result).
constexpr int kWinTilts = 4;
struct BPWinSet {
    BPWin w[kWinTilts];
    const float* img[kWinTilts];
    int n;
    float lambda;
    double distMin, distMax;
    int gated;   // 0: the dist gate passes every quad of the volume for every tilt (gate_passes_all)
};

// Block = 16 quads x 4 rows x 8 slices: each thread updates two quads, slices z and z + 4, so the
// block's per-tilt set-up (footprint, tile staging) serves twice the voxels.  Per tilt the block
// stages its footprint (one pixel of margin) in LDS.  Blocks whose footprint lies inside the image
// apply the quad footprint as a chord cap (or not at all where it covers the voxel boxes); blocks at
// the image edge clamp each quad's lattice to its clamped footprint.
template <int kOS, int kNI, int kNJ, typename T, bool kYCol>
__global__ void __launch_bounds__(256) bp_win_kernel(VolG v, BPWinSet S, T* __restrict__ vol, int W, int H,
                                                     size_t pitch) {
    __shared__ float tile[kTileFloats];
    const int tid = (int)threadIdx.x;
    const int xq = (int)blockIdx.x * 16 + (tid & 15);
    const int y = (int)blockIdx.y * 4 + ((tid >> 4) & 3);
    const int zq[2] = {(int)blockIdx.z * kWinBlockZ + (tid >> 6), (int)blockIdx.z * kWinBlockZ + (tid >> 6) + 4};
    const unsigned x = (unsigned)xq * 4u;
    bool inVol[2];
    T* vp[2];
    float out[2][4];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        inVol[q] = xq < v.nx / 4 && y < v.ny && zq[q] < v.nz;
        vp[q] = vol + ((size_t)(inVol[q] ? zq[q] : 0) * (size_t)v.ny + (size_t)(inVol[q] ? y : 0)) * (size_t)v.nx +
                (inVol[q] ? x : 0u);
        out[q][0] = out[q][1] = out[q][2] = out[q][3] = 0.0f;
        if (inVol[q]) {
            const float4 vb = load_quad(vp[q]);
            out[q][0] = vb.x;
            out[q][1] = vb.y;
            out[q][2] = vb.z;
            out[q][3] = vb.w;
        }
    }
    bool any[2] = {false, false};
    const double bx = (double)blockIdx.x * 64.0, by = (double)blockIdx.y * 4.0,
                 bz = (double)blockIdx.z * (double)kWinBlockZ;
    for (int t = 0; t < S.n; t++) {
        const BPWin& w = S.w[t];
        const float* __restrict__ img = S.img[t];
        const double base0 = w.G[0][0] * bx + w.G[0][1] * by + w.G[0][2] * bz;
        const double base1 = w.G[1][0] * bx + w.G[1][1] * by + w.G[1][2] * bz;
        const int ux0 = (int)floor(base0 + w.tlo[0]) - 1, ux1 = (int)ceil(base0 + w.thi[0]) + 1;
        const int uy0 = (int)floor(base1 + w.tlo[1]) - 1, uy1 = (int)ceil(base1 + w.thi[1]) + 1;
        const bool interior = ux0 >= 0 && uy0 >= 0 && ux1 <= W - 1 && uy1 <= H - 1;
        const int tx0 = max(0, ux0), ty0 = max(0, uy0);
        const int tw = max(0, min(W - 1, ux1) - tx0 + 1), th = max(0, min(H - 1, uy1) - ty0 + 1);
        // LDS row stride: odd where it fits, so that the four voxel rows of a wave, which read four
        // consecutive tile rows at the same columns, fall in different banks
        const int ts = (tw | 1) * th <= kTileFloats ? (tw | 1) : tw;
        if (t > 0) __syncthreads();   // the previous tilt's readers are done
        // tile entry idx = r * tw + c, strided by 256 over the block: one division per thread,
        // then the row and column are stepped (tiles are often a few columns wide and tall)
        if (tid < tw * th) {
            int r = tid / tw, c = tid - r * tw;
            const int dr = 256 / tw, dc = 256 - dr * tw;
            for (int idx = tid; idx < tw * th; idx += 256) {
                tile[r * ts + c] = img[(size_t)(ty0 + r) * pitch + (size_t)(tx0 + c)];
                r += dr;
                c += dc;
                if (c >= tw) {
                    c -= tw;
                    r++;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (!inVol[q]) continue;
            const double qx = (double)x, qy = (double)y, qz = (double)zq[q];
            if (S.gated) {
                const double gt = fabs(fma(w.gn[0], qx, fma(w.gn[1], qy, fma(w.gn[2], qz, w.g0))));
                if (!(gt >= S.distMin && gt < S.distMax)) continue;
            }
            any[q] |= window_quad<kOS, kNI, kNJ, kYCol>(w, tile, ts, tx0, ty0, interior, W, H, x, (unsigned)y,
                                                 (unsigned)zq[q], S.lambda, out[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < 2; q++)
        if (any[q]) store_quad(vp[q], out[q]);
}

// The CTF backprojection of one tilt in one launch: every voxel takes its slab's filtered image
// from the stack (slab k at stack + k * stride) instead of one gated launch per slab
// (Reconstructor.cpp:920-969).  A voxel lies in exactly one slab, so it receives the same single
// update, in the same arithmetic, as in the per-slab launches.  For LINEAR sampling without
// magnification anisotropy (affine footprint) the block stages its footprint of every slab its
// quads fall in, when that fits the LDS tile, and fetches from it with tex2d's arithmetic.  Eight
// waves per SIMD (61 registers, no spills) measured 5 % faster than the unbounded seven.
template <int kPM, typename T>
__global__ void __launch_bounds__(256, 8) bp_stack_kernel(VolG v, TiltG g, BPExact e, T* __restrict__ vol,
                                                       const float* __restrict__ stack, size_t stride, int W, int H,
                                                       size_t pitch, float lambda, int os, int filter, SlabArgs sl) {
    __shared__ float tile[kTileFloats];
    __shared__ int kr[2];
    const int tid = (int)threadIdx.x;
    const int xq = (int)blockIdx.x * 16 + (tid & 15);
    const int y = (int)blockIdx.y * 4 + ((tid >> 4) & 3);
    const int z = (int)blockIdx.z * 4 + (tid >> 6);
    const bool inVol = xq < v.nx / 4 && y < v.ny && z < v.nz;
    const unsigned x = (unsigned)xq * 4u;
    if (tid == 0) {
        kr[0] = INT_MAX;
        kr[1] = INT_MIN;
    }
    __syncthreads();
    const int slab = inVol ? find_slab(sl, gate_distance(e, v, g, x, (unsigned)y, (unsigned)z)) : -1;
    if (slab >= 0) {
        atomicMin(&kr[0], slab);
        atomicMax(&kr[1], slab);
    }
    __syncthreads();
    const int k0 = kr[0], k1 = kr[1];
    bool tiled = false;
    int tx0 = 0, ty0 = 0, tx1 = -1, ty1 = -1, tw = 0, plane = 0;
    if (e.affine && e.identityAniso && filter == ARTIA_FILTER_LINEAR && k0 <= k1) {
        // the block's footprint (affine, so corners suffice) with two pixels of margin: the
        // bilinear fetches reach one pixel below a quad's footprint, the other covers rounding
        const double bx = (double)blockIdx.x * 64.0, by = (double)blockIdx.y * 4.0, bz = (double)blockIdx.z * 4.0;
        const double base0 = e.G[0][0] * bx + e.G[0][1] * by + e.G[0][2] * bz;
        const double base1 = e.G[1][0] * bx + e.G[1][1] * by + e.G[1][2] * bz;
        tx0 = max(0, (int)floor(base0 + e.tlo[0]) - 2);
        ty0 = max(0, (int)floor(base1 + e.tlo[1]) - 2);
        tx1 = min(W - 1, (int)ceil(base0 + e.thi[0]) + 2);
        ty1 = min(H - 1, (int)ceil(base1 + e.thi[1]) + 2);
        tw = max(0, tx1 - tx0 + 1);
        plane = tw * max(0, ty1 - ty0 + 1);
        tiled = plane > 0 && (long long)plane * (long long)(k1 - k0 + 1) <= (long long)kTileFloats;
        if (tiled) {
            const int n = plane * (k1 - k0 + 1);
            for (int idx = tid; idx < n; idx += 256) {
                const int s = idx / plane, rem = idx - s * plane;
                const int r = rem / tw, c = rem - r * tw;
                tile[idx] = stack[(size_t)(k0 + s) * stride + (size_t)(ty0 + r) * pitch + (size_t)(tx0 + c)];
            }
            __syncthreads();
        }
    }
    if (slab < 0) return;
    const float* __restrict__ img = stack + (size_t)slab * stride;
    int px0, px1, py0, py1;
    if (e.affine) {
        if (!footprint_affine(e, x, (unsigned)y, (unsigned)z, W, H, -1.0, INFINITY, px0, px1, py0, py1)) return;
    } else {
        if (!footprint_literal(v, g, x, (unsigned)y, (unsigned)z, W, H, -1.0, INFINITY, px0, px1, py0, py1)) return;
    }
    const int nI = (px1 - px0) * os, nJ = (py1 - py0) * os;
    if (nI <= 0 || nJ <= 0) return;
    // the quad's fetches read pixels max(p0 - 1, 0) .. p1 on each axis
    const bool fromTile = tiled && max(px0 - 1, 0) >= tx0 && px1 <= tx1 && max(py0 - 1, 0) >= ty0 && py1 <= ty1;
    const bool inner = px0 >= 1 && px1 <= W - 1 && py0 >= 1 && py1 <= H - 1;   // no fetch clamps
    const float* __restrict__ slabTile = tile + (fromTile ? (slab - k0) * plane : 0);
    T* vp = vol + ((size_t)z * (size_t)v.ny + (size_t)y) * (size_t)v.nx + x;
    const float4 vb = load_quad(vp);
    float out[4] = {vb.x, vb.y, vb.z, vb.w};
    const float osf = (float)os, delta = 1.0f / osf;
    const float du0 = (float)(e.pu0 + (double)x * e.pui + (double)y * e.puj + (double)z * e.puk - (double)px0);
    const float dv0 = (float)(e.pv0 + (double)x * e.pvi + (double)y * e.pvj + (double)z * e.pvk - (double)py0);
    const float dui = (float)e.pui, dvi = (float)e.pvi;
    bool any = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float du = __builtin_fmaf((float)k, dui, du0);
        const float dv = __builtin_fmaf((float)k, dvi, dv0);
        const int i0 = max(0, (int)ceilf((du + e.uMin) * osf - 0.5f));
        const int i1 = min(nI - 1, (int)floorf((du + e.uMax) * osf - 0.5f));
        const int j0 = max(0, (int)ceilf((dv + e.vMin) * osf - 0.5f));
        const int j1 = min(nJ - 1, (int)floorf((dv + e.vMax) * osf - 0.5f));
        float num = 0.0f, den = 0.0f;
        if (fromTile && inner)
            voxel_sums_linear_tile<kPM, true>(e, slabTile, tw, tx0, ty0, W, H, px0, py0, delta, du, dv, i0, i1, j0,
                                              j1, num, den);
        else if (fromTile)
            voxel_sums_linear_tile<kPM, false>(e, slabTile, tw, tx0, ty0, W, H, px0, py0, delta, du, dv, i0, i1, j0,
                                               j1, num, den);
        else
            voxel_sums_global(e, g, img, W, H, pitch, filter, px0, py0, os, delta, du, dv, i0, i1, j0, j1, num, den);
        den = fmaxf(0.0f, den);
        if (den != 0.0f) {
            out[k] += (lambda * num) / den;
            any = true;
        }
    }
    if (any) store_quad(vp, out);
}

// bp_stack_kernel for LINEAR sampling at OS 1 or 2 with an affine footprint and no magnification
// anisotropy, blocks of 4 quads x 4 rows x 16 slices: thin along x, so that a block spans few slabs
// at every tilt.  For each slab its quads fall in, the block stages the LINEAR fetches at the
// lattice points of its footprint (tile_linear over the slab image: the values the per-box loop
// fetches), so a lattice point costs one LDS read instead of a bilinear fetch.  e holds the block
// offsets of this shape.  Quads whose footprint leaves the staged lattice fetch from global memory.
template <int kOS, int kPM, typename T>
__global__ void __launch_bounds__(256, 8) bp_stack_lat_kernel(VolG v, TiltG g, BPExact e, T* __restrict__ vol,
                                                           const float* __restrict__ stack, size_t stride, int W,
                                                           int H, size_t pitch, float lambda, SlabArgs sl) {
    __shared__ float tile[kTileFloats];
    __shared__ int kr[2];
    constexpr float kDelta = 1.0f / (float)kOS;
    const int tid = (int)threadIdx.x;
    const int xq = (int)blockIdx.x * 4 + (tid & 3);
    const int y = (int)blockIdx.y * 4 + ((tid >> 2) & 3);
    const int z = (int)blockIdx.z * 16 + (tid >> 4);
    const bool inVol = xq < v.nx / 4 && y < v.ny && z < v.nz;
    const unsigned x = (unsigned)xq * 4u;
    if (tid == 0) {
        kr[0] = INT_MAX;
        kr[1] = INT_MIN;
    }
    __syncthreads();
    const int slab = inVol ? find_slab(sl, gate_distance(e, v, g, x, (unsigned)y, (unsigned)z)) : -1;
    if (slab >= 0) {
        atomicMin(&kr[0], slab);
        atomicMax(&kr[1], slab);
    }
    __syncthreads();
    const int k0 = kr[0], k1 = kr[1];
    bool tiled = false;
    int tx0 = 0, ty0 = 0, tx1 = -1, ty1 = -1, ls = 0, planeL = 0;
    if (k0 <= k1) {
        // the block's footprint (one pixel of margin for rounding): its lattice points
        // [tx0 * OS, tx1 * OS) x [ty0 * OS, ty1 * OS) hold every quad footprint's lattice
        const double bx = (double)blockIdx.x * 16.0, by = (double)blockIdx.y * 4.0, bz = (double)blockIdx.z * 16.0;
        const double base0 = e.G[0][0] * bx + e.G[0][1] * by + e.G[0][2] * bz;
        const double base1 = e.G[1][0] * bx + e.G[1][1] * by + e.G[1][2] * bz;
        tx0 = max(0, (int)floor(base0 + e.tlo[0]) - 1);
        ty0 = max(0, (int)floor(base1 + e.tlo[1]) - 1);
        tx1 = min(W - 1, (int)ceil(base0 + e.thi[0]) + 1);
        ty1 = min(H - 1, (int)ceil(base1 + e.thi[1]) + 1);
        const int lw = max(0, tx1 - tx0) * kOS, lh = max(0, ty1 - ty0) * kOS;
        ls = lw | 1;   // odd LDS row stride
        planeL = ls * lh;
        tiled = lw > 0 && lh > 0 && (long long)planeL * (long long)(k1 - k0 + 1) <= (long long)kTileFloats;
        if (tiled) {
            // lattice point (c, r) at pixel position tx0 + (c + 0.5) / OS, ty0 + (r + 0.5) / OS (exact
            // in float, the per-box loop's px0 + (i + 0.5) / OS of the same point)
            const int per = lw * lh, n = per * (k1 - k0 + 1);
            for (int idx = tid; idx < n; idx += 256) {
                const int s = idx / per, rem = idx - s * per;
                const int r = rem / lw, c = rem - r * lw;
                const float px = (float)tx0 + ((float)c + 0.5f) * kDelta, py = (float)ty0 + ((float)r + 0.5f) * kDelta;
                tile[s * planeL + r * ls + c] =
                    tile_linear(stack + (size_t)(k0 + s) * stride, (int)pitch, 0, 0, W, H, px, py);
            }
        }
    }
    __syncthreads();
    if (slab < 0) return;
    int px0, px1, py0, py1;
    if (!footprint_affine(e, x, (unsigned)y, (unsigned)z, W, H, -1.0, INFINITY, px0, px1, py0, py1)) return;
    const int nI = (px1 - px0) * kOS, nJ = (py1 - py0) * kOS;
    if (nI <= 0 || nJ <= 0) return;
    const bool fromLat = tiled && px0 >= tx0 && px1 <= tx1 && py0 >= ty0 && py1 <= ty1;
    const float* __restrict__ lat = tile + (fromLat ? (slab - k0) * planeL : 0);
    const int oi = (px0 - tx0) * kOS, oj = (py0 - ty0) * kOS;
    const float* __restrict__ img = stack + (size_t)slab * stride;
    T* vp = vol + ((size_t)z * (size_t)v.ny + (size_t)y) * (size_t)v.nx + x;
    const float4 vb = load_quad(vp);
    float out[4] = {vb.x, vb.y, vb.z, vb.w};
    const float osf = (float)kOS;
    const float du0 = (float)(e.pu0 + (double)x * e.pui + (double)y * e.puj + (double)z * e.puk - (double)px0);
    const float dv0 = (float)(e.pv0 + (double)x * e.pvi + (double)y * e.pvj + (double)z * e.pvk - (double)py0);
    const float dui = (float)e.pui, dvi = (float)e.pvi;
    bool any = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float du = __builtin_fmaf((float)k, dui, du0);
        const float dv = __builtin_fmaf((float)k, dvi, dv0);
        const int i0 = max(0, (int)ceilf((du + e.uMin) * osf - 0.5f));
        const int i1 = min(nI - 1, (int)floorf((du + e.uMax) * osf - 0.5f));
        const int j0 = max(0, (int)ceilf((dv + e.vMin) * osf - 0.5f));
        const int j1 = min(nJ - 1, (int)floorf((dv + e.vMax) * osf - 0.5f));
        float num = 0.0f, den = 0.0f;
        if (fromLat)
            voxel_sums_lattice<kPM>(e, lat, ls, oi, oj, kDelta, du, dv, i0, i1, j0, j1, num, den);
        else
            voxel_sums_global(e, g, img, W, H, pitch, ARTIA_FILTER_LINEAR, px0, py0, kOS, kDelta, du, dv, i0, i1,
                              j0, j1, num, den);
        den = fmaxf(0.0f, den);
        if (den != 0.0f) {
            out[k] += (lambda * num) / den;
            any = true;
        }
    }
    if (any) store_quad(vp, out);
}

template <int kOS, typename T>
void launch_stack_lat_os(artia_ctx* ctx, dim3 grid, const VolG& v, const TiltG& g, const BPExact& e, T* vol,
                         const float* stack, size_t stride, int W, int H, size_t pitch, float lambda,
                         const SlabArgs& sl) {
    if (e.parallel == 0)
        hipLaunchKernelGGL((bp_stack_lat_kernel<kOS, 0, T>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, stack,
                           stride, W, H, pitch, lambda, sl);
    else if (e.parallel == 2)
        hipLaunchKernelGGL((bp_stack_lat_kernel<kOS, 2, T>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, stack,
                           stride, W, H, pitch, lambda, sl);
    else
        hipLaunchKernelGGL((bp_stack_lat_kernel<kOS, -1, T>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, stack,
                           stride, W, H, pitch, lambda, sl);
}

template <typename T>
void launch_stack_lat(artia_ctx* ctx, const VolG& v, const TiltG& g, const BPExact& e, T* vol, const float* stack,
                      size_t stride, int W, int H, size_t pitch, float lambda, int os, const SlabArgs& sl) {
    const dim3 grid((unsigned)((v.nx / 4 + 3) / 4), (unsigned)((v.ny + 3) / 4), (unsigned)((v.nz + 15) / 16));
    if (os == 1)
        launch_stack_lat_os<1>(ctx, grid, v, g, e, vol, stack, stride, W, H, pitch, lambda, sl);
    else
        launch_stack_lat_os<2>(ctx, grid, v, g, e, vol, stack, stride, W, H, pitch, lambda, sl);
}

template <int kOS, int kNI, bool kYCol, typename T>
void launch_win_nj(artia_ctx* ctx, dim3 grid, const VolG& v, const BPWinSet& S, int nj, T* vol, int W, int H,
                   size_t pitch) {
    constexpr int kLo = kOS == 1 ? 1 : 2;
    if (nj == kLo)
        hipLaunchKernelGGL((bp_win_kernel<kOS, kNI, kLo, T, kYCol>), grid, dim3(256), 0, ctx->stream, v, S, vol, W,
                           H, pitch);
    else if (nj == kLo + 1)
        hipLaunchKernelGGL((bp_win_kernel<kOS, kNI, kLo + 1, T, kYCol>), grid, dim3(256), 0, ctx->stream, v, S, vol,
                           W, H, pitch);
    else
        hipLaunchKernelGGL((bp_win_kernel<kOS, kNI, kLo + 2, T, kYCol>), grid, dim3(256), 0, ctx->stream, v, S, vol,
                           W, H, pitch);
}

template <int kOS, bool kYCol, typename T>
void launch_win_os(artia_ctx* ctx, dim3 grid, const VolG& v, const BPWinSet& S, int ni, int nj, T* vol, int W, int H,
                   size_t pitch) {
    constexpr int kLo = kOS == 1 ? 1 : 2;
    if (ni == kLo)
        launch_win_nj<kOS, kLo, kYCol>(ctx, grid, v, S, nj, vol, W, H, pitch);
    else if (ni == kLo + 1)
        launch_win_nj<kOS, kLo + 1, kYCol>(ctx, grid, v, S, nj, vol, W, H, pitch);
    else
        launch_win_nj<kOS, kLo + 2, kYCol>(ctx, grid, v, S, nj, vol, W, H, pitch);
}

// One launch of the window form over S.n tilts sharing the window size (ni, nj); the y slab term
// is taken once per window column when no tilt's detector v axis has a y component.
template <typename T>
void launch_win_t(artia_ctx* ctx, const VolG& v, const BPWinSet& S, int os, int ni, int nj, T* vol, int W, int H,
                  size_t pitch) {
    const int nq = v.nx / 4;
    const dim3 grid((unsigned)((nq + 15) / 16), (unsigned)((v.ny + 3) / 4),
                    (unsigned)((v.nz + kWinBlockZ - 1) / kWinBlockZ));
    bool ycol = true;
    for (int t = 0; t < S.n; t++) ycol = ycol && S.w[t].YVd == 0.0f;
    if (os == 1) {
        if (ycol)
            launch_win_os<1, true>(ctx, grid, v, S, ni, nj, vol, W, H, pitch);
        else
            launch_win_os<1, false>(ctx, grid, v, S, ni, nj, vol, W, H, pitch);
    } else {
        if (ycol)
            launch_win_os<2, true>(ctx, grid, v, S, ni, nj, vol, W, H, pitch);
        else
            launch_win_os<2, false>(ctx, grid, v, S, ni, nj, vol, W, H, pitch);
    }
}

template <int kOS, typename T>
void launch_tile(artia_ctx* ctx, dim3 grid, const VolG& v, const TiltG& g, const BPExact& e, T* vol,
                 const float* img, int W, int H, size_t pitch, float lambda, int os, double dmin, double dmax,
                 int filter) {
    if (e.parallel == 0)
        hipLaunchKernelGGL((bp_exact_kernel<kOS, 0, T>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, img, W, H,
                           pitch, lambda, os, dmin, dmax, filter);
    else if (e.parallel == 2)
        hipLaunchKernelGGL((bp_exact_kernel<kOS, 2, T>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, img, W, H,
                           pitch, lambda, os, dmin, dmax, filter);
    else
        hipLaunchKernelGGL((bp_exact_kernel<kOS, -1, T>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, img, W, H,
                           pitch, lambda, os, dmin, dmax, filter);
}

template <typename T>
void launch_bp_t(artia_ctx* ctx, const VolG& v, const TiltG& g, T* vol, const float* img, int W, int H, size_t pitch,
                 float lambda, int os, float distMin, float distMax, int filter) {
    const int nq = v.nx / 4;
    if (ctx->numerics == ARTIA_NUMERICS_REFERENCE) {
        const dim3 block(64, 4), grid((unsigned)((nq + 63) / 64), (unsigned)((v.ny + 3) / 4), (unsigned)v.nz);
        hipLaunchKernelGGL((bp_ref_kernel<T>), grid, block, 0, ctx->stream, v, g, vol, img, W, H, pitch, lambda, os,
                           1.0f / (float)os, distMin, distMax, filter);
        return;
    }
    BPWinSet S{};
    if (filter == ARTIA_FILTER_POINT && window_plan(v, g, os, S.w[0])) {
        S.img[0] = img;
        S.n = 1;
        S.lambda = lambda;
        S.distMin = distMin;
        S.distMax = distMax;
        S.gated = gate_passes_all(v, S.w[0], S.distMin, S.distMax) ? 0 : 1;
        launch_win_t(ctx, v, S, os, S.w[0].ni, S.w[0].nj, vol, W, H, pitch);
        return;
    }
    const BPExact e = prepare_exact(v, g);
    const dim3 grid((unsigned)((nq + 15) / 16), (unsigned)((v.ny + 3) / 4), (unsigned)((v.nz + 3) / 4));
    const bool tile = e.affine && filter == ARTIA_FILTER_POINT && e.identityAniso &&
                      (long long)(e.tileW + 2) * (long long)(e.tileH + 2) <= kTileFloats;
    const double dmin = distMin, dmax = distMax;
    if (!tile)
        hipLaunchKernelGGL((bp_exact_kernel<-1, -1, T>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, img, W, H,
                           pitch, lambda, os, dmin, dmax, filter);
    else if (os == 1)
        launch_tile<1>(ctx, grid, v, g, e, vol, img, W, H, pitch, lambda, os, dmin, dmax, filter);
    else if (os == 2)
        launch_tile<2>(ctx, grid, v, g, e, vol, img, W, H, pitch, lambda, os, dmin, dmax, filter);
    else if (os == 4)
        launch_tile<4>(ctx, grid, v, g, e, vol, img, W, H, pitch, lambda, os, dmin, dmax, filter);
    else
        launch_tile<0>(ctx, grid, v, g, e, vol, img, W, H, pitch, lambda, os, dmin, dmax, filter);
}

// bp_stack_kernel with the parallel-axis mask fixed at compile time where it is 0 or 2 (launch_tile)
template <typename T>
void launch_stack_t(artia_ctx* ctx, dim3 grid, const VolG& v, const TiltG& g, const BPExact& e, T* vol,
                    const float* stack, size_t stride, int W, int H, size_t pitch, float lambda, int os, int filter,
                    const SlabArgs& sl) {
    if (e.parallel == 0)
        hipLaunchKernelGGL((bp_stack_kernel<0, T>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, stack, stride, W, H,
                           pitch, lambda, os, filter, sl);
    else if (e.parallel == 2)
        hipLaunchKernelGGL((bp_stack_kernel<2, T>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, stack, stride, W, H,
                           pitch, lambda, os, filter, sl);
    else
        hipLaunchKernelGGL((bp_stack_kernel<-1, T>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, stack, stride, W,
                           H, pitch, lambda, os, filter, sl);
}

}  // namespace

void launch_bp_stack(artia_ctx* ctx, const VolG& v, const TiltG& g, VolData vol, const float* stack, size_t stride,
                     int W, int H, size_t pitch, float lambda, int os, int filter, const float* bounds, int nslabs,
                     float last_end) {
    const SlabArgs sl{bounds, nslabs, last_end};
    if ((os == 1 || os == 2) && filter == ARTIA_FILTER_LINEAR && (size_t)H * pitch < ((size_t)1 << 31)) {
        // the staged-lattice form, with the block offsets of its 16 x 4 x 16-voxel blocks
        const BPExact e16 = prepare_exact(v, g, 16, 4, 16);
        if (e16.affine && e16.identityAniso) {
            if (vol.half)
                launch_stack_lat(ctx, v, g, e16, static_cast<vhalf*>(vol.p), stack, stride, W, H, pitch, lambda, os,
                                 sl);
            else
                launch_stack_lat(ctx, v, g, e16, static_cast<float*>(vol.p), stack, stride, W, H, pitch, lambda, os,
                                 sl);
            return;
        }
    }
    const BPExact e = prepare_exact(v, g);
    const int nq = v.nx / 4;
    const dim3 grid((unsigned)((nq + 15) / 16), (unsigned)((v.ny + 3) / 4), (unsigned)((v.nz + 3) / 4));
    if (vol.half)
        launch_stack_t(ctx, grid, v, g, e, static_cast<vhalf*>(vol.p), stack, stride, W, H, pitch, lambda, os, filter,
                       sl);
    else
        launch_stack_t(ctx, grid, v, g, e, static_cast<float*>(vol.p), stack, stride, W, H, pitch, lambda, os, filter,
                       sl);
}

void launch_bp(artia_ctx* ctx, const VolG& v, const TiltG& g, VolData vol, const float* img, int W, int H,
               size_t pitch, float lambda, int os, float distMin, float distMax, int filter) {
    if (vol.half)
        launch_bp_t(ctx, v, g, static_cast<vhalf*>(vol.p), img, W, H, pitch, lambda, os, distMin, distMax, filter);
    else
        launch_bp_t(ctx, v, g, static_cast<float*>(vol.p), img, W, H, pitch, lambda, os, distMin, distMax, filter);
}

void launch_bp_subset(artia_ctx* ctx, const VolG& v, const TiltG* g, const float* const* imgs, int n, int W, int H,
                      size_t pitch, float lambda, int os, VolData vol) {
    int t = 0;
    while (t < n) {
        BPWinSet S{};
        if (ctx->numerics == ARTIA_NUMERICS_EXACT && window_plan(v, g[t], os, S.w[0])) {
            S.img[0] = imgs[t];
            S.n = 1;
            // the following tilts with the same window size join the launch
            while (S.n < kWinTilts && t + S.n < n && window_plan(v, g[t + S.n], os, S.w[S.n]) &&
                   S.w[S.n].ni == S.w[0].ni && S.w[S.n].nj == S.w[0].nj) {
                S.img[S.n] = imgs[t + S.n];
                S.n++;
            }
            S.lambda = lambda;
            S.distMin = 0.0;
            S.distMax = 9999999999999.0;
            S.gated = 0;
            for (int k = 0; k < S.n; k++)
                if (!gate_passes_all(v, S.w[k], S.distMin, S.distMax)) S.gated = 1;
            if (vol.half)
                launch_win_t(ctx, v, S, os, S.w[0].ni, S.w[0].nj, static_cast<vhalf*>(vol.p), W, H, pitch);
            else
                launch_win_t(ctx, v, S, os, S.w[0].ni, S.w[0].nj, static_cast<float*>(vol.p), W, H, pitch);
            t += S.n;
        } else {
            launch_bp(ctx, v, g[t], vol, imgs[t], W, H, pitch, lambda, os, 0.0f, 9999999999999.0f, ARTIA_FILTER_POINT);
            t++;
        }
    }
}

}  // namespace artia

using namespace artia;

namespace {
artia_status bp_entry(artia_ctx* ctx, const char* fn, const artia_vol_geom* vg, const artia_tilt_geom* tg,
                      VolData vol, const float* img, int width, int height, size_t pitch, float lambda,
                      int oversampling, float dist_min, float dist_max, int filter, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!valid_geometry(vg, tg) || !vol.p || !img)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: null pointer or invalid geometry", fn);
    if (vg->nx % 4 != 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: volume x dimension %d is not a multiple of 4", fn, vg->nx);
    // a quad is one 16-byte (fp32) or 8-byte (half) access
    const uintptr_t align = vol.half ? 7 : 15;
    if ((reinterpret_cast<uintptr_t>(vol.p) & align) != 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: volume pointer must be %d-byte aligned", fn,
                    (int)align + 1);
    if (width <= 0 || height <= 0 || pitch < (size_t)width || oversampling <= 0 || oversampling > 64)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: bad image size or oversampling", fn);
    if (filter != ARTIA_FILTER_POINT && filter != ARTIA_FILTER_LINEAR)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: unknown filter %d", fn, filter);
    Timing t(ctx, ms);
    launch_bp(ctx, to_device(*vg), to_device(*tg), vol, img, width, height, pitch, lambda, oversampling, dist_min,
              dist_max, filter);
    return t.finish();
}
}  // namespace

extern "C" {

artia_status artia_bp(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, float* vol,
                      const float* img, int width, int height, size_t pitch, float lambda, int oversampling,
                      float dist_min, float dist_max, int filter, float* ms) {
    return bp_entry(ctx, "artia_bp", vg, tg, VolData(vol), img, width, height, pitch, lambda, oversampling, dist_min,
                    dist_max, filter, ms);
}

artia_status artia_bp_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, uint16_t* vol,
                          const float* img, int width, int height, size_t pitch, float lambda, int oversampling,
                          float dist_min, float dist_max, int filter, float* ms) {
    return bp_entry(ctx, "artia_bp_f16", vg, tg, VolData::f16(vol), img, width, height, pitch, lambda, oversampling,
                    dist_min, dist_max, filter, ms);
}

}  // extern "C"
