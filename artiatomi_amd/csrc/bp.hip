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

// Per-tilt constants for the exact kernels, prepared on the host in double.
struct BPExact {
    // chord through a voxel for a sub-ray offset (a, b) in detector pitch units from the
    // projection of the voxel's min corner: entry_c = a*U[c] + b*V[c] + Smin[c],
    // exit_c = entry_c + Sabs[c] along the unit ray direction; chord = min exit - max entry.
    float U[3], V[3], Smin[3], Sabs[3];
    float uc[3], vc[3], sc[3];   // axes parallel to the rays: inside test on a*u + b*v
    int parallel;                // bit c set when n_c == 0
    // projection of the voxel (i,j,k) min corner: P = p0 + i*pi + j*pj + k*pk
    double pu0, pui, puj, puk, pv0, pvi, pvj, pvk;
    // projected box of one voxel relative to its corner projection
    float uMin, uMax, vMin, vMax;
    int identityAniso;
    // quad footprint as affine forms of (x, y, z): lower_r = G[r].q + lo[r], upper_r = G[r].q + hi[r]
    int affine;                  // 0: the volume straddles the detector-plane sign; use the literal form
    double G[2][3], lo[2], hi[2];
    // macro-cell first-corner distance t = |gn . q + g0| for the dist gate
    double gn[3], g0;
    // LDS tile bound for a 64x4x4-voxel block, in pixels (host upper bound), and the block's
    // footprint offsets relative to G.(block origin): [tlo, thi] per detector axis
    int tileW, tileH;
    double tlo[2], thi[2];
};

BPExact prepare_exact(const VolG& v, const TiltG& g) {
    BPExact e{};
    const double u[3] = {g.u[0], g.u[1], g.u[2]}, w[3] = {g.v[0], g.v[1], g.v[2]};
    const double nf[3] = {g.n[0], g.n[1], g.n[2]};
    const double nl = std::sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]);
    const double n[3] = {nf[0] / nl, nf[1] / nl, nf[2] / nl};
    const double vs[3] = {v.vs[0], v.vs[1], v.vs[2]};
    const double bb[3] = {v.bbMin[0], v.bbMin[1], v.bbMin[2]};
    const double d[3] = {g.d[0], g.d[1], g.d[2]};
    e.parallel = 0;
    for (int c = 0; c < 3; c++) {
        e.uc[c] = (float)u[c];
        e.vc[c] = (float)w[c];
        e.sc[c] = (float)vs[c];
        if (n[c] == 0.0) {
            e.parallel |= 1 << c;
            e.U[c] = 0.0f;
            e.V[c] = 0.0f;
            e.Smin[c] = -INFINITY;
            e.Sabs[c] = INFINITY;
        } else {
            e.U[c] = (float)(-u[c] / n[c]);
            e.V[c] = (float)(-w[c] / n[c]);
            const double S = vs[c] / n[c];
            e.Smin[c] = (float)std::fmin(0.0, S);
            e.Sabs[c] = (float)std::fabs(S);
        }
    }
    // inverse of the column basis [u v n]: p - d = alpha*u + beta*v + gamma*n
    const double B[9] = {u[0], w[0], nf[0], u[1], w[1], nf[1], u[2], w[2], nf[2]};
    const double det = B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                       B[2] * (B[3] * B[7] - B[4] * B[6]);
    const double r0[3] = {(B[4] * B[8] - B[5] * B[7]) / det, -(B[1] * B[8] - B[2] * B[7]) / det,
                          (B[1] * B[5] - B[2] * B[4]) / det};
    const double r1[3] = {-(B[3] * B[8] - B[5] * B[6]) / det, (B[0] * B[8] - B[2] * B[6]) / det,
                          -(B[0] * B[5] - B[2] * B[3]) / det};
    const double o[3] = {bb[0] - d[0], bb[1] - d[1], bb[2] - d[2]};
    e.pu0 = r0[0] * o[0] + r0[1] * o[1] + r0[2] * o[2];
    e.pv0 = r1[0] * o[0] + r1[1] * o[1] + r1[2] * o[2];
    e.pui = r0[0] * vs[0];
    e.puj = r0[1] * vs[1];
    e.puk = r0[2] * vs[2];
    e.pvi = r1[0] * vs[0];
    e.pvj = r1[1] * vs[1];
    e.pvk = r1[2] * vs[2];
    const double guard = 1e-4;
    e.uMin = (float)(std::fmin(0.0, e.pui) + std::fmin(0.0, e.puj) + std::fmin(0.0, e.puk) - guard);
    e.uMax = (float)(std::fmax(0.0, e.pui) + std::fmax(0.0, e.puj) + std::fmax(0.0, e.puk) + guard);
    e.vMin = (float)(std::fmin(0.0, e.pvi) + std::fmin(0.0, e.pvj) + std::fmin(0.0, e.pvk) - guard);
    e.vMax = (float)(std::fmax(0.0, e.pvi) + std::fmax(0.0, e.pvj) + std::fmax(0.0, e.pvk) + guard);
    const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    e.identityAniso = 1;
    for (int i = 0; i < 9; i++)
        if (g.Ainv[i] != I[i]) e.identityAniso = 0;

    // Quad footprint (BackProjectionSquareOS.cu:365-509) as affine forms.  A corner c maps to
    // h(c) = c + |n.c - n.d| (-n); with a single sign s of n.c - n.d over the whole volume,
    // h(c) = P c + s (n.d) n with P = I - s n n^T, so the corner-wise min/max of the macro-cell
    // split into a per-quad affine part and per-tilt constants.
    const double nd = ((-nf[0]) * d[0] - nf[1] * d[1]) - nf[2] * d[2];
    int sPos = 0, sNeg = 0;
    for (int c = 0; c < 8; c++) {
        const double p[3] = {(c & 4) ? (double)v.bbMax[0] : bb[0], (c & 2) ? (double)v.bbMax[1] : bb[1],
                             (c & 1) ? (double)v.bbMax[2] : bb[2]};
        const double t = nf[0] * p[0] + nf[1] * p[1] + nf[2] * p[2] + nd;
        (t >= 0 ? sPos : sNeg)++;
    }
    e.affine = (sPos == 8 || sNeg == 8) ? 1 : 0;
    const double s = sPos == 8 ? 1.0 : -1.0;
    double P[3][3];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) P[i][k] = (i == k ? 1.0 : 0.0) - s * nf[i] * nf[k];
    const double E[3] = {5.0 * vs[0], vs[1], vs[2]};
    double h0[3], mn[3], mx[3], A[3][3];
    for (int i = 0; i < 3; i++) {
        h0[i] = -s * nf[i] * nd;   // + s (n.d) n, with nd = -(n.d)
        mn[i] = mx[i] = 0.0;
        for (int k = 0; k < 3; k++) {
            h0[i] += P[i][k] * bb[k];
            A[i][k] = P[i][k] * vs[k];
            mn[i] += std::fmin(0.0, P[i][k] * E[k]);
            mx[i] += std::fmax(0.0, P[i][k] * E[k]);
        }
    }
    double loC[3] = {h0[0] + mn[0], h0[1] + mn[1], h0[2] + mn[2]};
    double hiC[3] = {h0[0] + mx[0], h0[1] + mx[1], h0[2] + mx[2]};
    if (nf[0] <= 0.0) std::swap(loC[0], hiC[0]);
    for (int r = 0; r < 2; r++) {
        const double m[4] = {g.M[4 * r], g.M[4 * r + 1], g.M[4 * r + 2], g.M[4 * r + 3]};
        for (int k = 0; k < 3; k++) e.G[r][k] = m[0] * A[0][k] + m[1] * A[1][k] + m[2] * A[2][k];
        const double pl = m[0] * loC[0] + m[1] * loC[1] + m[2] * loC[2] + m[3];
        const double ph = m[0] * hiC[0] + m[1] * hiC[1] + m[2] * hiC[2] + m[3];
        e.lo[r] = std::fmin(pl, ph);
        e.hi[r] = std::fmax(pl, ph);
    }
    for (int k = 0; k < 3; k++) e.gn[k] = s * nf[k] * vs[k];
    e.g0 = s * (nf[0] * bb[0] + nf[1] * bb[1] + nf[2] * bb[2] + nd);
    const double ext[3] = {60.0, 3.0, 3.0};   // quad origins within a 64 x 4 x 4 voxel block
    for (int r = 0; r < 2; r++) {
        e.tlo[r] = e.lo[r];
        e.thi[r] = e.hi[r];
        for (int k = 0; k < 3; k++) {
            e.tlo[r] += std::fmin(0.0, e.G[r][k] * ext[k]);
            e.thi[r] += std::fmax(0.0, e.G[r][k] * ext[k]);
        }
    }
    // pixel extent of a 64 x 4 x 4 voxel block's footprint, plus the per-quad margins
    e.tileW = (int)std::ceil(std::fabs(e.G[0][0]) * 64 + std::fabs(e.G[0][1]) * 4 + std::fabs(e.G[0][2]) * 4 +
                             (e.hi[0] - e.lo[0])) + 3;
    e.tileH = (int)std::ceil(std::fabs(e.G[1][0]) * 64 + std::fabs(e.G[1][1]) * 4 + std::fabs(e.G[1][2]) * 4 +
                             (e.hi[1] - e.lo[1])) + 3;
    return e;
}

// Footprint of a quad's macro-cell in double with the reference's operation order; returns
// false when the gate [distMin, distMax) rejects the quad.  Used when the affine form does not
// apply (the volume straddles the detector-plane sign).
__device__ __forceinline__ bool footprint_literal(const VolG& v, const TiltG& g, unsigned x, unsigned y, unsigned z,
                                                  int W, int H, double distMin, double distMax, int& px0, int& px1,
                                                  int& py0, int& py1) {
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

__device__ __forceinline__ bool footprint_affine(const BPExact& e, unsigned x, unsigned y, unsigned z, int W, int H,
                                                 double distMin, double distMax, int& px0, int& px1, int& py0,
                                                 int& py1) {
    const double qx = (double)x, qy = (double)y, qz = (double)z;
    const double t = fabs(fma(e.gn[0], qx, fma(e.gn[1], qy, fma(e.gn[2], qz, e.g0))));
    if (!(t >= distMin && t < distMax)) return false;
    const double a0 = fma(e.G[0][0], qx, fma(e.G[0][1], qy, e.G[0][2] * qz));
    const double a1 = fma(e.G[1][0], qx, fma(e.G[1][1], qy, e.G[1][2] * qz));
    px0 = (int)fmin(fmax(floor(a0 + e.lo[0]), 0.0), (double)(W - 1));
    px1 = (int)fmin(fmax(ceil(a0 + e.hi[0]), 0.0), (double)(W - 1));
    py0 = (int)fmin(fmax(floor(a1 + e.lo[1]), 0.0), (double)(H - 1));
    py1 = (int)fmin(fmax(ceil(a1 + e.hi[1]), 0.0), (double)(H - 1));
    return true;
}

constexpr int kTileFloats = 4096;   // 16 KiB of LDS for the block's projection tile

// Chord-weighted sums of one voxel over its sub-ray window [i0, i1] x [j0, j1]: entry/exit of
// the three slabs are linear in the sub-ray offset (a, b); a chord <= 0 adds exactly zero, so the
// loop is branch free and the per-voxel accumulation order (py outer, px inner) is kept.
// kOS > 0 fixes the oversampling at compile time (the pixel of lattice index i is i / kOS);
// kPar adds the inside test for axes parallel to the rays.
template <int kOS, bool kPar>
__device__ __forceinline__ void voxel_sums_tile(const BPExact& e, const float* __restrict__ tile, int tw, int tileOff,
                                                int os, float delta, float du, float dv, int i0, int i1, int j0,
                                                int j1, float& num, float& den) {
    const int O = kOS > 0 ? kOS : os;
    for (int j = j0; j <= j1; j++) {
        const float b = __builtin_fmaf((float)j + 0.5f, delta, -dv);
        const float bV0 = __builtin_fmaf(b, e.V[0], e.Smin[0]);
        const float bV1 = __builtin_fmaf(b, e.V[1], e.Smin[1]);
        const float bV2 = __builtin_fmaf(b, e.V[2], e.Smin[2]);
        float bw[3];
        if (kPar)
            for (int c = 0; c < 3; c++) bw[c] = b * e.vc[c];
        const float* row = tile + tileOff + (j / O) * tw;
        float a = __builtin_fmaf((float)i0 + 0.5f, delta, -du);
        for (int i = i0; i <= i1; i++) {
            const float e0 = __builtin_fmaf(a, e.U[0], bV0);
            const float e1 = __builtin_fmaf(a, e.U[1], bV1);
            const float e2 = __builtin_fmaf(a, e.U[2], bV2);
            const float tin = fmaxf(fmaxf(e0, e1), e2);
            const float tout = fminf(fminf(e0 + e.Sabs[0], e1 + e.Sabs[1]), e2 + e.Sabs[2]);
            float chord = fmaxf(tout - tin, 0.0f);
            if (kPar) {
                for (int c = 0; c < 3; c++) {
                    if (!((e.parallel >> c) & 1)) continue;
                    const float wc = __builtin_fmaf(a, e.uc[c], bw[c]);
                    chord = (wc > 0.0f && wc < e.sc[c]) ? chord : 0.0f;
                }
            }
            const float val = row[i / O];
            num = __builtin_fmaf(val, chord, num);
            den += chord;
            a += delta;
        }
    }
}

// Generic per-voxel sums reading the image from global memory (linear filtering, magnification
// anisotropy, or a footprint larger than the LDS tile).
__device__ __forceinline__ void voxel_sums_global(const BPExact& e, const TiltG& g, const float* __restrict__ img, int W,
                                                  int H, size_t pitch, int filter, int px0, int py0, int os,
                                                  float delta, float du, float dv, int i0, int i1, int j0, int j1,
                                                  float& num, float& den) {
    for (int j = j0; j <= j1; j++) {
        const float b = __builtin_fmaf((float)j + 0.5f, delta, -dv);
        const float bV0 = __builtin_fmaf(b, e.V[0], e.Smin[0]);
        const float bV1 = __builtin_fmaf(b, e.V[1], e.Smin[1]);
        const float bV2 = __builtin_fmaf(b, e.V[2], e.Smin[2]);
        for (int i = i0; i <= i1; i++) {
            const float a = __builtin_fmaf((float)i + 0.5f, delta, -du);
            const float e0 = __builtin_fmaf(a, e.U[0], bV0);
            const float e1 = __builtin_fmaf(a, e.U[1], bV1);
            const float e2 = __builtin_fmaf(a, e.U[2], bV2);
            const float tin = fmaxf(fmaxf(e0, e1), e2);
            const float tout = fminf(fminf(e0 + e.Sabs[0], e1 + e.Sabs[1]), e2 + e.Sabs[2]);
            float chord = tout - tin;
            for (int c = 0; c < 3; c++) {
                if (!((e.parallel >> c) & 1)) continue;
                const float wc = __builtin_fmaf(a, e.uc[c], b * e.vc[c]);
                if (!(wc > 0.0f && wc < e.sc[c])) chord = 0.0f;
            }
            if (chord > 0.0f) {
                const float pxf = (float)px0 + ((float)i + 0.5f) * delta;
                const float pyf = (float)py0 + ((float)j + 0.5f) * delta;
                float val;
                if (filter == ARTIA_FILTER_POINT && e.identityAniso) {
                    val = img[(size_t)(py0 + j / os) * pitch + (size_t)(px0 + i / os)];
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
}

// One thread per quad; a 256-thread block covers 16 quads (64 voxels) along x, 4 rows in y and
// 4 slices in z.  kOS >= 0 selects the LDS-tile path (POINT sampling, no magnification
// anisotropy, affine footprint): the block first stages the projection pixels its quads can
// reach.  kOS < 0 is the global-memory path.
template <int kOS, bool kPar>
__global__ void __launch_bounds__(256) bp_exact_kernel(VolG v, TiltG g, BPExact e, float* __restrict__ vol,
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
    float* vp = vol + ((size_t)z * (size_t)v.ny + (size_t)y) * (size_t)v.nx + x;
    const float4 vb = *reinterpret_cast<const float4*>(vp);
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
            voxel_sums_tile<(kOS > 0 ? kOS : 0), kPar>(e, tile, tw, tileOff, O, delta, du, dv, i0, i1, j0, j1, num,
                                                       den);
        else
            voxel_sums_global(e, g, img, W, H, pitch, filter, px0, py0, O, delta, du, dv, i0, i1, j0, j1, num, den);
        den = fmaxf(0.0f, den);
        if (den != 0.0f) {
            out[k] += (lambda * num) / den;
            any = true;
        }
    }
    if (any) *reinterpret_cast<float4*>(vp) = make_float4(out[0], out[1], out[2], out[3]);
}

template <int kOS>
void launch_tile(artia_ctx* ctx, dim3 grid, const VolG& v, const TiltG& g, const BPExact& e, float* vol,
                 const float* img, int W, int H, size_t pitch, float lambda, int os, double dmin, double dmax,
                 int filter) {
    if (e.parallel)
        hipLaunchKernelGGL((bp_exact_kernel<kOS, true>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, img, W, H,
                           pitch, lambda, os, dmin, dmax, filter);
    else
        hipLaunchKernelGGL((bp_exact_kernel<kOS, false>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, img, W, H,
                           pitch, lambda, os, dmin, dmax, filter);
}

}  // namespace

void launch_bp(artia_ctx* ctx, const VolG& v, const TiltG& g, float* vol, const float* img, int W, int H, size_t pitch,
               float lambda, int os, float distMin, float distMax, int filter) {
    const int nq = v.nx / 4;
    if (ctx->numerics == ARTIA_NUMERICS_REFERENCE) {
        const dim3 block(64, 4), grid((unsigned)((nq + 63) / 64), (unsigned)((v.ny + 3) / 4), (unsigned)v.nz);
        hipLaunchKernelGGL(bp_ref_kernel, grid, block, 0, ctx->stream, v, g, vol, img, W, H, pitch, lambda, os,
                           1.0f / (float)os, distMin, distMax, filter);
        return;
    }
    const BPExact e = prepare_exact(v, g);
    const dim3 grid((unsigned)((nq + 15) / 16), (unsigned)((v.ny + 3) / 4), (unsigned)((v.nz + 3) / 4));
    const bool tile = e.affine && filter == ARTIA_FILTER_POINT && e.identityAniso &&
                      (long long)(e.tileW + 2) * (long long)(e.tileH + 2) <= kTileFloats;
    const double dmin = distMin, dmax = distMax;
    if (!tile)
        hipLaunchKernelGGL((bp_exact_kernel<-1, true>), grid, dim3(256), 0, ctx->stream, v, g, e, vol, img, W, H,
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
