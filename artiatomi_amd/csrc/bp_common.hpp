// Exact-numerics backprojection building blocks shared by the per-call kernel (bp.hip) and the
// batched sub-volume kernel (subvol.hip): per-tilt footprint constants prepared on the host in
// double, the quad footprint and dist gate, and the per-voxel chord-weighted sums
// (BackProjectionSquareOS.cu:263-782).
#pragma once

#include <algorithm>
#include <cmath>

#include "device_common.hpp"

namespace artia {

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
    // LDS tile bound for one block of voxels, in pixels (host upper bound), and the block's
    // footprint offsets relative to G.(block origin): [tlo, thi] per detector axis
    int tileW, tileH;
    double tlo[2], thi[2];
};

inline BPExact prepare_exact(const VolG& v, const TiltG& g, int blockX = 64, int blockY = 4, int blockZ = 4) {
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
    // quad origins within a blockX x blockY x blockZ voxel block
    const double ext[3] = {(double)(blockX - 4), (double)(blockY - 1), (double)(blockZ - 1)};
    for (int r = 0; r < 2; r++) {
        e.tlo[r] = e.lo[r];
        e.thi[r] = e.hi[r];
        for (int k = 0; k < 3; k++) {
            e.tlo[r] += std::fmin(0.0, e.G[r][k] * ext[k]);
            e.thi[r] += std::fmax(0.0, e.G[r][k] * ext[k]);
        }
    }
    // pixel extent of the block's footprint, plus the per-quad margins
    e.tileW = (int)std::ceil(std::fabs(e.G[0][0]) * blockX + std::fabs(e.G[0][1]) * blockY +
                             std::fabs(e.G[0][2]) * blockZ + (e.hi[0] - e.lo[0])) + 3;
    e.tileH = (int)std::ceil(std::fabs(e.G[1][0]) * blockX + std::fabs(e.G[1][1]) * blockY +
                             std::fabs(e.G[1][2]) * blockZ + (e.hi[1] - e.lo[1])) + 3;
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

// CTF slab gate of one tilt: bounds[k] = start of slab k (the reference's float loop values,
// Reconstructor.cpp:920-969), slab k covers [bounds[k], bounds[k+1]) and the last one
// [bounds[K-1], last_end).
struct SlabArgs {
    const float* bounds;
    int count;
    float last_end;
};

// The slab holding distance t (the gate's distance of a voxel's first macro-cell corner), or -1.
__device__ __forceinline__ int find_slab(const SlabArgs& sl, double t) {
    if (sl.count <= 0 || t < (double)sl.bounds[0]) return -1;
    int lo = 0, hi = sl.count - 1;   // largest k with bounds[k] <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((double)sl.bounds[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    if (lo == sl.count - 1 && !(t < (double)sl.last_end)) return -1;
    return lo;
}

// The gate's distance of voxel (x, y, z): |n . c - n . d| at its first macro-cell corner, as
// footprint_affine / footprint_literal compute it.
__device__ __forceinline__ double gate_distance(const BPExact& e, const VolG& v, const TiltG& g, unsigned x, unsigned y,
                                                unsigned z) {
    if (e.affine) return fabs(fma(e.gn[0], (double)x, fma(e.gn[1], (double)y, fma(e.gn[2], (double)z, e.g0))));
    const double n[3] = {g.n[0], g.n[1], g.n[2]};
    const double d[3] = {g.d[0], g.d[1], g.d[2]};
    const double c[3] = {(double)v.bbMin[0] + (double)x * (double)v.vs[0],
                         (double)v.bbMin[1] + (double)y * (double)v.vs[1],
                         (double)v.bbMin[2] + (double)z * (double)v.vs[2]};
    const double nd = ((-n[0]) * d[0] - n[1] * d[1]) - n[2] * d[2];
    return fabs(((n[0] * c[0] + n[1] * c[1]) + n[2] * c[2]) + nd);
}

// Chord of the sub-ray at offset (a, b) through one voxel: entry/exit of the three slabs are linear
// in the offset; axes parallel to the rays (n_c == 0) contribute an inside test instead of a slab.
// kPM is the parallel-axis mask when known at compile time: 0 (none) or 2 (y only, every tilt about
// the y axis); -1 reads the mask at run time.  With n_y == 0 the y slab is (-inf, +inf) and drops
// out of the max/min exactly, so the kPM = 2 form gives the generic form's value bit for bit.
template <int kPM>
__device__ __forceinline__ float chord_at(const BPExact& e, float a, float b, float bV0, float bV1, float bV2) {
    if (kPM == 2) {
        const float e0 = __builtin_fmaf(a, e.U[0], bV0);
        const float e2 = __builtin_fmaf(a, e.U[2], bV2);
        const float tin = fmaxf(e0, e2);
        const float tout = fminf(e0 + e.Sabs[0], e2 + e.Sabs[2]);
        const float chord = fmaxf(tout - tin, 0.0f);
        const float wc = __builtin_fmaf(a, e.uc[1], b * e.vc[1]);
        return (wc > 0.0f && wc < e.sc[1]) ? chord : 0.0f;
    }
    const float e0 = __builtin_fmaf(a, e.U[0], bV0);
    const float e1 = __builtin_fmaf(a, e.U[1], bV1);
    const float e2 = __builtin_fmaf(a, e.U[2], bV2);
    const float tin = fmaxf(fmaxf(e0, e1), e2);
    const float tout = fminf(fminf(e0 + e.Sabs[0], e1 + e.Sabs[1]), e2 + e.Sabs[2]);
    float chord = fmaxf(tout - tin, 0.0f);
    if (kPM != 0) {
        for (int c = 0; c < 3; c++) {
            if (!((e.parallel >> c) & 1)) continue;
            const float wc = __builtin_fmaf(a, e.uc[c], b * e.vc[c]);
            chord = (wc > 0.0f && wc < e.sc[c]) ? chord : 0.0f;
        }
    }
    return chord;
}

// Sub-ray offsets of a window row: a_i = (i - i0) * delta + a0 with a0 = (i0 + 0.5) * delta - du,
// one rounding per lattice point, so every path below sees the same offsets.
__device__ __forceinline__ float lattice_a(int i, int i0, float delta, float a0) {
    return __builtin_fmaf((float)(i - i0), delta, a0);
}

// Chord-weighted sums of one voxel over its sub-ray window [i0, i1] x [j0, j1] from the LDS tile.
// A chord <= 0 adds exactly zero, so the loop is branch free; the per-voxel accumulation order
// (py outer, px inner) is the reference's.  kOS > 0 fixes the oversampling at compile time (the
// pixel of lattice index i is i / kOS).
template <int kOS, int kPM>
__device__ __forceinline__ void voxel_sums_tile(const BPExact& e, const float* __restrict__ tile, int tw, int tileOff,
                                                int os, float delta, float du, float dv, int i0, int i1, int j0,
                                                int j1, float& num, float& den) {
    const int O = kOS > 0 ? kOS : os;
    const float a0 = __builtin_fmaf((float)i0 + 0.5f, delta, -du);
    for (int j = j0; j <= j1; j++) {
        const float b = __builtin_fmaf((float)j + 0.5f, delta, -dv);
        const float bV0 = __builtin_fmaf(b, e.V[0], e.Smin[0]);
        const float bV1 = __builtin_fmaf(b, e.V[1], e.Smin[1]);
        const float bV2 = __builtin_fmaf(b, e.V[2], e.Smin[2]);
        const float* row = tile + tileOff + (j / O) * tw;
        for (int i = i0; i <= i1; i++) {
            const float chord = chord_at<kPM>(e, lattice_a(i, i0, delta, a0), b, bV0, bV1, bV2);
            const float val = row[i / O];
            num = __builtin_fmaf(val, chord, num);
            den += chord;
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
        // the same offsets and chords as voxel_sums_tile, so both paths agree bit for bit
        const float a0 = __builtin_fmaf((float)i0 + 0.5f, delta, -du);
        for (int i = i0; i <= i1; i++) {
            const float chord = chord_at<-1>(e, lattice_a(i, i0, delta, a0), b, bV0, bV1, bV2);
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

// tex2d's LINEAR fetch at (x, y) from an LDS copy of image pixels [tx0, tx0 + tw) x [ty0, ..) that
// holds every pixel the fetch reads: the same clamps, weights and lerps, so the same value.
__device__ __forceinline__ float tile_linear(const float* __restrict__ t, int tw, int tx0, int ty0, int W, int H,
                                             float x, float y) {
    x = fminf(fmaxf(x, -1.0f), (float)(W + 1));
    y = fminf(fmaxf(y, -1.0f), (float)(H + 1));
    const float xb = x - 0.5f, yb = y - 0.5f;
    const float fx = floorf(xb), fy = floorf(yb);
    const float a = xb - fx, b = yb - fy;
    int i0 = (int)fx, j0 = (int)fy;
    const int i1 = clampi(i0 + 1, 0, W - 1) - tx0, j1 = clampi(j0 + 1, 0, H - 1) - ty0;
    i0 = clampi(i0, 0, W - 1) - tx0;
    j0 = clampi(j0, 0, H - 1) - ty0;
    const float r0 = lerpf(t[j0 * tw + i0], t[j0 * tw + i1], a);
    const float r1 = lerpf(t[j1 * tw + i0], t[j1 * tw + i1], a);
    return lerpf(r0, r1, b);
}

// tile_linear where every clamp is a no-op: fetch points with x in [1, W - 1] and y in [1, H - 1],
// so the corner pixels floor(x - 0.5) and floor(x - 0.5) + 1 lie inside the image (same values).
__device__ __forceinline__ float tile_linear_inner(const float* __restrict__ t, int tw, int tx0, int ty0, float x,
                                                   float y) {
    const float xb = x - 0.5f, yb = y - 0.5f;
    const float fx = floorf(xb), fy = floorf(yb);
    const float a = xb - fx, b = yb - fy;
    const float* __restrict__ r = t + ((int)fy - ty0) * tw + ((int)fx - tx0);
    const float r0 = lerpf(r[0], r[1], a);
    const float r1 = lerpf(r[tw], r[tw + 1], a);
    return lerpf(r0, r1, b);
}

// voxel_sums_global for LINEAR sampling without magnification anisotropy (the fetch point is the
// sub-ray's pixel position), reading the image from an LDS tile (tile_linear).  kInner: the quad's
// footprint satisfies 1 <= px0, px1 <= W - 1, 1 <= py0, py1 <= H - 1, so every fetch point (in
// [p0 + delta/2, p1 - delta/2]) is one tile_linear_inner takes.
template <int kPM, bool kInner>
__device__ __forceinline__ void voxel_sums_linear_tile(const BPExact& e, const float* __restrict__ t, int tw, int tx0,
                                                       int ty0, int W, int H, int px0, int py0, float delta, float du,
                                                       float dv, int i0, int i1, int j0, int j1, float& num,
                                                       float& den) {
    for (int j = j0; j <= j1; j++) {
        const float b = __builtin_fmaf((float)j + 0.5f, delta, -dv);
        const float bV0 = __builtin_fmaf(b, e.V[0], e.Smin[0]);
        const float bV1 = __builtin_fmaf(b, e.V[1], e.Smin[1]);
        const float bV2 = __builtin_fmaf(b, e.V[2], e.Smin[2]);
        const float a0 = __builtin_fmaf((float)i0 + 0.5f, delta, -du);
        for (int i = i0; i <= i1; i++) {
            const float chord = chord_at<kPM>(e, lattice_a(i, i0, delta, a0), b, bV0, bV1, bV2);
            if (chord > 0.0f) {
                const float pxf = (float)px0 + ((float)i + 0.5f) * delta;
                const float pyf = (float)py0 + ((float)j + 0.5f) * delta;
                const float val = kInner ? tile_linear_inner(t, tw, tx0, ty0, pxf, pyf)
                                         : tile_linear(t, tw, tx0, ty0, W, H, pxf, pyf);
                num = __builtin_fmaf(val, chord, num);
                den += chord;
            }
        }
    }
}

// voxel_sums_linear_tile with each lattice point's LINEAR fetch read from a staged lattice plane:
// lat[(oj + j) * ls + (oi + i)] holds tile_linear's value at the pixel position of column i and row
// j of the quad's footprint lattice (px0 + (i + 0.5) delta, py0 + (j + 0.5) delta).  The same
// chords, values and sums in the same order, so the same bits.
template <int kPM>
__device__ __forceinline__ void voxel_sums_lattice(const BPExact& e, const float* __restrict__ lat, int ls, int oi,
                                                   int oj, float delta, float du, float dv, int i0, int i1, int j0,
                                                   int j1, float& num, float& den) {
    for (int j = j0; j <= j1; j++) {
        const float b = __builtin_fmaf((float)j + 0.5f, delta, -dv);
        const float bV0 = __builtin_fmaf(b, e.V[0], e.Smin[0]);
        const float bV1 = __builtin_fmaf(b, e.V[1], e.Smin[1]);
        const float bV2 = __builtin_fmaf(b, e.V[2], e.Smin[2]);
        const float a0 = __builtin_fmaf((float)i0 + 0.5f, delta, -du);
        const float* __restrict__ row = lat + (oj + j) * ls + oi;
        for (int i = i0; i <= i1; i++) {
            const float chord = chord_at<kPM>(e, lattice_a(i, i0, delta, a0), b, bV0, bV1, bV2);
            if (chord > 0.0f) {
                num = __builtin_fmaf(row[i], chord, num);
                den += chord;
            }
        }
    }
}

}  // namespace artia
