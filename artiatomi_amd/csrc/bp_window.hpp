// Fixed-window form of the exact backprojection (BackProjectionSquareOS.cu:263-782) for tilts
// whose beam has no y component (every tilt about the y axis: n_y is zero or a float residue of
// the rotation), POINT sampling, no magnification anisotropy.
//
// A voxel's sub-rays are the OS lattice points (i + 0.5) / OS of the detector inside the voxel's
// projected box.  The box holds at most kNI x kNJ lattice points for every voxel of a tilt, so the
// window starting at the box's first lattice point is walked fully unrolled, without per-voxel
// bounds; lattice points of the window outside the voxel's projection get a chord <= 0 and add
// nothing, as in the per-box loop.  With the sub-ray offset (a, b) from the projection of the
// voxel's min corner (detector pixels along u and v), the chord through the voxel is
//     chord = max(0, min(m, P, SP - P, Y, SY - Y))
// with P and Y affine in (a, b).  Evaluated as med3(min(SP/2 - |P - SP/2|, SY/2 - |Y - SY/2|), 0, m),
// which is the same clamp (m > 0), in 8 operations per lattice point:
//   - n_x != 0: the x and z slabs' entries e_x, e_z are affine in (a, b) and the chord
//     min(e_x + S_x, e_z + S_z) - max(e_x, e_z) equals min(min(S_x, S_z), S_x - d, S_z + d) for
//     d = e_z - e_x; P = S_x - d, SP = S_x + S_z, m = min(S_x, S_z);
//   - n_x == 0 (tilt 0): the ray runs inside the x slab when 0 < w_x < s_x (w_x the sub-ray's x
//     offset a u_x + b v_x); P = B w_x, SP = B s_x, m = S_z, with B = 2^100;
//   - y (parallel): Y = B w_y, SY = B s_y; min(Y, SY - Y) > 0 exactly when 0 < w_y < s_y, and
//     B is large enough that a sub-ray inside the slab by more than 1e-30 voxels is not clipped.
// For a float-residue n_y (~1e-9) the sub-ray's y offset drifts by less than 1e-8 voxels across a
// voxel, so the parallel form differs from the slab form only for sub-rays within rounding of a y
// face, where any evaluation is decided by rounding.
#pragma once

#include <cmath>

#include "bp_common.hpp"
#include "device_common.hpp"

namespace artia {

constexpr float kWinBig = 1.2676506002282294e30f;   // 2^100
// A y coordinate beyond every slab (|Y| - SY/2 stays above 1e35 for |Y| moved by any window step):
// a row placed there has every chord 0.
constexpr float kWinFar = 1e36f;

// Per-tilt constants of the window form, prepared on the host in double.
struct BPWin {
    float PU, PV, PC;        // Pc = P - SP/2 = a PU + b PV + PC
    float YU, YV, YC;        // Yc = Y - SY/2 = a YU + b YV + YC
    float PUd, PVd, YUd, YVd;   // the same per lattice step (times 1/OS)
    float SPh, SYh, m;       // SP/2, SY/2, m
    float uMin, vMin;        // the voxel's projected box starts (with guard) relative to its corner
    double pu0, pui, puj, puk, pv0, pvi, pvj, pvk;   // corner projection (detector pixels)
    double G[2][3], lo[2], hi[2], tlo[2], thi[2];     // quad and block footprints (see BPExact)
    double gn[3], g0;                                 // dist gate: |gn . q + g0|
    int ni, nj;              // window size in lattice points
    // The reference sums only the sub-rays inside a quad's footprint, the box spanned by two
    // projected corners of its macro-cell (BackProjectionSquareOS.cu:365-509, x bounds swapped
    // when n_x <= 0), which need not cover a voxel's projected box.  covers != 0: it covers the
    // projected boxes of all four voxels of every quad of the volume, so a quad whose footprint
    // the image edges do not clamp may walk its windows without the per-point footprint test.
    int covers;
    // masked != 0: the voxel boxes stick out of the quad footprints by less than 1 - 1/os pixels,
    // so every window of a block lies in its footprint tile with a one-pixel margin, and a block
    // away from the image edges can apply the footprint test as a per-point chord cap (0 outside).
    int masked;
};

// Whether the window form applies to a tilt and its constants, for oversampling os and a window of
// at most maxN lattice points per axis.  Requires the affine footprint of prepare_exact (passed in).
inline bool prepare_window(const VolG& v, const TiltG& g, const BPExact& e, int os, int maxN, BPWin& w) {
    if (!e.affine || !e.identityAniso || os <= 0) return false;
    const double nf[3] = {g.n[0], g.n[1], g.n[2]};
    const double nl = std::sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]);
    const double n[3] = {nf[0] / nl, nf[1] / nl, nf[2] / nl};
    if (!(std::fabs(n[1]) <= 1e-6) || !(std::fabs(n[2]) > 1e-3)) return false;
    const double u[3] = {g.u[0], g.u[1], g.u[2]}, wv[3] = {g.v[0], g.v[1], g.v[2]};
    const double vs[3] = {v.vs[0], v.vs[1], v.vs[2]};
    const double B = (double)kWinBig;
    auto slab = [&](int c, double& U, double& V, double& Smin, double& Sabs) {
        U = -u[c] / n[c];
        V = -wv[c] / n[c];
        const double S = vs[c] / n[c];
        Smin = std::fmin(0.0, S);
        Sabs = std::fabs(S);
    };
    double Uz, Vz, Sminz, Sabsz;
    slab(2, Uz, Vz, Sminz, Sabsz);
    if (std::fabs(n[0]) > 1e-6) {
        double Ux, Vx, Sminx, Sabsx;
        slab(0, Ux, Vx, Sminx, Sabsx);
        // P = S_x - d, d = e_z - e_x
        w.PU = (float)(-(Uz - Ux));
        w.PV = (float)(-(Vz - Vx));
        w.PC = (float)((Sabsx - (Sminz - Sminx)) - 0.5 * (Sabsx + Sabsz));
        w.SPh = (float)(0.5 * (Sabsx + Sabsz));
        w.m = (float)std::fmin(Sabsx, Sabsz);
    } else {
        w.PU = (float)(B * u[0]);
        w.PV = (float)(B * wv[0]);
        w.PC = (float)(-0.5 * B * vs[0]);
        w.SPh = (float)(0.5 * B * vs[0]);
        w.m = (float)Sabsz;
    }
    w.YU = (float)(B * u[1]);
    w.YV = (float)(B * wv[1]);
    w.YC = (float)(-0.5 * B * vs[1]);
    w.SYh = (float)(0.5 * B * vs[1]);
    w.PUd = (float)((double)w.PU / os);
    w.PVd = (float)((double)w.PV / os);
    w.YUd = (float)((double)w.YU / os);
    w.YVd = (float)((double)w.YV / os);
    w.uMin = e.uMin;
    w.vMin = e.vMin;
    w.pu0 = e.pu0;
    w.pui = e.pui;
    w.puj = e.puj;
    w.puk = e.puk;
    w.pv0 = e.pv0;
    w.pvi = e.pvi;
    w.pvj = e.pvj;
    w.pvk = e.pvk;
    for (int r = 0; r < 2; r++) {
        for (int k = 0; k < 3; k++) w.G[r][k] = e.G[r][k];
        w.lo[r] = e.lo[r];
        w.hi[r] = e.hi[r];
        w.tlo[r] = e.tlo[r];
        w.thi[r] = e.thi[r];
    }
    for (int k = 0; k < 3; k++) w.gn[k] = e.gn[k];
    w.g0 = e.g0;
    // lattice points in a box of length L: at most floor(L * os) + 1 (the box's float rounding is
    // far inside the 1e-4 guard already in uMin/uMax)
    w.ni = (int)std::floor(((double)e.uMax - (double)e.uMin) * os) + 1;
    w.nj = (int)std::floor(((double)e.vMax - (double)e.vMin) * os) + 1;
    // Footprint [G_r . q + lo_r, G_r . q + hi_r] against the union of the quad's voxel boxes
    // [p_r(q) + min(0, 3 p_ri) + min_r, p_r(q) + max(0, 3 p_ri) + max_r] (guarded, so conservative):
    // both ends are affine in the quad origin q, so the volume's corner quads decide.  A lattice
    // point inside a box then lies in [floor(lower) * os, ceil(upper) * os) as the clamp requires.
    // The window of a voxel ends less than 1/os past its box, so with boxes sticking out of the
    // footprint by less than 1 - 1/os pixels its pixels lie within one pixel of the footprint's
    // floor and ceiling: the block tile's margin.
    w.covers = 1;
    double stick = 0.0;   // largest distance a voxel box sticks out of its quad footprint
    const double qx[2] = {0.0, (double)std::max(0, v.nx - 4)}, qy[2] = {0.0, (double)(v.ny - 1)},
                 qz[2] = {0.0, (double)(v.nz - 1)};
    for (int c = 0; c < 8; c++) {
        const double q[3] = {qx[c & 1], qy[(c >> 1) & 1], qz[(c >> 2) & 1]};
        for (int r = 0; r < 2; r++) {
            const double gq = e.G[r][0] * q[0] + e.G[r][1] * q[1] + e.G[r][2] * q[2];
            const double p = r == 0 ? e.pu0 + e.pui * q[0] + e.puj * q[1] + e.puk * q[2]
                                    : e.pv0 + e.pvi * q[0] + e.pvj * q[1] + e.pvk * q[2];
            const double pi = r == 0 ? e.pui : e.pvi;
            const double boxLo = p + std::fmin(0.0, 3.0 * pi) + (double)(r == 0 ? e.uMin : e.vMin);
            const double boxHi = p + std::fmax(0.0, 3.0 * pi) + (double)(r == 0 ? e.uMax : e.vMax);
            if (!(gq + e.lo[r] <= boxLo && boxHi <= gq + e.hi[r])) w.covers = 0;
            stick = std::fmax(stick, std::fmax((gq + e.lo[r]) - boxLo, boxHi - (gq + e.hi[r])));
        }
    }
    w.masked = stick + 1.0 / os < 0.95 ? 1 : 0;
    return w.ni >= 1 && w.nj >= 1 && w.ni <= maxN && w.nj <= maxN;
}

// Whether the dist gate [distMin, distMax) passes every quad of the volume: |gn . q + g0| is the
// absolute value of an affine form of one sign over the volume (the affine footprint), so its
// extremes lie at the corner quads.
inline bool gate_passes_all(const VolG& v, const BPWin& w, double distMin, double distMax) {
    const double qx[2] = {0.0, (double)std::max(0, v.nx - 4)}, qy[2] = {0.0, (double)(v.ny - 1)},
                 qz[2] = {0.0, (double)(v.nz - 1)};
    for (int c = 0; c < 8; c++) {
        const double q[3] = {qx[c & 1], qy[(c >> 1) & 1], qz[(c >> 2) & 1]};
        const double t = std::fabs(std::fma(w.gn[0], q[0], std::fma(w.gn[1], q[1], std::fma(w.gn[2], q[2], w.g0))));
        if (!(t >= distMin && t < distMax)) return false;
    }
    return true;
}

// The chord-weighted sums of one voxel over its kNI x kNJ window, read from the block's LDS tile
// of image pixels (row pitch tw).  (i0, j0): the window's first lattice point, global; (a0, b0):
// its offset from the voxel's min-corner projection; tile origin pixel (tx0, ty0).  Only lattice
// points in the quad footprint [li0, li1) x [lj0, lj1) count, which kMode applies as:
//   kWinOpen:   not at all (the footprint covers every voxel box: w.covers);
//   kWinMasked: as a chord cap of 0 outside it (every window pixel lies in the tile);
//   kWinClamp:  by skipping the point and its read (blocks at the image edges).
// The three give the same sums for finite images: a point outside adds fma(val, 0, n) = n.
// kYCol: the y slab term does not change down a window column (YVd == 0: every tilt about y whose
// detector v axis has no y component, the u axis being the volume's y axis), so it is evaluated
// once per column, and the points whose chord is 0 for a whole column (outside the voxel's y slab,
// or outside the footprint's columns) or a whole row (outside the footprint's rows) are skipped
// with their reads.  The sums start at -0, the additive identity, so that the first point taken
// adds what the window's first point would.
constexpr int kWinOpen = 0, kWinMasked = 1, kWinClamp = 2;
template <int kOS, int kNI, int kNJ, int kMode, bool kYCol = false>
__device__ __forceinline__ void window_sums(const BPWin& w, const float* __restrict__ tile, int tw, int tx0, int ty0,
                                            int i0, int j0, float a0, float b0, int li0, int li1, int lj0, int lj1,
                                            float& num, float& den) {
    const float P0 = __builtin_fmaf(a0, w.PU, __builtin_fmaf(b0, w.PV, w.PC));
    const float Y0 = __builtin_fmaf(a0, w.YU, __builtin_fmaf(b0, w.YV, w.YC));
    // pixel of lattice point i: (i0 + di) / kOS = c0 + (ph + di) / kOS
    const int c0 = kOS == 1 ? i0 : (i0 >> 1), ph = kOS == 1 ? 0 : (i0 & 1);
    const int r0 = kOS == 1 ? j0 : (j0 >> 1), qh = kOS == 1 ? 0 : (j0 & 1);
    // tile rows are small: 24-bit multiplies (full rate; the 32-bit one is a quarter)
    const int base = __mul24(r0 - ty0, tw) + (c0 - tx0);
    const int qtw = qh ? tw : 0;
    // masked: the chord cap of each window column (m inside the footprint's columns, else 0)
    float cap[kNI];
#pragma unroll
    for (int di = 0; di < kNI; di++)
        cap[di] = (kMode == kWinMasked && !((unsigned)(i0 + di - li0) < (unsigned)(li1 - li0))) ? 0.0f : w.m;
    // kYCol: Yc = (Y0 + dj * 0) + di * YUd has the absolute value of Y0 + di * YUd at every point of
    // column di; a chord is positive only where min(A, Bq) > 0 and the cap is positive
    float Bcol[kNI];
    bool live[kNI];
#pragma unroll
    for (int di = 0; di < kNI; di++) {
        const float Yc = di == 0 ? Y0 : (di == 1 ? Y0 + w.YUd : __builtin_fmaf((float)di, w.YUd, Y0));
        Bcol[di] = w.SYh - fabsf(Yc);
        live[di] = Bcol[di] > 0.0f;
        if (kMode != kWinOpen) live[di] = live[di] && (unsigned)(i0 + di - li0) < (unsigned)(li1 - li0);
    }
    float n = kYCol ? -0.0f : 0.0f, d = kYCol ? -0.0f : 0.0f;
#pragma unroll
    for (int dj = 0; dj < kNJ; dj++) {
        const float Pr = dj == 0 ? P0 : (dj == 1 ? P0 + w.PVd : __builtin_fmaf((float)dj, w.PVd, P0));
        float Yr = dj == 0 ? Y0 : (dj == 1 ? Y0 + w.YVd : __builtin_fmaf((float)dj, w.YVd, Y0));
        const int rowOff = kOS == 1 ? dj * tw : (dj >> 1) * tw + ((dj & 1) ? qtw : 0);
        const int rowBase = base + rowOff;
        const int rowOdd = rowBase + ph;   // odd di of the 2x lattice: pixel c0 + ph + (di - 1) / 2
        bool rowOk = true;
        if (kMode != kWinOpen) rowOk = (unsigned)(j0 + dj - lj0) < (unsigned)(lj1 - lj0);
        // masked: a row outside the footprint moves beyond every y slab, so its chords are 0 (the
        // med3 of a negative value, 0 and the column cap), as a cap of 0 would make them
        if (kMode == kWinMasked) Yr = rowOk ? Yr : kWinFar;
        if (kYCol && !rowOk) continue;
#pragma unroll
        for (int di = 0; di < kNI; di++) {
            if (kYCol && !live[di]) continue;
            const float Pc = di == 0 ? Pr : (di == 1 ? Pr + w.PUd : __builtin_fmaf((float)di, w.PUd, Pr));
            const float Yc = di == 0 ? Yr : (di == 1 ? Yr + w.YUd : __builtin_fmaf((float)di, w.YUd, Yr));
            // min(m, P, SP - P, Y, SY - Y) clamped at 0 (cap 0: exactly 0)
            const float A = w.SPh - fabsf(Pc), Bq = kYCol ? Bcol[di] : w.SYh - fabsf(Yc);
            const float mcap = (kMode == kWinMasked && !kYCol) ? cap[di] : w.m;
            float chord = __builtin_amdgcn_fmed3f(fminf(A, Bq), 0.0f, mcap);
            int addr = kOS == 1 ? rowBase + di : ((di & 1) ? rowOdd + (di >> 1) : rowBase + (di >> 1));
            if (kMode == kWinClamp && !kYCol) {
                const bool ok = rowOk && (unsigned)(i0 + di - li0) < (unsigned)(li1 - li0);
                chord = ok ? chord : 0.0f;
                addr = ok ? addr : 0;
            }
            const float val = tile[addr];
            if (!kYCol && di == 0 && dj == 0) {
                n = val * chord;
                d = chord;
            } else {
                n = __builtin_fmaf(val, chord, n);
                d += chord;
            }
        }
    }
    num = n;
    den = d;
}

// The window form's plan for one tilt: the form applies, the footprint of a 64 x 4 x 8-voxel block
// fits the LDS tile and the window size has a compiled instance (OS 1: 1-3, OS 2: 2-4 lattice
// points per axis).  The per-call and the batched sub-volume backprojections take the same decision,
// so they stay bit-identical.
constexpr int kWinBlockZ = 8;   // slices per block of the per-call kernel (two quads per thread)
inline bool window_plan(const VolG& v, const TiltG& g, int os, BPWin& w) {
    if (os != 1 && os != 2) return false;
    const BPExact e = prepare_exact(v, g, 64, 4, kWinBlockZ);
    if ((long long)(e.tileW + 2) * (long long)(e.tileH + 2) > kTileFloats) return false;
    if (!prepare_window(v, g, e, os, os == 1 ? 3 : 4, w)) return false;
    return os == 1 ? true : (w.ni >= 2 && w.nj >= 2);
}

// The update of one quad (voxels x .. x+3 of row y, slice z) for one tilt: out[k] += lambda *
// num / den over each voxel's window, from an image tile of row pitch tw whose pixel (tx0, ty0) is
// tile[0] (the whole image with tw = pitch and tx0 = ty0 = 0).  Only lattice points in the quad's
// footprint clamped to [0, W-1] x [0, H-1] count, as in the per-box loop.  interior: the block's
// footprint lies in the image with a pixel to spare and the tile holds it with a one-pixel margin;
// then the footprint test is skipped (w.covers) or applied as a chord cap (w.masked), otherwise
// each point outside is skipped with its read.  Returns whether a voxel changed.
template <int kOS, int kNI, int kNJ, bool kYCol = false>
__device__ __forceinline__ bool window_quad(const BPWin& w, const float* __restrict__ tile, int tw, int tx0, int ty0,
                                            bool interior, int W, int H, unsigned x, unsigned y, unsigned z,
                                            float lambda, float out[4]) {
    constexpr float kDelta = 1.0f / (float)kOS;
    const double qx = (double)x, qy = (double)y, qz = (double)z;
    int li0 = 0, li1 = 0, lj0 = 0, lj1 = 0;
    const int mode = interior && w.covers ? kWinOpen : (interior && w.masked ? kWinMasked : kWinClamp);
    if (mode != kWinOpen) {
        // the quad's footprint clamped to [0, W-1] x [0, H-1] (footprint_affine): lattice points
        // [px0 * OS, px1 * OS) x [py0 * OS, py1 * OS); in an interior block the clamps are no-ops
        const double a0 = fma(w.G[0][0], qx, fma(w.G[0][1], qy, w.G[0][2] * qz));
        const double a1 = fma(w.G[1][0], qx, fma(w.G[1][1], qy, w.G[1][2] * qz));
        if (mode == kWinMasked) {
            li0 = kOS * (int)floor(a0 + w.lo[0]);
            li1 = kOS * (int)ceil(a0 + w.hi[0]);
            lj0 = kOS * (int)floor(a1 + w.lo[1]);
            lj1 = kOS * (int)ceil(a1 + w.hi[1]);
        } else {
            li0 = kOS * (int)fmin(fmax(floor(a0 + w.lo[0]), 0.0), (double)(W - 1));
            li1 = kOS * (int)fmin(fmax(ceil(a0 + w.hi[0]), 0.0), (double)(W - 1));
            lj0 = kOS * (int)fmin(fmax(floor(a1 + w.lo[1]), 0.0), (double)(H - 1));
            lj1 = kOS * (int)fmin(fmax(ceil(a1 + w.hi[1]), 0.0), (double)(H - 1));
        }
        if (li1 <= li0 || lj1 <= lj0) return false;
    }
    // corner projection of the quad's first voxel, split into an integer pixel and a float offset
    const double Uq = fma(w.pui, qx, fma(w.puj, qy, fma(w.puk, qz, w.pu0)));
    const double Vq = fma(w.pvi, qx, fma(w.pvj, qy, fma(w.pvk, qz, w.pv0)));
    const double fUq = floor(Uq), fVq = floor(Vq);
    const int iq = (int)fUq * kOS, jq = (int)fVq * kOS;
    const float fu = (float)(Uq - fUq), fv = (float)(Vq - fVq);
    const float dui = (float)w.pui, dvi = (float)w.pvi;
    bool any = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float du = __builtin_fmaf((float)k, dui, fu);
        const float dv = __builtin_fmaf((float)k, dvi, fv);
        // the window's first lattice point, as a float (an integer) and an int
        const float fi0 = ceilf((du + w.uMin) * (float)kOS - 0.5f);
        const float fj0 = ceilf((dv + w.vMin) * (float)kOS - 0.5f);
        const int i0 = (int)fi0, j0 = (int)fj0;
        const float a0 = __builtin_fmaf(fi0 + 0.5f, kDelta, -du);
        const float b0 = __builtin_fmaf(fj0 + 0.5f, kDelta, -dv);
        float num, den;
        if (mode == kWinOpen)
            window_sums<kOS, kNI, kNJ, kWinOpen, kYCol>(w, tile, tw, tx0, ty0, iq + i0, jq + j0, a0, b0, 0, 0, 0, 0, num,
                                                 den);
        else if (mode == kWinMasked)
            window_sums<kOS, kNI, kNJ, kWinMasked, kYCol>(w, tile, tw, tx0, ty0, iq + i0, jq + j0, a0, b0, li0, li1, lj0,
                                                   lj1, num, den);
        else
            window_sums<kOS, kNI, kNJ, kWinClamp, kYCol>(w, tile, tw, tx0, ty0, iq + i0, jq + j0, a0, b0, li0, li1, lj0,
                                                  lj1, num, den);
        if (den > 0.0f) {
            out[k] += (lambda * num) * __builtin_amdgcn_rcpf(den);
            any = true;
        }
    }
    return any;
}

// window_quad with the window size chosen at run time (w.ni, w.nj) among the compiled instances.
template <int kOS>
__device__ __forceinline__ bool window_quad_any(const BPWin& w, const float* __restrict__ tile, int tw, int tx0,
                                                int ty0, bool interior, int W, int H, unsigned x, unsigned y,
                                                unsigned z, float lambda, float out[4]) {
    constexpr int kLo = kOS == 1 ? 1 : 2;
#define ARTIA_WIN_CASE(NI, NJ)                                                                                   \
    if (w.ni == (NI) && w.nj == (NJ))                                                                           \
        return window_quad<kOS, (NI), (NJ)>(w, tile, tw, tx0, ty0, interior, W, H, x, y, z, lambda, out);
    ARTIA_WIN_CASE(kLo + 1, kLo + 1)
    ARTIA_WIN_CASE(kLo, kLo)
    ARTIA_WIN_CASE(kLo, kLo + 1)
    ARTIA_WIN_CASE(kLo + 1, kLo)
    ARTIA_WIN_CASE(kLo + 2, kLo + 1)
    ARTIA_WIN_CASE(kLo + 1, kLo + 2)
    ARTIA_WIN_CASE(kLo + 2, kLo + 2)
    ARTIA_WIN_CASE(kLo, kLo + 2)
    ARTIA_WIN_CASE(kLo + 2, kLo)
#undef ARTIA_WIN_CASE
    return false;
}

}  // namespace artia
