// Sub-volume WBP (EmSARTSubVolumes, EmSartSubVolumes.cpp:611-707 with
// Reconstructor::BackProjection{NoCTF,CTF}(vol, subVolumes, ...), Reconstructor.cpp:894-1034).
//
// Every (sub-volume s, tilt t) pair is one backprojection with lambda 1 of tilt t's (WBP-filtered)
// projection into sub-volume s, seen through tilt t's geometry shifted by the pair's extra shift.
// Without CTF the gate is [0, 1e13) and the image is POINT sampled; with CTF each tilt's projection
// is filtered once per defocus slab (no border crop) and every quad takes the slab image whose gate
// holds its first macro-cell corner, LINEAR sampled.
//
// plain:   the reference's schedule, one backprojection launch per active (s, t) and slab.
// batched: one persistent launch over a work queue of (sub-volume, voxel tile) items, sorted by
//          sub-volume size (largest first) and popped through a device counter.  A thread keeps its
//          voxels in registers across all tilts and writes them once.  The per-tilt update is the
//          plain form's arithmetic, added in the same order, so the two forms agree bit for bit.
//          Without CTF the counter reset and the persistent kernel are captured in one hipGraph.
//          With CTF the tilts go in chunks whose filtered slab images fit a stack budget: every
//          tilt of a chunk is filtered (one forward transform per tilt, batched inverse transforms
//          per slab), then one persistent launch adds the chunk's tilts in order.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "artia_internal.hpp"
#include "bp_common.hpp"
#include "bp_window.hpp"

namespace artia {
namespace {

struct SubvolDev {
    float* vol;
    int nx, ny, nz;
};

// Per (s, t) pair.  mode 0: inactive; 1: LDS-tile eligible (affine footprint, POINT, no
// magnification anisotropy); 2: global-memory path; 3: CTF slab stack, LDS-tile eligible for
// LINEAR fetches (affine footprint, no magnification anisotropy).
// Default bytes of filtered slab images a CTF plan keeps at once (a chunk of tilts); at most half
// the free device memory, at least one tilt's stack.
constexpr size_t kSubvolStackBytes = (size_t)16 << 30;

struct PairDev {
    BPExact e;
    BPWin w;     // the window form (bp_window.hpp) when win != 0, as the per-call backprojection takes it
    TiltG g;
    VolG v;
    int mode;
    int win;
};

// CTF: tilt t's slab bounds and where its filtered slab images start in the stack (in slab images).
struct TiltSlab {
    SlabArgs s;
    long long stack_off;
};

// One block's work: sub-volume s, tile origin (quad x, y, z) and tile shape (quads x rows x slices
// per quad layer); a thread handles kQPT quads stacked along z.
struct Item {
    int s, qx0, y0, z0;
    int qx, ty, tz, pad;
};

__global__ void reset_counter_kernel(unsigned* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0u;
}

// Update (lambda = 1) of one quad for one tilt; returns false when the quad is gated out.
template <int kOS, bool kTile, int kPM>
__device__ __forceinline__ void quad_update(const PairDev& p, const float* __restrict__ tile, int tw, int tx0, int ty0,
                                            const float* __restrict__ img, int W, int H, size_t pitch, int filter,
                                            int os, unsigned x, unsigned y, unsigned z, const SlabArgs& sl,
                                            size_t slabStride, float out[4]) {
    const BPExact& e = p.e;
    int px0, px1, py0, py1;
    const float* src = img;
    if (sl.count > 0) {
        // first macro-cell corner distance, as the gate of the per-slab launches computes it
        const int k = find_slab(sl, gate_distance(e, p.v, p.g, x, y, z));
        if (k < 0) return;
        src = img + (size_t)k * slabStride;
    }
    if (e.affine) {
        if (!footprint_affine(e, x, y, z, W, H, -1.0, INFINITY, px0, px1, py0, py1)) return;
    } else {
        if (!footprint_literal(p.v, p.g, x, y, z, W, H, -1.0, INFINITY, px0, px1, py0, py1)) return;
    }
    const int O = kOS > 0 ? kOS : os;
    const int nI = (px1 - px0) * O, nJ = (py1 - py0) * O;
    if (nI <= 0 || nJ <= 0) return;
    const float osf = (float)O, delta = 1.0f / osf;
    const float du0 = (float)(e.pu0 + (double)x * e.pui + (double)y * e.puj + (double)z * e.puk - (double)px0);
    const float dv0 = (float)(e.pv0 + (double)x * e.pvi + (double)y * e.pvj + (double)z * e.pvk - (double)py0);
    const float dui = (float)e.pui, dvi = (float)e.pvi;
    const int tileOff = kTile ? (py0 - ty0) * tw + (px0 - tx0) : 0;
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
            voxel_sums_global(e, p.g, src, W, H, pitch, filter, px0, py0, O, delta, du, dv, i0, i1, j0, j1, num, den);
        den = fmaxf(0.0f, den);
        if (den != 0.0f) out[k] += (1.0f * num) / den;
    }
}

// quad_update for a CTF pair whose quad lies in slab k: LINEAR fetches from the slab's LDS tile
// (pixels [tx0, tx1] x [ty0, ty1]) when the quad's fetches stay inside it, else from global memory.
template <int kPM>
__device__ __forceinline__ void quad_update_stack(const PairDev& p, const float* __restrict__ slabTile, bool tiled,
                                                  int tw, int tx0, int ty0, int tx1, int ty1,
                                                  const float* __restrict__ src, int W, int H, size_t pitch, int os,
                                                  unsigned x, unsigned y, unsigned z, float out[4]) {
    const BPExact& e = p.e;
    int px0, px1, py0, py1;
    if (!footprint_affine(e, x, y, z, W, H, -1.0, INFINITY, px0, px1, py0, py1)) return;
    const int nI = (px1 - px0) * os, nJ = (py1 - py0) * os;
    if (nI <= 0 || nJ <= 0) return;
    // the quad's fetches read pixels max(p0 - 1, 0) .. p1 on each axis
    const bool fromTile = tiled && max(px0 - 1, 0) >= tx0 && px1 <= tx1 && max(py0 - 1, 0) >= ty0 && py1 <= ty1;
    const bool inner = px0 >= 1 && px1 <= W - 1 && py0 >= 1 && py1 <= H - 1;   // no fetch clamps
    const float osf = (float)os, delta = 1.0f / osf;
    const float du0 = (float)(e.pu0 + (double)x * e.pui + (double)y * e.puj + (double)z * e.puk - (double)px0);
    const float dv0 = (float)(e.pv0 + (double)x * e.pvi + (double)y * e.pvj + (double)z * e.pvk - (double)py0);
    const float dui = (float)e.pui, dvi = (float)e.pvi;
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
            voxel_sums_global(e, p.g, src, W, H, pitch, ARTIA_FILTER_LINEAR, px0, py0, os, delta, du, dv, i0, i1, j0,
                              j1, num, den);
        den = fmaxf(0.0f, den);
        if (den != 0.0f) out[k] += (1.0f * num) / den;
    }
}

// Persistent backprojection over the work queue.  Tilts [t0, t1) of every item, added in tilt order
// to the voxels held in registers; image of tilt t at images + (t - t0) * imgStride (no CTF), or
// slab k of tilt t at images + (tslab[t].stack_off + k) * imgStride (CTF).  kCtf: compiled for the
// CTF plans (pairs of modes 3 and 2) or for the others (modes 1 and 2), so that each holds only its
// own paths' registers.  Six waves per SIMD (80 registers, a few dozen spilled) measured fastest:
// 0.246 -> 0.202 s for 2000 cubes without CTF, 2.86 -> 2.69 s with CTF, against the unbounded
// allocation (three to four waves); five, seven and eight waves were slower.
template <int kOS, int kQPT, bool kCtf>
__global__ void __launch_bounds__(256, 6) subvol_bp_kernel(const Item* __restrict__ items, int nitems,
                                                        unsigned* __restrict__ counter,
                                                        const SubvolDev* __restrict__ subs,
                                                        const PairDev* __restrict__ pairs, int ntilts, int t0, int t1,
                                                        const float* __restrict__ images, size_t imgStride, int W,
                                                        int H, size_t pitch, int os, int filter, SlabArgs slabs,
                                                        const TiltSlab* __restrict__ tslab) {
    __shared__ float tile[kTileFloats];
    __shared__ int cur;
    __shared__ int kr[2];
    const int tid = (int)threadIdx.x;
    for (;;) {
        if (tid == 0) cur = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = cur;
        __syncthreads();
        if (it >= nitems) break;
        const Item item = items[it];
        const SubvolDev sv = subs[item.s];
        const int qx = tid % item.qx, ty = (tid / item.qx) % item.ty, tz = tid / (item.qx * item.ty);
        const bool lane = tz < item.tz;   // tile shapes may leave threads idle
        const int xq = item.qx0 + qx, y = item.y0 + ty;
        const unsigned x = (unsigned)xq * 4u;
        float acc[kQPT][4];
        bool in[kQPT];
#pragma unroll
        for (int q = 0; q < kQPT; q++) {
            const int z = item.z0 + tz + q * item.tz;
            in[q] = lane && xq < sv.nx / 4 && y < sv.ny && z < sv.nz;
            if (in[q]) {
                const float4 vb = *reinterpret_cast<const float4*>(
                    sv.vol + ((size_t)z * (size_t)sv.ny + (size_t)y) * (size_t)sv.nx + x);
                acc[q][0] = vb.x;
                acc[q][1] = vb.y;
                acc[q][2] = vb.z;
                acc[q][3] = vb.w;
            }
        }
        const int bX = item.qx * 4, bY = item.ty, bZ = item.tz * kQPT;
        for (int t = t0; t < t1; t++) {
            const PairDev& p = pairs[(size_t)item.s * (size_t)ntilts + (size_t)t];
            if (p.mode == 0) continue;
            // no CTF: tilt t's image; CTF: the tilt's slab stack
            const float* img = images + (size_t)(t - t0) * imgStride;
            if (kCtf && p.mode == 3) {
                const TiltSlab ts = tslab[t];
                img = images + (size_t)ts.stack_off * imgStride;
                // slab of each quad, and the item's slab range
                if (tid == 0) {
                    kr[0] = INT_MAX;
                    kr[1] = INT_MIN;
                }
                __syncthreads();
                int ks[kQPT];
#pragma unroll
                for (int q = 0; q < kQPT; q++) {
                    const unsigned z = (unsigned)(item.z0 + tz + q * item.tz);
                    ks[q] = in[q] ? find_slab(ts.s, gate_distance(p.e, p.v, p.g, x, (unsigned)y, z)) : -1;
                    if (ks[q] >= 0) {
                        atomicMin(&kr[0], ks[q]);
                        atomicMax(&kr[1], ks[q]);
                    }
                }
                __syncthreads();
                const int k0 = kr[0], k1 = kr[1];
                // the item's footprint with two pixels of margin (see bp_stack_kernel), per slab
                const BPExact& e = p.e;
                const double ox = (double)(item.qx0 * 4), oy = (double)item.y0, oz = (double)item.z0;
                const double ext[3] = {(double)(bX - 4), (double)(bY - 1), (double)(bZ - 1)};
                double lo[2], hi[2];
                for (int r = 0; r < 2; r++) {
                    const double base = e.G[r][0] * ox + e.G[r][1] * oy + e.G[r][2] * oz;
                    lo[r] = base + e.lo[r];
                    hi[r] = base + e.hi[r];
                    for (int k = 0; k < 3; k++) {
                        lo[r] += fmin(0.0, e.G[r][k] * ext[k]);
                        hi[r] += fmax(0.0, e.G[r][k] * ext[k]);
                    }
                }
                const int tx0 = max(0, (int)floor(lo[0]) - 2), ty0 = max(0, (int)floor(lo[1]) - 2);
                const int tx1 = min(W - 1, (int)ceil(hi[0]) + 2), ty1 = min(H - 1, (int)ceil(hi[1]) + 2);
                const int tw = max(0, tx1 - tx0 + 1), plane = tw * max(0, ty1 - ty0 + 1);
                const bool tiled = k0 <= k1 && plane > 0 &&
                                   (long long)plane * (long long)(k1 - k0 + 1) <= (long long)kTileFloats;
                if (tiled) {
                    const int n = plane * (k1 - k0 + 1);
                    for (int idx = tid; idx < n; idx += 256) {
                        const int sl = idx / plane, rem = idx - sl * plane;
                        const int r = rem / tw, c = rem - r * tw;
                        tile[idx] = img[(size_t)(k0 + sl) * imgStride + (size_t)(ty0 + r) * pitch + (size_t)(tx0 + c)];
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int q = 0; q < kQPT; q++) {
                    if (ks[q] < 0) continue;
                    const unsigned z = (unsigned)(item.z0 + tz + q * item.tz);
                    const float* st = tile + (tiled ? (ks[q] - k0) * plane : 0);
                    const float* src = img + (size_t)ks[q] * imgStride;
                    if (p.e.parallel == 0)
                        quad_update_stack<0>(p, st, tiled, tw, tx0, ty0, tx1, ty1, src, W, H, pitch, os, x,
                                             (unsigned)y, z, acc[q]);
                    else if (p.e.parallel == 2)
                        quad_update_stack<2>(p, st, tiled, tw, tx0, ty0, tx1, ty1, src, W, H, pitch, os, x,
                                             (unsigned)y, z, acc[q]);
                    else
                        quad_update_stack<-1>(p, st, tiled, tw, tx0, ty0, tx1, ty1, src, W, H, pitch, os, x,
                                              (unsigned)y, z, acc[q]);
                }
                __syncthreads();   // the tile and kr are reused by the next tilt or item
                continue;
            }
            bool useTile = false;
            int tx0 = 0, ty0 = 0, tw = 0, th = 0;
            if (!kCtf && p.mode == 1) {
                const BPExact& e = p.e;
                const double ox = (double)(item.qx0 * 4), oy = (double)item.y0, oz = (double)item.z0;
                double lo[2], hi[2];
                const double ext[3] = {(double)(bX - 4), (double)(bY - 1), (double)(bZ - 1)};
                for (int r = 0; r < 2; r++) {
                    const double base = e.G[r][0] * ox + e.G[r][1] * oy + e.G[r][2] * oz;
                    lo[r] = base + e.lo[r];
                    hi[r] = base + e.hi[r];
                    for (int k = 0; k < 3; k++) {
                        lo[r] += fmin(0.0, e.G[r][k] * ext[k]);
                        hi[r] += fmax(0.0, e.G[r][k] * ext[k]);
                    }
                }
                tx0 = max(0, (int)floor(lo[0]) - 1);
                ty0 = max(0, (int)floor(lo[1]) - 1);
                const int tx1 = min(W - 1, (int)ceil(hi[0]) + 1);
                const int ty1 = min(H - 1, (int)ceil(hi[1]) + 1);
                tw = max(0, tx1 - tx0 + 1);
                th = max(0, ty1 - ty0 + 1);
                useTile = tw * th <= kTileFloats;
            }
            if ((kOS == 1 || kOS == 2) && !kCtf && p.win) {
                // the per-call backprojection's window form (POINT, lambda 1), from the LDS tile or,
                // when the item's footprint does not fit it, from the image with clamped lattices
                const bool interior = useTile && tx0 >= 1 && ty0 >= 1 && tx0 + tw <= W - 1 && ty0 + th <= H - 1;
                // odd LDS row stride where it fits (rows of a wave in different banks)
                const int ts = (tw | 1) * th <= kTileFloats ? (tw | 1) : tw;
                if (useTile) {
                    for (int idx = tid; idx < tw * th; idx += 256) {
                        const int r = idx / tw, c = idx - r * tw;
                        tile[r * ts + c] = img[(size_t)(ty0 + r) * pitch + (size_t)(tx0 + c)];
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int q = 0; q < kQPT; q++) {
                    if (!in[q]) continue;
                    const unsigned z = (unsigned)(item.z0 + tz + q * item.tz);
                    if (useTile)
                        window_quad_any<(kOS == 1 ? 1 : 2)>(p.w, tile, ts, tx0, ty0, interior, W, H, x, (unsigned)y, z,
                                                           1.0f, acc[q]);
                    else
                        window_quad_any<(kOS == 1 ? 1 : 2)>(p.w, img, (int)pitch, 0, 0, false, W, H, x, (unsigned)y,
                                                           z, 1.0f, acc[q]);
                }
                if (useTile) __syncthreads();
                continue;
            }
            if (useTile) {
                for (int idx = tid; idx < tw * th; idx += 256) {
                    const int r = idx / tw, c = idx - r * tw;
                    tile[idx] = img[(size_t)(ty0 + r) * pitch + (size_t)(tx0 + c)];
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < kQPT; q++) {
                    if (!in[q]) continue;
                    const unsigned z = (unsigned)(item.z0 + tz + q * item.tz);
                    if (p.e.parallel == 0)
                        quad_update<kOS, true, 0>(p, tile, tw, tx0, ty0, img, W, H, pitch, filter, os, x, (unsigned)y,
                                                  z, slabs, imgStride, acc[q]);
                    else if (p.e.parallel == 2)
                        quad_update<kOS, true, 2>(p, tile, tw, tx0, ty0, img, W, H, pitch, filter, os, x, (unsigned)y,
                                                  z, slabs, imgStride, acc[q]);
                    else
                        quad_update<kOS, true, -1>(p, tile, tw, tx0, ty0, img, W, H, pitch, filter, os, x,
                                                   (unsigned)y, z, slabs, imgStride, acc[q]);
                }
                __syncthreads();
            } else {
#pragma unroll
                for (int q = 0; q < kQPT; q++) {
                    if (!in[q]) continue;
                    const unsigned z = (unsigned)(item.z0 + tz + q * item.tz);
                    quad_update<kOS, false, -1>(p, tile, 0, 0, 0, img, W, H, pitch, filter, os, x, (unsigned)y, z,
                                                slabs, imgStride, acc[q]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kQPT; q++) {
            if (!in[q]) continue;
            const int z = item.z0 + tz + q * item.tz;
            *reinterpret_cast<float4*>(sv.vol + ((size_t)z * (size_t)sv.ny + (size_t)y) * (size_t)sv.nx + x) =
                make_float4(acc[q][0], acc[q][1], acc[q][2], acc[q][3]);
        }
    }
}

// Largest power of two <= cap dividing n (>= 1).
int pow2_divisor(int n, int cap) {
    int q = 1;
    while (q * 2 <= cap && n % (q * 2) == 0) q *= 2;
    return q;
}

// Tile shape of a sub-volume for a policy: quads along x, rows along y, slices along z per quad
// layer (quads * rows * slices == 256 threads).
void tile_shape(int nq, int policy, int& qx, int& ty, int& tz) {
    if (policy == 0) {
        qx = pow2_divisor(nq, 16);
        ty = 4;
    } else {
        qx = std::max(1, pow2_divisor(nq, 16) / 2);
        ty = 8;
    }
    tz = std::max(1, 256 / (qx * ty));
}

void build_items(const std::vector<SubvolDev>& subs, const std::vector<int>& order, int policy, int qpt,
                 std::vector<Item>& items) {
    items.clear();
    for (int s : order) {
        const SubvolDev& sv = subs[(size_t)s];
        int qx, ty, tz;
        tile_shape(sv.nx / 4, policy, qx, ty, tz);
        const int zext = tz * qpt;
        for (int z0 = 0; z0 < sv.nz; z0 += zext)
            for (int y0 = 0; y0 < sv.ny; y0 += ty)
                for (int q0 = 0; q0 < sv.nx / 4; q0 += qx) items.push_back(Item{s, q0, y0, z0, qx, ty, tz, 0});
    }
}

void launch_subvol(artia_ctx* ctx, hipStream_t stream, int qpt, int grid, const Item* items, int nitems,
                   unsigned* counter, const SubvolDev* subs, const PairDev* pairs, int ntilts, int t0, int t1,
                   const float* images, size_t imgStride, int W, int H, size_t pitch, int os, int filter,
                   SlabArgs slabs, const TiltSlab* tslab) {
    (void)ctx;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
#define ARTIA_SUBVOL_LAUNCH(OS, QPT)                                                                                 \
    if (tslab)                                                                                                      \
        hipLaunchKernelGGL((subvol_bp_kernel<OS, QPT, true>), dim3((unsigned)grid), dim3(256), 0, stream, items,    \
                           nitems, counter, subs, pairs, ntilts, t0, t1, images, imgStride, W, H, pitch, os, filter,  \
                           slabs, tslab);                                                                           \
    else                                                                                                            \
        hipLaunchKernelGGL((subvol_bp_kernel<OS, QPT, false>), dim3((unsigned)grid), dim3(256), 0, stream, items,   \
                           nitems, counter, subs, pairs, ntilts, t0, t1, images, imgStride, W, H, pitch, os, filter,  \
                           slabs, tslab)
    const int o = (os == 1 || os == 2 || os == 4) ? os : 0;
    if (qpt == 2) {
        if (o == 1) ARTIA_SUBVOL_LAUNCH(1, 2);
        else if (o == 2) ARTIA_SUBVOL_LAUNCH(2, 2);
        else if (o == 4) ARTIA_SUBVOL_LAUNCH(4, 2);
        else ARTIA_SUBVOL_LAUNCH(0, 2);
    } else {
        if (o == 1) ARTIA_SUBVOL_LAUNCH(1, 1);
        else if (o == 2) ARTIA_SUBVOL_LAUNCH(2, 1);
        else if (o == 4) ARTIA_SUBVOL_LAUNCH(4, 1);
        else ARTIA_SUBVOL_LAUNCH(0, 1);
    }
#undef ARTIA_SUBVOL_LAUNCH
}

// Slab starts of one tilt: the reference's `for (float ray = t_in; ray < t_out; ray += step)`.
void slab_starts(float t_in, float t_out, float step, std::vector<float>& out) {
    out.clear();
    for (float ray = t_in; ray < t_out; ray += step) out.push_back(ray);
}

struct Inputs {
    const artia_subvol* subs;
    int nsub;
    const artia_tilt_geom* geoms;
    const unsigned char* active;
    int ntilts;
    const float* images;
    int W, H;
    size_t pitch;
    int os;
    const artia_ctf_params* ctf;
    const float* slab_range;
    const artia_tilt_defocus* defocus;
};

artia_status validate(artia_ctx* ctx, const char* fn, const Inputs& in) {
    if (!in.subs || !in.geoms || !in.images || in.nsub <= 0 || in.ntilts <= 0 || in.W <= 0 || in.H <= 0 ||
        in.pitch < (size_t)in.W || in.os <= 0 || in.os > 64)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: invalid arguments", fn);
    if (in.ctf && (!in.slab_range || !in.defocus || !(in.ctf->slice_thickness_nm > 0.0f) ||
                   !(in.ctf->pixelsize_nm > 0.0f) || !(in.ctf->voltage_kV > 0.0f) || !(in.ctf->wiener >= 0.0f)))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: incomplete CTF arguments", fn);
    for (int s = 0; s < in.nsub; s++) {
        const artia_subvol& sv = in.subs[s];
        if (!sv.vol || sv.geom.nx <= 0 || sv.geom.ny <= 0 || sv.geom.nz <= 0 || sv.geom.nx % 4 != 0 ||
            (reinterpret_cast<uintptr_t>(sv.vol) & 15) != 0)
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT,
                        "%s: sub-volume %d needs a 16-byte aligned volume with x dimension a multiple of 4", fn, s);
        for (int t = 0; t < in.ntilts; t++) {
            const size_t i = (size_t)s * (size_t)in.ntilts + (size_t)t;
            if ((!in.active || in.active[i]) && !valid_geometry(&sv.geom, &in.geoms[i]))
                return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: invalid geometry for sub-volume %d, tilt %d", fn, s,
                            t);
        }
    }
    if (in.ctf) {
        const float step = in.ctf->slice_thickness_nm / in.ctf->pixelsize_nm;
        for (int t = 0; t < in.ntilts; t++) {
            const float a = in.slab_range[2 * t], b = in.slab_range[2 * t + 1];
            if (!std::isfinite(a) || !std::isfinite(b) || !(step > 0.0f) || (double)(b - a) / step > 65536.0 ||
                (b > a && a + step == a))
                return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: slab range of tilt %d holds too many slabs", fn, t);
        }
    }
    return ARTIA_OK;
}

}  // namespace
}  // namespace artia

using namespace artia;

struct artia_subvol_plan {
    artia_ctx* ctx = nullptr;
    artia::Inputs in{};
    std::vector<artia_tilt_defocus> defocus;
    std::vector<float> slab_range;
    artia_ctf_params ctfp{};
    bool ctf = false;
    std::vector<SubvolDev> subs;
    std::vector<int> order;   // sub-volumes by size, largest first
    SubvolDev* d_subs = nullptr;
    PairDev* d_pairs = nullptr;
    Item* d_items = nullptr;
    unsigned* d_counter = nullptr;
    float* d_bounds = nullptr;     // CTF: slab starts of every tilt, concatenated
    float* d_stack = nullptr;      // CTF: the filtered slab images of the current chunk of tilts
    TiltSlab* d_tslab = nullptr;   // CTF: per tilt, its slab bounds and stack offset
    std::vector<int> bound_count;
    std::vector<long long> stack_off;   // per tilt, within its chunk (in slab images)
    std::vector<int> chunks;            // tilt chunk starts, then ntilts
    int max_slabs = 0;
    int nitems = 0;
    artia_subvol_tuning tuning{};
    int grid = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~artia_subvol_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        for (void* p : {(void*)d_subs, (void*)d_pairs, (void*)d_items, (void*)d_counter, (void*)d_bounds,
                        (void*)d_stack, (void*)d_tslab})
            if (p) (void)hipFree(p);
    }
};

namespace {

artia_status upload(artia_ctx* ctx, void** dst, const void* src, size_t bytes, const char* what) {
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    artia_status st = check_hip(ctx, hipMalloc(dst, bytes > 0 ? bytes : 16), what);
    if (st != ARTIA_OK) return st;
    if (bytes > 0) st = check_hip(ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), what);
    return st;
}

artia_status set_items(artia_subvol_plan* p, int policy, int qpt) {
    std::vector<Item> items;
    build_items(p->subs, p->order, policy, qpt, items);
    p->nitems = (int)items.size();
    return upload(p->ctx, (void**)&p->d_items, items.data(), items.size() * sizeof(Item), "sub-volume queue");
}

// CTF: tilt t's slab images, filtered into its place in the stack (batched FFTs, no border crop).
artia_status filter_tilt(artia_subvol_plan* p, int t) {
    const Inputs& in = p->in;
    const int count = p->bound_count[(size_t)t];
    std::vector<float> starts;
    slab_starts(p->slab_range[2 * (size_t)t], p->slab_range[2 * (size_t)t + 1],
                p->ctfp.slice_thickness_nm / p->ctfp.pixelsize_nm, starts);
    std::vector<float> dmin((size_t)count), dmax((size_t)count);
    for (int k = 0; k < count; k++)
        artia_geom_defocus_minmax(starts[(size_t)k], &p->defocus[(size_t)t], &p->ctfp, &dmin[(size_t)k],
                                  &dmax[(size_t)k]);
    const size_t slabStride = (size_t)in.H * in.pitch;
    return ctf_filter_slabs(p->ctx, in.images + (size_t)t * slabStride, 0, count, in.W, in.H, in.pitch,
                            p->d_stack + (size_t)p->stack_off[(size_t)t] * slabStride, slabStride, nullptr, nullptr, nullptr, p->ctfp, dmin.data(), dmax.data(),
                            p->defocus[(size_t)t].astig_angle_deg, 0);
}

// CTF: the persistent backprojection of tilts [t0, t1) from their filtered stacks.
void launch_tilts(artia_subvol_plan* p, int t0, int t1, hipStream_t stream) {
    const Inputs& in = p->in;
    const size_t slabStride = (size_t)in.H * in.pitch;
    launch_subvol(p->ctx, stream, p->tuning.quads_per_thread, p->grid, p->d_items, p->nitems, p->d_counter,
                  p->d_subs, p->d_pairs, in.ntilts, t0, t1, p->d_stack, slabStride, in.W, in.H, in.pitch, in.os,
                  ARTIA_FILTER_LINEAR, SlabArgs{nullptr, 0, 0.0f}, p->d_tslab);
}

// Enqueues one full backprojection of the plan on `stream` (the graph body for the no-CTF case).
artia_status enqueue(artia_subvol_plan* p, hipStream_t stream) {
    const Inputs& in = p->in;
    if (!p->ctf) {
        launch_subvol(p->ctx, stream, p->tuning.quads_per_thread, p->grid, p->d_items, p->nitems, p->d_counter,
                      p->d_subs, p->d_pairs, in.ntilts, 0, in.ntilts, in.images, (size_t)in.H * in.pitch, in.W, in.H,
                      in.pitch, in.os, ARTIA_FILTER_POINT, SlabArgs{nullptr, 0, 0.0f}, nullptr);
        return check_hip(p->ctx, hipGetLastError(), "sub-volume backprojection launch");
    }
    // per chunk of tilts whose slab stacks fit together: filter them all, then one persistent
    // launch adds the chunk's tilts to the voxels in registers
    artia_ctx* ctx = p->ctx;
    hipStream_t saved = ctx->stream;
    ctx->stream = stream;
    for (size_t c = 0; c + 1 < p->chunks.size(); c++) {
        const int t0 = p->chunks[c], t1 = p->chunks[c + 1];
        bool any = false;
        for (int t = t0; t < t1; t++) {
            if (p->bound_count[(size_t)t] == 0) continue;
            any = true;
            const artia_status st = filter_tilt(p, t);
            if (st != ARTIA_OK) {
                ctx->stream = saved;
                return st;
            }
        }
        if (any) launch_tilts(p, t0, t1, stream);
    }
    ctx->stream = saved;
    return check_hip(ctx, hipGetLastError(), "sub-volume CTF backprojection launch");
}

int cu_count(artia_ctx* ctx) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess || prop.multiProcessorCount <= 0) return 256;
    return prop.multiProcessorCount;
}

// Times the candidate tunings on a sample of the queue (up to 8 sub-volumes per size, their
// contents saved and restored) and keeps the fastest.
artia_status autotune(artia_subvol_plan* p) {
    artia_ctx* ctx = p->ctx;
    std::vector<int> sample;
    {
        std::vector<std::pair<long long, int>> seen;
        for (int s : p->order) {
            const SubvolDev& sv = p->subs[(size_t)s];
            const long long key = (long long)sv.nx * 1000000LL + (long long)sv.ny * 1000 + sv.nz;
            int c = 0;
            for (auto& kv : seen)
                if (kv.first == key) c = ++kv.second;
            if (c == 0) seen.push_back({key, 1});
            if (c <= 8) sample.push_back(s);
        }
    }
    size_t backupFloats = 0;
    for (int s : sample) backupFloats += (size_t)p->subs[(size_t)s].nx * p->subs[(size_t)s].ny * p->subs[(size_t)s].nz;
    float* backup = nullptr;
    artia_status st = check_hip(ctx, hipMalloc(&backup, backupFloats * sizeof(float) + 16), "autotune backup");
    if (st != ARTIA_OK) return st;
    auto copy_all = [&](bool save) {
        size_t off = 0;
        for (int s : sample) {
            const SubvolDev& sv = p->subs[(size_t)s];
            const size_t n = (size_t)sv.nx * sv.ny * sv.nz;
            if (save)
                (void)hipMemcpyAsync(backup + off, sv.vol, n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream);
            else
                (void)hipMemcpyAsync(sv.vol, backup + off, n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream);
            off += n;
        }
    };
    copy_all(true);
    const std::vector<int> fullOrder = p->order;
    p->order = sample;
    const int cus = cu_count(ctx);
    const int policies[2] = {0, 1}, qpts[2] = {1, 2}, bpcs[3] = {2, 4, 8};
    float best = INFINITY;
    artia_subvol_tuning bestT = p->tuning;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // With CTF only the backprojection launches are timed, on up to four tilts spread over the
    // series, each filtered once; without CTF the whole (single-launch) pass.
    std::vector<int> tuneTilts;
    if (p->ctf) {
        std::vector<int> withSlabs;
        for (int t = 0; t < p->in.ntilts; t++)
            if (p->bound_count[(size_t)t] > 0) withSlabs.push_back(t);
        const int m = (int)withSlabs.size(), want = std::min(4, m);
        for (int i = 0; i < want; i++) tuneTilts.push_back(withSlabs[(size_t)((2 * i + 1) * m / (2 * want))]);
    } else {
        tuneTilts.push_back(-1);
    }
    float total[2][2][3] = {};
    for (int tt : tuneTilts) {
        if (tt >= 0 && (st = filter_tilt(p, tt)) != ARTIA_OK) break;
        for (int a = 0; a < 2 && st == ARTIA_OK; a++)
            for (int c = 0; c < 2 && st == ARTIA_OK; c++)
                for (int d = 0; d < 3 && st == ARTIA_OK; d++) {
                    p->tuning = artia_subvol_tuning{0, policies[a], qpts[c], bpcs[d]};
                    p->grid = cus * bpcs[d];
                    if ((st = set_items(p, policies[a], qpts[c])) != ARTIA_OK) break;
                    float ms = 0.0f;
                    for (int rep = 0; rep < 2 && st == ARTIA_OK; rep++) {
                        (void)hipEventRecord(e0, ctx->stream);
                        if (tt >= 0) {
                            launch_tilts(p, tt, tt + 1, ctx->stream);
                            st = check_hip(ctx, hipGetLastError(), "autotune launch");
                        } else {
                            st = enqueue(p, ctx->stream);
                        }
                        (void)hipEventRecord(e1, ctx->stream);
                        copy_all(false);
                        if (st == ARTIA_OK) st = check_hip(ctx, hipEventSynchronize(e1), "autotune run");
                        if (st == ARTIA_OK) (void)hipEventElapsedTime(&ms, e0, e1);
                    }
                    total[a][c][d] += ms;
                }
    }
    for (int a = 0; a < 2 && st == ARTIA_OK; a++)
        for (int c = 0; c < 2; c++)
            for (int d = 0; d < 3; d++)
                if (total[a][c][d] < best) {
                    best = total[a][c][d];
                    bestT = artia_subvol_tuning{0, policies[a], qpts[c], bpcs[d]};
                }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(backup);
    p->order = fullOrder;
    if (st != ARTIA_OK) return st;
    p->tuning = bestT;
    p->tuning.autotune = 1;
    return ARTIA_OK;
}

}  // namespace

extern "C" {

artia_status artia_subvol_wbp_plain(artia_ctx* ctx, const artia_subvol* subs, int nsub, const artia_tilt_geom* geoms,
                                    const unsigned char* active, int ntilts, const float* images, int width,
                                    int height, size_t pitch, int oversampling, const artia_ctf_params* ctf,
                                    const float* slab_range, const artia_tilt_defocus* defocus, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    const Inputs in{subs, nsub, geoms, active, ntilts, images, width, height, pitch, oversampling, ctf, slab_range,
                    defocus};
    artia_status st = validate(ctx, "artia_subvol_wbp_plain", in);
    if (st != ARTIA_OK) return st;
    const size_t npix = (size_t)height * pitch;
    if (ctf && (st = ensure_workspace(ctx, npix)) != ARTIA_OK) return st;
    Timing tm(ctx, ms);
    for (int t = 0; t < ntilts; t++) {
        const float* img = images + (size_t)t * npix;
        if (!ctf) {
            for (int s = 0; s < nsub; s++) {
                const size_t i = (size_t)s * (size_t)ntilts + (size_t)t;
                if (active && !active[i]) continue;
                launch_bp(ctx, to_device(subs[s].geom), to_device(geoms[i]), subs[s].vol, img, width, height, pitch,
                          1.0f, oversampling, 0.0f, 9999999999999.0f, ARTIA_FILTER_POINT);
            }
            continue;
        }
        // Reconstructor.cpp:993-1031: per slab, filter the projection, then every sub-volume
        const float step = ctf->slice_thickness_nm / ctf->pixelsize_nm;
        for (float ray = slab_range[2 * t]; ray < slab_range[2 * t + 1]; ray += step) {
            float dmin, dmax;
            artia_geom_defocus_minmax(ray, &defocus[t], ctf, &dmin, &dmax);
            st = ctf_slab_filter(ctx, const_cast<float*>(img), width, height, pitch, ctx->ws, pitch, 0, nullptr,
                                 nullptr, nullptr, *ctf, dmin, dmax, defocus[t].astig_angle_deg, 0);
            if (st != ARTIA_OK) return st;
            for (int s = 0; s < nsub; s++) {
                const size_t i = (size_t)s * (size_t)ntilts + (size_t)t;
                if (active && !active[i]) continue;
                launch_bp(ctx, to_device(subs[s].geom), to_device(geoms[i]), subs[s].vol, ctx->ws, width, height,
                          pitch, 1.0f, oversampling, ray, ray + step, ARTIA_FILTER_LINEAR);
            }
        }
    }
    st = check_hip(ctx, hipGetLastError(), "artia_subvol_wbp_plain launch");
    if (st != ARTIA_OK) return st;
    return tm.finish();
}

artia_status artia_subvol_plan_create(artia_ctx* ctx, const artia_subvol* subs, int nsub,
                                      const artia_tilt_geom* geoms, const unsigned char* active, int ntilts,
                                      const float* images, int width, int height, size_t pitch, int oversampling,
                                      const artia_ctf_params* ctf, const float* slab_range,
                                      const artia_tilt_defocus* defocus, const artia_subvol_tuning* tuning,
                                      artia_subvol_plan** out) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!out) return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_subvol_plan_create: out is NULL");
    *out = nullptr;
    if (ctx->numerics != ARTIA_NUMERICS_EXACT)
        return fail(ctx, ARTIA_ERR_UNSUPPORTED,
                    "artia_subvol_plan_create: the batched form evaluates exact numerics only; use "
                    "artia_subvol_wbp_plain for reference numerics");
    const Inputs in{subs, nsub, geoms, active, ntilts, images, width, height, pitch, oversampling, ctf, slab_range,
                    defocus};
    artia_status st = validate(ctx, "artia_subvol_plan_create", in);
    if (st != ARTIA_OK) return st;
    if (tuning && (tuning->quads_per_thread < 1 || tuning->quads_per_thread > 2 || tuning->tile_policy < 0 ||
                   tuning->tile_policy > 1 || tuning->blocks_per_cu < 1 || tuning->blocks_per_cu > 16))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_subvol_plan_create: bad tuning");
    artia_subvol_plan* p = new artia_subvol_plan();
    p->ctx = ctx;
    p->in = in;
    p->ctf = ctf != nullptr;
    if (ctf) {
        p->ctfp = *ctf;
        p->defocus.assign(defocus, defocus + ntilts);
        p->slab_range.assign(slab_range, slab_range + 2 * (size_t)ntilts);
        p->in.ctf = &p->ctfp;
        p->in.slab_range = p->slab_range.data();
        p->in.defocus = p->defocus.data();
    }
    p->subs.resize((size_t)nsub);
    for (int s = 0; s < nsub; s++)
        p->subs[(size_t)s] = SubvolDev{subs[s].vol, subs[s].geom.nx, subs[s].geom.ny, subs[s].geom.nz};
    p->order.resize((size_t)nsub);
    for (int s = 0; s < nsub; s++) p->order[(size_t)s] = s;
    std::stable_sort(p->order.begin(), p->order.end(), [&](int a, int b) {
        const SubvolDev &A = p->subs[(size_t)a], &B = p->subs[(size_t)b];
        return (long long)A.nx * A.ny * A.nz > (long long)B.nx * B.ny * B.nz;
    });
    // per-pair constants (extra shifts are already folded into geoms)
    std::vector<PairDev> pairs((size_t)nsub * (size_t)ntilts);
    for (int s = 0; s < nsub; s++) {
        const VolG v = to_device(subs[s].geom);
        for (int t = 0; t < ntilts; t++) {
            const size_t i = (size_t)s * (size_t)ntilts + (size_t)t;
            PairDev& pd = pairs[i];
            std::memset(&pd, 0, sizeof(PairDev));
            pd.v = v;
            if (active && !active[i]) continue;
            pd.g = to_device(geoms[i]);
            pd.e = prepare_exact(v, pd.g);
            pd.mode = (pd.e.affine && pd.e.identityAniso) ? (ctf ? 3 : 1) : 2;
            // (the image read with 32-bit offsets when an item's footprint does not fit the LDS tile)
            pd.win = (!ctf && pd.mode == 1 && (size_t)height * pitch < ((size_t)1 << 31) &&
                      window_plan(v, pd.g, oversampling, pd.w)) ? 1 : 0;
        }
    }
    if ((st = upload(ctx, (void**)&p->d_pairs, pairs.data(), pairs.size() * sizeof(PairDev), "pair table")) !=
            ARTIA_OK ||
        (st = upload(ctx, (void**)&p->d_subs, p->subs.data(), p->subs.size() * sizeof(SubvolDev), "sub-volumes")) !=
            ARTIA_OK ||
        (st = upload(ctx, (void**)&p->d_counter, nullptr, 0, "queue counter")) != ARTIA_OK) {
        delete p;
        return st;
    }
    if (ctf) {
        std::vector<float> all, starts;
        std::vector<int> boff;
        std::vector<float> last;
        const float step = ctf->slice_thickness_nm / ctf->pixelsize_nm;
        for (int t = 0; t < ntilts; t++) {
            slab_starts(slab_range[2 * t], slab_range[2 * t + 1], step, starts);
            boff.push_back((int)all.size());
            p->bound_count.push_back((int)starts.size());
            last.push_back(starts.empty() ? 0.0f : starts.back() + step);
            all.insert(all.end(), starts.begin(), starts.end());
            p->max_slabs = std::max(p->max_slabs, (int)starts.size());
        }
        // tilt chunks: consecutive tilts whose slab images fit the stack budget together
        const size_t slabBytes = (size_t)height * pitch * sizeof(float);
        size_t budget = ctx->stack_limit ? ctx->stack_limit : kSubvolStackBytes;
        size_t freeB = 0, totalB = 0;
        if (hipMemGetInfo(&freeB, &totalB) == hipSuccess && freeB / 2 < budget) budget = freeB / 2;
        const long long capSlabs = std::max((long long)p->max_slabs, (long long)(budget / slabBytes));
        long long used = 0, peak = 1;
        p->chunks.push_back(0);
        for (int t = 0; t < ntilts; t++) {
            const long long c = p->bound_count[(size_t)t];
            if (used > 0 && used + c > capSlabs) {
                p->chunks.push_back(t);
                used = 0;
            }
            p->stack_off.push_back(used);
            used += c;
            peak = std::max(peak, used);
        }
        p->chunks.push_back(ntilts);
        if ((st = upload(ctx, (void**)&p->d_bounds, all.data(), all.size() * sizeof(float), "slab bounds")) !=
            ARTIA_OK) {
            delete p;
            return st;
        }
        std::vector<TiltSlab> ts((size_t)ntilts);
        for (int t = 0; t < ntilts; t++)
            ts[(size_t)t] = TiltSlab{SlabArgs{p->d_bounds + boff[(size_t)t], p->bound_count[(size_t)t], last[(size_t)t]},
                                     p->stack_off[(size_t)t]};
        if ((st = upload(ctx, (void**)&p->d_tslab, ts.data(), ts.size() * sizeof(TiltSlab), "tilt slab table")) !=
                ARTIA_OK ||
            (st = check_hip(ctx, hipMalloc(&p->d_stack, (size_t)peak * slabBytes), "slab image stack")) != ARTIA_OK) {
            delete p;
            return st;
        }
    }
    const int cus = cu_count(ctx);
    if (tuning) {
        p->tuning = *tuning;
        p->tuning.autotune = 0;
    } else {
        p->tuning = artia_subvol_tuning{0, 0, 1, 4};
        if ((st = autotune(p)) != ARTIA_OK) {
            delete p;
            return st;
        }
    }
    p->grid = cus * p->tuning.blocks_per_cu;
    if ((st = set_items(p, p->tuning.tile_policy, p->tuning.quads_per_thread)) != ARTIA_OK) {
        delete p;
        return st;
    }
    if (!p->ctf) {
        // capture the counter reset and the persistent kernel once; execute replays the graph
        if ((st = check_hip(ctx, hipStreamCreate(&p->cap), "capture stream")) != ARTIA_OK ||
            (st = check_hip(ctx, hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal), "begin capture")) !=
                ARTIA_OK) {
            delete p;
            return st;
        }
        st = enqueue(p, p->cap);
        const artia_status st2 = check_hip(ctx, hipStreamEndCapture(p->cap, &p->graph), "end capture");
        if (st == ARTIA_OK) st = st2;
        if (st == ARTIA_OK)
            st = check_hip(ctx, hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0), "graph instantiate");
        if (st != ARTIA_OK) {
            delete p;
            return st;
        }
    }
    *out = p;
    return ARTIA_OK;
}

artia_status artia_subvol_plan_execute(artia_subvol_plan* p, float* ms) {
    if (!p) return ARTIA_ERR_INVALID_ARGUMENT;
    artia_ctx* ctx = p->ctx;
    Timing tm(ctx, ms);
    artia_status st;
    if (p->exec)
        st = check_hip(ctx, hipGraphLaunch(p->exec, ctx->stream), "graph launch");
    else
        st = enqueue(p, ctx->stream);
    if (st != ARTIA_OK) return st;
    return tm.finish();
}

artia_status artia_subvol_plan_tuning(const artia_subvol_plan* p, artia_subvol_tuning* out) {
    if (!p || !out) return ARTIA_ERR_INVALID_ARGUMENT;
    *out = p->tuning;
    return ARTIA_OK;
}

artia_status artia_subvol_plan_destroy(artia_subvol_plan* p) {
    if (!p) return ARTIA_ERR_INVALID_ARGUMENT;
    (void)hipStreamSynchronize(p->ctx->stream);
    delete p;
    return ARTIA_OK;
}

}  // extern "C"
