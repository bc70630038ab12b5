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
//   exact      the ray/box intersection in double, samples at t_in + k*0.15*s generated in
//              float relative to the entry point, values summed in a double accumulator;
//              this tracks the double-precision evaluation of the same algorithm.
#include "device_common.hpp"

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

template <bool kSlicer>
__global__ void __launch_bounds__(256) march_ref_kernel(VolG v, TiltG g, const float* __restrict__ vol, Roi roi,
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

// ---------------------------------------------------------------- exact numerics

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

// Sums the samples of one sub-ray between t_first and t_last (inclusive) at spacing step,
// returning sum(values).  Positions are generated in float voxel coordinates relative to the
// first sample; the partial sums go to double every 64 samples.
__device__ __forceinline__ double march_samples_exact(const VolG& v, const TiltG& g, const float* __restrict__ vol,
                                                      const double c[3], double t_first, int count, double step) {
    // voxel coordinate = (p - bbMin) * rcp * dim, with the float-rounded rcp the reference uploads
    float q0[3], dq[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double scale = (double)v.rcp[i] * (double)v.dim[i];
        const double p = (c[i] + t_first * (double)g.n[i]) - (double)v.bbMin[i];
        q0[i] = (float)(p * scale);
        dq[i] = (float)(step * (double)g.n[i] * scale);
    }
    q0[2] -= v.zshift;
    double total = 0.0;
    for (int k0 = 0; k0 < count; k0 += 64) {
        const int k1 = min(count, k0 + 64);
        float part = 0.0f;
        for (int k = k0; k < k1; k++) {
            const float fk = (float)k;
            const float qx = __builtin_fmaf(fk, dq[0], q0[0]);
            const float qy = __builtin_fmaf(fk, dq[1], q0[1]);
            const float qz = __builtin_fmaf(fk, dq[2], q0[2]);
            part += tex3d_linear(vol, v.nx, v.ny, v.nz, qx, qy, qz);
        }
        total += (double)part;
    }
    return total;
}

template <bool kSlicer>
__global__ void __launch_bounds__(256) march_exact_kernel(VolG v, TiltG g, const float* __restrict__ vol, Roi roi,
                                                          size_t pitch, float* __restrict__ proj,
                                                          float* __restrict__ dist, float tminDef, float tmaxDef) {
    const int x = roi.x0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = roi.y0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= roi.x1 || y >= roi.y1) return;
    const double step = (double)v.vs[0] * 0.15;
    double acc = 0.0, len = 0.0;
    int hits = 0;
#pragma unroll 1
    for (int s = 0; s < 4; s++) {
        const double px = (double)x + ((s & 2) ? 0.75 : 0.25);
        const double py = (double)y + ((s & 1) ? 0.75 : 0.25);
        RayD r = ray_setup_d(v, g, px, py);
        if (!((r.t_out - r.t_in) > 0.0)) continue;
        double t_first = r.t_in, t_in = r.t_in, t_out = r.t_out;
        if (kSlicer) {
            t_in = fmax(t_in, (double)tminDef);
            t_out = fmin(t_out, (double)tmaxDef);
            if (!(t_in <= t_out)) continue;
            // first sample at the detector point (ForwardProjectionSlicer.cu:129-133)
            acc += (double)tex3d_linear(
                vol, v.nx, v.ny, v.nz, (float)((r.c[0] - (double)v.bbMin[0]) * ((double)v.rcp[0] * (double)v.dim[0])),
                (float)((r.c[1] - (double)v.bbMin[1]) * ((double)v.rcp[1] * (double)v.dim[1])),
                (float)((r.c[2] - (double)v.bbMin[2]) * ((double)v.rcp[2] * (double)v.dim[2]) - (double)v.zshift));
            t_first = t_in + step;
        } else {
            hits++;
            len += t_out - t_in;
        }
        if (t_first > t_out) continue;
        const double span = floor((t_out - t_first) / step);
        const int count = (int)fmin(span + 1.0, (double)kMaxSamplesPerRay);
        acc += march_samples_exact(v, g, vol, r.c, t_first, count, step);
    }
    const size_t i = (size_t)y * pitch + (size_t)x;
    proj[i] += (float)(acc * step * 0.25);
    if (!kSlicer) dist[i] = hits ? (float)(len / (double)hits) : 0.0f;
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

void launch_march_impl(artia_ctx* ctx, const VolG& v, const TiltG& g, const float* vol, const Roi& r, size_t pitch,
                       float* proj, float* dist, bool slicer, float tmin, float tmax) {
    if (r.x1 <= r.x0 || r.y1 <= r.y0) return;
    const dim3 grid = grid_for(r), block(64, 4);
    if (ctx->numerics == ARTIA_NUMERICS_REFERENCE) {
        if (slicer)
            hipLaunchKernelGGL(march_ref_kernel<true>, grid, block, 0, ctx->stream, v, g, vol, r, pitch, proj, dist,
                               tmin, tmax);
        else
            hipLaunchKernelGGL(march_ref_kernel<false>, grid, block, 0, ctx->stream, v, g, vol, r, pitch, proj, dist,
                               tmin, tmax);
    } else {
        if (slicer)
            hipLaunchKernelGGL(march_exact_kernel<true>, grid, block, 0, ctx->stream, v, g, vol, r, pitch, proj, dist,
                               tmin, tmax);
        else
            hipLaunchKernelGGL(march_exact_kernel<false>, grid, block, 0, ctx->stream, v, g, vol, r, pitch, proj,
                               dist, tmin, tmax);
    }
}

}  // namespace

void launch_march(artia_ctx* ctx, const VolG& v, const TiltG& g, const float* vol, int W, int H, size_t pitch,
                  float* proj, float* dist, int x0, int y0, int x1, int y1) {
    const int roi[4] = {x0, y0, x1, y1};
    launch_march_impl(ctx, v, g, vol, make_roi(roi, W, H), pitch, proj, dist, false, 0.0f, 0.0f);
}

void launch_vol_traversal(artia_ctx* ctx, const VolG& v, const TiltG& g, int W, int H, size_t pitch, float* dist,
                          int x0, int y0, int x1, int y1) {
    const int roiA[4] = {x0, y0, x1, y1};
    const Roi r = make_roi(roiA, W, H);
    if (r.x1 <= r.x0 || r.y1 <= r.y0) return;
    if (ctx->numerics == ARTIA_NUMERICS_REFERENCE)
        hipLaunchKernelGGL(vol_traversal_ref_kernel, grid_for(r), dim3(64, 4), 0, ctx->stream, v, g, r, pitch, dist);
    else
        hipLaunchKernelGGL(vol_traversal_exact_kernel, grid_for(r), dim3(64, 4), 0, ctx->stream, v, g, r, pitch,
                           dist);
}

}  // namespace artia

using namespace artia;

namespace {
artia_status check_image(artia_ctx* ctx, const char* fn, int W, int H, size_t pitch) {
    if (W <= 0 || H <= 0 || pitch < (size_t)W)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: bad image size %dx%d pitch %zu", fn, W, H, pitch);
    return ARTIA_OK;
}
}  // namespace

extern "C" {

artia_status artia_fp_march(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, const float* vol,
                            int width, int height, size_t pitch, float* proj, float* dist, const int* roi, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!valid_geometry(vg, tg) || !vol || !proj || !dist)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_fp_march: null pointer or invalid geometry");
    artia_status st = check_image(ctx, "artia_fp_march", width, height, pitch);
    if (st != ARTIA_OK) return st;
    Timing t(ctx, ms);
    launch_march_impl(ctx, to_device(*vg), to_device(*tg), vol, make_roi(roi, width, height), pitch, proj, dist,
                      false, 0.0f, 0.0f);
    return t.finish();
}

artia_status artia_fp_slicer(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, const float* vol,
                             int width, int height, size_t pitch, float* proj, float tmin, float tmax, const int* roi,
                             float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!valid_geometry(vg, tg) || !vol || !proj)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_fp_slicer: null pointer or invalid geometry");
    artia_status st = check_image(ctx, "artia_fp_slicer", width, height, pitch);
    if (st != ARTIA_OK) return st;
    Timing t(ctx, ms);
    launch_march_impl(ctx, to_device(*vg), to_device(*tg), vol, make_roi(roi, width, height), pitch, proj, nullptr,
                      true, tmin, tmax);
    return t.finish();
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
