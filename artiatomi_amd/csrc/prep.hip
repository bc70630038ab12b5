// Projection preparation on gfx950: PrepareProjection (Reconstructor.cpp:1158-1383) over a stack of
// projections.  The reference prepares one projection at a time with NPP statistics copied to the
// host between steps; here a chunk of projections goes through every step together (batched hipFFT
// plans, one launch per step and chunk), and the per-projection statistics stay on the device
// until the end of the call.
//
// Steps per projection (see artia_prepare_projections in artia.h):
//   convert + mirror square -> mean of the square -> bad pixels -> [WBP: crop, statistics,
//   normalise, mirror square] -> [R2C -> fourierFilter -> doseWeighting -> wbpWeighting -> C2R]
//   -> crop / n^2 -> [SART: statistics, normalise].
// Sums are in double, in a fixed order (block partials, then one block per projection), so a
// projection's statistics do not depend on the chunk it is in.
#include <cmath>
#include <vector>

#include "artia_internal.hpp"
#include "device_common.hpp"
#include "image_ops.hpp"

namespace artia {
namespace {

constexpr int kRedBlocks = 256;   // partial sums per image and reduction

// Per-projection state, on the device for the whole call.
struct PrepStat {
    double sq_mean;          // mean of the converted square (the bad-pixel reference)
    double mean, std;        // statistics of the normalisation region
    float mean_hf, std_hf;   // the normalisation's subtrahend and divisor magnitude
    float dose, cos_tilt;    // AccumulatedDose; cosf(|tilt| in radians) for DownWeightTiltsForWBP
    int bad;                 // bad pixels replaced
    int pad;
};

dim3 grid2(int w, int h, int z) { return dim3((unsigned)((w + 63) / 64), (unsigned)((h + 3) / 4), (unsigned)z); }

// nppiConvert_*32f + CopyToSquare (mirror padding), image z of the chunk
template <class T>
__global__ void __launch_bounds__(256) prep_square_kernel(const T* __restrict__ in, size_t pitch, size_t istride, int W,
                                                          int H, int N, int bx, int by, int mirrorY,
                                                          float* __restrict__ sq) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= N || y >= N) return;
    int xi, yi;
    mirror_source(x, y, W, H, bx, by, mirrorY, xi, yi);
    sq[(size_t)blockIdx.z * (size_t)N * (size_t)N + (size_t)y * N + x] =
        (float)in[(size_t)blockIdx.z * istride + (size_t)yi * pitch + (size_t)xi];
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Partial sums over the w x h region at (x0, y0) of image z: pass 0 sums the values, pass 1 the
// squared deviations from st[z].mean.  Block b takes rows b, b + kRedBlocks, ..
__global__ void __launch_bounds__(256) region_sums_kernel(const float* __restrict__ img, size_t istride, size_t pitch,
                                                          int x0, int y0, int w, int h, const PrepStat* __restrict__ st,
                                                          int pass, double* __restrict__ partial) {
    __shared__ double sh[256];
    const int z = (int)blockIdx.y;
    const float* base = img + (size_t)z * istride;
    const double m = pass ? st[z].mean : 0.0;
    double acc = 0.0;
    for (int r = (int)blockIdx.x; r < h; r += kRedBlocks) {
        const float* row = base + (size_t)(y0 + r) * pitch + x0;
        for (int c = (int)threadIdx.x; c < w; c += 256) {
            const double v = (double)row[c];
            acc += pass ? (v - m) * (v - m) : v;
        }
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[(size_t)z * kRedBlocks + blockIdx.x] = s;
}

// which: 0 -> sq_mean = sum / count, 1 -> mean = sum / count, 2 -> std = sqrt(sum / count)
__global__ void __launch_bounds__(256) region_finish_kernel(const double* __restrict__ partial, double count, int which,
                                                            PrepStat* __restrict__ st) {
    __shared__ double sh[256];
    const int z = (int)blockIdx.x;
    const double s = block_sum(partial[(size_t)z * kRedBlocks + threadIdx.x], sh);
    if (threadIdx.x == 0) {
        if (which == 0) st[z].sq_mean = s / count;
        else if (which == 1) st[z].mean = s / count;
        else st[z].std = sqrt(s / count);
    }
}

// nppiCompareC (> bad_pixel_value * mean) + nppiSum_8u / 255 + nppiSet_32f_C1MR(mean)
__global__ void __launch_bounds__(256) bad_pixel_kernel(float* __restrict__ sq, int N, float bad_value,
                                                        PrepStat* __restrict__ st) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int z = (int)blockIdx.y;
    const float mean = (float)st[z].sq_mean;
    const float thr = bad_value * mean;
    const size_t plane = (size_t)N * (size_t)N;
    float* p = sq + (size_t)z * plane;
    int local = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < plane; i += (size_t)gridDim.x * 256) {
        if (p[i] > thr) {
            p[i] = mean;
            local++;
        }
    }
    if (local) atomicAdd(&cnt, local);
    __syncthreads();
    if (threadIdx.x == 0 && cnt) atomicAdd(&st[z].bad, cnt);
}

// The normalisation's m and s (Reconstructor.cpp:1232-1260 for WBP, :1346-1368 for SART)
__global__ void norm_params_kernel(PrepStat* __restrict__ st, int nz, int wbp, int norm, int down_weight) {
    const int z = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (z >= nz) return;
    PrepStat& s = st[z];
    float std_hf = (float)s.std;
    const float meanValue = (float)s.mean;
    float mean_hf = meanValue;
    if (norm == ARTIA_NORM_MEAN) std_hf = meanValue;
    if (norm == ARTIA_NORM_NONE) {
        std_hf = 1.0f;
        if (wbp) mean_hf = 0.0f;
    }
    if (wbp && down_weight) std_hf = std_hf / s.cos_tilt;
    s.mean_hf = mean_hf;
    s.std_hf = std_hf;
}

// out = square[by + y][bx + x] / norm (nppiDivC from the square's inner region)
__global__ void __launch_bounds__(256) crop_div_kernel(const float* __restrict__ sq, int N, int bx, int by, int W, int H,
                                                       float norm, float* __restrict__ out, size_t pitch,
                                                       size_t ostride) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= W || y >= H) return;
    out[(size_t)blockIdx.z * ostride + (size_t)y * pitch + x] =
        sq[(size_t)blockIdx.z * (size_t)N * (size_t)N + (size_t)(by + y) * N + (bx + x)] / norm;
}

// nppiSubC(m) then nppiDivC(-s), in place
__global__ void __launch_bounds__(256) normalise_kernel(float* __restrict__ img, int W, int H, size_t pitch,
                                                        size_t istride, const PrepStat* __restrict__ st) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= W || y >= H) return;
    const PrepStat& s = st[blockIdx.z];
    float* p = img + (size_t)blockIdx.z * istride + (size_t)y * pitch + x;
    float v = *p - s.mean_hf;
    v = v / -s.std_hf;
    *p = v;
}

// fourierFilter (wbpWeighting.cu:146-212) and doseWeighting (:214-240), each a float factor applied
// in the reference's order, on the half spectra of the chunk
__global__ void __launch_bounds__(256) prep_spectrum_kernel(float2* __restrict__ spec, size_t pitch, int N, int band,
                                                            float lp, float hp, float lps, float hps, int dose_on,
                                                            float pixA, const PrepStat* __restrict__ st) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= N / 2 + 1 || y >= N) return;
    float mx = (float)x, my = (float)y;
    if (my > N * 0.5f) my = (N - my) * -1.0f;
    const float dist = sqrtf(mx * mx + my * my);
    float2* p = spec + (size_t)blockIdx.z * pitch * (size_t)N + (size_t)y * pitch + x;
    float2 e = *p;
    if (band) {
        float fil = 0;
        lp = lp - lps;
        hp = hp + hps;
        if (lp > 0) {
            if (dist <= lp) fil = 1;
        } else {
            if (dist <= N / 2 - 1) fil = 1;
        }
        if (lps > 0) {
            const float fil2 = (-fil + 1.0f) * expf(-((dist - lp) * (dist - lp) / (2 * lps * lps)));
            if (fil2 > 0.001f) fil = fil2;
        }
        if (lps > 0 && lp == 0 && hp == 0 && hps == 0) fil = expf(-((dist - lp) * (dist - lp) / (2 * lps * lps)));
        if (hp > 0) {
            float fil2 = 0;
            if (dist >= hp) fil2 = 1;
            fil *= fil2;
            if (hps > 0) {
                const float fil3 = (-fil2 + 1) * expf(-((dist - hp) * (dist - hp) / (2 * hps * hps)));
                if (fil3 > 0.001f) fil = fil3;
            }
        }
        e.x *= fil;
        e.y *= fil;
    }
    if (dose_on) {
        const float d = dist / (N / 2 / pixA);
        const float fil = expf(-st[blockIdx.z].dose * d);
        e.x *= fil;
        e.y *= fil;
    }
    *p = e;
}

template <class T>
void launch_square(artia_ctx* ctx, const void* in, size_t pitch, size_t istride, int W, int H, int N, int mirrorY,
                   float* sq, int b) {
    hipLaunchKernelGGL(prep_square_kernel<T>, grid2(N, N, b), dim3(64, 4), 0, ctx->stream, static_cast<const T*>(in),
                       pitch, istride, W, H, N, (N - W) / 2, (N - H) / 2, mirrorY, sq);
}

void square_of(artia_ctx* ctx, int dtype, const void* in, size_t pitch, size_t istride, int W, int H, int N, int mirrorY,
               float* sq, int b) {
    switch (dtype) {
        case ARTIA_DT_SHORT: launch_square<short>(ctx, in, pitch, istride, W, H, N, mirrorY, sq, b); break;
        case ARTIA_DT_USHORT: launch_square<unsigned short>(ctx, in, pitch, istride, W, H, N, mirrorY, sq, b); break;
        case ARTIA_DT_INT: launch_square<int>(ctx, in, pitch, istride, W, H, N, mirrorY, sq, b); break;
        case ARTIA_DT_UINT: launch_square<unsigned>(ctx, in, pitch, istride, W, H, N, mirrorY, sq, b); break;
        default: launch_square<float>(ctx, in, pitch, istride, W, H, N, mirrorY, sq, b); break;
    }
}

size_t dtype_size(int dtype) {
    return dtype == ARTIA_DT_SHORT || dtype == ARTIA_DT_USHORT ? 2 : 4;
}

// mean (which 1) or sq_mean (which 0), or std (which 2, needs mean), of a region of b images
void region_stat(artia_ctx* ctx, const float* img, size_t istride, size_t pitch, int x0, int y0, int w, int h,
                 PrepStat* st, int which, double* partial, int b) {
    hipLaunchKernelGGL(region_sums_kernel, dim3(kRedBlocks, (unsigned)b), dim3(256), 0, ctx->stream, img, istride,
                       pitch, x0, y0, w, h, st, which == 2 ? 1 : 0, partial);
    hipLaunchKernelGGL(region_finish_kernel, dim3((unsigned)b), dim3(256), 0, ctx->stream, partial,
                       (double)w * (double)h, which, st);
}

}  // namespace
}  // namespace artia

using namespace artia;

extern "C" {

artia_status artia_prepare_projections(artia_ctx* ctx, const artia_prep_params* pp, const void* images, int n,
                                       int width, int height, size_t pitch, const float* tilt_deg,
                                       const float* psi_deg, const float* dose, float* out, size_t out_pitch,
                                       float* mean, float* std, int* bad_pixels, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!pp || !images || !out || n <= 0 || width <= 0 || height <= 0 || pitch < (size_t)width ||
        out_pitch < (size_t)width || !mean || !std || !bad_pixels)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_prepare_projections: invalid arguments");
    if (pp->dtype < ARTIA_DT_SHORT || pp->dtype > ARTIA_DT_FLOAT)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_prepare_projections: unsupported data type %d", pp->dtype);
    if (pp->normalization < ARTIA_NORM_MEAN || pp->normalization > ARTIA_NORM_NONE)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_prepare_projections: unknown normalisation %d",
                    pp->normalization);
    if (pp->wbp && (pp->wbp_method < 0 || pp->wbp_method > 4 || !psi_deg || (pp->down_weight_tilts && !tilt_deg)))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT,
                    "artia_prepare_projections: WBP needs a method 0..4, psi angles and, to down-weight, tilt angles");
    if (pp->dose_weighting && (!dose || !(pp->pixelsize_nm > 0.0f)))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT,
                    "artia_prepare_projections: dose weighting needs doses and a pixel size");
    const int N = width > height ? width : height;
    const bool fft = !pp->skip_filter || pp->wbp;
    FftBatch& fb = ctx->fftb;
    artia_status st = fb.ensure(ctx, N);
    if (st != ARTIA_OK) return st;
    // workspace: partial sums of a chunk, then the per-projection state
    const size_t partFloats = (size_t)fb.cap * kRedBlocks * 2;
    const size_t statFloats = ((size_t)n * sizeof(PrepStat) + sizeof(float) - 1) / sizeof(float);
    if ((st = ensure_workspace(ctx, partFloats + statFloats + 2)) != ARTIA_OK) return st;
    double* partial = reinterpret_cast<double*>(ctx->ws);
    PrepStat* d_stat = reinterpret_cast<PrepStat*>(ctx->ws + partFloats);
    std::vector<PrepStat> hs((size_t)n);
    for (int i = 0; i < n; i++) {
        PrepStat& s = hs[(size_t)i];
        s = PrepStat{};
        s.dose = pp->dose_weighting ? dose[i] : 0.0f;
        // Reconstructor.cpp:1250: cosf(abs(tilt) / 180.0f * M_PI), the product in double
        s.cos_tilt = (pp->wbp && pp->down_weight_tilts)
                         ? cosf((float)((double)(std::fabs(tilt_deg[i]) / 180.0f) * 3.14159265358979323846))
                         : 1.0f;
    }
    Timing t(ctx, ms);
    st = check_hip(ctx, hipMemcpyAsync(d_stat, hs.data(), hs.size() * sizeof(PrepStat), hipMemcpyHostToDevice,
                                       ctx->stream),
                   "hipMemcpyAsync(projection state)");
    if (st != ARTIA_OK) return st;
    const size_t plane = (size_t)N * N, half = (size_t)(N / 2 + 1);
    const size_t istride = (size_t)height * pitch, ostride = (size_t)height * out_pitch;
    const int bx = (N - width) / 2, by = (N - height) / 2;
    const float pixA = pp->pixelsize_nm * 10.0f;
    for (int k0 = 0; k0 < n;) {
        int b = fb.cap, pi = 0;
        while (b > n - k0) b /= 2;
        for (int q = b; q > 1; q /= 2) pi++;
        PrepStat* s = d_stat + k0;
        float* o = out + (size_t)k0 * ostride;
        square_of(ctx, pp->dtype, static_cast<const char*>(images) + (size_t)k0 * istride * dtype_size(pp->dtype),
                  pitch, istride, width, height, N, pp->mirror_y, fb.square, b);
        region_stat(ctx, fb.square, plane, (size_t)N, 0, 0, N, N, s, 0, partial, b);
        if (pp->correct_bad_pixels)
            hipLaunchKernelGGL(bad_pixel_kernel, dim3(512, (unsigned)b), dim3(256), 0, ctx->stream, fb.square, N,
                               pp->bad_pixel_value, s);
        if (pp->wbp) {
            // statistics and normalisation of the raw image, then the mirror square again
            hipLaunchKernelGGL(crop_div_kernel, grid2(width, height, b), dim3(64, 4), 0, ctx->stream, fb.square, N, bx,
                               by, width, height, 1.0f, o, out_pitch, ostride);
            region_stat(ctx, o, ostride, out_pitch, 0, 0, width, height, s, 1, partial, b);
            region_stat(ctx, o, ostride, out_pitch, 0, 0, width, height, s, 2, partial, b);
            hipLaunchKernelGGL(norm_params_kernel, dim3((unsigned)((b + 63) / 64)), dim3(64), 0, ctx->stream, s, b, 1,
                               pp->normalization, pp->down_weight_tilts);
            hipLaunchKernelGGL(normalise_kernel, grid2(width, height, b), dim3(64, 4), 0, ctx->stream, o, width, height,
                               out_pitch, ostride, s);
            launch_square<float>(ctx, o, out_pitch, ostride, width, height, N, 0, fb.square, b);
        }
        if (fft) {
            if (hipfftSetStream(fb.r2c[pi], ctx->stream) != HIPFFT_SUCCESS ||
                hipfftSetStream(fb.c2r[pi], ctx->stream) != HIPFFT_SUCCESS)
                return fail(ctx, ARTIA_ERR_FFT, "hipfftSetStream failed");
            if (hipfftExecR2C(fb.r2c[pi], fb.square, (hipfftComplex*)fb.spec) != HIPFFT_SUCCESS)
                return fail(ctx, ARTIA_ERR_FFT, "hipfftExecR2C (batch %d) failed", b);
            if (!pp->skip_filter || pp->dose_weighting)
                hipLaunchKernelGGL(prep_spectrum_kernel, grid2(N / 2 + 1, N, b), dim3(64, 4), 0, ctx->stream, fb.spec,
                                   half, N, pp->skip_filter ? 0 : 1, pp->lp, pp->hp, pp->lps, pp->hps,
                                   pp->dose_weighting, pixA, s);
            if (pp->wbp)
                for (int i = 0; i < b; i++)
                    if ((st = run_wbp_weight(ctx, fb.spec + (size_t)i * half * N, N, half, psi_deg[k0 + i],
                                             pp->wbp_method)) != ARTIA_OK)
                        return st;
            if (hipfftExecC2R(fb.c2r[pi], (hipfftComplex*)fb.spec, fb.square) != HIPFFT_SUCCESS)
                return fail(ctx, ARTIA_ERR_FFT, "hipfftExecC2R (batch %d) failed", b);
        }
        hipLaunchKernelGGL(crop_div_kernel, grid2(width, height, b), dim3(64, 4), 0, ctx->stream, fb.square, N, bx, by,
                           width, height, fft ? (float)N * (float)N : 1.0f, o, out_pitch, ostride);
        if (!pp->wbp) {
            // SART: statistics of the filtered image (its centre for IMOD-aligned stacks), normalised
            const bool centre = pp->normalization == ARTIA_NORM_NONE;
            const int x0 = centre ? width / 4 : 0, y0 = centre ? height / 4 : 0;
            const int w = centre ? width / 2 : width, h = centre ? height / 2 : height;
            region_stat(ctx, o, ostride, out_pitch, x0, y0, w, h, s, 1, partial, b);
            region_stat(ctx, o, ostride, out_pitch, x0, y0, w, h, s, 2, partial, b);
            hipLaunchKernelGGL(norm_params_kernel, dim3((unsigned)((b + 63) / 64)), dim3(64), 0, ctx->stream, s, b, 0,
                               pp->normalization, 0);
            hipLaunchKernelGGL(normalise_kernel, grid2(width, height, b), dim3(64, 4), 0, ctx->stream, o, width, height,
                               out_pitch, ostride, s);
        }
        if ((st = check_hip(ctx, hipGetLastError(), "projection preparation launch")) != ARTIA_OK) return st;
        k0 += b;
    }
    st = check_hip(ctx, hipMemcpyAsync(hs.data(), d_stat, hs.size() * sizeof(PrepStat), hipMemcpyDeviceToHost,
                                       ctx->stream),
                   "hipMemcpyAsync(projection statistics)");
    if (st == ARTIA_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (st != ARTIA_OK) return st;
    for (int i = 0; i < n; i++) {
        mean[i] = (float)hs[(size_t)i].mean;
        std[i] = (float)hs[(size_t)i].std;
        bad_pixels[i] = hs[(size_t)i].bad;
    }
    return t.finish();
}

}  // extern "C"
