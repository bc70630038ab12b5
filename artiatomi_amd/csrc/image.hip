// Per-pixel image operators on gfx950: residual (compare), projected-volume border crop
// (cropBorder), pad-to-square (makeSquare) and the half-volume slice export (convertVolume).  These are bandwidth-trivial; one thread per
// pixel, 64x4 blocks so a wave covers 64 consecutive pixels of a row.
#include "device_common.hpp"
#include "image_ops.hpp"

namespace artia {
namespace {

// Compare.cu:49-130 (arithmetic in image_ops.hpp, shared with the fused forward-projection epilogue)
__global__ void __launch_bounds__(256) compare_kernel(int W, int H, size_t pitch, const float* __restrict__ real,
                                                      float* __restrict__ virt, const float* __restrict__ dist,
                                                      float L, F4 cut, F4 dim, float S) {
    const unsigned x = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= (unsigned)W || y >= (unsigned)H) return;
    const size_t i = (size_t)y * pitch + x;
    virt[i] = compare_value(x, y, W, H, real[i], virt[i], dist[i], L, cut, dim, S);
}

// Compare.cu:170-277 (weight in image_ops.hpp), applied `repeat` times in place as that many
// launches would apply it
__global__ void __launch_bounds__(256) crop_border_kernel(int W, int H, size_t pitch, float* __restrict__ img,
                                                          CropParams c, int repeat) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * pitch + x;
    const float w = crop_weight(x, y, W, H, c);
    float val = img[i];
    for (int r = 0; r < repeat; r++) val = w * val;
    img[i] = val;
}

// CopyToSquare.cu:37-92: W x H -> N x N, centred zero fill or mirror reflection, optional y flip.
__global__ void __launch_bounds__(256) make_square_kernel(int W, int H, size_t pitch, const float* __restrict__ in,
                                                          int N, float* __restrict__ out, int bx, int by, int mirrorY,
                                                          int fillZero) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= N || y >= N) return;
    int xi = x - bx, yi = y - by;
    float val = 0.0f;
    if (fillZero) {
        if (xi >= 0 && xi < W && yi >= 0 && yi < H) {
            if (mirrorY) yi = H - yi - 1;
            val = in[(size_t)yi * pitch + xi];
        }
    } else {
        mirror_source(x, y, W, H, bx, by, mirrorY, xi, yi);
        val = in[(size_t)yi * pitch + xi];
    }
    out[(size_t)y * N + x] = val;
}

// convertVolumeFP16ToFP32 (BackProjectionSquareOS.cu:1294-1323): slices z0 .. z0 + gridDim.z - 1
// of a half volume as negated floats (the sign EmSART writes volumes with), slice z0 + k at
// out + k * ny * out_pitch.
__global__ void __launch_bounds__(256) volume_f16_to_f32_kernel(const vhalf* __restrict__ vol, int nx, int ny, int z0,
                                                                float* __restrict__ out, size_t out_pitch) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)blockIdx.y, k = (int)blockIdx.z;
    if (x >= nx) return;
    const size_t src = ((size_t)(z0 + k) * (size_t)ny + (size_t)y) * (size_t)nx + (size_t)x;
    out[((size_t)k * (size_t)ny + (size_t)y) * out_pitch + (size_t)x] = -(float)vol[src];
}

dim3 grid2d(int w, int h) { return dim3((unsigned)((w + 63) / 64), (unsigned)((h + 3) / 4)); }

}  // namespace

void launch_compare(artia_ctx* ctx, int W, int H, size_t pitch, const float* real, float* virt, const float* dist,
                    float realLength, const float cut[4], const float dim[4], float scale) {
    const F4 c{cut[0], cut[1], cut[2], cut[3]}, d{dim[0], dim[1], dim[2], dim[3]};
    hipLaunchKernelGGL(compare_kernel, grid2d(W, H), dim3(64, 4), 0, ctx->stream, W, H, pitch, real, virt, dist,
                       realLength, c, d, scale);
}

void launch_crop_border(artia_ctx* ctx, int W, int H, size_t pitch, float* img, const float cut[2], const float dim[2],
                        const int p[8], int repeat) {
    if (repeat <= 0) return;
    hipLaunchKernelGGL(crop_border_kernel, grid2d(W, H), dim3(64, 4), 0, ctx->stream, W, H, pitch, img,
                       make_crop_params(cut, dim, p), repeat);
}

void launch_make_square(artia_ctx* ctx, int W, int H, size_t pitch, const float* in, int N, float* out, int bx, int by,
                        int mirrorY, int fillZero) {
    hipLaunchKernelGGL(make_square_kernel, grid2d(N, N), dim3(64, 4), 0, ctx->stream, W, H, pitch, in, N, out, bx, by,
                       mirrorY, fillZero);
}

}  // namespace artia

using namespace artia;

extern "C" {

artia_status artia_compare(artia_ctx* ctx, int width, int height, size_t pitch, const float* real, float* virt,
                           const float* dist, float real_length, const float crop[4], const float crop_dim[4],
                           float proj_scale, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!real || !virt || !dist || !crop || !crop_dim || width <= 0 || height <= 0 || pitch < (size_t)width)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_compare: invalid arguments");
    Timing t(ctx, ms);
    launch_compare(ctx, width, height, pitch, real, virt, dist, real_length, crop, crop_dim, proj_scale);
    return t.finish();
}

artia_status artia_crop_border(artia_ctx* ctx, int width, int height, size_t pitch, float* img,
                               const float cut_length[2], const float dim_length[2], const int p[8], float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!img || !cut_length || !dim_length || !p || width <= 0 || height <= 0 || pitch < (size_t)width)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_crop_border: invalid arguments");
    Timing t(ctx, ms);
    launch_crop_border(ctx, width, height, pitch, img, cut_length, dim_length, p);
    return t.finish();
}

artia_status artia_make_square(artia_ctx* ctx, int width, int height, size_t pitch, const float* in, int n, float* out,
                               int border_x, int border_y, int mirror_y, int fill_zero, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!in || !out || width <= 0 || height <= 0 || pitch < (size_t)width || n <= 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_make_square: invalid arguments");
    Timing t(ctx, ms);
    launch_make_square(ctx, width, height, pitch, in, n, out, border_x, border_y, mirror_y, fill_zero);
    return t.finish();
}

artia_status artia_volume_f16_to_f32(artia_ctx* ctx, const uint16_t* vol, int nx, int ny, int nz, int z0, int count,
                                     float* out, size_t out_pitch, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!vol || !out || nx <= 0 || ny <= 0 || nz <= 0 || out_pitch < (size_t)nx || z0 < 0 || count < 0 ||
        count > nz - z0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_volume_f16_to_f32: invalid arguments");
    if (ny > 65535 || count > 65535)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_volume_f16_to_f32: at most 65535 rows and slices per call");
    Timing t(ctx, ms);
    if (count > 0)
        hipLaunchKernelGGL(volume_f16_to_f32_kernel, dim3((unsigned)((nx + 255) / 256), (unsigned)ny, (unsigned)count),
                           dim3(256), 0, ctx->stream, reinterpret_cast<const vhalf*>(vol), nx, ny, z0, out, out_pitch);
    return t.finish();
}

}  // extern "C"
