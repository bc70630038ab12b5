// Device helpers shared by the image kernels: the border dimming of compare/cropBorder and the
// quadrilateral crop weight (Compare.cu:138-277).
#pragma once

#include <hip/hip_runtime.h>

namespace artia {

struct F4 {
    float x, y, z, w;
};

// Crop parameters: cut and dim lengths and the corners p1..p4 as (x, y) pairs.
struct CropParams {
    float cx, cy, dx, dy;
    int p[8];
};

inline CropParams make_crop_params(const float cut[2], const float dim[2], const int p[8]) {
    CropParams c;
    c.cx = cut[0];
    c.cy = cut[1];
    c.dx = dim[0];
    c.dy = dim[1];
    for (int i = 0; i < 8; i++) c.p[i] = p[i];
    return c;
}

// CopyToSquare.cu:37-92 mirror padding: source pixel of square pixel (x, y) for a W x H image
// centred at (bx, by), reflected at the edges, optionally flipped in y; a pad wider than the image
// would reflect out of it, so the indices are clamped.
__device__ __forceinline__ void mirror_source(int x, int y, int W, int H, int bx, int by, int mirrorY, int& xi,
                                              int& yi) {
    xi = x - bx;
    yi = y - by;
    if (xi < 0) xi = -xi - 1;
    if (xi >= W) xi = W - (xi - W) - 1;
    if (yi < 0) yi = -yi - 1;
    if (yi >= H) yi = H - (yi - H) - 1;
    if (mirrorY) yi = H - yi - 1;
    xi = xi < 0 ? 0 : (xi > W - 1 ? W - 1 : xi);
    yi = yi < 0 ? 0 : (yi > H - 1 ? H - 1 : yi);
}

__device__ __forceinline__ float border_weight(float w) { return 1.0f - expf(-((w * w) * 9.0f)); }

__device__ __forceinline__ float dist2(float ax, float ay, float bx, float by) {
    return sqrtf((ax - bx) * (ax - bx) + (ay - by) * (ay - by));
}

// Distance from (x, y) to the segment v-w (Compare.cu:138-167)
__device__ __forceinline__ float segment_distance(int vx, int vy, int wx, int wy, int x, int y) {
    const float l2 = (float)((vx - wx) * (vx - wx) + (vy - wy) * (vy - wy));
    if (l2 == 0) return dist2((float)vx, (float)vy, (float)x, (float)y);
    const float x1 = (float)(x - vx), y1 = (float)(y - vy);
    const float x2 = (float)(wx - vx), y2 = (float)(wy - vy);
    const float t = (x1 * x2 + y1 * y2) / l2;
    if (t < 0.0f) return dist2((float)x, (float)y, (float)vx, (float)vy);
    if (t > 1.0f) return dist2((float)x, (float)y, (float)wx, (float)wy);
    return dist2((float)x, (float)y, (float)vx + t * (float)(wx - vx), (float)vy + t * (float)(wy - vy));
}

// Compare.cu:170-277: 0 outside the quadrilateral p1 p2 p4 p3 (strict integer cross products);
// inside, wx * wy with each factor dimmed within cut + dim of its nearest edge.
__device__ __forceinline__ float crop_weight(int x, int y, int W, int H, const CropParams& c) {
    const int p1x = c.p[0], p1y = c.p[1], p2x = c.p[2], p2y = c.p[3];
    const int p3x = c.p[4], p3y = c.p[5], p4x = c.p[6], p4y = c.p[7];
    float wx = 0.0f, wy = 0.0f;
    if ((p2x - p1x) * (y - p1y) - (p2y - p1y) * (x - p1x) < 0 && (p3x - p4x) * (y - p4y) - (p3y - p4y) * (x - p4x) < 0 &&
        (p1x - p3x) * (y - p3y) - (p1y - p3y) * (x - p3x) < 0 && (p4x - p2x) * (y - p2y) - (p4y - p2y) * (x - p2x) < 0) {
        wx = 1.0f;
        wy = 1.0f;
        float mx = (float)(3 * W), my = (float)(3 * H);
        mx = fminf(mx, segment_distance(p1x, p1y, p2x, p2y, x, y));
        mx = fminf(mx, segment_distance(p4x, p4y, p3x, p3y, x, y));
        mx = fminf(mx, (float)x);
        mx = fminf(mx, (float)(W - x - 1));
        my = fminf(my, segment_distance(p3x, p3y, p1x, p1y, x, y));
        my = fminf(my, segment_distance(p2x, p2y, p4x, p4y, x, y));
        my = fminf(my, (float)y);
        my = fminf(my, (float)(H - y - 1));
        if (mx < c.cx + c.dx) {
            float w = (mx - c.cx) / c.dx;
            if (w < 0) w = 0;
            wx = border_weight(w);
        }
        if (my < c.cy + c.dy) {
            float w = (my - c.cy) / c.dy;
            if (w < 0) w = 0;
            wy = border_weight(w);
        }
    }
    return wx * wy;
}

// Compare.cu:49-130: r = dist >= 1 ? ((b - (p/s)/dist*L)/L)*s : 0, times the border dimming.
// The right/top branch evaluates a non-positive w and so zeroes a band of CropDim + Crop pixels
// (reference behaviour kept).  Shared by compare_kernel and the forward projector's residual
// epilogue, so both give the same bits.
__device__ __forceinline__ float compare_value(unsigned x, unsigned y, int W, int H, float real, float virt, float val,
                                               float L, const F4& cut, const F4& dim, float S) {
    float error = 0.0f;
    if (val >= 1.0f) error = ((real - ((virt / S) / val * L)) / L) * S;
    float yb = 1.0f, ya = 1.0f, xb = 1.0f, xa = 1.0f;
    if ((float)y < cut.z + dim.z) {
        float w = ((float)y - cut.z) / dim.z;
        if (w < 0) w = 0;
        yb = border_weight(w);
    }
    if ((float)y > (((float)H - dim.w) - cut.w) - 1.0f) {
        float w = (((float)((unsigned)H - y - 1u) - cut.w) - (float)(H - 1)) / dim.w;
        if (w < 0) w = 0.0f;
        ya = border_weight(w);
    }
    if ((float)x < cut.y + dim.y) {
        float w = ((float)x - cut.y) / dim.y;
        if (w < 0) w = 0;
        xb = border_weight(w);
    }
    if ((float)x > (((float)W - dim.x) - cut.x) - 1.0f) {
        float w = (((float)((unsigned)W - x - 1u) - cut.x) - (float)(W - 1)) / dim.x;
        if (w < 0) w = 0.0f;
        xa = border_weight(w);
    }
    return (((xa * xb) * ya) * yb) * error;
}

// The SART residual of one tilt as the forward projector's epilogue computes it: compare of the
// measured image against the projection and its mean chord, then (when crop_on) cropBorder's
// weight, each in the arithmetic of the separate kernels.
struct ResidEpi {
    const float* real;   // measured projection, same pitch as the output
    float L;             // real traversal length (Reconstructor.cpp:1390-1393)
    F4 cut, dim;         // Crop, CropDim
    float S;             // ProjectionScaleFactor
    CropParams crop;     // cropBorder (CutLength, DimLength, hit points)
    int crop_on;
};

__device__ __forceinline__ float residual_value(const ResidEpi& e, int x, int y, int W, int H, size_t pitch, float p,
                                                float dist) {
    const size_t i = (size_t)y * pitch + (size_t)x;
    const float r = compare_value((unsigned)x, (unsigned)y, W, H, e.real[i], p, dist, e.L, e.cut, e.dim, e.S);
    return e.crop_on ? crop_weight(x, y, W, H, e.crop) * r : r;
}

}  // namespace artia
