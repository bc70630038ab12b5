// Device helpers shared by the image kernels: the border dimming of compare/cropBorder and the
// quadrilateral crop weight (Compare.cu:138-277).
#pragma once

#include <hip/hip_runtime.h>

namespace artia {

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

}  // namespace artia
