// Internal declarations shared by the HIP translation units of libartia.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/artia.h"

struct artia_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int numerics = ARTIA_NUMERICS_EXACT;
    std::string last_error;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    // driver workspace (projection-sized scratch images), grown on demand
    float* ws = nullptr;
    size_t ws_floats = 0;
    // per-kernel-class event timing of driver launches (artia_profile_enable)
    bool prof = false;
    std::vector<hipEvent_t> prof_events;   // pool, reused across runs
    std::vector<int> prof_marks;           // kernel class of each recorded start/stop pair
    double prof_ms[ARTIA_PROF_COUNT] = {};
    long long prof_launches[ARTIA_PROF_COUNT] = {};
};

namespace artia {

// Device-side geometry (by-value kernel arguments; uniform, so they live in SGPRs)
struct TiltG {
    float d[3], u[3], v[3], n[3];
    float M[16];
    float A[9], Ainv[9];
};
struct VolG {
    float bbMin[3], bbMax[3], rcp[3], dim[3], vs[3];
    float zshift;
    int nx, ny, nz;
};

TiltG to_device(const artia_tilt_geom& g);
VolG to_device(const artia_vol_geom& g);

artia_status fail(artia_ctx* ctx, artia_status code, const char* fmt, ...);
artia_status check_hip(artia_ctx* ctx, hipError_t err, const char* what);

// Brackets a launch with events on the context stream when the caller asked for timing.
class Timing {
public:
    Timing(artia_ctx* ctx, float* ms);
    artia_status finish();   // records the stop event, waits, writes *ms
private:
    artia_ctx* ctx_;
    float* ms_;
};

artia_status ensure_workspace(artia_ctx* ctx, size_t floats);

// Records a start event for kernel class `kind` when profiling is on; returns the pair index
// (or -1).  prof_stop records the matching stop event; prof_collect waits for the last event
// and folds all pairs into the per-class totals.
int prof_start(artia_ctx* ctx, int kind);
void prof_stop(artia_ctx* ctx, int pair);
artia_status prof_collect(artia_ctx* ctx);
bool valid_geometry(const artia_vol_geom* vg, const artia_tilt_geom* tg);

// Kernel launchers used by the drivers (no argument validation, stream = ctx->stream)
void launch_march(artia_ctx* ctx, const VolG& vg, const TiltG& tg, const float* vol, int W, int H,
                  size_t pitch, float* proj, float* dist, int x0, int y0, int x1, int y1);
void launch_vol_traversal(artia_ctx* ctx, const VolG& vg, const TiltG& tg, int W, int H, size_t pitch,
                          float* dist, int x0, int y0, int x1, int y1);
void launch_compare(artia_ctx* ctx, int W, int H, size_t pitch, const float* real, float* virt,
                    const float* dist, float realLength, const float cut[4], const float dim[4], float scale);
void launch_crop_border(artia_ctx* ctx, int W, int H, size_t pitch, float* img, const float cut[2],
                        const float dim[2], const int p[8]);
void launch_bp(artia_ctx* ctx, const VolG& vg, const TiltG& tg, float* vol, const float* img, int W, int H,
               size_t pitch, float lambda, int os, float distMin, float distMax, int filter);
void launch_fill(artia_ctx* ctx, float* p, size_t n, float value);

}  // namespace artia
