// Internal declarations shared by the HIP translation units of libartia.
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/artia.h"

struct artia_ctx;

namespace artia {
// Square N x N single-precision R2C/C2R plans with their square image and half-spectrum
// buffers ((N/2+1) x N complex), rebuilt when N changes.
struct FftPlans {
    int N = 0;
    hipfftHandle r2c = 0, c2r = 0;
    float* square = nullptr;
    float2* spec = nullptr;
    artia_status ensure(artia_ctx* ctx, int n);
    void release();
};

// Batched plans for stacks of N x N slab images: one R2C and one C2R plan per power-of-two batch
// 1, 2, 4, .., cap sharing one work area, with `cap` squares and half spectra.  cap is the largest
// power of two up to kFftBatchMax whose squares and spectra fit in 2 GiB.
constexpr int kFftBatchMax = 32;
constexpr int kFftBatchPlans = 6;   // log2(kFftBatchMax) + 1
struct FftBatch {
    int N = 0, cap = 0;
    hipfftHandle r2c[kFftBatchPlans] = {}, c2r[kFftBatchPlans] = {};
    float* square = nullptr;
    float2* spec = nullptr;
    void* work = nullptr;
    artia_status ensure(artia_ctx* ctx, int n);
    void release();
};
}  // namespace artia

struct artia_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int numerics = ARTIA_NUMERICS_EXACT;
    bool march_lds = true;   // exact forward projection through LDS-staged slabs where it applies
    size_t stack_limit = 0;  // bytes the CTF slab stacks may take (0: the built-in caps)
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
    artia::FftPlans fft;
    artia::FftBatch fftb;
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

// A volume as the launchers take it: dense [z][y][x] fp32 voxels, or IEEE binary16 voxels (half,
// the reference's FP16Volume storage).  Converts implicitly from a float pointer.
struct VolData {
    void* p = nullptr;
    bool half = false;
    VolData() = default;
    VolData(float* f) : p(f) {}
    VolData(const float* f) : p(const_cast<float*>(f)) {}
    static VolData f16(const uint16_t* h) {
        VolData d;
        d.p = const_cast<uint16_t*>(h);
        d.half = true;
        return d;
    }
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

// The allocation-free SART entry point's workspace: a device array of kNoAllocScratchFloats floats
// (64 MiB) per GPU, held for the life of the process.  noalloc_scratch gives the context device's
// address (looked up by artia_create), or an error when the lookup failed.
constexpr size_t kNoAllocScratchFloats = (size_t)16 << 20;
artia_status noalloc_scratch(artia_ctx* ctx, float** out);

// Records a start event for kernel class `kind` when profiling is on; returns the pair index
// (or -1).  prof_stop records the matching stop event; prof_collect waits for the last event
// and folds all pairs into the per-class totals.
int prof_start(artia_ctx* ctx, int kind);
void prof_stop(artia_ctx* ctx, int pair);
artia_status prof_collect(artia_ctx* ctx);
bool valid_geometry(const artia_vol_geom* vg, const artia_tilt_geom* tg);

// Kernel launchers used by the drivers (no argument validation, stream = ctx->stream)
void launch_march(artia_ctx* ctx, const VolG& vg, const TiltG& tg, VolData vol, int W, int H,
                  size_t pitch, float* proj, float* dist, int x0, int y0, int x1, int y1);
// Forward projection from zero straight into the SART residual: compare (real, Crop, CropDim,
// ProjectionScaleFactor) and, when corners != nullptr, cropBorder (cut, dim, corners), bit-identical
// to fill + march + compare + crop_border.  Exact numerics need no dist image (dist may be null);
// reference numerics run those launches and need dist.
void launch_march_residual(artia_ctx* ctx, const VolG& vg, const TiltG& tg, VolData vol, int W, int H, size_t pitch,
                           const float* real, float realLength, const float crop[4], const float cropDim[4],
                           float scale, const float cut[2], const float dim[2], const int* corners, float* resid,
                           float* dist);
void launch_vol_traversal(artia_ctx* ctx, const VolG& vg, const TiltG& tg, int W, int H, size_t pitch,
                          float* dist, int x0, int y0, int x1, int y1);
void launch_compare(artia_ctx* ctx, int W, int H, size_t pitch, const float* real, float* virt,
                    const float* dist, float realLength, const float cut[4], const float dim[4], float scale);
// cropBorder in place, `repeat` times over (the weight multiplied in that many times, as that many
// launches would)
void launch_crop_border(artia_ctx* ctx, int W, int H, size_t pitch, float* img, const float cut[2],
                        const float dim[2], const int p[8], int repeat = 1);
void launch_bp(artia_ctx* ctx, const VolG& vg, const TiltG& tg, VolData vol, const float* img, int W, int H,
               size_t pitch, float lambda, int os, float distMin, float distMax, int filter);
// The SART backprojections of one ordered subset (POINT sampling, gate [0, 1e13)): tilt k's residual at
// imgs[k], in tilt order.  Exact numerics run up to four tilts per launch, adding their updates in
// registers; the result equals n launch_bp calls.
void launch_bp_subset(artia_ctx* ctx, const VolG& vg, const TiltG* g, const float* const* imgs, int n, int W, int H,
                      size_t pitch, float lambda, int os, VolData vol);
// exact-numerics BP of one tilt's CTF slabs in one launch: slab k's image at stack + k * stride, slab k
// covering distances [bounds[k], bounds[k+1]) (the last [bounds[n-1], last_end)); bounds on the device
void launch_bp_stack(artia_ctx* ctx, const VolG& vg, const TiltG& tg, VolData vol, const float* stack, size_t stride,
                     int W, int H, size_t pitch, float lambda, int os, int filter, const float* bounds, int nslabs,
                     float last_end);
void launch_fill(artia_ctx* ctx, float* p, size_t n, float value);
void launch_slicer(artia_ctx* ctx, const VolG& vg, const TiltG& tg, VolData vol, int W, int H, size_t pitch,
                   float* proj, float tmin, float tmax);
// The slicer over slabs bounds[k] .. bounds[k] + step (k < nslab) into stack + k * stride; bounds
// on the host and, as d_bounds, on the device.  With scratch of slicer_stack_scratch_bytes(W, H, n)
// bytes the slabs go in passes of n (one volume pass each), otherwise one slicer launch per slab.
size_t slicer_stack_scratch_bytes(int W, int H, int nslab);
void launch_slicer_stack(artia_ctx* ctx, const VolG& v, const TiltG& g, VolData vol, int W, int H, size_t pitch,
                         float* stack, size_t stride, const float* bounds, const float* d_bounds, int nslab,
                         float step, void* scratch, size_t scratch_bytes);
void launch_make_square(artia_ctx* ctx, int W, int H, size_t pitch, const float* in, int N, float* out, int bx, int by,
                        int mirrorY, int fillZero);

// Fourier-space operators (spectral.hip)
float ctf_wavelength(float voltage_kV);
artia_status run_ctf(artia_ctx* ctx, float2* spec, int N, size_t pitch, const artia_ctf_params& p, float dmin_nm,
                     float dmax_nm, float angle_deg, int forFP);
artia_status run_wbp_weight(artia_ctx* ctx, float2* spec, int N, size_t pitch, float psi_deg, int method);
// cropBorder(img) -> makeSquare -> R2C -> ctf -> C2R -> out (+)= cropBorder(square / N^2)
artia_status ctf_slab_filter(artia_ctx* ctx, float* img, int W, int H, size_t pitch, float* out, size_t out_pitch,
                             int accumulate, const int corners[8], const float cut[2], const float dim[2],
                             const artia_ctf_params& p, float dmin_nm, float dmax_nm, float angle_deg, int forFP);
// The slab filter over a stack of n slabs, batched (exact numerics), bit-identical to n calls of
// ctf_slab_filter in slab order:
//   forFP = 1: slab s reads src + s * src_stride (cropped once); out (one image) += every slab's
//              cropped result in slab order
//   forFP = 0: every slab reads src (src_stride 0), cropped s + 1 times as the repeated in-place
//              crop leaves it; slab s's result is stored at out + s * out_stride; src is not changed
// corners == nullptr: no border crop (the sub-volume WBP's slab filter)
artia_status ctf_filter_slabs(artia_ctx* ctx, const float* src, size_t src_stride, int n, int W, int H, size_t pitch,
                              float* out, size_t out_stride, const int corners[8], const float cut[2],
                              const float dim[2], const artia_ctf_params& p, const float* dmin_nm,
                              const float* dmax_nm, float angle_deg, int forFP);
// makeSquare (reflect) -> R2C -> wbpWeighting -> C2R -> out = square / N^2
artia_status wbp_filter(artia_ctx* ctx, const float* img, int W, int H, size_t pitch, float* out, size_t out_pitch,
                        float psi_deg, int method);

}  // namespace artia
