// Context lifecycle, error reporting, timing and workspace for libartia.
#include <cmath>
#include <cstdio>
#include <mutex>
#include <new>

#include "artia_internal.hpp"

namespace {
// Workspace of the allocation-free SART entry point (artia_sart_run_noalloc): a file-scope device
// array, so that entry point never allocates.  Every GPU the library runs on holds its own copy
// until the process ends.  Host code never uses the array's name (a host-side stand-in, not a
// device address): artia_create looks up each device's address with hipGetSymbolAddress, with that
// device current, and everything else takes the pointer kept here.
static __device__ __align__(256) float scratch[artia::kNoAllocScratchFloats];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_done[kMaxScratchDevices] = {};
hipError_t g_scratch_err[kMaxScratchDevices] = {};
float* g_scratch_ptr[kMaxScratchDevices] = {};

// Once per device, with that device current (artia_create's set-up).
void resolve_scratch(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = err;
    g_scratch_ptr[device] = err == hipSuccess ? static_cast<float*>(p) : nullptr;
    g_scratch_done[device] = true;
}
}  // namespace

namespace artia {

artia_status noalloc_scratch(artia_ctx* ctx, float** out) {
    const int d = ctx->device;
    if (d < 0 || d >= kMaxScratchDevices)
        return fail(ctx, ARTIA_ERR_UNSUPPORTED, "allocation-free workspace: device %d beyond the %d supported", d,
                    kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[d] || g_scratch_err[d] != hipSuccess || !g_scratch_ptr[d])
        return fail(ctx, ARTIA_ERR_DEVICE, "allocation-free workspace of device %d unavailable: %s", d,
                    g_scratch_done[d] ? hipGetErrorString(g_scratch_err[d]) : "not looked up");
    *out = g_scratch_ptr[d];
    return ARTIA_OK;
}

TiltG to_device(const artia_tilt_geom& g) {
    TiltG t;
    for (int i = 0; i < 3; i++) {
        t.d[i] = g.d[i];
        t.u[i] = g.u[i];
        t.v[i] = g.v[i];
        t.n[i] = g.n[i];
    }
    for (int i = 0; i < 16; i++) t.M[i] = g.det_matrix[i];
    for (int i = 0; i < 9; i++) {
        t.A[i] = g.mag_aniso[i];
        t.Ainv[i] = g.mag_aniso_inv[i];
    }
    return t;
}

VolG to_device(const artia_vol_geom& g) {
    VolG v;
    for (int i = 0; i < 3; i++) {
        v.bbMin[i] = g.bbox_min[i];
        v.bbMax[i] = g.bbox_max[i];
        v.rcp[i] = g.bbox_rcp[i];
        v.dim[i] = g.dim[i];
        v.vs[i] = g.voxel_size[i];
    }
    v.zshift = g.z_shift;
    v.nx = g.nx;
    v.ny = g.ny;
    v.nz = g.nz;
    return v;
}

artia_status fail(artia_ctx* ctx, artia_status code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return code;
}

artia_status check_hip(artia_ctx* ctx, hipError_t err, const char* what) {
    if (err == hipSuccess) return ARTIA_OK;
    artia_status code = err == hipErrorOutOfMemory ? ARTIA_ERR_OUT_OF_MEMORY : ARTIA_ERR_DEVICE;
    return fail(ctx, code, "%s: %s", what, hipGetErrorString(err));
}

Timing::Timing(artia_ctx* ctx, float* ms) : ctx_(ctx), ms_(ms) {
    if (ms_) (void)hipEventRecord(ctx_->ev_start, ctx_->stream);
}

artia_status Timing::finish() {
    artia_status st = check_hip(ctx_, hipGetLastError(), "kernel launch");
    if (st != ARTIA_OK) return st;
    if (!ms_) return ARTIA_OK;
    st = check_hip(ctx_, hipEventRecord(ctx_->ev_stop, ctx_->stream), "hipEventRecord");
    if (st != ARTIA_OK) return st;
    st = check_hip(ctx_, hipEventSynchronize(ctx_->ev_stop), "hipEventSynchronize");
    if (st != ARTIA_OK) return st;
    return check_hip(ctx_, hipEventElapsedTime(ms_, ctx_->ev_start, ctx_->ev_stop), "hipEventElapsedTime");
}

artia_status ensure_workspace(artia_ctx* ctx, size_t floats) {
    if (ctx->ws_floats >= floats) return ARTIA_OK;
    if (ctx->ws) {
        artia_status st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        if (st != ARTIA_OK) return st;
        (void)hipFree(ctx->ws);
        ctx->ws = nullptr;
        ctx->ws_floats = 0;
    }
    artia_status st = check_hip(ctx, hipMalloc(&ctx->ws, floats * sizeof(float)), "hipMalloc(workspace)");
    if (st != ARTIA_OK) return st;
    ctx->ws_floats = floats;
    return ARTIA_OK;
}

bool valid_geometry(const artia_vol_geom* vg, const artia_tilt_geom* tg) {
    if (!vg || !tg) return false;
    if (vg->nx <= 0 || vg->ny <= 0 || vg->nz <= 0) return false;
    for (int i = 0; i < 3; i++) {
        if (!(vg->voxel_size[i] > 0.0f) || !std::isfinite(vg->bbox_min[i]) || !std::isfinite(vg->bbox_max[i]))
            return false;
        if (!std::isfinite(tg->d[i]) || !std::isfinite(tg->u[i]) || !std::isfinite(tg->v[i]) ||
            !std::isfinite(tg->n[i]))
            return false;
    }
    // the ray direction must be a unit vector (up to float rounding) for the sample loops to be bounded
    const double nn = (double)tg->n[0] * tg->n[0] + (double)tg->n[1] * tg->n[1] + (double)tg->n[2] * tg->n[2];
    if (!(nn > 0.98 && nn < 1.02)) return false;
    return true;
}

int prof_start(artia_ctx* ctx, int kind) {
    if (!ctx->prof) return -1;
    const size_t pair = ctx->prof_marks.size();
    while (ctx->prof_events.size() < 2 * (pair + 1)) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->prof_events.push_back(e);
    }
    ctx->prof_marks.push_back(kind);
    (void)hipEventRecord(ctx->prof_events[2 * pair], ctx->stream);
    return (int)pair;
}

void prof_stop(artia_ctx* ctx, int pair) {
    if (pair >= 0) (void)hipEventRecord(ctx->prof_events[2 * (size_t)pair + 1], ctx->stream);
}

artia_status prof_collect(artia_ctx* ctx) {
    if (!ctx->prof || ctx->prof_marks.empty()) return ARTIA_OK;
    const size_t n = ctx->prof_marks.size();
    artia_status st = check_hip(ctx, hipEventSynchronize(ctx->prof_events[2 * n - 1]), "hipEventSynchronize");
    if (st != ARTIA_OK) return st;
    for (size_t i = 0; i < n; i++) {
        float ms = 0.0f;
        st = check_hip(ctx, hipEventElapsedTime(&ms, ctx->prof_events[2 * i], ctx->prof_events[2 * i + 1]),
                       "hipEventElapsedTime");
        if (st != ARTIA_OK) return st;
        ctx->prof_ms[ctx->prof_marks[i]] += ms;
        ctx->prof_launches[ctx->prof_marks[i]] += 1;
    }
    ctx->prof_marks.clear();
    return ARTIA_OK;
}

__global__ void fill_kernel(float* __restrict__ p, size_t n, float value) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = value;
}

void launch_fill(artia_ctx* ctx, float* p, size_t n, float value) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, p, n, value);
}

}  // namespace artia

using namespace artia;

extern "C" {

artia_status artia_create(int device, artia_ctx** out) {
    if (!out) return ARTIA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return ARTIA_ERR_DEVICE;
    artia_ctx* ctx = new (std::nothrow) artia_ctx();
    if (!ctx) return ARTIA_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipEventCreate(&ctx->ev_start) != hipSuccess ||
        hipEventCreate(&ctx->ev_stop) != hipSuccess) {
        delete ctx;
        return ARTIA_ERR_DEVICE;
    }
    // the allocation-free entry point's workspace address on this device (device current)
    resolve_scratch(device);
    ctx->stream = nullptr;   // the null stream until the caller sets one
    *out = ctx;
    return ARTIA_OK;
}

artia_status artia_destroy(artia_ctx* ctx) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    else (void)hipDeviceSynchronize();
    if (ctx->ws) (void)hipFree(ctx->ws);
    ctx->fft.release();
    ctx->fftb.release();
    for (hipEvent_t e : ctx->prof_events) (void)hipEventDestroy(e);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    delete ctx;
    return ARTIA_OK;
}

artia_status artia_set_stream(artia_ctx* ctx, void* hip_stream) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    ctx->stream = (hipStream_t)hip_stream;
    return ARTIA_OK;
}

artia_status artia_get_stream(artia_ctx* ctx, void** hip_stream) {
    if (!ctx || !hip_stream) return ARTIA_ERR_INVALID_ARGUMENT;
    *hip_stream = (void*)ctx->stream;
    return ARTIA_OK;
}

artia_status artia_set_numerics(artia_ctx* ctx, int numerics) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (numerics != ARTIA_NUMERICS_EXACT && numerics != ARTIA_NUMERICS_REFERENCE)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_set_numerics: unknown mode %d", numerics);
    ctx->numerics = numerics;
    return ARTIA_OK;
}

artia_status artia_set_fp_staging(artia_ctx* ctx, int enable) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    ctx->march_lds = enable != 0;
    return ARTIA_OK;
}

artia_status artia_set_stack_limit(artia_ctx* ctx, size_t bytes) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    ctx->stack_limit = bytes;
    return ARTIA_OK;
}

artia_status artia_sync(artia_ctx* ctx) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

artia_status artia_profile_enable(artia_ctx* ctx, int enable) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    ctx->prof = enable != 0;
    ctx->prof_marks.clear();
    for (int i = 0; i < ARTIA_PROF_COUNT; i++) {
        ctx->prof_ms[i] = 0.0;
        ctx->prof_launches[i] = 0;
    }
    return ARTIA_OK;
}

artia_status artia_profile_read(artia_ctx* ctx, int kernel, double* total_ms, long long* launches) {
    if (!ctx || kernel < 0 || kernel >= ARTIA_PROF_COUNT || !total_ms || !launches)
        return ARTIA_ERR_INVALID_ARGUMENT;
    *total_ms = ctx->prof_ms[kernel];
    *launches = ctx->prof_launches[kernel];
    return ARTIA_OK;
}

const char* artia_last_error(artia_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

const char* artia_version(void) { return "artia-mi355x 0.1 (gfx950)"; }

}  // extern "C"
