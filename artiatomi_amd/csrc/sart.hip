// Device-resident SART driver: EmSART's ordered-subset loop (EmSart.cpp:579-675) over the
// operators of this library, without the per-tilt host round trips of the reference
// (residual D2H/H2D at EmSart.cpp:632,654 and the projection upload of
// Reconstructor.cpp:1395).  All launches go to the context stream back to back.
#include <cmath>

#include <algorithm>
#include <vector>

#include "artia_internal.hpp"

using namespace artia;

namespace {
// Upper bound on CTF slabs per tilt: a slab loop longer than this means a degenerate range/step.
constexpr double kMaxSlabs = 1 << 16;

// One tilt's CTF slabs: [bounds[k], bounds[k] + step) for k < n, stepped in float as the
// reference's loops step them (Reconstructor.cpp:514, 934), with each slab's defocus range
// (GetDefocusMinMax, Reconstructor.cpp:522, 942); the last slab ends at last_end.
struct CtfSlabs {
    const float* bounds;     // host
    const float* d_bounds;   // device copy (stacked forms)
    const float* dmin;
    const float* dmax;
    int n;
    float step, last_end;
};

// Workspace of the stacked forms: a tilt's slab images at stack + k * stride and the stacked
// slicer's sub-ray sums.  on == false: one launch per slab (reference numerics, or no room).
struct CtfStackWs {
    bool on;
    float* stack;
    size_t stride;
    void* slicer_scratch;
    size_t slicer_bytes;
};

// Reconstructor::ForwardProjectionCTF's slab loop (Reconstructor.cpp:514-584) for one tilt:
// proj += cropBorder(ctf-filter(cropBorder(slicer(slab)))) over the slabs in slab order, with the
// image corners as the crop quadrilateral.  Stacked: every slab's slicer image in one volume pass,
// then the batched filter sums them in slab order; otherwise one slicer launch and one filter per
// slab through `scratch` (W x H).  Both leave the same bits in proj.
artia_status ctf_forward_slabs(artia_ctx* ctx, const VolG& v, const TiltG& g, VolData vol, int W, int H, size_t pitch,
                               float* proj, float* scratch, const CtfSlabs& s, const CtfStackWs& ws,
                               const artia_ctf_params& ctf, float astig_deg, const float cut[2], const float dim[2]) {
    // image corners (Reconstructor.cpp:557-565): pA, pB, pC, pD
    const int frame[8] = {0, H - 1, W - 1, H - 1, 0, 0, W - 1, 0};
    if (ws.on) {
        int pr = prof_start(ctx, ARTIA_PROF_FP);
        launch_slicer_stack(ctx, v, g, vol, W, H, pitch, ws.stack, ws.stride, s.bounds, s.d_bounds, s.n, s.step,
                            ws.slicer_scratch, ws.slicer_bytes);
        prof_stop(ctx, pr);
        pr = prof_start(ctx, ARTIA_PROF_FFT);
        const artia_status st = ctf_filter_slabs(ctx, ws.stack, ws.stride, s.n, W, H, pitch, proj, 0, frame, cut, dim,
                                                 ctf, s.dmin, s.dmax, astig_deg, 1);
        prof_stop(ctx, pr);
        return st;
    }
    const size_t plane = (size_t)(H - 1) * pitch + (size_t)W;
    for (int k = 0; k < s.n; k++) {
        int pr = prof_start(ctx, ARTIA_PROF_FILL);
        launch_fill(ctx, scratch, plane, 0.0f);
        prof_stop(ctx, pr);
        pr = prof_start(ctx, ARTIA_PROF_FP);
        launch_slicer(ctx, v, g, vol, W, H, pitch, scratch, s.bounds[k], s.bounds[k] + s.step);
        prof_stop(ctx, pr);
        pr = prof_start(ctx, ARTIA_PROF_FFT);
        const artia_status st = ctf_slab_filter(ctx, scratch, W, H, pitch, proj, pitch, 1, frame, cut, dim, ctf,
                                                s.dmin[k], s.dmax[k], astig_deg, 1);
        prof_stop(ctx, pr);
        if (st != ARTIA_OK) return st;
    }
    return ARTIA_OK;
}

// Reconstructor::BackProjectionCTF's slab loop (Reconstructor.cpp:934-966) for one tilt: per slab,
// cropBorder(res) in place, the Wiener/phase-flip filter, cropBorder, and the slab-gated LINEAR
// backprojection.  Stacked: the slabs are filtered in batches into the stack (res unchanged) and
// backprojected in one launch; otherwise one filter (res cropped in place) and one gated launch
// per slab through `scratch`.  Both leave the same bits in the volume.
artia_status ctf_backward_slabs(artia_ctx* ctx, const VolG& v, const TiltG& g, VolData vol, int W, int H, size_t pitch,
                                float* res, float* scratch, const int* corners, const CtfSlabs& s,
                                const CtfStackWs& ws, const artia_ctf_params& ctf, float astig_deg, const float cut[2],
                                const float dim[2], float lambda, int os) {
    if (ws.on) {
        int pr = prof_start(ctx, ARTIA_PROF_FFT);
        const artia_status st = ctf_filter_slabs(ctx, res, 0, s.n, W, H, pitch, ws.stack, ws.stride, corners, cut,
                                                 dim, ctf, s.dmin, s.dmax, astig_deg, 0);
        prof_stop(ctx, pr);
        if (st != ARTIA_OK || s.n == 0) return st;
        pr = prof_start(ctx, ARTIA_PROF_BP);
        launch_bp_stack(ctx, v, g, vol, ws.stack, ws.stride, W, H, pitch, lambda, os, ARTIA_FILTER_LINEAR, s.d_bounds,
                        s.n, s.last_end);
        prof_stop(ctx, pr);
        return ARTIA_OK;
    }
    for (int k = 0; k < s.n; k++) {
        int pr = prof_start(ctx, ARTIA_PROF_FFT);
        const artia_status st = ctf_slab_filter(ctx, res, W, H, pitch, scratch, pitch, 0, corners, cut, dim, ctf,
                                                s.dmin[k], s.dmax[k], astig_deg, 0);
        prof_stop(ctx, pr);
        if (st != ARTIA_OK) return st;
        pr = prof_start(ctx, ARTIA_PROF_BP);
        launch_bp(ctx, v, g, vol, scratch, W, H, pitch, lambda, os, s.bounds[k], s.bounds[k] + s.step,
                  ARTIA_FILTER_LINEAR);
        prof_stop(ctx, pr);
    }
    return ARTIA_OK;
}

// Bytes for the stacked slicer's sub-ray sums next to a stack of stack_floats: every slab of a tilt
// in one volume pass while they fit 16 GiB (or what the context's stack limit leaves), passes of
// fewer slabs otherwise; 0 (one slicer launch per slab) when not even one slab's sums fit.
size_t slicer_scratch_for(const artia_ctx* ctx, int W, int H, int max_slabs, size_t stack_floats) {
    const size_t cap = ctx->stack_limit ? ctx->stack_limit - stack_floats * sizeof(float) : ((size_t)16 << 30);
    int pass = max_slabs;
    while (pass > 1 && slicer_stack_scratch_bytes(W, H, pass) > cap) pass = (pass + 1) / 2;
    return slicer_stack_scratch_bytes(W, H, pass) <= cap ? slicer_stack_scratch_bytes(W, H, pass) : 0;
}

// Whether a stack of max_slabs images of `plane` floats takes the stacked forms: exact numerics and
// the stack within the context's stack limit (32 GiB by default).
bool stack_fits(const artia_ctx* ctx, int max_slabs, size_t plane) {
    const double cap = ctx->stack_limit ? (double)ctx->stack_limit : 34359738368.0;
    return ctx->numerics == ARTIA_NUMERICS_EXACT && max_slabs > 0 && (double)max_slabs * (double)plane * 4.0 <= cap;
}

// The slabs and workspace of one per-tilt CTF call: the slab loop of GetDefocusDistances' range
// [a, b) stepped by CTFSliceThickness / pixel size, each slab's defocus range, and (stacked) the
// slab stack, its device bounds and the slicer's sums; `scratch` is one image (with_scratch).
struct TiltCtfPlan {
    std::vector<float> bounds, dmin, dmax;
    CtfSlabs s{};
    CtfStackWs ws{};
    float* scratch = nullptr;
};

artia_status plan_tilt_ctf(artia_ctx* ctx, const char* fn, float a, float b, const artia_tilt_defocus& d,
                           const artia_ctf_params& ctf, int W, int H, size_t pitch, bool with_scratch,
                           TiltCtfPlan& p) {
    const float step = ctf.slice_thickness_nm / ctf.pixelsize_nm;
    if (!std::isfinite(a) || !std::isfinite(b) || !(step > 0.0f) || (double)(b - a) / step > kMaxSlabs ||
        (b > a && a + step == a))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: the slab range holds too many slabs", fn);
    float ray = a;
    for (; ray < b; ray += step) p.bounds.push_back(ray);
    const int n = (int)p.bounds.size();
    p.dmin.resize((size_t)n);
    p.dmax.resize((size_t)n);
    for (int k = 0; k < n; k++) artia_geom_defocus_minmax(p.bounds[(size_t)k], &d, &ctf, &p.dmin[(size_t)k], &p.dmax[(size_t)k]);
    const size_t plane = pitch * (size_t)H;
    const bool stacked = stack_fits(ctx, n, plane);
    const size_t stack_floats = stacked ? (size_t)n * plane : 0;
    const size_t slicer_bytes = stacked ? slicer_scratch_for(ctx, W, H, n, stack_floats) : 0;
    const size_t scratch_floats = with_scratch ? plane : 0;
    const size_t head = (scratch_floats + stack_floats + (stacked ? (size_t)n : 0) + 63) & ~(size_t)63;
    artia_status st = ensure_workspace(ctx, head + (slicer_bytes + sizeof(float) - 1) / sizeof(float));
    if (st != ARTIA_OK) return st;
    p.scratch = with_scratch ? ctx->ws : nullptr;
    float* stack = ctx->ws + scratch_floats;
    float* d_bounds = stack + stack_floats;
    if (stacked && n > 0) {
        // earlier work on the stream may still read the workspace
        st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        if (st == ARTIA_OK)
            st = check_hip(ctx, hipMemcpy(d_bounds, p.bounds.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice),
                           "hipMemcpy(slab bounds)");
        if (st != ARTIA_OK) return st;
    }
    p.s = CtfSlabs{p.bounds.data(), d_bounds, p.dmin.data(), p.dmax.data(), n, step, ray};
    p.ws = CtfStackWs{stacked, stack, plane, ctx->ws + head, slicer_bytes};
    return ARTIA_OK;
}

bool valid_ctf_params(const artia_ctf_params* ctf) {
    return ctf && ctf->slice_thickness_nm > 0.0f && ctf->pixelsize_nm > 0.0f && ctf->voltage_kV > 0.0f &&
           ctf->wiener >= 0.0f;
}

artia_status fp_ctf(artia_ctx* ctx, const char* fn, const artia_vol_geom* vg, const artia_tilt_geom* tg, VolData vol,
                    int volume_is_empty, int width, int height, size_t pitch, float slab_start, float slab_end,
                    const artia_tilt_defocus* defocus, const artia_ctf_params* ctf, const float cut_length[2],
                    const float dim_length[2], float* proj, float* dist, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!vg || !tg || (!vol.p && !volume_is_empty) || !defocus || !cut_length || !dim_length || !proj || !dist ||
        width <= 0 || height <= 0 || pitch < (size_t)width || !valid_ctf_params(ctf))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: invalid arguments", fn);
    if (!valid_geometry(vg, tg)) return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: invalid geometry", fn);
    const VolG v = to_device(*vg);
    const TiltG g = to_device(*tg);
    TiltCtfPlan p;
    if (!volume_is_empty) {
        const artia_status st = plan_tilt_ctf(ctx, fn, slab_start, slab_end, *defocus, *ctf, width, height, pitch,
                                              false, p);
        if (st != ARTIA_OK) return st;
    }
    Timing t(ctx, ms);
    if (!volume_is_empty) {
        // dist is the per-slab form's slab image
        const artia_status st = ctf_forward_slabs(ctx, v, g, vol, width, height, pitch, proj, dist, p.s, p.ws, *ctf,
                                                  defocus->astig_angle_deg, cut_length, dim_length);
        if (st != ARTIA_OK) return st;
    }
    launch_fill(ctx, dist, (size_t)(height - 1) * pitch + (size_t)width, 0.0f);
    launch_vol_traversal(ctx, v, g, width, height, pitch, dist, 0, 0, width, height);
    return t.finish();
}

artia_status bp_ctf(artia_ctx* ctx, const char* fn, const artia_vol_geom* vg, const artia_tilt_geom* tg, VolData vol,
                    float* residual, int width, int height, size_t pitch, const int hitpoints[8], float slab_start,
                    float slab_end, const artia_tilt_defocus* defocus, const artia_ctf_params* ctf,
                    const float cut_length[2], const float dim_length[2], float lambda, int oversampling, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!vg || !tg || !vol.p || !residual || !hitpoints || !defocus || !cut_length || !dim_length || width <= 0 ||
        height <= 0 || pitch < (size_t)width || oversampling <= 0 || oversampling > 64 || !valid_ctf_params(ctf))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: invalid arguments", fn);
    if (!valid_geometry(vg, tg)) return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: invalid geometry", fn);
    if (vg->nx % 4 != 0 || (reinterpret_cast<uintptr_t>(vol.p) & (vol.half ? 7u : 15u)) != 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: volume x dimension must be a multiple of 4 "
                                                     "and the volume 16-byte (half: 8-byte) aligned", fn);
    TiltCtfPlan p;
    artia_status st = plan_tilt_ctf(ctx, fn, slab_start, slab_end, *defocus, *ctf, width, height, pitch, true, p);
    if (st != ARTIA_OK) return st;
    const VolG v = to_device(*vg);
    const TiltG g = to_device(*tg);
    Timing t(ctx, ms);
    st = ctf_backward_slabs(ctx, v, g, vol, width, height, pitch, residual, p.scratch, hitpoints, p.s, p.ws, *ctf,
                            defocus->astig_angle_deg, cut_length, dim_length, lambda, oversampling);
    if (st != ARTIA_OK) return st;
    // the per-slab loop crops the residual in place once per slab; the stacked form reads it unchanged
    if (p.ws.on) launch_crop_border(ctx, width, height, pitch, residual, cut_length, dim_length, hitpoints, p.s.n);
    return t.finish();
}

// The subset's tilts go to the backprojection in chunks of this many (a fixed host array: the
// driver's loop allocates nothing); chunking does not change the result.
constexpr int kSubsetChunk = 16;

// fixed_ws != nullptr: the allocation-free form, with (1 + SIRTCount) images of workspace taken
// from fixed_ws (fixed_floats floats) instead of the context's workspace.
artia_status sart_run(artia_ctx* ctx, const char* fn, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                      const int* hitpoints, const float* real_lengths, int ntilts, const int* index_list,
                      const float* projections, int width, int height, VolData vol, const artia_sart_params* sp,
                      int iterations, int* volume_is_empty, float* ms, float* fixed_ws = nullptr,
                      size_t fixed_floats = 0) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!vg || !tilts || !hitpoints || !real_lengths || !index_list || !projections || !vol.p || !sp ||
        !volume_is_empty || ntilts <= 0 || width <= 0 || height <= 0 || iterations < 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: invalid arguments", fn);
    if (vg->nx % 4 != 0 || (reinterpret_cast<uintptr_t>(vol.p) & (vol.half ? 7u : 15u)) != 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: volume x dimension must be a multiple of 4 "
                                                     "and the volume 16-byte (half: 8-byte) aligned", fn);
    if (sp->sirt_count <= 0 || sp->oversampling <= 0 || sp->oversampling > 64)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: bad SIRTCount or OverSampling", fn);
    for (int i = 0; i < ntilts; i++) {
        if (index_list[i] < 0 || index_list[i] >= ntilts)
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: index list entry %d out of range", fn, i);
        if (!valid_geometry(vg, &tilts[i]))
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: invalid geometry for tilt %d", fn, i);
    }
    const int S = sp->sirt_count;
    const size_t npix = (size_t)width * (size_t)height;
    const size_t need = npix * (size_t)(1 + S);
    artia_status st = ARTIA_OK;
    float* ws = nullptr;
    if (fixed_ws) {
        if (need > fixed_floats)
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: %d x %d images with SIRTCount %d need %zu floats of "
                        "workspace; the allocation-free entry point holds at most %zu", fn, width, height, S, need,
                        fixed_floats);
        if (ctx->prof)
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: kernel profiling is on, and its event pool "
                                                         "allocates; turn it off first", fn);
        ws = fixed_ws;
    } else {
        st = ensure_workspace(ctx, need);
        if (st != ARTIA_OK) return st;
        ws = ctx->ws;
    }
    float* dist = ws;
    float* residual = ws + npix;   // S images; the forward projection is built in place
    const VolG v = to_device(*vg);
    const size_t pitch = (size_t)width;
    const float lam = sp->lambda / (float)S;
    bool empty = *volume_is_empty != 0;
    TiltG sub_geoms[kSubsetChunk];
    const float* sub_imgs[kSubsetChunk];

    Timing t(ctx, ms);
    for (int iter = 0; iter < iterations; iter++) {
        for (int step = 0; step < (ntilts + S - 1) / S; step++) {
            for (int i = 0; i < S; i++) {
                if (step * S + i >= ntilts) continue;
                const int idx = index_list[step * S + i];
                const TiltG g = to_device(tilts[idx]);
                float* proj = residual + (size_t)i * npix;
                const float* real = projections + (size_t)idx * npix;
                const int* corners = hitpoints + 8 * (size_t)idx;
                if (!empty) {
                    // projection, compare and cropBorder in one launch (the residual epilogue)
                    const int pr = prof_start(ctx, ARTIA_PROF_FP);
                    launch_march_residual(ctx, v, g, vol, width, height, pitch, real, real_lengths[idx], sp->crop,
                                          sp->crop_dim, sp->proj_scale, sp->cut_length, sp->dim_length, corners,
                                          proj, dist);
                    prof_stop(ctx, pr);
                    continue;
                }
                int pr = prof_start(ctx, ARTIA_PROF_FILL);
                launch_fill(ctx, proj, npix, 0.0f);
                launch_fill(ctx, dist, npix, 0.0f);
                prof_stop(ctx, pr);
                pr = prof_start(ctx, ARTIA_PROF_FP);
                launch_vol_traversal(ctx, v, g, width, height, pitch, dist, 0, 0, width, height);
                prof_stop(ctx, pr);
                pr = prof_start(ctx, ARTIA_PROF_COMPARE);
                launch_compare(ctx, width, height, pitch, real, proj, dist, real_lengths[idx], sp->crop, sp->crop_dim,
                               sp->proj_scale);
                prof_stop(ctx, pr);
                pr = prof_start(ctx, ARTIA_PROF_CROP);
                launch_crop_border(ctx, width, height, pitch, proj, sp->cut_length, sp->dim_length, corners);
                prof_stop(ctx, pr);
            }
            // the subset's backprojections (residuals cropped already), in subset order
            const int nsub = std::min(S, ntilts - step * S);
            for (int c0 = 0; c0 < nsub; c0 += kSubsetChunk) {
                const int nc = std::min(kSubsetChunk, nsub - c0);
                for (int i = 0; i < nc; i++) {
                    sub_geoms[i] = to_device(tilts[index_list[step * S + c0 + i]]);
                    sub_imgs[i] = residual + (size_t)(c0 + i) * npix;
                }
                empty = false;
                const int pr = prof_start(ctx, ARTIA_PROF_BP);
                launch_bp_subset(ctx, v, sub_geoms, sub_imgs, nc, width, height, pitch, lam, sp->oversampling, vol);
                prof_stop(ctx, pr);
            }
        }
    }
    *volume_is_empty = empty ? 1 : 0;
    st = t.finish();
    if (st != ARTIA_OK) return st;
    return prof_collect(ctx);
}

artia_status sart_ctf_run(artia_ctx* ctx, const char* fn, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                          const int* hitpoints, const float* real_lengths, const float* slab_range,
                          const artia_tilt_defocus* defocus, int ntilts, const int* index_list,
                          const float* projections, int width, int height, VolData vol, const artia_sart_params* sp,
                          const artia_ctf_params* ctf, int iterations, int* volume_is_empty, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!vg || !tilts || !hitpoints || !real_lengths || !slab_range || !defocus || !index_list || !projections ||
        !vol.p || !sp || !ctf || !volume_is_empty || ntilts <= 0 || width <= 0 || height <= 0 || iterations < 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: invalid arguments", fn);
    if (vg->nx % 4 != 0 || (reinterpret_cast<uintptr_t>(vol.p) & (vol.half ? 7u : 15u)) != 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: volume x dimension must be a multiple "
                                                     "of 4 and the volume 16-byte (half: 8-byte) aligned", fn);
    if (sp->sirt_count <= 0 || sp->oversampling <= 0 || sp->oversampling > 64)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: bad SIRTCount or OverSampling", fn);
    if (!(ctf->slice_thickness_nm > 0.0f) || !(ctf->pixelsize_nm > 0.0f) || !(ctf->voltage_kV > 0.0f) ||
        !(ctf->wiener >= 0.0f))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: bad CTF parameters", fn);
    // Reconstructor.cpp:514: the slab step in pixels, accumulated in float as the reference does
    const float step = ctf->slice_thickness_nm / ctf->pixelsize_nm;
    for (int i = 0; i < ntilts; i++) {
        if (index_list[i] < 0 || index_list[i] >= ntilts)
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: index list entry %d out of range", fn, i);
        if (!valid_geometry(vg, &tilts[i]))
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: invalid geometry for tilt %d", fn, i);
        const float a = slab_range[2 * i], b = slab_range[2 * i + 1];
        if (!std::isfinite(a) || !std::isfinite(b) || !(step > 0.0f) || (double)(b - a) / step > (double)kMaxSlabs ||
            (b > a && a + step == a))
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "%s: slab range of tilt %d holds too many "
                                                         "slabs", fn, i);
    }
    const int S = sp->sirt_count;
    const size_t npix = (size_t)width * (size_t)height;
    // the backprojection slabs of every tilt, as the reference's float loop steps through them
    std::vector<float> bounds;
    std::vector<size_t> first((size_t)ntilts);
    std::vector<int> nslab((size_t)ntilts);
    std::vector<float> last_end((size_t)ntilts);
    int max_slabs = 0;
    for (int i = 0; i < ntilts; i++) {
        first[(size_t)i] = bounds.size();
        float ray = slab_range[2 * i];
        for (; ray < slab_range[2 * i + 1]; ray += step) bounds.push_back(ray);
        nslab[(size_t)i] = (int)(bounds.size() - first[(size_t)i]);
        last_end[(size_t)i] = ray;
        max_slabs = std::max(max_slabs, nslab[(size_t)i]);
    }
    // one launch per tilt over a stack of its filtered slab images while the stack fits, one launch
    // per slab otherwise
    const bool stacked = stack_fits(ctx, max_slabs, npix);
    const size_t stack_floats = stacked ? (size_t)max_slabs * npix : 0;
    const size_t slicer_bytes = stacked ? slicer_scratch_for(ctx, width, height, max_slabs, stack_floats) : 0;
    const size_t head = (npix * (size_t)(1 + S) + stack_floats + (stacked ? bounds.size() : 0) + 63) & ~(size_t)63;
    artia_status st = ensure_workspace(ctx, head + (slicer_bytes + sizeof(float) - 1) / sizeof(float));
    if (st != ARTIA_OK) return st;
    float* dist = ctx->ws;
    float* residual = ctx->ws + npix;
    float* stack = ctx->ws + npix * (size_t)(1 + S);
    float* d_bounds = stack + stack_floats;
    void* slicer_scratch = ctx->ws + head;
    if (stacked) {
        // earlier work on the stream may still read the workspace
        st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        if (st == ARTIA_OK)
            st = check_hip(ctx, hipMemcpy(d_bounds, bounds.data(), bounds.size() * sizeof(float),
                                          hipMemcpyHostToDevice),
                           "hipMemcpy(slab bounds)");
        if (st != ARTIA_OK) return st;
    }
    // the defocus range of every slab (Reconstructor.cpp:522, 942)
    std::vector<float> sdmin(bounds.size()), sdmax(bounds.size());
    for (int i = 0; i < ntilts; i++)
        for (int k = 0; k < nslab[(size_t)i]; k++) {
            const size_t j = first[(size_t)i] + (size_t)k;
            artia_geom_defocus_minmax(bounds[j], &defocus[i], ctf, &sdmin[j], &sdmax[j]);
        }
    const VolG v = to_device(*vg);
    const size_t pitch = (size_t)width;
    const float lam = sp->lambda / (float)S;
    const CtfStackWs ws{stacked, stack, npix, slicer_scratch, slicer_bytes};
    auto slabs_of = [&](int i) {
        const size_t f = first[(size_t)i];
        return CtfSlabs{bounds.data() + f, d_bounds + f, sdmin.data() + f, sdmax.data() + f, nslab[(size_t)i], step,
                        last_end[(size_t)i]};
    };
    bool empty = *volume_is_empty != 0;

    Timing t(ctx, ms);
    for (int iter = 0; iter < iterations; iter++) {
        for (int sstep = 0; sstep < (ntilts + S - 1) / S; sstep++) {
            for (int i = 0; i < S; i++) {
                if (sstep * S + i >= ntilts) continue;
                const int idx = index_list[sstep * S + i];
                const TiltG g = to_device(tilts[idx]);
                float* proj = residual + (size_t)i * npix;
                int pr = prof_start(ctx, ARTIA_PROF_FILL);
                launch_fill(ctx, proj, npix, 0.0f);
                launch_fill(ctx, dist, npix, 0.0f);
                prof_stop(ctx, pr);
                if (!empty) {
                    st = ctf_forward_slabs(ctx, v, g, vol, width, height, pitch, proj, dist, slabs_of(idx), ws, *ctf,
                                           defocus[idx].astig_angle_deg, sp->cut_length, sp->dim_length);
                    if (st != ARTIA_OK) return st;
                    pr = prof_start(ctx, ARTIA_PROF_FILL);
                    launch_fill(ctx, dist, npix, 0.0f);
                    prof_stop(ctx, pr);
                }
                pr = prof_start(ctx, ARTIA_PROF_FP);
                launch_vol_traversal(ctx, v, g, width, height, pitch, dist, 0, 0, width, height);
                prof_stop(ctx, pr);
                pr = prof_start(ctx, ARTIA_PROF_COMPARE);
                launch_compare(ctx, width, height, pitch, projections + (size_t)idx * npix, proj, dist,
                               real_lengths[idx], sp->crop, sp->crop_dim, sp->proj_scale);
                prof_stop(ctx, pr);
            }
            for (int i = 0; i < S; i++) {
                if (sstep * S + i >= ntilts) continue;
                empty = false;
                const int idx = index_list[sstep * S + i];
                const TiltG g = to_device(tilts[idx]);
                float* res = residual + (size_t)i * npix;
                st = ctf_backward_slabs(ctx, v, g, vol, width, height, pitch, res, dist, hitpoints + 8 * (size_t)idx,
                                        slabs_of(idx), ws, *ctf, defocus[idx].astig_angle_deg, sp->cut_length,
                                        sp->dim_length, lam, sp->oversampling);
                if (st != ARTIA_OK) return st;
            }
        }
    }
    *volume_is_empty = empty ? 1 : 0;
    st = t.finish();
    if (st != ARTIA_OK) return st;
    return prof_collect(ctx);
}

}  // namespace

extern "C" {

artia_status artia_sart_run(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                            const int* hitpoints, const float* real_lengths, int ntilts, const int* index_list,
                            const float* projections, int width, int height, float* vol, const artia_sart_params* sp,
                            int iterations, int* volume_is_empty, float* ms) {
    return sart_run(ctx, "artia_sart_run", vg, tilts, hitpoints, real_lengths, ntilts, index_list, projections, width,
                    height, VolData(vol), sp, iterations, volume_is_empty, ms);
}

artia_status artia_sart_run_noalloc(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                                    const int* hitpoints, const float* real_lengths, int ntilts,
                                    const int* index_list, const float* projections, int width, int height,
                                    float* vol, const artia_sart_params* sp, int iterations, int* volume_is_empty,
                                    float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    float* ws = nullptr;
    const artia_status st = noalloc_scratch(ctx, &ws);
    if (st != ARTIA_OK) return st;
    return sart_run(ctx, "artia_sart_run_noalloc", vg, tilts, hitpoints, real_lengths, ntilts, index_list,
                    projections, width, height, VolData(vol), sp, iterations, volume_is_empty, ms, ws,
                    kNoAllocScratchFloats);
}

artia_status artia_sart_run_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                                const int* hitpoints, const float* real_lengths, int ntilts, const int* index_list,
                                const float* projections, int width, int height, uint16_t* vol,
                                const artia_sart_params* sp, int iterations, int* volume_is_empty, float* ms) {
    return sart_run(ctx, "artia_sart_run_f16", vg, tilts, hitpoints, real_lengths, ntilts, index_list, projections,
                    width, height, VolData::f16(vol), sp, iterations, volume_is_empty, ms);
}

artia_status artia_sart_ctf_run(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                                const int* hitpoints, const float* real_lengths, const float* slab_range,
                                const artia_tilt_defocus* defocus, int ntilts, const int* index_list,
                                const float* projections, int width, int height, float* vol,
                                const artia_sart_params* sp, const artia_ctf_params* ctf, int iterations,
                                int* volume_is_empty, float* ms) {
    return sart_ctf_run(ctx, "artia_sart_ctf_run", vg, tilts, hitpoints, real_lengths, slab_range, defocus, ntilts,
                        index_list, projections, width, height, VolData(vol), sp, ctf, iterations, volume_is_empty,
                        ms);
}

artia_status artia_sart_ctf_run_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                                    const int* hitpoints, const float* real_lengths, const float* slab_range,
                                    const artia_tilt_defocus* defocus, int ntilts, const int* index_list,
                                    const float* projections, int width, int height, uint16_t* vol,
                                    const artia_sart_params* sp, const artia_ctf_params* ctf, int iterations,
                                    int* volume_is_empty, float* ms) {
    return sart_ctf_run(ctx, "artia_sart_ctf_run_f16", vg, tilts, hitpoints, real_lengths, slab_range, defocus,
                        ntilts, index_list, projections, width, height, VolData::f16(vol), sp, ctf, iterations,
                        volume_is_empty, ms);
}

}  // extern "C"

extern "C" {

artia_status artia_fp_ctf(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, const float* vol,
                          int volume_is_empty, int width, int height, size_t pitch, float slab_start, float slab_end,
                          const artia_tilt_defocus* defocus, const artia_ctf_params* ctf, const float cut_length[2],
                          const float dim_length[2], float* proj, float* dist, float* ms) {
    return fp_ctf(ctx, "artia_fp_ctf", vg, tg, VolData(vol), volume_is_empty, width, height, pitch, slab_start,
                  slab_end, defocus, ctf, cut_length, dim_length, proj, dist, ms);
}

artia_status artia_fp_ctf_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, const uint16_t* vol,
                              int volume_is_empty, int width, int height, size_t pitch, float slab_start,
                              float slab_end, const artia_tilt_defocus* defocus, const artia_ctf_params* ctf,
                              const float cut_length[2], const float dim_length[2], float* proj, float* dist,
                              float* ms) {
    return fp_ctf(ctx, "artia_fp_ctf_f16", vg, tg, VolData::f16(vol), volume_is_empty, width, height, pitch,
                  slab_start, slab_end, defocus, ctf, cut_length, dim_length, proj, dist, ms);
}

artia_status artia_bp_ctf(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, float* vol,
                          float* residual, int width, int height, size_t pitch, const int hitpoints[8],
                          float slab_start, float slab_end, const artia_tilt_defocus* defocus,
                          const artia_ctf_params* ctf, const float cut_length[2], const float dim_length[2],
                          float lambda, int oversampling, float* ms) {
    return bp_ctf(ctx, "artia_bp_ctf", vg, tg, VolData(vol), residual, width, height, pitch, hitpoints, slab_start,
                  slab_end, defocus, ctf, cut_length, dim_length, lambda, oversampling, ms);
}

artia_status artia_bp_ctf_f16(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tg, uint16_t* vol,
                              float* residual, int width, int height, size_t pitch, const int hitpoints[8],
                              float slab_start, float slab_end, const artia_tilt_defocus* defocus,
                              const artia_ctf_params* ctf, const float cut_length[2], const float dim_length[2],
                              float lambda, int oversampling, float* ms) {
    return bp_ctf(ctx, "artia_bp_ctf_f16", vg, tg, VolData::f16(vol), residual, width, height, pitch, hitpoints,
                  slab_start, slab_end, defocus, ctf, cut_length, dim_length, lambda, oversampling, ms);
}

}  // extern "C"
