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

artia_status sart_run(artia_ctx* ctx, const char* fn, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                      const int* hitpoints, const float* real_lengths, int ntilts, const int* index_list,
                      const float* projections, int width, int height, VolData vol, const artia_sart_params* sp,
                      int iterations, int* volume_is_empty, float* ms) {
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
    artia_status st = ensure_workspace(ctx, npix * (size_t)(1 + S));
    if (st != ARTIA_OK) return st;
    float* dist = ctx->ws;
    float* residual = ctx->ws + npix;   // S images; the forward projection is built in place
    const VolG v = to_device(*vg);
    const size_t pitch = (size_t)width;
    const float lam = sp->lambda / (float)S;
    bool empty = *volume_is_empty != 0;
    std::vector<TiltG> sub_geoms((size_t)S);
    std::vector<const float*> sub_imgs((size_t)S);

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
            for (int i = 0; i < nsub; i++) {
                sub_geoms[(size_t)i] = to_device(tilts[index_list[step * S + i]]);
                sub_imgs[(size_t)i] = residual + (size_t)i * npix;
            }
            if (nsub > 0) {
                empty = false;
                const int pr = prof_start(ctx, ARTIA_PROF_BP);
                launch_bp_subset(ctx, v, sub_geoms.data(), sub_imgs.data(), nsub, width, height, pitch, lam,
                                 sp->oversampling, vol);
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
    // one launch per tilt over a stack of its filtered slab images, while the stack fits 32 GiB
    // the slab stack (32 GiB at most, or the context's stack limit) and the stacked slicer's sub-ray
    // sums: every slab of a tilt in one volume pass while they fit 16 GiB (or what the limit leaves),
    // in passes of fewer slabs otherwise; without room for the stack, one launch per slab
    const double stack_cap = ctx->stack_limit ? (double)ctx->stack_limit : 34359738368.0;
    const bool stacked = ctx->numerics == ARTIA_NUMERICS_EXACT && max_slabs > 0 &&
                         (double)max_slabs * (double)npix * 4.0 <= stack_cap;
    const size_t stack_floats = stacked ? (size_t)max_slabs * npix : 0;
    size_t slicer_bytes = 0;
    if (stacked) {
        const size_t cap = ctx->stack_limit ? ctx->stack_limit - stack_floats * sizeof(float) : ((size_t)16 << 30);
        int pass = max_slabs;
        while (pass > 1 && slicer_stack_scratch_bytes(width, height, pass) > cap) pass = (pass + 1) / 2;
        // no room for one slab's sums: the slicer goes slab by slab
        slicer_bytes = slicer_stack_scratch_bytes(width, height, pass) <= cap ? slicer_stack_scratch_bytes(width, height, pass)
                                                                              : 0;
    }
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
    // image corners for the forward slabs (Reconstructor.cpp:557-565): pA, pB, pC, pD
    const int frame[8] = {0, height - 1, width - 1, height - 1, 0, 0, width - 1, 0};
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
                if (!empty && stacked) {
                    // every slab's slicer image into the stack, then the batched filter sums them in
                    // slab order
                    const size_t f = first[(size_t)idx];
                    const int ns = nslab[(size_t)idx];
                    pr = prof_start(ctx, ARTIA_PROF_FP);
                    launch_slicer_stack(ctx, v, g, vol, width, height, pitch, stack, npix, &bounds[f], d_bounds + f, ns,
                                        step, slicer_scratch, slicer_bytes);
                    prof_stop(ctx, pr);
                    pr = prof_start(ctx, ARTIA_PROF_FFT);
                    st = ctf_filter_slabs(ctx, stack, npix, ns, width, height, pitch, proj, 0, frame, sp->cut_length,
                                          sp->dim_length, *ctf, &sdmin[f], &sdmax[f], defocus[idx].astig_angle_deg, 1);
                    prof_stop(ctx, pr);
                    if (st != ARTIA_OK) return st;
                    pr = prof_start(ctx, ARTIA_PROF_FILL);
                    launch_fill(ctx, dist, npix, 0.0f);
                    prof_stop(ctx, pr);
                } else if (!empty) {
                    for (float ray = slab_range[2 * idx]; ray < slab_range[2 * idx + 1]; ray += step) {
                        float dmin, dmax;
                        artia_geom_defocus_minmax(ray, &defocus[idx], ctf, &dmin, &dmax);
                        pr = prof_start(ctx, ARTIA_PROF_FILL);
                        launch_fill(ctx, dist, npix, 0.0f);
                        prof_stop(ctx, pr);
                        pr = prof_start(ctx, ARTIA_PROF_FP);
                        launch_slicer(ctx, v, g, vol, width, height, pitch, dist, ray, ray + step);
                        prof_stop(ctx, pr);
                        pr = prof_start(ctx, ARTIA_PROF_FFT);
                        st = ctf_slab_filter(ctx, dist, width, height, pitch, proj, pitch, 1, frame, sp->cut_length,
                                             sp->dim_length, *ctf, dmin, dmax, defocus[idx].astig_angle_deg, 1);
                        prof_stop(ctx, pr);
                        if (st != ARTIA_OK) return st;
                    }
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
                const int* hp = hitpoints + 8 * (size_t)idx;
                if (stacked) {
                    // every slab's filtered image into the stack, then one launch
                    const size_t f = first[(size_t)idx];
                    int pr = prof_start(ctx, ARTIA_PROF_FFT);
                    st = ctf_filter_slabs(ctx, res, 0, nslab[(size_t)idx], width, height, pitch, stack, npix, hp,
                                          sp->cut_length, sp->dim_length, *ctf, &sdmin[f], &sdmax[f],
                                          defocus[idx].astig_angle_deg, 0);
                    prof_stop(ctx, pr);
                    if (st != ARTIA_OK) return st;
                    if (nslab[(size_t)idx] > 0) {
                        pr = prof_start(ctx, ARTIA_PROF_BP);
                        launch_bp_stack(ctx, v, g, vol, stack, npix, width, height, pitch, lam, sp->oversampling,
                                        ARTIA_FILTER_LINEAR, d_bounds + f, nslab[(size_t)idx], last_end[(size_t)idx]);
                        prof_stop(ctx, pr);
                    }
                    continue;
                }
                for (float ray = slab_range[2 * idx]; ray < slab_range[2 * idx + 1]; ray += step) {
                    float dmin, dmax;
                    artia_geom_defocus_minmax(ray, &defocus[idx], ctf, &dmin, &dmax);
                    int pr = prof_start(ctx, ARTIA_PROF_FFT);
                    st = ctf_slab_filter(ctx, res, width, height, pitch, dist, pitch, 0, hp, sp->cut_length,
                                         sp->dim_length, *ctf, dmin, dmax, defocus[idx].astig_angle_deg, 0);
                    prof_stop(ctx, pr);
                    if (st != ARTIA_OK) return st;
                    pr = prof_start(ctx, ARTIA_PROF_BP);
                    launch_bp(ctx, v, g, vol, dist, width, height, pitch, lam, sp->oversampling, ray, ray + step,
                              ARTIA_FILTER_LINEAR);
                    prof_stop(ctx, pr);
                }
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
