// Device-resident SART driver: EmSART's ordered-subset loop (EmSart.cpp:579-675) over the
// operators of this library, without the per-tilt host round trips of the reference
// (residual D2H/H2D at EmSart.cpp:632,654 and the projection upload of
// Reconstructor.cpp:1395).  All launches go to the context stream back to back.
#include "artia_internal.hpp"

using namespace artia;

extern "C" {

artia_status artia_sart_run(artia_ctx* ctx, const artia_vol_geom* vg, const artia_tilt_geom* tilts,
                            const int* hitpoints, const float* real_lengths, int ntilts, const int* index_list,
                            const float* projections, int width, int height, float* vol, const artia_sart_params* sp,
                            int iterations, int* volume_is_empty, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!vg || !tilts || !hitpoints || !real_lengths || !index_list || !projections || !vol || !sp ||
        !volume_is_empty || ntilts <= 0 || width <= 0 || height <= 0 || iterations < 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_sart_run: invalid arguments");
    if (vg->nx % 4 != 0 || (reinterpret_cast<uintptr_t>(vol) & 15) != 0)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_sart_run: volume x dimension must be a multiple of 4 "
                                                     "and the volume 16-byte aligned");
    if (sp->sirt_count <= 0 || sp->oversampling <= 0 || sp->oversampling > 64)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_sart_run: bad SIRTCount or OverSampling");
    for (int i = 0; i < ntilts; i++) {
        if (index_list[i] < 0 || index_list[i] >= ntilts)
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_sart_run: index list entry %d out of range", i);
        if (!valid_geometry(vg, &tilts[i]))
            return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_sart_run: invalid geometry for tilt %d", i);
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

    Timing t(ctx, ms);
    for (int iter = 0; iter < iterations; iter++) {
        for (int step = 0; step < (ntilts + S - 1) / S; step++) {
            for (int i = 0; i < S; i++) {
                if (step * S + i >= ntilts) continue;
                const int idx = index_list[step * S + i];
                const TiltG g = to_device(tilts[idx]);
                float* proj = residual + (size_t)i * npix;
                int pr = prof_start(ctx, ARTIA_PROF_FILL);
                launch_fill(ctx, proj, npix, 0.0f);
                launch_fill(ctx, dist, npix, 0.0f);
                prof_stop(ctx, pr);
                pr = prof_start(ctx, ARTIA_PROF_FP);
                if (!empty)
                    launch_march(ctx, v, g, vol, width, height, pitch, proj, dist, 0, 0, width, height);
                else
                    launch_vol_traversal(ctx, v, g, width, height, pitch, dist, 0, 0, width, height);
                prof_stop(ctx, pr);
                pr = prof_start(ctx, ARTIA_PROF_COMPARE);
                launch_compare(ctx, width, height, pitch, projections + (size_t)idx * npix, proj, dist,
                               real_lengths[idx], sp->crop, sp->crop_dim, sp->proj_scale);
                prof_stop(ctx, pr);
            }
            for (int i = 0; i < S; i++) {
                if (step * S + i >= ntilts) continue;
                empty = false;
                const int idx = index_list[step * S + i];
                const TiltG g = to_device(tilts[idx]);
                float* res = residual + (size_t)i * npix;
                int pr = prof_start(ctx, ARTIA_PROF_CROP);
                launch_crop_border(ctx, width, height, pitch, res, sp->cut_length, sp->dim_length,
                                   hitpoints + 8 * (size_t)idx);
                prof_stop(ctx, pr);
                pr = prof_start(ctx, ARTIA_PROF_BP);
                launch_bp(ctx, v, g, vol, res, width, height, pitch, lam, sp->oversampling, 0.0f, 9999999999999.0f,
                          ARTIA_FILTER_POINT);
                prof_stop(ctx, pr);
            }
        }
    }
    *volume_is_empty = empty ? 1 : 0;
    st = t.finish();
    if (st != ARTIA_OK) return st;
    return prof_collect(ctx);
}

}  // extern "C"
