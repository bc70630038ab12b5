// artia_kernels.hpp - C++ functors over the C ABI of artia.h, shaped like the reference's
// operator classes (src/EmSART/Kernels.h:37-245) so that Reconstructor-style host code can swap
// its CUDA functors for these one for one.
//
// Differences from the reference, all deliberate:
//   - geometry is an argument (artia_vol_geom / artia_tilt_geom) instead of module __constant__
//     memory filled by SetConstantValues (Kernels.cpp:949-1132);
//   - images are device pointers with a pitch in elements, not CudaPitchedDeviceVariable;
//   - operator() returns the elapsed milliseconds and synchronises, like the reference
//     (Kernels.cpp:65-86); launch() enqueues on the context stream without waiting;
//   - failures throw artia::Error (a std::runtime_error), as the reference throws
//     CudaException (CudaException.h:34-36).
#ifndef ARTIA_KERNELS_HPP
#define ARTIA_KERNELS_HPP

#include <stdexcept>
#include <string>

#include "artia.h"

namespace artia {

class Error : public std::runtime_error {
public:
    Error(artia_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    artia_status status;
};

// One library context on one device (the reference's CUcontext + stream 0).
class Context {
public:
    explicit Context(int device = 0, void* hip_stream = nullptr) {
        check(artia_create(device, &ctx_), "artia_create");
        if (hip_stream) check(artia_set_stream(ctx_, hip_stream), "artia_set_stream");
    }
    ~Context() {
        if (ctx_) artia_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    artia_ctx* get() const { return ctx_; }
    void set_numerics(artia_numerics m) { check(artia_set_numerics(ctx_, m), "artia_set_numerics"); }
    void sync() { check(artia_sync(ctx_), "artia_sync"); }
    void check(artia_status s, const char* what) const {
        if (s != ARTIA_OK) {
            std::string msg = what;
            if (ctx_) {
                msg += ": ";
                msg += artia_last_error(ctx_);
            }
            throw Error(s, msg);
        }
    }

private:
    artia_ctx* ctx_ = nullptr;
};

// FPKernel (Kernels.h:37-45): proj += march, dist = mean chord.
class FPKernel {
public:
    explicit FPKernel(Context& c) : c_(c) {}
    float operator()(const artia_vol_geom& vg, const artia_tilt_geom& tg, const float* vol, int w, int h,
                     size_t pitch, float* proj, float* dist, const int* roi = nullptr) {
        float ms = 0.0f;
        c_.check(artia_fp_march(c_.get(), &vg, &tg, vol, w, h, pitch, proj, dist, roi, &ms), "FPKernel");
        return ms;
    }
    void launch(const artia_vol_geom& vg, const artia_tilt_geom& tg, const float* vol, int w, int h, size_t pitch,
                float* proj, float* dist, const int* roi = nullptr) {
        c_.check(artia_fp_march(c_.get(), &vg, &tg, vol, w, h, pitch, proj, dist, roi, nullptr), "FPKernel");
    }
    // a half volume (FP16Volume): the volume pointer holds IEEE half bits
    float operator()(const artia_vol_geom& vg, const artia_tilt_geom& tg, const uint16_t* vol, int w, int h,
                     size_t pitch, float* proj, float* dist, const int* roi = nullptr) {
        float ms = 0.0f;
        c_.check(artia_fp_march_f16(c_.get(), &vg, &tg, vol, w, h, pitch, proj, dist, roi, &ms), "FPKernel");
        return ms;
    }

private:
    Context& c_;
};

// SlicerKernel (Kernels.h:47-55): proj += march restricted to [tmin, tmax).
class SlicerKernel {
public:
    explicit SlicerKernel(Context& c) : c_(c) {}
    float operator()(const artia_vol_geom& vg, const artia_tilt_geom& tg, const float* vol, int w, int h,
                     size_t pitch, float* proj, float tmin, float tmax, const int* roi = nullptr) {
        float ms = 0.0f;
        c_.check(artia_fp_slicer(c_.get(), &vg, &tg, vol, w, h, pitch, proj, tmin, tmax, roi, &ms), "SlicerKernel");
        return ms;
    }
    float operator()(const artia_vol_geom& vg, const artia_tilt_geom& tg, const uint16_t* vol, int w, int h,
                     size_t pitch, float* proj, float tmin, float tmax, const int* roi = nullptr) {
        float ms = 0.0f;
        c_.check(artia_fp_slicer_f16(c_.get(), &vg, &tg, vol, w, h, pitch, proj, tmin, tmax, roi, &ms),
                 "SlicerKernel");
        return ms;
    }

private:
    Context& c_;
};

// VolTravLengthKernel (Kernels.h:57-65)
class VolTravLengthKernel {
public:
    explicit VolTravLengthKernel(Context& c) : c_(c) {}
    float operator()(const artia_vol_geom& vg, const artia_tilt_geom& tg, int w, int h, size_t pitch, float* dist,
                     const int* roi = nullptr) {
        float ms = 0.0f;
        c_.check(artia_vol_traversal(c_.get(), &vg, &tg, w, h, pitch, dist, roi, &ms), "VolTravLengthKernel");
        return ms;
    }

private:
    Context& c_;
};

// CompKernel (Kernels.h:67-74)
class CompKernel {
public:
    explicit CompKernel(Context& c) : c_(c) {}
    float operator()(int w, int h, size_t pitch, const float* real, float* virt, const float* dist,
                     float real_length, const float crop[4], const float crop_dim[4], float proj_scale) {
        float ms = 0.0f;
        c_.check(artia_compare(c_.get(), w, h, pitch, real, virt, dist, real_length, crop, crop_dim, proj_scale, &ms),
                 "CompKernel");
        return ms;
    }

private:
    Context& c_;
};

// CropBorderKernel (Kernels.h:76-83); p = p1..p4 as (x, y) pairs
class CropBorderKernel {
public:
    explicit CropBorderKernel(Context& c) : c_(c) {}
    float operator()(int w, int h, size_t pitch, float* img, const float cut[2], const float dim[2], const int p[8]) {
        float ms = 0.0f;
        c_.check(artia_crop_border(c_.get(), w, h, pitch, img, cut, dim, p, &ms), "CropBorderKernel");
        return ms;
    }

private:
    Context& c_;
};

// BPKernel (Kernels.h:85-92): the image is an explicit argument (the reference binds a texref)
class BPKernel {
public:
    explicit BPKernel(Context& c) : c_(c) {}
    float operator()(const artia_vol_geom& vg, const artia_tilt_geom& tg, float* vol, const float* img, int w,
                     int h, size_t pitch, float lambda, int oversampling, float dist_min, float dist_max,
                     artia_filter filter) {
        float ms = 0.0f;
        c_.check(artia_bp(c_.get(), &vg, &tg, vol, img, w, h, pitch, lambda, oversampling, dist_min, dist_max,
                          filter, &ms),
                 "BPKernel");
        return ms;
    }
    // BPKernel(module, fp16 = true), backProjectionFP16 (Kernels.cpp:327-334): a half volume
    float operator()(const artia_vol_geom& vg, const artia_tilt_geom& tg, uint16_t* vol, const float* img, int w,
                     int h, size_t pitch, float lambda, int oversampling, float dist_min, float dist_max,
                     artia_filter filter) {
        float ms = 0.0f;
        c_.check(artia_bp_f16(c_.get(), &vg, &tg, vol, img, w, h, pitch, lambda, oversampling, dist_min, dist_max,
                              filter, &ms),
                 "BPKernel");
        return ms;
    }

private:
    Context& c_;
};

// Reconstructor::ForwardProjectionCTF / BackProjectionCTF for one tilt (Reconstructor.cpp:496-644,
// 920-969): the slab loop over [slab_start, slab_end) with the slicer, the CTF filters and the
// slab-gated backprojection, as one call each.
class CtfTiltSteps {
public:
    explicit CtfTiltSteps(Context& c) : c_(c) {}
    float forward(const artia_vol_geom& vg, const artia_tilt_geom& tg, const float* vol, bool volume_is_empty, int w,
                  int h, size_t pitch, float slab_start, float slab_end, const artia_tilt_defocus& defocus,
                  const artia_ctf_params& ctf, const float cut[2], const float dim[2], float* proj, float* dist) {
        float ms = 0.0f;
        c_.check(artia_fp_ctf(c_.get(), &vg, &tg, vol, volume_is_empty, w, h, pitch, slab_start, slab_end, &defocus,
                              &ctf, cut, dim, proj, dist, &ms),
                 "ForwardProjectionCTF");
        return ms;
    }
    float backward(const artia_vol_geom& vg, const artia_tilt_geom& tg, float* vol, float* residual, int w, int h,
                   size_t pitch, const int hitpoints[8], float slab_start, float slab_end,
                   const artia_tilt_defocus& defocus, const artia_ctf_params& ctf, const float cut[2],
                   const float dim[2], float lambda, int oversampling) {
        float ms = 0.0f;
        c_.check(artia_bp_ctf(c_.get(), &vg, &tg, vol, residual, w, h, pitch, hitpoints, slab_start, slab_end,
                              &defocus, &ctf, cut, dim, lambda, oversampling, &ms),
                 "BackProjectionCTF");
        return ms;
    }

private:
    Context& c_;
};

// CopyToSquareKernel (Kernels.h:104-111)
class CopyToSquareKernel {
public:
    explicit CopyToSquareKernel(Context& c) : c_(c) {}
    float operator()(int w, int h, size_t pitch, const float* in, int n, float* out, int bx, int by, bool mirror_y,
                     bool fill_zero) {
        float ms = 0.0f;
        c_.check(artia_make_square(c_.get(), w, h, pitch, in, n, out, bx, by, mirror_y, fill_zero, &ms),
                 "CopyToSquareKernel");
        return ms;
    }

private:
    Context& c_;
};

// CTFKernel (Kernels.h:94-102) on an R2C half spectrum, pitch in complex elements
class CTFKernel {
public:
    explicit CTFKernel(Context& c) : c_(c) {}
    float operator()(void* spectrum, int n, size_t pitch, const artia_ctf_params& p, float defocus_min_nm,
                     float defocus_max_nm, float angle_deg, bool for_fp) {
        float ms = 0.0f;
        c_.check(artia_ctf_apply(c_.get(), spectrum, n, pitch, &p, defocus_min_nm, defocus_max_nm, angle_deg,
                                 for_fp, &ms),
                 "CTFKernel");
        return ms;
    }

private:
    Context& c_;
};

// WbpWeightingKernel (Kernels.h:159-166)
class WbpWeightingKernel {
public:
    explicit WbpWeightingKernel(Context& c) : c_(c) {}
    float operator()(void* spectrum, int n, size_t pitch, float psi_deg, int method) {
        float ms = 0.0f;
        c_.check(artia_wbp_weight(c_.get(), spectrum, n, pitch, psi_deg, method, &ms), "WbpWeightingKernel");
        return ms;
    }

private:
    Context& c_;
};

}  // namespace artia

#endif  // ARTIA_KERNELS_HPP
