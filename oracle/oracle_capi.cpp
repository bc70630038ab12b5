// C entry points of the CPU oracle for ctypes (tests, smoke check, CPU-baseline timing).
// TEST INFRASTRUCTURE ONLY - never linked into the product library.
#include <omp.h>

#include "oracle.hpp"

using namespace artia_oracle;

extern "C" {

void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_get_threads(void) { return omp_get_max_threads(); }

void oracle_tilt_geometry(const TiltParams* tp, const GlobalParams* gp, int W, int H, float magAmount,
                          float magAngleDeg, TiltGeom* out) {
    tilt_geometry(*tp, *gp, W, H, magAmount, magAngleDeg, *out);
}
void oracle_rotate_vec(const TiltParams* tp, const GlobalParams* gp, const float* in, float* out) {
    rotate_vec(*tp, *gp, in, out);
}
void oracle_mag_aniso_matrix(float amount, float angleDeg, float dimX, float dimY, float* out) {
    mag_aniso_matrix(amount, angleDeg, dimX, dimY, out);
}
void oracle_volume_geometry(int nx, int ny, int nz, const float* vs, const float* shift, VolGeom* out) {
    volume_geometry(nx, ny, nz, vs, shift, *out);
}
void oracle_subvolume_geometry(int n, float vsSub, const float* vsFull, const VolGeom* full, const float* pos,
                               const float* shift, VolGeom* out) {
    subvolume_geometry(n, vsSub, vsFull, *full, pos, shift, *out);
}
void oracle_hit_points(const VolGeom* full, const TiltGeom* g, int* out) { hit_points(*full, *g, out); }
float oracle_real_length(const VolGeom* full, const TiltGeom* g) { return real_length(*full, *g); }
void oracle_index_list(const float* tilts, int count, int* out) { projection_index_list(tilts, count, out); }
void oracle_glibc_rand(unsigned seed, int count, int* out) { glibc_rand_sequence(seed, count, out); }

#define ORACLE_KERNELS(SUF, T)                                                                                  \
    void oracle_march_##SUF(const VolGeom* vg, const TiltGeom* tg, const T* vol, int W, int H, size_t pitch,    \
                            T* proj, T* dist, const int* roi) {                                                 \
        march<T>(*vg, *tg, vol, W, H, pitch, proj, dist, Roi{roi[0], roi[1], roi[2], roi[3]});                  \
    }                                                                                                           \
    void oracle_slicer_##SUF(const VolGeom* vg, const TiltGeom* tg, const T* vol, int W, int H, size_t pitch,   \
                             T* proj, double tmin, double tmax, const int* roi) {                               \
        slicer<T>(*vg, *tg, vol, W, H, pitch, proj, (T)(float)tmin, (T)(float)tmax,                             \
                  Roi{roi[0], roi[1], roi[2], roi[3]});                                                         \
    }                                                                                                           \
    void oracle_vol_traversal_##SUF(const VolGeom* vg, const TiltGeom* tg, int W, int H, size_t pitch, T* dist, \
                                    const int* roi) {                                                           \
        vol_traversal<T>(*vg, *tg, W, H, pitch, dist, Roi{roi[0], roi[1], roi[2], roi[3]});                     \
    }                                                                                                           \
    void oracle_compare_##SUF(int W, int H, size_t pitch, const T* real, T* virt, const T* dist, float realLen, \
                              const float* cut, const float* dim, float scale) {                                \
        compare<T>(W, H, pitch, real, virt, dist, realLen, cut, dim, scale);                                    \
    }                                                                                                           \
    void oracle_crop_border_##SUF(int W, int H, size_t pitch, T* img, const float* cut, const float* dim,       \
                                  const int* p) {                                                               \
        crop_border<T>(W, H, pitch, img, cut, dim, p);                                                          \
    }                                                                                                           \
    void oracle_back_project_##SUF(const VolGeom* vg, const TiltGeom* tg, T* vol, const T* img, int W, int H,   \
                                   size_t pitch, float lambda, int os, float distMin, float distMax,            \
                                   int filter) {                                                                \
        back_project<T>(*vg, *tg, vol, img, W, H, pitch, lambda, os, distMin, distMax, filter);                 \
    }                                                                                                           \
    void oracle_march_adjoint_##SUF(const VolGeom* vg, const TiltGeom* tg, const T* resid, int W, int H,      \
                                    size_t pitch, T* grad, const int* roi, double scale) {                  \
        march_adjoint<T>(*vg, *tg, resid, W, H, pitch, grad, Roi{roi[0], roi[1], roi[2], roi[3]}, (T)scale); \
    }                                                                                                       \
    void oracle_march_shifted_##SUF(const VolGeom* vg, const TiltGeom* tg, const T* vol, int W, int H,        \
                                    size_t pitch, T* proj, const int* roi, double sx, double sy) {          \
        march_shifted<T>(*vg, *tg, vol, W, H, pitch, proj, Roi{roi[0], roi[1], roi[2], roi[3]}, sx, sy);    \
    }                                                                                                       \
    void oracle_make_square_##SUF(int W, int H, size_t pitch, const T* in, int N, T* out, int bx, int by,       \
                                  int mirrorY, int fillZero) {                                                  \
        make_square<T>(W, H, pitch, in, N, out, bx, by, mirrorY != 0, fillZero != 0);                          \
    }                                                                                                           \
    void oracle_sart_##SUF(const VolGeom* vg, const VolGeom* full, const TiltGeom* tilts, int ntilts,           \
                           const int* indexList, const T* projections, int W, int H, T* vol,                    \
                           const SartParams* sp, int iterations, int* volumeIsEmpty) {                          \
        std::vector<TiltGeom> tv(tilts, tilts + ntilts);                                                        \
        std::vector<int> il(indexList, indexList + ntilts);                                                     \
        bool empty = *volumeIsEmpty != 0;                                                                       \
        sart<T>(*vg, *full, tv, il, projections, W, H, vol, *sp, iterations, &empty);                           \
        *volumeIsEmpty = empty ? 1 : 0;                                                                         \
    }

float oracle_ctf_wavelength(float kV) { return ctf_wavelength(kV); }
void oracle_ctf_consts(float cs_mm, float kV, float ps_nm, int N, float amp, CtfConsts* out) {
    *out = ctf_consts(cs_mm, kV, ps_nm, N, amp);
}
void oracle_defocus_distances(const VolGeom* full, const TiltGeom* g, int W, int H, float* tin, float* tout) {
    defocus_distances(*full, *g, W, H, *tin, *tout);
}
void oracle_defocus_min_max(float ray, float tilt_deg, float dmin, float dmax, const CtfSlabParams* p, float* omin,
                            float* omax) {
    defocus_min_max(ray, tilt_deg, dmin, dmax, *p, *omin, *omax);
}
void oracle_ctf_apply_f32(float* spec, int N, size_t pitch, const CtfConsts* k, float dmin, float dmax, float ang,
                          int fp, int pf, float w, const float* beta) {
    ctf_apply<float>(spec, N, pitch, *k, dmin, dmax, ang, fp != 0, pf != 0, w, beta);
}
void oracle_ctf_apply_f64(double* spec, int N, size_t pitch, const CtfConsts* k, float dmin, float dmax, float ang,
                          int fp, int pf, float w, const float* beta) {
    ctf_apply<double>(spec, N, pitch, *k, dmin, dmax, ang, fp != 0, pf != 0, w, beta);
}
void oracle_wbp_weight_f32(float* spec, int N, size_t pitch, float psi, int method) {
    wbp_weight<float>(spec, N, pitch, psi, method);
}
void oracle_wbp_weight_f64(double* spec, int N, size_t pitch, float psi, int method) {
    wbp_weight<double>(spec, N, pitch, psi, method);
}

void oracle_fourier_filter_f32(float* spec, int N, size_t pitch, float lp, float hp, float lps, float hps) {
    fourier_filter<float>(spec, N, pitch, lp, hp, lps, hps);
}
void oracle_fourier_filter_f64(double* spec, int N, size_t pitch, float lp, float hp, float lps, float hps) {
    fourier_filter<double>(spec, N, pitch, lp, hp, lps, hps);
}
void oracle_dose_weighting_f32(float* spec, int N, size_t pitch, float dose, float pixA) {
    dose_weighting<float>(spec, N, pitch, dose, pixA);
}
void oracle_dose_weighting_f64(double* spec, int N, size_t pitch, float dose, float pixA) {
    dose_weighting<double>(spec, N, pitch, dose, pixA);
}

ORACLE_KERNELS(f32, float)
ORACLE_KERNELS(f64, double)

}  // extern "C"
