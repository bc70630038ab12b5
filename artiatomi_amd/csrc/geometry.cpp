// Host geometry of the reconstruction core, exported through the C ABI so EmSART-style host
// code can build the per-tilt constants it used to upload with SetConstantValues
// (Kernels.cpp:949-1132).  Arithmetic follows the reference host code's float/double mix:
// rotations in double rounded to float (Projection.cpp:80-143), float vector algebra with
// host normalize = v * (1 / sqrt(dot)), the detector inverse by 4x4 cofactors in double
// (Projection.cpp:289-404).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/artia.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr float kDetectorDistance = -10000.0f;   // DIST, Constants.h:33-34

// Row-major 3x3 product with the reference's accumulation order (Matrix<T>::Mul,
// FilterGraph/Matrix.cpp:625-640).
template <class T>
void mul33(const T* a, const T* b, T* out) {
    T r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T acc = 0;
            for (int k = 0; k < 3; ++k) acc += a[3 * i + k] * b[3 * k + j];
            r[3 * i + j] = acc;
        }
    std::copy(r, r + 9, out);
}

template <class T>
void mul3v(const T* a, const T* v, T* out) {
    T r[3];
    for (int i = 0; i < 3; ++i) {
        T acc = 0;
        for (int k = 0; k < 3; ++k) acc += a[3 * i + k] * v[k];
        r[i] = acc;
    }
    std::copy(r, r + 3, out);
}

void unit(float* v) {
    const float inv = 1.0f / std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) v[i] *= inv;
}

// Rotation of a detector-frame vector into the volume frame for one tilt.
void rotate(const artia_tilt_params& tp, const artia_global_params& gp, const float* in, float* out) {
    const double theta = (double)(float)(tp.tilt_deg + gp.add_tilt_deg) / 180.0 * kPi;
    const double tau = (double)gp.add_tilt_x_deg / 180.0 * kPi;
    const double psi = (double)(gp.use_fix_psi ? -gp.fix_psi_deg : -tp.psi_deg) / 180.0 * kPi;
    const double phi = (double)gp.phi_deg / 180.0 * kPi;
    const double ct = std::cos(theta), st = std::sin(theta);
    const double cp = std::cos(phi), sp = std::sin(phi);
    const double cs = std::cos(psi), ss = std::sin(psi);
    const double ca = std::cos(tau), sa = std::sin(tau);
    const double flip[9] = {0, -1, 0, 1, 0, 0, 0, 0, -1};
    const double rz[9] = {cs, -ss, 0, ss, cs, 0, 0, 0, 1};
    const double rx[9] = {1, 0, 0, 0, cp, -sp, 0, sp, cp};
    const double rxT[9] = {1, 0, 0, 0, cp, sp, 0, -sp, cp};
    const double ry[9] = {ct, 0, st, 0, 1, 0, -st, 0, ct};
    const double rxTilt[9] = {1, 0, 0, 0, ca, -sa, 0, sa, ca};
    double a[9], b[9];
    mul33(ry, rxT, a);
    mul33(rx, a, b);   // tilt axis turned by phi: Rx(phi) Ry(theta) Rx(phi)^T
    double v[3] = {in[0], in[1], in[2]}, w[3];
    mul3v(flip, v, w);
    mul3v(rz, w, v);
    mul3v(b, v, w);
    mul3v(rxTilt, w, v);
    for (int i = 0; i < 3; ++i) out[i] = (float)v[i];
}

void aniso(float amount, float angleDeg, float dimX, float dimY, float* out) {
    const float ang = (float)((double)(angleDeg / 180.0f) * kPi);
    const float c1 = std::cos(ang), s1 = std::sin(ang), c2 = std::cos(-ang), s2 = std::sin(-ang);
    const float r1[9] = {c1, -s1, 0, s1, c1, 0, 0, 0, 1};
    const float r2[9] = {c2, -s2, 0, s2, c2, 0, 0, 0, 1};
    const float toCentre[9] = {1, 0, -dimX / 2.0f, 0, 1, -dimY / 2.0f, 0, 0, 1};
    const float back[9] = {1, 0, dimX / 2.0f, 0, 1, dimY / 2.0f, 0, 0, 1};
    const float stretch[9] = {amount, 0, 0, 0, 1, 0, 0, 0, 1};
    float a[9], b[9], c[9];
    mul33(back, r2, a);
    mul33(a, stretch, b);
    mul33(b, r1, c);
    mul33(c, toCentre, out);
}

void invert_detector(const float* u, const float* v, const float* n, const float* d, float* M) {
    const double e[16] = {u[0], u[1], u[2], 0.0, v[0], v[1], v[2], 0.0, n[0], n[1], n[2], 0.0, d[0], d[1], d[2], 1.0};
    const double a0 = e[0] * e[5] - e[1] * e[4], a1 = e[0] * e[6] - e[2] * e[4], a2 = e[0] * e[7] - e[3] * e[4];
    const double a3 = e[1] * e[6] - e[2] * e[5], a4 = e[1] * e[7] - e[3] * e[5], a5 = e[2] * e[7] - e[3] * e[6];
    const double b0 = e[8] * e[13] - e[9] * e[12], b1 = e[8] * e[14] - e[10] * e[12], b2 = e[8] * e[15] - e[11] * e[12];
    const double b3 = e[9] * e[14] - e[10] * e[13], b4 = e[9] * e[15] - e[11] * e[13], b5 = e[10] * e[15] - e[11] * e[14];
    const double det = a0 * b5 - a1 * b4 + a2 * b3 + a3 * b2 - a4 * b1 + a5 * b0;
    const double adj[16] = {
        +e[5] * b5 - e[6] * b4 + e[7] * b3,   -e[4] * b5 + e[6] * b2 - e[7] * b1,
        +e[4] * b4 - e[5] * b2 + e[7] * b0,   -e[4] * b3 + e[5] * b1 - e[6] * b0,
        -e[1] * b5 + e[2] * b4 - e[3] * b3,   +e[0] * b5 - e[2] * b2 + e[3] * b1,
        -e[0] * b4 + e[1] * b2 - e[3] * b0,   +e[0] * b3 - e[1] * b1 + e[2] * b0,
        +e[13] * a5 - e[14] * a4 + e[15] * a3, -e[12] * a5 + e[14] * a2 - e[15] * a1,
        +e[12] * a4 - e[13] * a2 + e[15] * a0, -e[12] * a3 + e[13] * a1 - e[14] * a0,
        -e[9] * a5 + e[10] * a4 - e[11] * a3,  +e[8] * a5 - e[10] * a2 + e[11] * a1,
        -e[8] * a4 + e[9] * a2 - e[11] * a0,   +e[8] * a3 - e[9] * a1 + e[10] * a0};
    const double inv = 1.0 / det;
    for (int i = 0; i < 16; ++i) M[i] = (float)(adj[i] * inv);
}

}  // namespace

extern "C" {

artia_status artia_geom_tilt(const artia_tilt_params* tp, const artia_global_params* gp, int width, int height,
                             float mag_aniso_amount, float mag_aniso_angle_deg, artia_tilt_geom* out) {
    if (!tp || !gp || !out || width <= 0 || height <= 0 || !(mag_aniso_amount != 0.0f))
        return ARTIA_ERR_INVALID_ARGUMENT;
    const bool mag = tp->mag > 0.8f;
    const float scale = mag ? tp->mag : 1.0f;
    const float ex[3] = {scale, 0, 0}, ey[3] = {0, scale, 0};
    float corner[3] = {-0.5f * (float)width, -0.5f * (float)height, kDetectorDistance};
    if (mag) {
        corner[0] = corner[0] * tp->mag;
        corner[1] = corner[1] * tp->mag;
    }
    float u[3], v[3], pos[3];
    rotate(*tp, *gp, ex, u);
    rotate(*tp, *gp, ey, v);
    rotate(*tp, *gp, corner, pos);
    for (int i = 0; i < 3; ++i) {
        out->u[i] = u[i];
        out->v[i] = v[i];
        out->d[i] = (pos[i] - u[i] * tp->shift_x) - v[i] * tp->shift_y;
    }
    float uh[3] = {u[0], u[1], u[2]}, vh[3] = {v[0], v[1], v[2]};
    unit(uh);
    unit(vh);
    float n[3] = {uh[2] * vh[1] - uh[1] * vh[2], uh[0] * vh[2] - uh[2] * vh[0], uh[1] * vh[0] - uh[0] * vh[1]};
    unit(n);
    for (int i = 0; i < 3; ++i) out->n[i] = n[i];
    invert_detector(u, v, n, out->d, out->det_matrix);
    aniso(mag_aniso_amount, mag_aniso_angle_deg, (float)width, (float)height, out->mag_aniso);
    aniso(1.0f / mag_aniso_amount, mag_aniso_angle_deg, (float)width, (float)height, out->mag_aniso_inv);
    return ARTIA_OK;
}

artia_status artia_geom_volume(int nx, int ny, int nz, const float voxel_size[3], const float volume_shift[3],
                               artia_vol_geom* out) {
    if (!voxel_size || !volume_shift || !out || nx <= 0 || ny <= 0 || nz <= 0) return ARTIA_ERR_INVALID_ARGUMENT;
    const int dims[3] = {nx, ny, nz};
    out->bbox_min[0] = (voxel_size[0] * (-0.5f * (float)nx) + 0.0f) - volume_shift[0];
    out->bbox_min[1] = (voxel_size[1] * (-0.5f * (float)ny) + 0.0f) - volume_shift[1];
    out->bbox_min[2] = voxel_size[2] * (-0.5f * (float)nz) - volume_shift[2];
    for (int i = 0; i < 3; ++i) {
        const float extent = (float)dims[i] * voxel_size[i];
        out->bbox_max[i] = out->bbox_min[i] + extent;
        out->bbox_rcp[i] = 1.0f / extent;
        out->dim[i] = (float)dims[i];
        out->voxel_size[i] = voxel_size[i];
    }
    out->z_shift = 0.0f;
    out->nx = nx;
    out->ny = ny;
    out->nz = nz;
    return ARTIA_OK;
}

artia_status artia_geom_subvolume(int n, float voxel_size_sub, const float voxel_size_full[3],
                                  const artia_vol_geom* full, const float position[3], const float shift[3],
                                  artia_vol_geom* out) {
    if (!voxel_size_full || !full || !position || !shift || !out || n <= 0) return ARTIA_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < 3; ++i) {
        const float base = voxel_size_sub * (-0.5f * (float)n) + voxel_size_full[i] * (position[i] - 1);
        out->bbox_min[i] = (base + shift[i] * voxel_size_full[i]) + full->bbox_min[i];
        const float extent = (float)n * voxel_size_sub;
        out->bbox_max[i] = out->bbox_min[i] + extent;
        out->bbox_rcp[i] = 1.0f / ((float)n * voxel_size_sub);
        out->dim[i] = (float)n;
        out->voxel_size[i] = voxel_size_sub;
    }
    out->z_shift = 0.0f;
    out->nx = out->ny = out->nz = n;
    return ARTIA_OK;
}

artia_status artia_geom_hitpoints(const artia_vol_geom* full, const artia_tilt_geom* g, int out[8]) {
    if (!full || !g || !out) return ARTIA_ERR_INVALID_ARGUMENT;
    const float* lo = full->bbox_min;
    const float* hi = full->bbox_max;
    const float nd = ((-g->n[0]) * g->d[0] - g->n[1] * g->d[1]) - g->n[2] * g->d[2];
    float px[8][2];
    for (int c = 0; c < 8; ++c) {
        const int bx = (c >> 2) & 1, by = (c >> 1) & 1, bz = c & 1;
        const float p[3] = {bx ? lo[0] + (hi[0] - lo[0]) : lo[0], by ? lo[1] + (hi[1] - lo[1]) : lo[1],
                            bz ? lo[2] + (hi[2] - lo[2]) : lo[2]};
        float t = (g->n[0] * p[0] + g->n[1] * p[1]) + g->n[2] * p[2];
        t += nd;
        t = std::fabs(t);
        float h[3];
        for (int i = 0; i < 3; ++i) h[i] = t * (-g->n[i]) + p[i];
        for (int r = 0; r < 2; ++r) {
            const float* m = g->det_matrix + 4 * r;
            px[c][r] = ((m[0] * h[0] + m[1] * h[1]) + m[2] * h[2]) + 1.f * m[3];
        }
    }
    for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 2; ++r) {
            const float a = px[2 * k][r], b = px[2 * k + 1][r];
            out[2 * k + r] = (int)(std::fabs(a) < std::fabs(b) ? a : b);
        }
    return ARTIA_OK;
}

artia_status artia_geom_real_length(const artia_vol_geom* full, const artia_tilt_geom* g, float* out) {
    if (!full || !g || !out) return ARTIA_ERR_INVALID_ARGUMENT;
    const float nz = g->n[2];
    *out = std::fabs((kDetectorDistance - full->bbox_min[2]) / nz - (kDetectorDistance - full->bbox_max[2]) / nz);
    return ARTIA_OK;
}

// The detector pixel nearest to the volume centre's foot point on the detector plane fixes the
// plane from which the eight bounding-box corners' distances along n are measured.
artia_status artia_geom_defocus_distances(const artia_vol_geom* full, const artia_tilt_geom* g, float* t_in,
                                          float* t_out) {
    if (!full || !g || !t_in || !t_out) return ARTIA_ERR_INVALID_ARGUMENT;
    const float* n = g->n;
    float centre[3];
    for (int i = 0; i < 3; ++i) centre[i] = full->bbox_min[i] + (full->dim[i] * full->voxel_size[i] * 0.5f);
    const float nd = ((-n[0]) * g->d[0] - n[1] * g->d[1]) - n[2] * g->d[2];
    float t = (n[0] * centre[0] + n[1] * centre[1]) + n[2] * centre[2];
    t += nd;
    t = std::fabs(t);
    float foot[3];
    for (int i = 0; i < 3; ++i) foot[i] = t * (-n[i]) + centre[i];
    int pix[2];
    for (int r = 0; r < 2; ++r) {
        const float* m = g->det_matrix + 4 * r;
        pix[r] = (int)std::round(((m[0] * foot[0] + m[1] * foot[1]) + m[2] * foot[2]) + 1.f * m[3]);
    }
    float origin[3];
    for (int i = 0; i < 3; ++i) origin[i] = (g->d[i] + (float)pix[0] * g->u[i]) + (float)pix[1] * g->v[i];
    const float no = ((-n[0]) * origin[0] - n[1] * origin[1]) - n[2] * origin[2];
    float lo = 2 * -kDetectorDistance, hi = 2 * kDetectorDistance;
    for (int c = 0; c < 8; ++c) {
        const int bx = (c >> 2) & 1, by = (c >> 1) & 1, bz = c & 1;
        const float* a = full->bbox_min;
        const float* b = full->bbox_max;
        float tc = (n[0] * (a[0] + bx * (b[0] - a[0])) + n[1] * (a[1] + by * (b[1] - a[1]))) +
                   n[2] * (a[2] + bz * (b[2] - a[2]));
        tc += no;
        lo = std::min(lo, tc);
        hi = std::max(hi, tc);
    }
    *t_in = lo;
    *t_out = hi;
    return ARTIA_OK;
}

// Defocus of the slab starting at `ray` (pixels along n): the slab's distance from the
// detector in nm, quantised to the slice thickness, optionally corrected for the volume's
// z shift projected on the tilt, and its sign flipped for IMOD's convention.
artia_status artia_geom_defocus_minmax(float ray, const artia_tilt_defocus* d, const artia_ctf_params* p,
                                       float* defocus_min_nm, float* defocus_max_nm) {
    if (!d || !p || !defocus_min_nm || !defocus_max_nm || !(p->slice_thickness_nm > 0.0f))
        return ARTIA_ERR_INVALID_ARGUMENT;
    const float tilt = d->tilt_deg / 180.0f * (float)kPi;
    const float T = p->slice_thickness_nm;
    float z = ray + kDetectorDistance;
    z = (std::round(z * p->pixelsize_nm * T) / T) + T / 2.0f;
    if (!p->ignore_z_shift) z = z - (p->volume_shift_z * p->pixelsize_nm * std::cos(tilt));
    if (p->switch_direction) z *= -1;
    *defocus_min_nm = d->defocus_min_nm + z;
    *defocus_max_nm = d->defocus_max_nm + z;
    return ARTIA_OK;
}

// Fisher-Yates with glibc's rand() after srand(0) (additive feedback generator, degree 31,
// separation 3, 310 warm-up draws), then the smallest |tilt| moved to the front.
artia_status artia_marker_tilt_params(const float* data, int dim_x, int dim_y, int dim_z, int ref_marker,
                                      int proj, artia_tilt_params* out, int* good) {
    if (!data || !out || !good || dim_x < 10 || dim_y <= 0 || dim_z <= 0 || ref_marker < 0 || ref_marker >= dim_z ||
        proj < 0 || proj >= dim_y)
        return ARTIA_ERR_INVALID_ARGUMENT;
    const float* m = data + (size_t)ref_marker * (size_t)dim_x * (size_t)dim_y + (size_t)proj * (size_t)dim_x;
    // MarkerFileItem_enum: 0 tilt, 1 x, 2 y, 3 deviation, 4 x shift, 5 y shift, 6/7 projected x/y,
    // 8 magnification, 9 psi
    out->tilt_deg = m[0];
    out->shift_x = m[4];
    out->shift_y = m[5];
    out->mag = m[8];
    out->psi_deg = m[9];
    *good = (m[1] > -1.0f && m[2] > -1.0f) ? 1 : 0;
    return ARTIA_OK;
}

artia_status artia_ctf_file_defocus(const float* data, int dim_x, int dim_y, int index, float tilt_deg,
                                    artia_tilt_defocus* out) {
    if (!data || !out || dim_x <= 0 || dim_y < 5 || index < 0 || index >= dim_x) return ARTIA_ERR_INVALID_ARGUMENT;
    out->defocus_min_nm = data[1 * (size_t)dim_x + (size_t)index];
    out->defocus_max_nm = data[2 * (size_t)dim_x + (size_t)index];
    // fdata[4 * DimX + index] * 180.0f / M_PI: the product in float, the quotient in double
    out->astig_angle_deg = (float)((double)(data[4 * (size_t)dim_x + (size_t)index] * 180.0f) / 3.14159265358979323846);
    out->tilt_deg = tilt_deg;
    return ARTIA_OK;
}

artia_status artia_projection_index_list(const float* tilt_deg, int count, int* out) {
    if (!tilt_deg || !out || count <= 0) return ARTIA_ERR_INVALID_ARGUMENT;
    std::vector<uint32_t> s(344 + (size_t)count);
    s[0] = 1;   // srand(0) seeds with 1
    for (int i = 1; i < 31; ++i) {
        const int64_t prev = (int32_t)s[i - 1];
        int64_t next = 16807 * (prev % 127773) - 2836 * (prev / 127773);
        if (next < 0) next += 2147483647;
        s[i] = (uint32_t)next;
    }
    for (int i = 31; i < 34; ++i) s[i] = s[i - 31];
    for (size_t i = 34; i < s.size(); ++i) s[i] = s[i - 31] + s[i - 3];
    for (int i = 0; i < count; ++i) out[i] = i;
    for (int i = 0; i + 1 < count; ++i) {
        const int draw = (int)(s[344 + (size_t)i] >> 1);
        std::swap(out[i], out[i + draw % (count - i)]);
    }
    float best = FLT_MAX;
    int zero = 0;
    for (int i = 0; i < count; ++i)
        if (std::fabs(tilt_deg[i]) < best) {
            best = std::fabs(tilt_deg[i]);
            zero = i;
        }
    for (int i = 0; i < count; ++i)
        if (out[i] == zero) {
            std::swap(out[0], out[i]);
            break;
        }
    return ARTIA_OK;
}

}  // extern "C"
