// CPU oracle for the EmSART reconstruction core - implementation.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp for the contract and the pinning of each function).
// Build: g++ -O2 -mfma -ffp-contract=off -fopenmp -fPIC -shared (oracle/build.sh).
#include "oracle.hpp"

#include <cfloat>
#include <climits>
#include <cstdio>

namespace artia_oracle {

// ========================================================================================
// Host geometry
// ========================================================================================

namespace {
// Matrix<T>::Mul (FilterGraph/Matrix.cpp:625-640): row-major, val = 0; val += a*b.
template <class T>
void matmul3(const T a[9], const T b[9], T out[9]) {
    T r[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            T val = 0;
            for (int k = 0; k < 3; k++) val += a[i * 3 + k] * b[k * 3 + j];
            r[i * 3 + j] = val;
        }
    std::memcpy(out, r, sizeof(r));
}
template <class T>
void matvec3(const T a[9], const T v[3], T out[3]) {
    T r[3];
    for (int i = 0; i < 3; i++) {
        T val = 0;
        for (int k = 0; k < 3; k++) val += a[i * 3 + k] * v[k];
        r[i] = val;
    }
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
}
// cutil_math.h:1207-1211 host normalize: v * (1/sqrtf(dot(v,v)))
void normalize3f(float v[3]) {
    float dot = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    float inv = 1.0f / std::sqrt(dot);
    v[0] *= inv; v[1] *= inv; v[2] *= inv;
}
}  // namespace

// Projection::RotateMatrix (Projection.cpp:80-143)
void rotate_vec(const TiltParams& tp, const GlobalParams& gp, const float in[3], float out[3]) {
    const double pi = 3.14159265358979323846;
    double tilt = (double)(float)(tp.tilt_deg + gp.add_tilt_deg) / 180.0 * pi;
    double tiltX = (double)gp.add_tilt_x_deg / 180.0 * pi;
    double psi = (double)(-tp.psi_deg) / 180.0 * pi;
    if (gp.use_fix_psi) psi = (double)(-gp.fix_psi_deg) / 180.0 * pi;
    double phi = (double)gp.phi_deg / 180.0 * pi;

    const double adj[9] = {0, -1, 0, 1, 0, 0, 0, 0, -1};
    const double mPsi[9] = {std::cos(psi), -std::sin(psi), 0, std::sin(psi), std::cos(psi), 0, 0, 0, 1};
    const double mPhi[9] = {1, 0, 0, 0, std::cos(phi), -std::sin(phi), 0, std::sin(phi), std::cos(phi)};
    const double mPhiT[9] = {mPhi[0], mPhi[3], mPhi[6], mPhi[1], mPhi[4], mPhi[7], mPhi[2], mPhi[5], mPhi[8]};
    const double mTheta[9] = {std::cos(tilt), 0, std::sin(tilt), 0, 1, 0, -std::sin(tilt), 0, std::cos(tilt)};
    const double mX[9] = {1, 0, 0, 0, std::cos(tiltX), -std::sin(tiltX), 0, std::sin(tiltX), std::cos(tiltX)};

    double tmp[9], th[9];
    matmul3(mTheta, mPhiT, tmp);
    matmul3(mPhi, tmp, th);                       // mTheta = mPhi * (mTheta * mPhi^T)
    double v[3] = {(double)in[0], (double)in[1], (double)in[2]};
    double r1[3], r2[3], r3[3], r4[3];
    matvec3(adj, v, r1);                          // mAdjust * v
    matvec3(mPsi, r1, r2);                        // mPsi * (...)
    matvec3(th, r2, r3);                          // mTheta * (...)
    matvec3(mX, r3, r4);                          // mXTilt * (...)
    out[0] = (float)r4[0]; out[1] = (float)r4[1]; out[2] = (float)r4[2];
}

// Reconstructor::GetMagAnistropyMatrix (Reconstructor.cpp:386-426) with
// Matrix<float>::GetRotationMatrix3DZ (FilterGraph/Matrix.cpp:835-844); float arithmetic.
void mag_aniso_matrix(float amount, float angleDeg, float dimX, float dimY, float out[9]) {
    float angle = (float)((double)(angleDeg / 180.0f) * 3.14159265358979323846);
    auto rotz = [](float a, float r[9]) {
        float c = std::cos(a), s = std::sin(a);
        float m[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
        std::memcpy(r, m, sizeof(m));
    };
    float rot1[9], rot2[9];
    rotz(angle, rot1);
    rotz(-angle, rot2);
    float shiftC[9] = {1, 0, -dimX / 2.0f, 0, 1, -dimY / 2.0f, 0, 0, 1};
    float shiftB[9] = {1, 0, dimX / 2.0f, 0, 1, dimY / 2.0f, 0, 0, 1};
    float stretch[9] = {amount, 0, 0, 0, 1, 0, 0, 0, 1};
    float t1[9], t2[9], t3[9];
    matmul3(shiftB, rot2, t1);
    matmul3(t1, stretch, t2);
    matmul3(t2, rot1, t3);
    matmul3(t3, shiftC, out);
}

// Projection::GetPosition / GetPixelUPitch / GetPixelVPitch / GetNormalVector /
// GetDetectorMatrix (Projection.cpp:167-404), plus the mag-anisotropy pair.
void tilt_geometry(const TiltParams& tp, const GlobalParams& gp, int W, int H, float magAmount,
                   float magAngleDeg, TiltGeom& g) {
    const bool useMag = tp.mag > 0.8f;
    float uin[3] = {useMag ? 1 * tp.mag : 1.0f, 0, 0};
    float vin[3] = {0, useMag ? 1 * tp.mag : 1.0f, 0};
    float u[3], v[3];
    rotate_vec(tp, gp, uin, u);
    rotate_vec(tp, gp, vin, v);
    float pin[3];
    if (useMag) {
        pin[0] = -0.5f * (float)W * tp.mag;
        pin[1] = -0.5f * (float)H * tp.mag;
    } else {
        pin[0] = -0.5f * (float)W;
        pin[1] = -0.5f * (float)H;
    }
    pin[2] = kDist;
    float pos[3];
    rotate_vec(tp, gp, pin, pos);
    for (int i = 0; i < 3; i++) {
        g.u[i] = u[i];
        g.v[i] = v[i];
        g.d[i] = (pos[i] - u[i] * tp.shift_x) - v[i] * tp.shift_y;
    }
    float un[3] = {u[0], u[1], u[2]}, vn[3] = {v[0], v[1], v[2]};
    normalize3f(un);
    normalize3f(vn);
    float nn[3] = {un[2] * vn[1] - un[1] * vn[2], un[0] * vn[2] - un[2] * vn[0], un[1] * vn[0] - un[0] * vn[1]};
    normalize3f(nn);
    for (int i = 0; i < 3; i++) g.n[i] = nn[i];

    // 4x4 cofactor inverse in double (Projection.cpp:289-404), os = 1
    double e[16] = {(double)u[0] / 1.0, (double)u[1] / 1.0, (double)u[2] / 1.0, 0.0,
                    (double)v[0] / 1.0, (double)v[1] / 1.0, (double)v[2] / 1.0, 0.0,
                    (double)nn[0], (double)nn[1], (double)nn[2], 0.0,
                    (double)g.d[0], (double)g.d[1], (double)g.d[2], 1.0};
    double A0 = e[0] * e[5] - e[1] * e[4], A1 = e[0] * e[6] - e[2] * e[4], A2 = e[0] * e[7] - e[3] * e[4];
    double A3 = e[1] * e[6] - e[2] * e[5], A4 = e[1] * e[7] - e[3] * e[5], A5 = e[2] * e[7] - e[3] * e[6];
    double B0 = e[8] * e[13] - e[9] * e[12], B1 = e[8] * e[14] - e[10] * e[12], B2 = e[8] * e[15] - e[11] * e[12];
    double B3 = e[9] * e[14] - e[10] * e[13], B4 = e[9] * e[15] - e[11] * e[13], B5 = e[10] * e[15] - e[11] * e[14];
    double det = A0 * B5 - A1 * B4 + A2 * B3 + A3 * B2 - A4 * B1 + A5 * B0;
    double m[16];
    m[0] = +e[5] * B5 - e[6] * B4 + e[7] * B3;
    m[1] = -e[4] * B5 + e[6] * B2 - e[7] * B1;
    m[2] = +e[4] * B4 - e[5] * B2 + e[7] * B0;
    m[3] = -e[4] * B3 + e[5] * B1 - e[6] * B0;
    m[4] = -e[1] * B5 + e[2] * B4 - e[3] * B3;
    m[5] = +e[0] * B5 - e[2] * B2 + e[3] * B1;
    m[6] = -e[0] * B4 + e[1] * B2 - e[3] * B0;
    m[7] = +e[0] * B3 - e[1] * B1 + e[2] * B0;
    m[8] = +e[13] * A5 - e[14] * A4 + e[15] * A3;
    m[9] = -e[12] * A5 + e[14] * A2 - e[15] * A1;
    m[10] = +e[12] * A4 - e[13] * A2 + e[15] * A0;
    m[11] = -e[12] * A3 + e[13] * A1 - e[14] * A0;
    m[12] = -e[9] * A5 + e[10] * A4 - e[11] * A3;
    m[13] = +e[8] * A5 - e[10] * A2 + e[11] * A1;
    m[14] = -e[8] * A4 + e[9] * A2 - e[11] * A0;
    m[15] = +e[8] * A3 - e[9] * A1 + e[10] * A0;
    double inv = 1.0 / det;
    for (int i = 0; i < 16; i++) g.M[i] = (float)(m[i] * inv);

    mag_aniso_matrix(magAmount, magAngleDeg, (float)W, (float)H, g.A);
    mag_aniso_matrix(1.0f / magAmount, magAngleDeg, (float)W, (float)H, g.Ainv);
}

// Volume::PositionInSpace (Volume.cpp:270-286) for a single sub-volume (one GPU) plus
// GetSubVolumeBBox*/GetSubVolumeBBoxRcp (Volume.cpp:216-261).
void volume_geometry(int nx, int ny, int nz, const float vs[3], const float volShift[3], VolGeom& vg) {
    const float shiftXY[2] = {0.0f, 0.0f};   // Projection::GetMinimumTiltShift returns 0 (Projection.cpp:434-437)
    vg.bbMin[0] = (vs[0] * (-0.5f * (float)nx) + shiftXY[0]) - volShift[0];
    vg.bbMin[1] = (vs[1] * (-0.5f * (float)ny) + shiftXY[1]) - volShift[1];
    vg.bbMin[2] = vs[2] * (-0.5f * (float)nz) - volShift[2];
    const int dims[3] = {nx, ny, nz};
    for (int i = 0; i < 3; i++) {
        float box = (float)dims[i] * vs[i];
        vg.bbMax[i] = vg.bbMin[i] + box;
        vg.rcp[i] = 1.0f / ((float)dims[i] * vs[i]);
        vg.dim[i] = (float)dims[i];
        vg.vs[i] = vs[i];
    }
    vg.zshift = 0.0f;
    vg.nx = nx; vg.ny = ny; vg.nz = nz;
}

// Volume::PositionInSpace for a cubic sub-volume of a motive list (Volume.cpp:288-306).
void subvolume_geometry(int n, float vsSub, const float vsFull[3], const VolGeom& full,
                        const float pos[3], const float shift[3], VolGeom& vg) {
    for (int i = 0; i < 3; i++) {
        vg.bbMin[i] = ((vsSub * (-0.5f * (float)n) + vsFull[i] * (pos[i] - 1)) + shift[i] * vsFull[i]) + full.bbMin[i];
        float box = (float)n * vsSub;
        vg.bbMax[i] = vg.bbMin[i] + box;
        vg.rcp[i] = 1.0f / ((float)n * vsSub);
        vg.dim[i] = (float)n;
        vg.vs[i] = vsSub;
    }
    vg.zshift = 0.0f;
    vg.nx = vg.ny = vg.nz = n;
}

// Projection::ComputeHitPoints (Projection.cpp:850-1049): corners of the full-volume box
// projected onto the detector; per pair the component with the smaller magnitude, float->int.
void hit_points(const VolGeom& full, const TiltGeom& g, int out[8]) {
    const float* mn = full.bbMin;
    const float* mx = full.bbMax;
    const float* n = g.n;
    const float* d = g.d;
    float p[8][3];
    int idx = 0;
    for (int cx = 0; cx < 2; cx++)
        for (int cy = 0; cy < 2; cy++)
            for (int cz = 0; cz < 2; cz++) {
                float c[3] = {cx ? mn[0] + (mx[0] - mn[0]) : mn[0], cy ? mn[1] + (mx[1] - mn[1]) : mn[1],
                              cz ? mn[2] + (mx[2] - mn[2]) : mn[2]};
                float t = (n[0] * c[0] + n[1] * c[1]) + n[2] * c[2];
                t += ((-n[0] * d[0]) - n[1] * d[1]) - n[2] * d[2];
                t = std::fabs(t);
                for (int i = 0; i < 3; i++) p[idx][i] = t * (-n[i]) + c[i];
                idx++;
            }
    float q[8][3];
    for (int k = 0; k < 8; k++)
        for (int r = 0; r < 3; r++)
            q[k][r] = ((g.M[r * 4 + 0] * p[k][0] + g.M[r * 4 + 1] * p[k][1]) + g.M[r * 4 + 2] * p[k][2]) + 1.f * g.M[r * 4 + 3];
    for (int pair = 0; pair < 4; pair++) {
        const float* a = q[2 * pair];
        const float* b = q[2 * pair + 1];
        out[pair * 2 + 0] = (int)(std::fabs(a[0]) < std::fabs(b[0]) ? a[0] : b[0]);
        out[pair * 2 + 1] = (int)(std::fabs(a[1]) < std::fabs(b[1]) ? a[1] : b[1]);
    }
}

// Reconstructor::Compare host part (Reconstructor.cpp:1390-1393)
float real_length(const VolGeom& full, const TiltGeom& g) {
    float nz = g.n[2];
    return std::fabs((kDist - full.bbMin[2]) / nz - (kDist - full.bbMax[2]) / nz);
}

// glibc random_r TYPE_3 (the generator behind srand/rand): r[i] = r[i-3] + r[i-31] mod 2^32,
// seeded by the Park-Miller LCG, first 310 outputs discarded, output = r >> 1.
int glibc_rand_sequence(unsigned seed, int count, int* out) {
    if (seed == 0) seed = 1;
    std::vector<int32_t> r(344 + (size_t)count);
    r[0] = (int32_t)seed;
    for (int i = 1; i < 31; i++) {
        int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
        int64_t word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        r[i] = (int32_t)word;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (size_t i = 34; i < r.size(); i++) r[i] = (int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
    for (int k = 0; k < count; k++) out[k] = (int32_t)((uint32_t)r[344 + k] >> 1);
    return count;
}

// Projection::CreateProjectionIndexList, PLT_RANDOM_START_ZERO_TILT (Projection.cpp:561-596)
// with GetMinimumTiltIndex (Projection.cpp:407-423).  All projections are "good".
void projection_index_list(const float* tilts, int count, int* out) {
    for (int i = 0; i < count; i++) out[i] = i;
    std::vector<int> rnd((size_t)std::max(count, 1));
    glibc_rand_sequence(0, count, rnd.data());
    for (int i = 0; i < count - 1; i++) {
        int r = i + (rnd[i] % (count - i));
        std::swap(out[i], out[r]);
    }
    float minTilt = FLT_MAX;
    int minIdx = 0;
    for (int i = 0; i < count; i++)
        if (std::fabs(tilts[i]) < minTilt) {
            minTilt = std::fabs(tilts[i]);
            minIdx = i;
        }
    for (int i = 0; i < count; i++)
        if (out[i] == minIdx) {
            std::swap(out[0], out[i]);
            break;
        }
}

// ========================================================================================
// Kernel bodies
// ========================================================================================

template <class T>
struct GeomT {
    T d[3], u[3], v[3], n[3], M[16], A[9], Ainv[9];
    T bbMin[3], bbMax[3], rcp[3], dim[3], vs[3], zshift;
    GeomT(const VolGeom& vg, const TiltGeom& tg) {
        for (int i = 0; i < 3; i++) {
            d[i] = tg.d[i]; u[i] = tg.u[i]; v[i] = tg.v[i]; n[i] = tg.n[i];
            bbMin[i] = vg.bbMin[i]; bbMax[i] = vg.bbMax[i]; rcp[i] = vg.rcp[i];
            dim[i] = vg.dim[i]; vs[i] = vg.vs[i];
        }
        for (int i = 0; i < 16; i++) M[i] = tg.M[i];
        for (int i = 0; i < 9; i++) { A[i] = tg.A[i]; Ainv[i] = tg.Ainv[i]; }
        zshift = vg.zshift;
    }
};

// Slab test of a ray c + t*n against [bbMin, bbMax] (ForwardProjectionRayMarcher_TL.cu:92-107)
template <class T>
inline void slab_test(const GeomT<T>& G, const T c[3], T& t_in, T& t_out) {
    T tmin[3], tmax[3];
    for (int i = 0; i < 3; i++) {
        T tE = (G.bbMin[i] - c[i]) / G.n[i];
        T tX = (G.bbMax[i] - c[i]) / G.n[i];
        tmin[i] = fminT(tE, tX);
        tmax[i] = fmaxT(tE, tX);
    }
    t_in = fmaxT(fmaxT(tmin[0], tmin[1]), tmin[2]);
    t_out = fminT(fminT(tmax[0], tmax[1]), tmax[2]);
}

template <class T>
inline void sub_ray_origin(const GeomT<T>& G, T px, T py, T c[3]) {
    T xa, ya;
    mag_apply(G.A, px, py, xa, ya);
    for (int i = 0; i < 3; i++) c[i] = (G.d[i] + xa * G.u[i]) + ya * G.v[i];
}

// Guard against non-terminating sample loops for degenerate geometry (never reached for a
// positive voxel size; the GPU kernels carry the same cap).
constexpr int kMaxSamplesPerRay = 1 << 22;

template <class T>
void march(const VolGeom& vg, const TiltGeom& tg, const T* vol, int W, int H, size_t pitch,
           T* proj, T* dist, Roi roi) {
    const GeomT<T> G(vg, tg);
    const T step = G.vs[0] * T(0.15);
    const int x0 = std::max(roi.x0, 0), y0 = std::max(roi.y0, 0);
    const int x1 = std::min(roi.x1, W), y1 = std::min(roi.y1, H);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = y0; y < y1; y++) {
        for (int x = x0; x < x1; x++) {
            T temp = 0, gcount = 0, glen = 0;
            for (T osx = T(0.25); osx < T(0.9); osx += T(0.5)) {
                for (T osy = T(0.25); osy < T(0.9); osy += T(0.5)) {
                    T c[3];
                    sub_ray_origin(G, (T)x + osx, (T)y + osy, c);
                    T t_in, t_out;
                    slab_test(G, c, t_in, t_out);
                    if ((t_out - t_in) > T(0)) {
                        gcount += T(1);
                        glen += (t_out - t_in);
                        T f[3] = {c[0] + t_in * G.n[0], c[1] + t_in * G.n[1], c[2] + t_in * G.n[2]};
                        int guard = 0;
                        while (t_in <= t_out && guard++ < kMaxSamplesPerRay) {
                            T qx = ((f[0] - G.bbMin[0]) * G.rcp[0]) * G.dim[0];
                            T qy = ((f[1] - G.bbMin[1]) * G.rcp[1]) * G.dim[1];
                            T qz = ((f[2] - G.bbMin[2]) * G.rcp[2]) * G.dim[2] - G.zshift;
                            T val = tex3d_linear(vol, vg.nx, vg.ny, vg.nz, qx, qy, qz);
                            temp += (val * G.vs[0]) * T(0.15);
                            t_in += step;
                            for (int i = 0; i < 3; i++) f[i] = c[i] + t_in * G.n[i];
                        }
                    }
                }
            }
            const size_t i = (size_t)y * pitch + (size_t)x;
            proj[i] += temp * T(0.25);
            dist[i] = fmaxT(T(0), glen / gcount);
        }
    }
}

template <class T>
void march_adjoint(const VolGeom& vg, const TiltGeom& tg, const T* resid, int W, int H, size_t pitch, T* grad,
                   Roi roi, T scale) {
    const GeomT<T> G(vg, tg);
    const T step = G.vs[0] * T(0.15);
    const int nx = vg.nx, ny = vg.ny, nz = vg.nz;
    const size_t sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
    const int x0 = std::max(roi.x0, 0), y0 = std::max(roi.y0, 0);
    const int x1 = std::min(roi.x1, W), y1 = std::min(roi.y1, H);
    // serial: scatter targets overlap between pixels
    for (int y = y0; y < y1; y++) {
        for (int x = x0; x < x1; x++) {
            const T r = resid[(size_t)y * pitch + (size_t)x] * scale * (G.vs[0] * T(0.15)) * T(0.25);
            if (r == T(0)) continue;
            for (T osx = T(0.25); osx < T(0.9); osx += T(0.5)) {
                for (T osy = T(0.25); osy < T(0.9); osy += T(0.5)) {
                    T c[3];
                    sub_ray_origin(G, (T)x + osx, (T)y + osy, c);
                    T t_in, t_out;
                    slab_test(G, c, t_in, t_out);
                    if (!((t_out - t_in) > T(0))) continue;
                    T f[3] = {c[0] + t_in * G.n[0], c[1] + t_in * G.n[1], c[2] + t_in * G.n[2]};
                    int guard = 0;
                    while (t_in <= t_out && guard++ < kMaxSamplesPerRay) {
                        T q[3] = {((f[0] - G.bbMin[0]) * G.rcp[0]) * G.dim[0],
                                  ((f[1] - G.bbMin[1]) * G.rcp[1]) * G.dim[1],
                                  ((f[2] - G.bbMin[2]) * G.rcp[2]) * G.dim[2] - G.zshift};
                        // tex3d_linear's weights and clamped indices
                        const int n3[3] = {nx, ny, nz};
                        int lo[3], hi[3];
                        T wlo[3], whi[3];
                        for (int a = 0; a < 3; a++) {
                            const T qc = fminT(fmaxT(q[a], T(-1)), T(n3[a] + 1)) - T(0.5);
                            const T fl = std::floor(qc);
                            const T w = qc - fl;
                            lo[a] = clampi((int)fl, 0, n3[a] - 1);
                            hi[a] = clampi((int)fl + 1, 0, n3[a] - 1);
                            wlo[a] = T(1) - w;
                            whi[a] = w;
                        }
                        for (int k = 0; k < 2; k++)
                            for (int j = 0; j < 2; j++)
                                for (int i = 0; i < 2; i++) {
                                    const T w = (i ? whi[0] : wlo[0]) * (j ? whi[1] : wlo[1]) * (k ? whi[2] : wlo[2]);
                                    grad[(size_t)(k ? hi[2] : lo[2]) * sz + (size_t)(j ? hi[1] : lo[1]) * sy +
                                         (size_t)(i ? hi[0] : lo[0])] += r * w;
                                }
                        t_in += step;
                        for (int i = 0; i < 3; i++) f[i] = c[i] + t_in * G.n[i];
                    }
                }
            }
        }
    }
}

template <class T>
void march_shifted(const VolGeom& vg, const TiltGeom& tg, const T* vol, int W, int H, size_t pitch, T* proj,
                   Roi roi, double sx, double sy) {
    const GeomT<T> G(vg, tg);
    const T step = G.vs[0] * T(0.15);
    const int x0 = std::max(roi.x0, 0), y0 = std::max(roi.y0, 0);
    const int x1 = std::min(roi.x1, W), y1 = std::min(roi.y1, H);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = y0; y < y1; y++) {
        for (int x = x0; x < x1; x++) {
            T temp = 0;
            for (T osx = T(0.25); osx < T(0.9); osx += T(0.5)) {
                for (T osy = T(0.25); osy < T(0.9); osy += T(0.5)) {
                    T c[3];
                    sub_ray_origin(G, (T)x + osx, (T)y + osy, c);
                    for (int i = 0; i < 3; i++) c[i] = c[i] - ((T)sx * G.u[i] + (T)sy * G.v[i]);
                    T t_in, t_out;
                    slab_test(G, c, t_in, t_out);
                    if (!((t_out - t_in) > T(0))) continue;
                    T f[3] = {c[0] + t_in * G.n[0], c[1] + t_in * G.n[1], c[2] + t_in * G.n[2]};
                    int guard = 0;
                    while (t_in <= t_out && guard++ < kMaxSamplesPerRay) {
                        T qx = ((f[0] - G.bbMin[0]) * G.rcp[0]) * G.dim[0];
                        T qy = ((f[1] - G.bbMin[1]) * G.rcp[1]) * G.dim[1];
                        T qz = ((f[2] - G.bbMin[2]) * G.rcp[2]) * G.dim[2] - G.zshift;
                        temp += (tex3d_linear(vol, vg.nx, vg.ny, vg.nz, qx, qy, qz) * G.vs[0]) * T(0.15);
                        t_in += step;
                        for (int i = 0; i < 3; i++) f[i] = c[i] + t_in * G.n[i];
                    }
                }
            }
            proj[(size_t)y * pitch + (size_t)x] += temp * T(0.25);
        }
    }
}

template <class T>
void slicer(const VolGeom& vg, const TiltGeom& tg, const T* vol, int W, int H, size_t pitch,
            T* proj, T tminDef, T tmaxDef, Roi roi) {
    const GeomT<T> G(vg, tg);
    const T step = G.vs[0] * T(0.15);
    const int x0 = std::max(roi.x0, 0), y0 = std::max(roi.y0, 0);
    const int x1 = std::min(roi.x1, W), y1 = std::min(roi.y1, H);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = y0; y < y1; y++) {
        for (int x = x0; x < x1; x++) {
            T temp = 0;
            for (T osx = T(0.25); osx < T(0.8); osx += T(0.5)) {
                for (T osy = T(0.25); osy < T(0.8); osy += T(0.5)) {
                    T c[3];
                    sub_ray_origin(G, (T)x + osx, (T)y + osy, c);
                    T t_in, t_out;
                    slab_test(G, c, t_in, t_out);
                    if ((t_out - t_in) > T(0)) {
                        t_in = fmaxT(t_in, tminDef);
                        t_out = fminT(t_out, tmaxDef);
                        // first sample is taken at the detector point c itself
                        // (ForwardProjectionSlicer.cu:129-133, f.w is never applied)
                        T f[3] = {c[0], c[1], c[2]};
                        int guard = 0;
                        while (t_in <= t_out && guard++ < kMaxSamplesPerRay) {
                            T qx = ((f[0] - G.bbMin[0]) * G.rcp[0]) * G.dim[0];
                            T qy = ((f[1] - G.bbMin[1]) * G.rcp[1]) * G.dim[1];
                            T qz = ((f[2] - G.bbMin[2]) * G.rcp[2]) * G.dim[2] - G.zshift;
                            T val = tex3d_linear(vol, vg.nx, vg.ny, vg.nz, qx, qy, qz);
                            temp += (val * G.vs[0]) * T(0.15);
                            t_in += step;
                            for (int i = 0; i < 3; i++) f[i] = c[i] + t_in * G.n[i];
                        }
                    }
                }
            }
            proj[(size_t)y * pitch + (size_t)x] += temp * T(0.25);
        }
    }
}

template <class T>
void vol_traversal(const VolGeom& vg, const TiltGeom& tg, int W, int H, size_t pitch, T* dist, Roi roi) {
    const GeomT<T> G(vg, tg);
    const int x0 = std::max(roi.x0, 0), y0 = std::max(roi.y0, 0);
    const int x1 = std::min(roi.x1, W), y1 = std::min(roi.y1, H);
#pragma omp parallel for schedule(static)
    for (int y = y0; y < y1; y++) {
        for (int x = x0; x < x1; x++) {
            T c[3];
            sub_ray_origin(G, (T)x + T(0.5), (T)y + T(0.5), c);
            T t_in, t_out;
            slab_test(G, c, t_in, t_out);
            T val = 0;
            if (t_out > t_in) val = t_out - t_in;
            dist[(size_t)y * pitch + (size_t)x] += val;
        }
    }
}

template <class T>
void compare(int W, int H, size_t pitch, const T* real, T* virt, const T* dist, float realLength,
             const float cut[4], const float dim[4], float scale) {
    const T L = realLength, S = scale;
    const T cx = cut[0], cy = cut[1], cz = cut[2], cw = cut[3];
    const T dx = dim[0], dy = dim[1], dz = dim[2], dw = dim[3];
#pragma omp parallel for schedule(static)
    for (int yy = 0; yy < H; yy++) {
        for (int xx = 0; xx < W; xx++) {
            const unsigned x = (unsigned)xx, y = (unsigned)yy;
            const size_t i = (size_t)y * pitch + x;
            T val = dist[i];
            T error = 0;
            if (val >= T(1)) error = ((real[i] - ((virt[i] / S) / val * L)) / L) * S;
            T dYB = 1, dYA = 1, dXB = 1, dXA = 1;
            if ((T)y < cz + dz) {
                T w = ((T)y - cz) / dz;
                if (w < 0) w = 0;
                dYB = T(1) - std::exp(-((w * w) * T(9)));
            }
            if ((T)y > (((T)H - dw) - cw) - T(1)) {
                T w = (((T)(unsigned)(H - y - 1) - cw) - (T)(H - 1)) / dw;
                if (w < 0) w = T(0);
                dYA = T(1) - std::exp(-((w * w) * T(9)));
            }
            if ((T)x < cy + dy) {
                T w = ((T)x - cy) / dy;
                if (w < 0) w = 0;
                dXB = T(1) - std::exp(-((w * w) * T(9)));
            }
            if ((T)x > (((T)W - dx) - cx) - T(1)) {
                T w = (((T)(unsigned)(W - x - 1) - cx) - (T)(W - 1)) / dx;
                if (w < 0) w = T(0);
                dXA = T(1) - std::exp(-((w * w) * T(9)));
            }
            virt[i] = (((dXA * dXB) * dYA) * dYB) * error;
        }
    }
}

namespace {
template <class T>
inline T pdist(T ax, T ay, T bx, T by) {
    return std::sqrt((ax - bx) * (ax - bx) + (ay - by) * (ay - by));
}
// Compare.cu:138-167 GetDistance (point-to-segment distance, float)
template <class T>
inline T seg_distance(int vx, int vy, int wx, int wy, int x, int y) {
    T l2 = (T)((vx - wx) * (vx - wx) + (vy - wy) * (vy - wy));
    if (l2 == 0) return pdist<T>((T)vx, (T)vy, (T)x, (T)y);
    T x1 = (T)(x - vx), y1 = (T)(y - vy);
    T x2 = (T)(wx - vx), y2 = (T)(wy - vy);
    T dot = x1 * x2 + y1 * y2;
    T t = dot / l2;
    if (t < T(0)) return pdist<T>((T)x, (T)y, (T)vx, (T)vy);
    if (t > T(1)) return pdist<T>((T)x, (T)y, (T)wx, (T)wy);
    T x3 = (T)vx + t * (T)(wx - vx);
    T y3 = (T)vy + t * (T)(wy - vy);
    return pdist<T>((T)x, (T)y, x3, y3);
}
}  // namespace

template <class T>
void crop_border(int W, int H, size_t pitch, T* img, const float cut[2], const float dim[2], const int p[8]) {
    const int p1x = p[0], p1y = p[1], p2x = p[2], p2y = p[3], p3x = p[4], p3y = p[5], p4x = p[6], p4y = p[7];
    const T cx = cut[0], cy = cut[1], dx = dim[0], dy = dim[1];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            T distX = 0, distY = 0;
            if ((p2x - p1x) * (y - p1y) - (p2y - p1y) * (x - p1x) < 0 &&
                (p3x - p4x) * (y - p4y) - (p3y - p4y) * (x - p4x) < 0 &&
                (p1x - p3x) * (y - p3y) - (p1y - p3y) * (x - p3x) < 0 &&
                (p4x - p2x) * (y - p2y) - (p4y - p2y) * (x - p2x) < 0) {
                distX = 1;
                distY = 1;
                T mdx = (T)(3 * W), mdy = (T)(3 * H);
                mdx = fminT(mdx, seg_distance<T>(p1x, p1y, p2x, p2y, x, y));
                mdx = fminT(mdx, seg_distance<T>(p4x, p4y, p3x, p3y, x, y));
                mdx = fminT(mdx, (T)x);
                mdx = fminT(mdx, (T)(W - x - 1));
                mdy = fminT(mdy, seg_distance<T>(p3x, p3y, p1x, p1y, x, y));
                mdy = fminT(mdy, seg_distance<T>(p2x, p2y, p4x, p4y, x, y));
                mdy = fminT(mdy, (T)y);
                mdy = fminT(mdy, (T)(H - y - 1));
                if (mdx < cx + dx) {
                    T w = (mdx - cx) / dx;
                    if (w < 0) w = 0;
                    distX = T(1) - std::exp(-((w * w) * T(9)));
                }
                if (mdy < cy + dy) {
                    T w = (mdy - cy) / dy;
                    if (w < 0) w = 0;
                    distY = T(1) - std::exp(-((w * w) * T(9)));
                }
            }
            const size_t i = (size_t)y * pitch + x;
            img[i] = (distX * distY) * img[i];
        }
    }
}

// Ray origin of a BP sub-ray: r0 + 100000*n evaluated in double and rounded once to T
// (BackProjectionSquareOS.cu:549-551: the literal 100000.0 is a double).
template <class T>
inline T bp_source(T r, T n) { return (T)((double)r + 100000.0 * (double)n); }

template <class T>
void back_project(const VolGeom& vg, const TiltGeom& tg, T* vol, const T* img, int W, int H,
                  size_t pitch, float lambdaF, int os, float distMinF, float distMaxF, int filter) {
    const GeomT<T> G(vg, tg);
    const T lambda = lambdaF, distMin = distMinF, distMax = distMaxF;
    const T osInv = (T)1 / (T)os;     // 1.0f / (float)OverSampling (Reconstructor.cpp:885)
    const int nq = vg.nx / 4;
    const T* n = G.n;
    const T* d = G.d;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int z = 0; z < vg.nz; z++) {
        for (int y = 0; y < vg.ny; y++) {
            for (int xq = 0; xq < nq; xq++) {
                const unsigned x = (unsigned)xq * 4u;
                T* vp = vol + ((size_t)z * vg.ny + y) * (size_t)vg.nx + x;
                T mcMin[3] = {G.bbMin[0] + (T)x * G.vs[0], G.bbMin[1] + (T)(unsigned)y * G.vs[1],
                              G.bbMin[2] + (T)(unsigned)z * G.vs[2]};
                T mcMax[3] = {G.bbMin[0] + (T)(x + 5) * G.vs[0], G.bbMin[1] + (T)(unsigned)(y + 1) * G.vs[1],
                              G.bbMin[2] + (T)(unsigned)(z + 1) * G.vs[2]};
                T bMin[3] = {(T)FLT_MAX, (T)FLT_MAX, (T)FLT_MAX};
                T bMax[3] = {T(-200000), T(-200000), T(-200000)};
                const T nd = ((-n[0]) * d[0] - n[1] * d[1]) - n[2] * d[2];
                bool gated = false;
                // corner order x-major, then y, then z (BackProjectionSquareOS.cu:359-453)
                for (int cx = 0; cx < 2 && !gated; cx++)
                    for (int cy = 0; cy < 2 && !gated; cy++)
                        for (int cz = 0; cz < 2; cz++) {
                            T c[3] = {cx ? mcMin[0] + (mcMax[0] - mcMin[0]) : mcMin[0],
                                      cy ? mcMin[1] + (mcMax[1] - mcMin[1]) : mcMin[1],
                                      cz ? mcMin[2] + (mcMax[2] - mcMin[2]) : mcMin[2]};
                            T t = (n[0] * c[0] + n[1] * c[1]) + n[2] * c[2];
                            t += nd;
                            t = std::fabs(t);
                            if (cx == 0 && cy == 0 && cz == 0 && !(t >= distMin && t < distMax)) {
                                gated = true;
                                break;
                            }
                            for (int i = 0; i < 3; i++) {
                                T h = t * (-n[i]) + c[i];
                                bMin[i] = fminT(h, bMin[i]);
                                bMax[i] = fmaxT(h, bMax[i]);
                            }
                        }
                if (gated) continue;
                T lo[3], hi[3];
                for (int i = 0; i < 3; i++) {
                    lo[i] = fminT(bMin[i], bMax[i]);
                    hi[i] = fmaxT(bMin[i], bMax[i]);
                }
                if (n[0] <= T(0)) std::swap(lo[0], hi[0]);
                T plo[2], phi[2];
                for (int r = 0; r < 2; r++) {
                    plo[r] = ((G.M[r * 4 + 0] * lo[0] + G.M[r * 4 + 1] * lo[1]) + G.M[r * 4 + 2] * lo[2]) + T(1) * G.M[r * 4 + 3];
                    phi[r] = ((G.M[r * 4 + 0] * hi[0] + G.M[r * 4 + 1] * hi[1]) + G.M[r * 4 + 2] * hi[2]) + T(1) * G.M[r * 4 + 3];
                }
                int px0 = (int)std::floor(fminT(plo[0], phi[0]));
                int py0 = (int)std::floor(fminT(plo[1], phi[1]));
                int px1 = (int)std::ceil(fmaxT(plo[0], phi[0]));
                int py1 = (int)std::ceil(fmaxT(plo[1], phi[1]));
                px0 = (int)fminT(fmaxT((T)px0, T(0)), (T)(W - 1));
                px1 = (int)fminT(fmaxT((T)px1, T(0)), (T)(W - 1));
                py0 = (int)fminT(fmaxT((T)py0, T(0)), (T)(H - 1));
                py1 = (int)fminT(fmaxT((T)py1, T(0)), (T)(H - 1));

                T num[4] = {0, 0, 0, 0}, den[4] = {0, 0, 0, 0};
                const T yLo = G.bbMin[1] + (T)(unsigned)y * G.vs[1], yHi = G.bbMin[1] + (T)(unsigned)(y + 1) * G.vs[1];
                const T zLo = G.bbMin[2] + (T)(unsigned)z * G.vs[2], zHi = G.bbMin[2] + (T)(unsigned)(z + 1) * G.vs[2];
                T xPl[5];
                for (int k = 0; k < 5; k++) xPl[k] = G.bbMin[0] + (T)(x + (unsigned)k) * G.vs[0];
                for (T pyv = (T)py0 + osInv * T(0.5); pyv < (T)py1; pyv += osInv) {
                    for (T pxv = (T)px0 + osInv * T(0.5); pxv < (T)px1; pxv += osInv) {
                        T xa, ya;
                        mag_apply(G.Ainv, pxv, pyv, xa, ya);
                        T ray[3], src[3];
                        for (int i = 0; i < 3; i++) {
                            ray[i] = d[i];
                            ray[i] = ray[i] + pxv * G.u[i];
                        }
                        for (int i = 0; i < 3; i++) ray[i] = ray[i] + pyv * G.v[i];
                        for (int i = 0; i < 3; i++) {
                            src[i] = bp_source(ray[i], n[i]);
                            ray[i] = ray[i] - src[i];
                        }
                        T dot = (ray[0] * ray[0] + ray[1] * ray[1]) + ray[2] * ray[2];
                        T inv = T(1) / std::sqrt(dot);
                        for (int i = 0; i < 3; i++) ray[i] = ray[i] * inv;
                        const T val = tex2d(img, W, H, pitch, filter, xa, ya);
                        const T tyE = (yLo - src[1]) / ray[1], tyX = (yHi - src[1]) / ray[1];
                        const T tzE = (zLo - src[2]) / ray[2], tzX = (zHi - src[2]) / ray[2];
                        for (int k = 0; k < 4; k++) {
                            T txE = (xPl[k] - src[0]) / ray[0];
                            T txX = (xPl[k + 1] - src[0]) / ray[0];
                            T t_in = fmaxT(fmaxT(fminT(txE, txX), fminT(tyE, tyX)), fminT(tzE, tzX));
                            T t_out = fminT(fminT(fmaxT(txE, txX), fmaxT(tyE, tyX)), fmaxT(tzE, tzX));
                            if ((t_out - t_in) > T(0)) {
                                num[k] += val * (t_out - t_in);
                                den[k] += (t_out - t_in);
                            }
                        }
                    }
                }
                for (int k = 0; k < 4; k++) {
                    den[k] = fmaxT(T(0), den[k]);
                    if (den[k] != T(0)) vp[k] += (lambda * num[k]) / den[k];
                }
            }
        }
    }
}

template <class T>
void make_square(int W, int H, size_t pitch, const T* in, int N, T* out, int bx, int by, bool mirrorY,
                 bool fillZero) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < N; y++) {
        for (int x = 0; x < N; x++) {
            if (fillZero) {
                T val = 0;
                int xi = x - bx, yi = y - by;
                if (xi >= 0 && xi < W && yi >= 0 && yi < H) {
                    if (mirrorY) yi = H - yi - 1;
                    val = in[(size_t)yi * pitch + xi];
                }
                out[(size_t)y * N + x] = val;
            } else {
                int xi = x - bx;
                if (xi < 0) xi = -xi - 1;
                if (xi >= W) xi = W - (xi - W) - 1;
                int yi = y - by;
                if (yi < 0) yi = -yi - 1;
                if (yi >= H) yi = H - (yi - H) - 1;
                if (mirrorY) yi = H - yi - 1;
                out[(size_t)y * N + x] = in[(size_t)yi * pitch + xi];
            }
        }
    }
}

// ---------------------------------------------------------------- CTF and WBP weighting

float ctf_wavelength(float voltage_kV) {
    const float h = (float)6.63E-34, c = (float)3.00E+08;
    const int E0 = 511;
    const float V = voltage_kV;
    const float a = (2 * E0 * V * 1000.0f * 1000.0f) + (V * V * 1000.0f * 1000.0f);
    const double arg = (double)a * 1.602E-19 * 1.602E-19;   // double literals in the macro
    return (h * c) / std::sqrt((float)arg);
}

CtfConsts ctf_consts(float cs_mm, float voltage_kV, float pixelsize_nm, int N, float amp_contrast) {
    CtfConsts k;
    k.cs_mm = cs_mm;
    k.voltage_kV = voltage_kV;
    k.pixelsize_m = pixelsize_nm * std::pow(10.0f, -9.0f);
    k.pixelcount = (float)N;
    const float maxFreq = (float)(1.0 / (double)(k.pixelsize_m * 2.0f));
    k.freq_step = maxFreq / (k.pixelcount / 2.0f);
    k.amp_contrast = amp_contrast;
    k.phase_contrast = std::sqrt(1 - amp_contrast * amp_contrast);
    return k;
}

template <class T>
void ctf_apply(T* spec, int N, size_t pitch, const CtfConsts& k, float defocusMin_nm, float defocusMax_nm,
               float angle_deg, bool forFP, bool phaseFlipOnly, float wiener, const float betaFac[4]) {
    const float pi_f = 3.14159265358979323846f;
    const T def0 = (T)(defocusMin_nm * 0.000000001f), def1 = (T)(defocusMax_nm * 0.000000001f);
    const T angle = (T)(angle_deg / 180.0f * (float)3.14159265358979323846);
    const T lam = (T)ctf_wavelength(k.voltage_kV);
    const T Cs = (T)(k.cs_mm * 0.001f);
    const T W = wiener;
    const T pc = k.phase_contrast, ac = k.amp_contrast;
    const int nx = N / 2 + 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < N; y++) {
        for (int x = 0; x < nx; x++) {
            T xpos = (T)x, ypos = (T)y;
            if (ypos > (T)k.pixelcount * T(0.5)) ypos = ((T)k.pixelcount - ypos) * T(-1);
            T alpha = (xpos == 0) ? (T)(pi_f * 0.5f) : std::atan2(ypos, xpos);
            T beta = alpha - angle;
            T defocus = def0 + (T(1) - std::cos(T(2) * beta)) * (def1 - def0);
            T len = std::sqrt(xpos * xpos + ypos * ypos);
            len *= (T)k.freq_step;
            T m = (T)(pi_f / 2.0f) * (Cs * lam * lam * lam * len * len * len * len - T(2) * defocus * lam * len * len);
            T n = pc * std::sin(m) + ac * std::cos(m);
            T* r = spec + 2 * ((size_t)y * pitch + (size_t)x);
            const T radius = std::sqrt(xpos * xpos + ypos * ypos);
            if (!phaseFlipOnly && radius > (T)betaFac[0]) {
                T l = len / T(100000000.0f);
                T e = std::exp((-(T)betaFac[1] * l - (T)betaFac[2] * l * l) - (T)betaFac[3] * l * l * l);
                if (forFP) {
                    e = std::max(e, (T)0.01f);
                    T val = n * e;
                    if (std::fabs(val) < (T)0.0001f && val >= 0) val = (T)0.0001f;
                    if (std::fabs(val) < (T)0.0001f && val < 0) val = (T)-0.0001f;
                    r[0] = r[0] * -val;
                    r[1] = r[1] * -val;
                } else {
                    e = std::max(e, W);
                    T val = n * e;
                    r[0] = r[0] * -val / (val * val + W);
                    r[1] = r[1] * -val / (val * val + W);
                }
            }
            if (phaseFlipOnly && n >= 0) {
                r[0] = -r[0];
                r[1] = -r[1];
            }
        }
    }
}

template <class T>
void wbp_weight(T* spec, int N, size_t pitch, float psi_deg, int method) {
    const T ang = (T)(-psi_deg / 180.0f * (float)3.14159265358979323846);
    const T sn = std::sin(ang), cs = std::cos(ang);
    const int nx = N / 2 + 1;
    const unsigned pc = (unsigned)N;
    struct Poly { float p1, p2, p3, p4; };
    const Poly polys[3] = {{1.000654227857550f, 0.006008581017124f, 0.004159659493151f, 0.000975903396538f},
                           {1.003784816598589f, 0.029016377161629f, 0.019582940715148f, 0.004559409669984f},
                           {0.999884616010943f, 0.000307646262566f, 0.004742915272196f, 0.004806551368900f}};
#pragma omp parallel for schedule(static)
    for (int y = 0; y < N; y++) {
        for (int x = 0; x < nx; x++) {
            T xpos = (T)x, ypos = (T)y;
            if (ypos > (T)pc * T(0.5)) ypos = ((T)pc - ypos) * T(-1);
            const T temp = xpos;
            xpos = cs * xpos - sn * ypos;
            ypos = sn * temp + cs * ypos;
            const T len = ypos / (T)pc / T(2);
            T w = 1;
            if (method == 0) {
                w = std::min(std::fabs(len), T(1));
            } else if (method >= 2 && method <= 4) {
                const Poly& q = polys[method - 2];
                if (len == 0) {
                    w = 0;
                } else {
                    const T L = std::log(std::fabs(len));
                    w = (((T)q.p1 + (T)q.p2 * L) + (T)q.p3 * L * L) + (T)q.p4 * L * L * L;
                }
                w = std::max(T(0), std::min(w, T(1)));
            }
            T* r = spec + 2 * ((size_t)y * pitch + (size_t)x);
            r[0] *= w;
            r[1] *= w;
        }
    }
}

template <class T>
void fourier_filter(T* spec, int N, size_t pitch, float lp, float hp, float lps, float hps) {
    // the kernel's arithmetic in T (float: the reference's float factor)
    const T LP = (T)lp - (T)lps, HP = (T)hp + (T)hps, LPS = (T)lps, HPS = (T)hps;
    const int nx = N / 2 + 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < N; y++) {
        for (int x = 0; x < nx; x++) {
            T mx = (T)x, my = (T)y;
            if (my > (T)N * T(0.5)) my = ((T)N - my) * T(-1);
            const T dist = std::sqrt(mx * mx + my * my);
            T fil = 0;
            if (LP > 0) {
                if (dist <= LP) fil = 1;
            } else {
                if (dist <= (T)(N / 2 - 1)) fil = 1;
            }
            if (LPS > 0) {
                const T fil2 = (-fil + T(1)) * std::exp(-((dist - LP) * (dist - LP) / (T(2) * LPS * LPS)));
                if (fil2 > (T)0.001f) fil = fil2;
            }
            if (LPS > 0 && LP == 0 && HP == 0 && HPS == 0)
                fil = std::exp(-((dist - LP) * (dist - LP) / (T(2) * LPS * LPS)));
            if (HP > 0) {
                T fil2 = 0;
                if (dist >= HP) fil2 = 1;
                fil *= fil2;
                if (HPS > 0) {
                    const T fil3 = (-fil2 + T(1)) * std::exp(-((dist - HP) * (dist - HP) / (T(2) * HPS * HPS)));
                    if (fil3 > (T)0.001f) fil = fil3;
                }
            }
            T* r = spec + 2 * ((size_t)y * pitch + (size_t)x);
            r[0] *= fil;
            r[1] *= fil;
        }
    }
}

template <class T>
void dose_weighting(T* spec, int N, size_t pitch, float dose, float pixelsize_A) {
    const int nx = N / 2 + 1;
    const T scale = (T)(N / 2) / (T)pixelsize_A;   // pixelcount / 2 is an integer division
#pragma omp parallel for schedule(static)
    for (int y = 0; y < N; y++) {
        for (int x = 0; x < nx; x++) {
            T mx = (T)x, my = (T)y;
            if (my > (T)N * T(0.5)) my = ((T)N - my) * T(-1);
            const T dist = std::sqrt(mx * mx + my * my) / scale;
            const T fil = std::exp(-(T)dose * dist);
            T* r = spec + 2 * ((size_t)y * pitch + (size_t)x);
            r[0] *= fil;
            r[1] *= fil;
        }
    }
}

void defocus_distances(const VolGeom& full, const TiltGeom& g, int W, int H, float& t_in, float& t_out) {
    const float* n = g.n;
    const float* d = g.d;
    const float* mn = full.bbMin;
    const float* mx = full.bbMax;
    const float cx = mn[0] + (full.dim[0] * full.vs[0] * 0.5f), cy = mn[1] + (full.dim[1] * full.vs[1] * 0.5f),
                cz = mn[2] + (full.dim[2] * full.vs[2] * 0.5f);
    float t = (n[0] * cx + n[1] * cy) + n[2] * cz;
    t += ((-n[0] * d[0]) - n[1] * d[1]) - n[2] * d[2];
    t = std::fabs(t);
    const float h[3] = {t * (-n[0]) + cx, t * (-n[1]) + cy, t * (-n[2]) + cz};
    float hp[2];
    for (int r = 0; r < 2; r++)
        hp[r] = ((g.M[4 * r] * h[0] + g.M[4 * r + 1] * h[1]) + g.M[4 * r + 2] * h[2]) + 1.f * g.M[4 * r + 3];
    const int hitX = (int)std::round(hp[0]), hitY = (int)std::round(hp[1]);
    float pos[3];
    for (int i = 0; i < 3; i++) pos[i] = (d[i] + (float)hitX * g.u[i]) + (float)hitY * g.v[i];
    (void)W;
    (void)H;
    t_in = 2 * -kDist;
    t_out = 2 * kDist;
    for (int x = 0; x <= 1; x++)
        for (int y = 0; y <= 1; y++)
            for (int z = 0; z <= 1; z++) {
                float tt = (n[0] * (mn[0] + x * (mx[0] - mn[0])) + n[1] * (mn[1] + y * (mx[1] - mn[1]))) +
                           n[2] * (mn[2] + z * (mx[2] - mn[2]));
                tt += ((-n[0] * pos[0]) - n[1] * pos[1]) - n[2] * pos[2];
                if (tt < t_in) t_in = tt;
                if (tt > t_out) t_out = tt;
            }
}

void defocus_min_max(float ray, float tilt_deg, float defMin_nm, float defMax_nm, const CtfSlabParams& p,
                     float& outMin, float& outMax) {
    const float tilt = tilt_deg / 180.0f * (float)3.14159265358979323846;
    float dist0 = ray + kDist;
    const float T = p.slice_thickness_nm;
    if (p.ignore_z_shift)
        dist0 = (std::round(dist0 * p.pixelsize_nm * T) / T) + T / 2.0f;
    else
        dist0 = (std::round(dist0 * p.pixelsize_nm * T) / T) + T / 2.0f - (p.volume_shift_z * p.pixelsize_nm * std::cos(tilt));
    if (p.switch_direction) dist0 *= -1;
    outMin = defMin_nm + dist0;
    outMax = defMax_nm + dist0;
}

template <class T>
void sart(const VolGeom& vg, const VolGeom& full, const std::vector<TiltGeom>& tilts,
          const std::vector<int>& indexList, const T* projections, int W, int H, T* vol,
          const SartParams& sp, int iterationsToRun, bool* volumeIsEmpty) {
    const size_t npix = (size_t)W * H;
    if (sp.sirt_count <= 0 || sp.sirt_count > (int)indexList.size() || sp.os <= 0 || sp.os > 64) {
        std::fprintf(stderr, "oracle sart: invalid SIRTCount %d or OverSampling %d\n", sp.sirt_count, sp.os);
        return;
    }
    const int S = sp.sirt_count;
    const int P = (int)indexList.size();
    std::vector<T> proj(npix), dist(npix);
    std::vector<std::vector<T>> residual((size_t)S, std::vector<T>(npix));
    const Roi all{0, 0, W, H};
    for (int iter = 0; iter < iterationsToRun; iter++) {
        for (int step = 0; step < (P + S - 1) / S; step++) {
            for (int i = 0; i < S; i++) {
                if (step * S + i >= P) continue;
                const int idx = indexList[(size_t)step * S + i];
                const TiltGeom& tg = tilts[(size_t)idx];
                std::fill(proj.begin(), proj.end(), T(0));   // ResetProjectionsDevice
                std::fill(dist.begin(), dist.end(), T(0));
                if (!*volumeIsEmpty)
                    march<T>(vg, tg, vol, W, H, (size_t)W, proj.data(), dist.data(), all);
                else
                    vol_traversal<T>(vg, tg, W, H, (size_t)W, dist.data(), all);
                compare<T>(W, H, (size_t)W, projections + (size_t)idx * npix, proj.data(), dist.data(),
                           real_length(full, tg), sp.crop, sp.crop_dim, sp.scale);
                residual[(size_t)i] = proj;
            }
            for (int i = 0; i < S; i++) {
                if (step * S + i >= P) continue;
                *volumeIsEmpty = false;
                const int idx = indexList[(size_t)step * S + i];
                const TiltGeom& tg = tilts[(size_t)idx];
                int hp[8];
                hit_points(full, tg, hp);
                crop_border<T>(W, H, (size_t)W, residual[(size_t)i].data(), sp.cut, sp.dim_len, hp);
                back_project<T>(vg, tg, vol, residual[(size_t)i].data(), W, H, (size_t)W, sp.lambda / (float)S,
                                sp.os, 0.0f, 9999999999999.0f, FILTER_POINT);
            }
        }
    }
}

#define ARTIA_ORACLE_INSTANTIATE(T)                                                                    \
    template void march<T>(const VolGeom&, const TiltGeom&, const T*, int, int, size_t, T*, T*, Roi);  \
    template void slicer<T>(const VolGeom&, const TiltGeom&, const T*, int, int, size_t, T*, T, T, Roi); \
    template void vol_traversal<T>(const VolGeom&, const TiltGeom&, int, int, size_t, T*, Roi);        \
    template void compare<T>(int, int, size_t, const T*, T*, const T*, float, const float*, const float*, float); \
    template void crop_border<T>(int, int, size_t, T*, const float*, const float*, const int*);         \
    template void back_project<T>(const VolGeom&, const TiltGeom&, T*, const T*, int, int, size_t, float, int, float, float, int); \
    template void make_square<T>(int, int, size_t, const T*, int, T*, int, int, bool, bool);          \
    template void march_adjoint<T>(const VolGeom&, const TiltGeom&, const T*, int, int, size_t, T*, Roi, T); \
    template void march_shifted<T>(const VolGeom&, const TiltGeom&, const T*, int, int, size_t, T*, Roi, double, \
                                   double);                                                            \
    template void sart<T>(const VolGeom&, const VolGeom&, const std::vector<TiltGeom>&, const std::vector<int>&, \
                          const T*, int, int, T*, const SartParams&, int, bool*);

#define ARTIA_ORACLE_INSTANTIATE_SPEC(T)                                                                  \
    template void ctf_apply<T>(T*, int, size_t, const CtfConsts&, float, float, float, bool, bool, float,   \
                               const float*);                                                              \
    template void wbp_weight<T>(T*, int, size_t, float, int);                                            \
    template void fourier_filter<T>(T*, int, size_t, float, float, float, float);                         \
    template void dose_weighting<T>(T*, int, size_t, float, float);

ARTIA_ORACLE_INSTANTIATE_SPEC(float)
ARTIA_ORACLE_INSTANTIATE_SPEC(double)
ARTIA_ORACLE_INSTANTIATE(float)
ARTIA_ORACLE_INSTANTIATE(double)

}  // namespace artia_oracle
