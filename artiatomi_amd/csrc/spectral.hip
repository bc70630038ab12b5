// Fourier-space operators on gfx950: the CTF multiply / Wiener inverse / phase flip (ctf.cu:80-240),
// the WBP weighting (wbpWeighting.cu:48-144), and the fused per-slab image filters that replace
// the reference's cuFFT/NPP glue (Reconstructor.cpp:557-583, 952-965, 1271-1319):
//   slab filter : cropBorder (in place) -> makeSquare (zero fill) -> R2C -> ctf -> C2R ->
//                 / N^2 with crop to W x H -> cropBorder -> store or accumulate
//   WBP filter  : makeSquare (reflect) -> R2C -> wbpWeighting -> C2R -> / N^2 with crop
// The FFTs are hipFFT single-precision plans cached per size in the context; the spectrum is the
// (N/2+1) x N half spectrum, rows of N/2+1 complex values.
#include <algorithm>
#include <cmath>

#include "device_common.hpp"
#include "image_ops.hpp"

namespace artia {
namespace {

constexpr float kPiF = 3.14159265358979323846f;

struct CtfK {
    float cs_m;         // Cs * 0.001 (mm -> m)
    float lambda;       // electron wavelength (m), the reference's float constants
    float freq_step;    // (1 / (2 * pixelsize)) / (N / 2)
    float pc, ac;       // phase and amplitude contrast
    float halfN;        // c_pixelcount * 0.5
    float def0, def1, angle;   // metres, metres, radians
    float wiener;
    float beta[4];
    int forFP, phaseFlip, N;
};

// Reference numerics: the kernel body's float operation order.
__global__ void __launch_bounds__(256) ctf_ref_kernel(float2* __restrict__ spec, size_t pitch, CtfK k) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= k.N / 2 + 1 || y >= k.N) return;
    float xpos = (float)x, ypos = (float)y;
    if (ypos > (float)k.N * 0.5f) ypos = ((float)k.N - ypos) * -1.0f;
    const float alpha = (xpos == 0) ? (kPiF * 0.5f) : atan2f(ypos, xpos);
    const float beta = alpha - k.angle;
    const float defocus = k.def0 + (1 - cosf(2 * beta)) * (k.def1 - k.def0);
    float len = sqrtf(xpos * xpos + ypos * ypos);
    len *= k.freq_step;
    const float lam = k.lambda, Cs = k.cs_m;
    const float m = (kPiF / 2.0f) * (Cs * lam * lam * lam * len * len * len * len - 2 * defocus * lam * len * len);
    const float n = k.pc * sinf(m) + k.ac * cosf(m);
    float2 r = spec[(size_t)y * pitch + x];
    const float radius = sqrtf(xpos * xpos + ypos * ypos);
    if (!k.phaseFlip && radius > k.beta[0]) {
        const float l = len / 100000000.0f;
        float e = expf((-k.beta[1] * l - k.beta[2] * l * l) - k.beta[3] * l * l * l);
        if (k.forFP) {
            e = fmaxf(e, 0.01f);
            float val = n * e;
            if (fabsf(val) < 0.0001f && val >= 0) val = 0.0001f;
            if (fabsf(val) < 0.0001f && val < 0) val = -0.0001f;
            r.x = r.x * -val;
            r.y = r.y * -val;
        } else {
            e = fmaxf(e, k.wiener);
            const float val = n * e;
            r.x = r.x * -val / (val * val + k.wiener);
            r.y = r.y * -val / (val * val + k.wiener);
        }
    }
    if (k.phaseFlip && n >= 0) {
        r.x = -r.x;
        r.y = -r.y;
    }
    spec[(size_t)y * pitch + x] = r;
}

// Exact numerics: the CTF factor in double (the phase reaches hundreds of radians at Nyquist, where
// a float phase carries ~1e-5 absolute error), applied to the float spectrum.
// The exact CTF factor of one spectrum point in two parts: what does not depend on the defocus
// (ctf_point), and the factor and its application for one defocus pair (ctf_apply_at); together the
// double arithmetic of the kernel body, in its order.
struct CtfPoint {
    double c2b;      // cos(2 beta)
    double len;      // spatial frequency
    double lam4;     // Cs lam^3 len^4, the defocus-free term of the phase
    double env;      // the envelope, clamped as the direction requires (used when damped)
    bool damped;     // outside the untouched low-frequency disk, and not phase flipping
};

__device__ __forceinline__ CtfPoint ctf_point(int x, int y, const CtfK& k) {
    CtfPoint p;
    double xpos = (double)x, ypos = (double)y;
    if (ypos > (double)k.N * 0.5) ypos = ((double)k.N - ypos) * -1.0;
    const double alpha = (xpos == 0) ? (double)(kPiF * 0.5f) : atan2(ypos, xpos);
    const double beta = alpha - (double)k.angle;
    p.c2b = cos(2 * beta);
    double len = sqrt(xpos * xpos + ypos * ypos);
    len *= (double)k.freq_step;
    p.len = len;
    const double lam = k.lambda, Cs = k.cs_m;
    p.lam4 = Cs * lam * lam * lam * len * len * len * len;
    const double radius = sqrt(xpos * xpos + ypos * ypos);
    p.damped = !k.phaseFlip && radius > (double)k.beta[0];
    p.env = 0.0;
    if (p.damped) {
        const double l = len / 100000000.0;
        const double e = exp((-(double)k.beta[1] * l - (double)k.beta[2] * l * l) - (double)k.beta[3] * l * l * l);
        p.env = k.forFP ? fmax(e, (double)0.01f) : fmax(e, (double)k.wiener);
    }
    return p;
}

__device__ __forceinline__ float2 ctf_apply_at(float2 r, const CtfPoint& p, const CtfK& k, float def0f, float def1f) {
    const double def0 = def0f, def1 = def1f;
    const double defocus = def0 + (1 - p.c2b) * (def1 - def0);
    const double lam = k.lambda, len = p.len;
    const double m = (double)(kPiF / 2.0f) * (p.lam4 - 2 * defocus * lam * len * len);
    // sin and cos from one argument reduction (the same values as the two calls)
    double sm, cm;
    sincos(m, &sm, &cm);
    const double n = (double)k.pc * sm + (double)k.ac * cm;
    double f = 1.0;
    if (p.damped) {
        if (k.forFP) {
            double val = n * p.env;
            if (fabs(val) < (double)0.0001f && val >= 0) val = (double)0.0001f;
            if (fabs(val) < (double)0.0001f && val < 0) val = -(double)0.0001f;
            f = -val;
        } else {
            const double W = k.wiener;
            const double val = n * p.env;
            f = -val / (val * val + W);
        }
    }
    if (k.phaseFlip && n >= 0) f = -f;
    r.x = (float)((double)r.x * f);
    r.y = (float)((double)r.y * f);
    return r;
}

__device__ __forceinline__ void ctf_exact_at(float2* __restrict__ spec, size_t pitch, int x, int y, const CtfK& k) {
    const CtfPoint p = ctf_point(x, y, k);
    float2* q = spec + (size_t)y * pitch + x;
    *q = ctf_apply_at(*q, p, k, k.def0, k.def1);
}

__global__ void __launch_bounds__(256) ctf_exact_kernel(float2* __restrict__ spec, size_t pitch, CtfK k) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= k.N / 2 + 1 || y >= k.N) return;
    ctf_exact_at(spec, pitch, x, y, k);
}

// The same over a stack of half spectra (stride N * (N/2+1)), slab blockIdx.z with its own defocus.
struct CtfStack {
    CtfK k;
    float def0[kFftBatchMax], def1[kFftBatchMax];
};

__global__ void __launch_bounds__(256) ctf_exact_stack_kernel(float2* __restrict__ spec, size_t pitch, CtfStack b,
                                                              int ns) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= b.k.N / 2 + 1 || y >= b.k.N) return;
    const CtfPoint p = ctf_point(x, y, b.k);   // once for the stack's ns spectra
    float2* q = spec + (size_t)y * pitch + x;
    const size_t plane = pitch * (size_t)b.k.N;
    for (int s = 0; s < ns; s++) q[(size_t)s * plane] = ctf_apply_at(q[(size_t)s * plane], p, b.k, b.def0[s], b.def1[s]);
}

// ns slabs of ONE image: spectrum s = base * CTF(slab s's defocus) * scale.  scale is 1, or 1 / N^2
// (a power of two, so the product is exact) when the inverse transform writes the slab images
// straight into their destination: C2R(scale X) = scale C2R(X) bit for bit, which is the
// division by N^2 of the finishing pass.
__global__ void __launch_bounds__(256) ctf_exact_fan_kernel(const float2* __restrict__ base,
                                                            float2* __restrict__ spec, size_t pitch, CtfStack b,
                                                            int ns, float scale) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= b.k.N / 2 + 1 || y >= b.k.N) return;
    const CtfPoint p = ctf_point(x, y, b.k);
    const size_t i = (size_t)y * pitch + x, plane = pitch * (size_t)b.k.N;
    const float2 r0 = base[i];
    for (int s = 0; s < ns; s++) {
        float2 r = ctf_apply_at(r0, p, b.k, b.def0[s], b.def1[s]);
        r.x = r.x * scale;
        r.y = r.y * scale;
        spec[(size_t)s * plane + i] = r;
    }
}

// wbpWeighting.cu:48-144 (sin/cos of the angle as IEEE functions rather than __sinf/__cosf)
__global__ void __launch_bounds__(256) wbp_weight_kernel(float2* __restrict__ spec, size_t pitch, int N, float ang,
                                                         int method, int exact) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= N / 2 + 1 || y >= N) return;
    const unsigned pc = (unsigned)N;
    float w = 1.0f;
    if (exact) {
        double xpos = (double)x, ypos = (double)y;
        if (ypos > (double)pc * 0.5) ypos = ((double)pc - ypos) * -1.0;
        const double sn = sin((double)ang), cs = cos((double)ang);
        const double yr = sn * xpos + cs * ypos;
        const double len = yr / (double)pc / 2.0;
        double wd = 1.0;
        if (method == 0) {
            wd = fmin(fabs(len), 1.0);
        } else if (method >= 2 && method <= 4) {
            const double P[3][4] = {{1.000654227857550f, 0.006008581017124f, 0.004159659493151f, 0.000975903396538f},
                                    {1.003784816598589f, 0.029016377161629f, 0.019582940715148f, 0.004559409669984f},
                                    {0.999884616010943f, 0.000307646262566f, 0.004742915272196f, 0.004806551368900f}};
            const double* q = P[method - 2];
            if (len == 0) {
                wd = 0;
            } else {
                const double L = log(fabs(len));
                wd = ((q[0] + q[1] * L) + q[2] * L * L) + q[3] * L * L * L;
            }
            wd = fmax(0.0, fmin(wd, 1.0));
        }
        w = (float)wd;
        float2 r = spec[(size_t)y * pitch + x];
        r.x = (float)((double)r.x * wd);
        r.y = (float)((double)r.y * wd);
        spec[(size_t)y * pitch + x] = r;
        return;
    }
    float xpos = (float)x, ypos = (float)y;
    if (ypos > (float)pc * 0.5f) ypos = ((float)pc - ypos) * -1.0f;
    const float temp = xpos;
    const float sn = sinf(ang), cs = cosf(ang);
    xpos = cs * xpos - sn * ypos;
    ypos = sn * temp + cs * ypos;
    const float len = ypos / (float)pc / 2.0f;
    if (method == 0) {
        w = fminf(fabsf(len), 1.0f);
    } else if (method >= 2 && method <= 4) {
        const float P[3][4] = {{1.000654227857550f, 0.006008581017124f, 0.004159659493151f, 0.000975903396538f},
                               {1.003784816598589f, 0.029016377161629f, 0.019582940715148f, 0.004559409669984f},
                               {0.999884616010943f, 0.000307646262566f, 0.004742915272196f, 0.004806551368900f}};
        const float* q = P[method - 2];
        if (len == 0) {
            w = 0;
        } else {
            const float L = logf(fabsf(len));
            w = ((q[0] + q[1] * L) + q[2] * L * L) + q[3] * L * L * L;
        }
        w = fmaxf(0.0f, fminf(w, 1.0f));
    }
    float2 r = spec[(size_t)y * pitch + x];
    r.x *= w;
    r.y *= w;
    spec[(size_t)y * pitch + x] = r;
}

// out (+)= crop(x, y) * (square[by + y][bx + x] / N^2)  (nppiDivC + cropBorder + nppiAdd); over a
// stack of ns squares, square s into out + s * out_stride
__global__ void __launch_bounds__(256) finish_kernel(const float* __restrict__ sq, int ns, int N, int bx, int by, int W,
                                                     int H, float* __restrict__ out, size_t pitch, float norm,
                                                     int accumulate, int crop, CropParams cp, size_t out_stride) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= W || y >= H) return;
    // the crop weight depends on the pixel only: evaluated once for the ns squares
    const float w = crop ? crop_weight(x, y, W, H, cp) : 1.0f;
    const size_t j = (size_t)(by + y) * N + (bx + x), plane = (size_t)N * (size_t)N;
    const size_t i = (size_t)y * pitch + x;
    for (int s = 0; s < ns; s++) {
        float v = sq[(size_t)s * plane + j] / norm;
        if (crop) v = w * v;
        float* o = out + (size_t)s * out_stride;
        o[i] = accumulate ? o[i] + v : v;
    }
}

// out += crop(x, y) * (square_s[by + y][bx + x] / N^2) for s = 0 .. ns-1 in turn: the sum, rounded
// after every slab, of ns accumulating finish_kernel calls
__global__ void __launch_bounds__(256) finish_sum_kernel(const float* __restrict__ sq, int ns, int N, int bx, int by,
                                                         int W, int H, float* __restrict__ out, size_t pitch,
                                                         float norm, int crop, CropParams cp) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= W || y >= H) return;
    const float w = crop ? crop_weight(x, y, W, H, cp) : 1.0f;
    const size_t i = (size_t)y * pitch + x;
    const size_t j = (size_t)(by + y) * N + (bx + x), plane = (size_t)N * (size_t)N;
    float acc = out[i];
    for (int s = 0; s < ns; s++) {
        float v = sq[(size_t)s * plane + j] / norm;
        if (crop) v = w * v;
        acc = acc + v;
    }
    out[i] = acc;
}

// Zero-filled centred squares of a stack of ns slab images: square_s = makeSquare(cropBorder^r(src_s))
// with r = rep0 + rep_step * s crops applied in turn (rep_step 1: the repeated in-place crop of one
// image, src_stride 0, carried from slab to slab with one multiplication each)
__global__ void __launch_bounds__(256) slab_square_kernel(const float* __restrict__ src, size_t src_stride, int W,
                                                          int H, size_t pitch, int N, int bx, int by,
                                                          float* __restrict__ sq, CropParams cp, int rep0,
                                                          int rep_step, int ns) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= N || y >= N) return;
    const size_t plane = (size_t)N * (size_t)N;
    float* o = sq + (size_t)y * N + x;
    const int xi = x - bx, yi = y - by;
    if (!(xi >= 0 && xi < W && yi >= 0 && yi < H)) {
        for (int s = 0; s < ns; s++) o[(size_t)s * plane] = 0.0f;
        return;
    }
    const float* in = src + (size_t)yi * pitch + xi;
    const float w = (rep0 > 0 || rep_step > 0) ? crop_weight(xi, yi, W, H, cp) : 1.0f;
    if (rep_step == 0) {
        for (int s = 0; s < ns; s++) {
            float val = in[(size_t)s * src_stride];
            for (int r = 0; r < rep0; r++) val = w * val;
            o[(size_t)s * plane] = val;
        }
    } else {
        float val = in[0];
        for (int r = 0; r < rep0; r++) val = w * val;
        for (int s = 0; s < ns; s++) {
            o[(size_t)s * plane] = val;
            val = w * val;
        }
    }
}

dim3 grid2d(int w, int h) { return dim3((unsigned)((w + 63) / 64), (unsigned)((h + 3) / 4)); }

CtfK make_ctf(int N, const artia_ctf_params& p, float dmin_nm, float dmax_nm, float angle_deg, int forFP) {
    CtfK k;
    k.cs_m = p.cs_mm * 0.001f;
    k.lambda = ctf_wavelength(p.voltage_kV);
    const float ps = p.pixelsize_nm * powf(10.0f, -9.0f);
    const float maxFreq = (float)(1.0 / (double)(ps * 2.0f));
    k.freq_step = maxFreq / ((float)N / 2.0f);
    k.ac = p.amp_contrast;
    k.pc = sqrtf(1 - p.amp_contrast * p.amp_contrast);
    k.halfN = (float)N * 0.5f;
    k.def0 = dmin_nm * 0.000000001f;
    k.def1 = dmax_nm * 0.000000001f;
    k.angle = angle_deg / 180.0f * kPiF;
    k.wiener = p.wiener;
    for (int i = 0; i < 4; i++) k.beta[i] = p.beta_fac[i];
    k.forFP = forFP;
    k.phaseFlip = p.phase_flip_only;
    k.N = N;
    return k;
}

}  // namespace

// ctf.cu:47-63 with the kernel's float constants (h = 6.63e-34, c = 3e8, E0 = 511 keV)
float ctf_wavelength(float voltage_kV) {
    const float h = (float)6.63E-34, c = (float)3.00E+08;
    const int E0 = 511;
    const float a = (2 * E0 * voltage_kV * 1000.0f * 1000.0f) + (voltage_kV * voltage_kV * 1000.0f * 1000.0f);
    const double arg = (double)a * 1.602E-19 * 1.602E-19;
    return (h * c) / sqrtf((float)arg);
}

artia_status FftPlans::ensure(artia_ctx* ctx, int n) {
    if (N == n && r2c && c2r) return ARTIA_OK;
    release();
    if (hipfftPlan2d(&r2c, n, n, HIPFFT_R2C) != HIPFFT_SUCCESS ||
        hipfftPlan2d(&c2r, n, n, HIPFFT_C2R) != HIPFFT_SUCCESS) {
        release();
        return fail(ctx, ARTIA_ERR_FFT, "hipfftPlan2d(%d x %d) failed", n, n);
    }
    artia_status st = check_hip(ctx, hipMalloc(&square, (size_t)n * n * sizeof(float)), "hipMalloc(square)");
    if (st != ARTIA_OK) {
        release();
        return st;
    }
    st = check_hip(ctx, hipMalloc(&spec, (size_t)(n / 2 + 1) * n * sizeof(float2)), "hipMalloc(spectrum)");
    if (st != ARTIA_OK) {
        release();
        return st;
    }
    N = n;
    return ARTIA_OK;
}

void FftPlans::release() {
    if (r2c) hipfftDestroy(r2c);
    if (c2r) hipfftDestroy(c2r);
    if (square) (void)hipFree(square);
    if (spec) (void)hipFree(spec);
    r2c = c2r = 0;
    square = nullptr;
    spec = nullptr;
    N = 0;
}

artia_status FftBatch::ensure(artia_ctx* ctx, int n) {
    if (N == n && cap > 0) return ARTIA_OK;
    release();
    const double per_slab = (double)n * n * sizeof(float) + (double)(n / 2 + 1) * n * sizeof(float2);
    int c = kFftBatchMax;
    while (c > 1 && (double)c * per_slab > 2147483648.0) c /= 2;
    size_t work_bytes = 0;
    int dims[2] = {n, n};
    for (int i = 0, b = 1; b <= c; i++, b *= 2) {
        size_t w1 = 0, w2 = 0;
        if (hipfftCreate(&r2c[i]) != HIPFFT_SUCCESS || hipfftCreate(&c2r[i]) != HIPFFT_SUCCESS ||
            hipfftSetAutoAllocation(r2c[i], 0) != HIPFFT_SUCCESS ||
            hipfftSetAutoAllocation(c2r[i], 0) != HIPFFT_SUCCESS ||
            hipfftMakePlanMany(r2c[i], 2, dims, nullptr, 1, n * n, nullptr, 1, n * (n / 2 + 1), HIPFFT_R2C, b,
                               &w1) != HIPFFT_SUCCESS ||
            hipfftMakePlanMany(c2r[i], 2, dims, nullptr, 1, n * (n / 2 + 1), nullptr, 1, n * n, HIPFFT_C2R, b,
                               &w2) != HIPFFT_SUCCESS) {
            release();
            return fail(ctx, ARTIA_ERR_FFT, "hipfftMakePlanMany(%d x %d, batch %d) failed", n, n, b);
        }
        work_bytes = std::max(work_bytes, std::max(w1, w2));
    }
    artia_status st = check_hip(ctx, hipMalloc(&square, (size_t)c * n * n * sizeof(float)), "hipMalloc(squares)");
    if (st == ARTIA_OK)
        st = check_hip(ctx, hipMalloc(&spec, (size_t)(c + 1) * (n / 2 + 1) * n * sizeof(float2)), "hipMalloc(spectra)");
    if (st == ARTIA_OK && work_bytes > 0)
        st = check_hip(ctx, hipMalloc(&work, work_bytes), "hipMalloc(fft work area)");
    if (st != ARTIA_OK) {
        release();
        return st;
    }
    for (int i = 0, b = 1; b <= c; i++, b *= 2) {
        if (hipfftSetWorkArea(r2c[i], work) != HIPFFT_SUCCESS || hipfftSetWorkArea(c2r[i], work) != HIPFFT_SUCCESS) {
            release();
            return fail(ctx, ARTIA_ERR_FFT, "hipfftSetWorkArea failed");
        }
    }
    N = n;
    cap = c;
    return ARTIA_OK;
}

void FftBatch::release() {
    for (int i = 0; i < kFftBatchPlans; i++) {
        if (r2c[i]) hipfftDestroy(r2c[i]);
        if (c2r[i]) hipfftDestroy(c2r[i]);
        r2c[i] = c2r[i] = 0;
    }
    if (square) (void)hipFree(square);
    if (spec) (void)hipFree(spec);
    if (work) (void)hipFree(work);
    square = nullptr;
    spec = nullptr;
    work = nullptr;
    N = cap = 0;
}

artia_status run_ctf(artia_ctx* ctx, float2* spec, int N, size_t pitch, const artia_ctf_params& p, float dmin_nm,
                     float dmax_nm, float angle_deg, int forFP) {
    const CtfK k = make_ctf(N, p, dmin_nm, dmax_nm, angle_deg, forFP);
    if (ctx->numerics == ARTIA_NUMERICS_REFERENCE)
        hipLaunchKernelGGL(ctf_ref_kernel, grid2d(N / 2 + 1, N), dim3(64, 4), 0, ctx->stream, spec, pitch, k);
    else
        hipLaunchKernelGGL(ctf_exact_kernel, grid2d(N / 2 + 1, N), dim3(64, 4), 0, ctx->stream, spec, pitch, k);
    return check_hip(ctx, hipGetLastError(), "ctf launch");
}

artia_status run_wbp_weight(artia_ctx* ctx, float2* spec, int N, size_t pitch, float psi_deg, int method) {
    const float ang = -psi_deg / 180.0f * kPiF;
    hipLaunchKernelGGL(wbp_weight_kernel, grid2d(N / 2 + 1, N), dim3(64, 4), 0, ctx->stream, spec, pitch, N, ang,
                       method, ctx->numerics == ARTIA_NUMERICS_EXACT ? 1 : 0);
    return check_hip(ctx, hipGetLastError(), "wbpWeighting launch");
}

artia_status ctf_slab_filter(artia_ctx* ctx, float* img, int W, int H, size_t pitch, float* out, size_t out_pitch,
                             int accumulate, const int corners[8], const float cut[2], const float dim[2],
                             const artia_ctf_params& p, float dmin_nm, float dmax_nm, float angle_deg, int forFP) {
    const int N = W > H ? W : H;
    artia_status st = ctx->fft.ensure(ctx, N);
    if (st != ARTIA_OK) return st;
    if (corners) launch_crop_border(ctx, W, H, pitch, img, cut, dim, corners);
    launch_make_square(ctx, W, H, pitch, img, N, ctx->fft.square, (N - W) / 2, (N - H) / 2, 0, 1);
    if (hipfftSetStream(ctx->fft.r2c, ctx->stream) != HIPFFT_SUCCESS ||
        hipfftSetStream(ctx->fft.c2r, ctx->stream) != HIPFFT_SUCCESS)
        return fail(ctx, ARTIA_ERR_FFT, "hipfftSetStream failed");
    if (hipfftExecR2C(ctx->fft.r2c, ctx->fft.square, (hipfftComplex*)ctx->fft.spec) != HIPFFT_SUCCESS)
        return fail(ctx, ARTIA_ERR_FFT, "hipfftExecR2C failed");
    st = run_ctf(ctx, ctx->fft.spec, N, (size_t)(N / 2 + 1), p, dmin_nm, dmax_nm, angle_deg, forFP);
    if (st != ARTIA_OK) return st;
    if (hipfftExecC2R(ctx->fft.c2r, (hipfftComplex*)ctx->fft.spec, ctx->fft.square) != HIPFFT_SUCCESS)
        return fail(ctx, ARTIA_ERR_FFT, "hipfftExecC2R failed");
    const CropParams cp = corners ? make_crop_params(cut, dim, corners) : CropParams{};
    hipLaunchKernelGGL(finish_kernel, grid2d(W, H), dim3(64, 4), 0, ctx->stream, ctx->fft.square, 1, N, (N - W) / 2,
                       (N - H) / 2, W, H, out, out_pitch, (float)(N * N), accumulate, corners ? 1 : 0, cp, (size_t)0);
    return check_hip(ctx, hipGetLastError(), "slab filter");
}

artia_status ctf_filter_slabs(artia_ctx* ctx, const float* src, size_t src_stride, int n, int W, int H, size_t pitch,
                              float* out, size_t out_stride, const int corners[8], const float cut[2],
                              const float dim[2], const artia_ctf_params& p, const float* dmin_nm,
                              const float* dmax_nm, float angle_deg, int forFP) {
    if (n <= 0) return ARTIA_OK;
    const int N = W > H ? W : H;
    FftBatch& fb = ctx->fftb;
    artia_status st = fb.ensure(ctx, N);
    if (st != ARTIA_OK) return st;
    const int crop = corners ? 1 : 0;
    const CropParams cp = crop ? make_crop_params(cut, dim, corners) : CropParams{};
    const int bx = (N - W) / 2, by = (N - H) / 2;
    const size_t half = (size_t)(N / 2 + 1);
    // Every slab filtering the same uncropped image (the sub-volume WBP): one forward transform,
    // kept in the spare spectrum slot, serves all slabs.  When the slab images are whole squares
    // of a power-of-two edge stored back to back, the inverse transform writes them in place with
    // 1 / N^2 folded into the CTF factor.
    const bool shared = src_stride == 0 && !crop;
    const bool direct = shared && !forFP && W == N && H == N && pitch == (size_t)N &&
                        out_stride == (size_t)N * (size_t)N && (N & (N - 1)) == 0;
    float2* base = fb.spec + (size_t)fb.cap * half * (size_t)N;
    if (shared) {
        hipLaunchKernelGGL(slab_square_kernel, dim3((unsigned)((N + 63) / 64), (unsigned)((N + 3) / 4)), dim3(64, 4), 0,
                           ctx->stream, src, (size_t)0, W, H, pitch, N, bx, by, fb.square, cp, 0, 0, 1);
        if (hipfftSetStream(fb.r2c[0], ctx->stream) != HIPFFT_SUCCESS)
            return fail(ctx, ARTIA_ERR_FFT, "hipfftSetStream failed");
        if (hipfftExecR2C(fb.r2c[0], fb.square, (hipfftComplex*)base) != HIPFFT_SUCCESS)
            return fail(ctx, ARTIA_ERR_FFT, "hipfftExecR2C failed");
    }
    for (int k0 = 0; k0 < n;) {
        // the largest power-of-two batch that fits: at most log2(cap) + 1 chunks past the full ones
        int b = fb.cap, pi = 0;
        while (b > n - k0) b /= 2;
        for (int t = b; t > 1; t /= 2) pi++;
        if (hipfftSetStream(fb.r2c[pi], ctx->stream) != HIPFFT_SUCCESS ||
            hipfftSetStream(fb.c2r[pi], ctx->stream) != HIPFFT_SUCCESS)
            return fail(ctx, ARTIA_ERR_FFT, "hipfftSetStream failed");
        if (!shared) {
            hipLaunchKernelGGL(slab_square_kernel, dim3((unsigned)((N + 63) / 64), (unsigned)((N + 3) / 4)),
                               dim3(64, 4), 0, ctx->stream, src + (size_t)k0 * src_stride, src_stride, W, H, pitch, N,
                               bx, by, fb.square, cp, !crop ? 0 : (forFP ? 1 : k0 + 1), crop && !forFP ? 1 : 0, b);
            if (hipfftExecR2C(fb.r2c[pi], fb.square, (hipfftComplex*)fb.spec) != HIPFFT_SUCCESS)
                return fail(ctx, ARTIA_ERR_FFT, "hipfftExecR2C (batch %d) failed", b);
        }
        CtfStack cs;
        cs.k = make_ctf(N, p, dmin_nm[k0], dmax_nm[k0], angle_deg, forFP);
        for (int s = 0; s < b; s++) {
            const CtfK ks = make_ctf(N, p, dmin_nm[k0 + s], dmax_nm[k0 + s], angle_deg, forFP);
            cs.def0[s] = ks.def0;
            cs.def1[s] = ks.def1;
        }
        const dim3 sgrid((unsigned)((N / 2 + 1 + 63) / 64), (unsigned)((N + 3) / 4));
        if (shared)
            hipLaunchKernelGGL(ctf_exact_fan_kernel, sgrid, dim3(64, 4), 0, ctx->stream, base, fb.spec, half, cs, b,
                               direct ? 1.0f / ((float)N * (float)N) : 1.0f);
        else
            hipLaunchKernelGGL(ctf_exact_stack_kernel, sgrid, dim3(64, 4), 0, ctx->stream, fb.spec, half, cs, b);
        float* inv = direct ? out + (size_t)k0 * out_stride : fb.square;
        if (hipfftExecC2R(fb.c2r[pi], (hipfftComplex*)fb.spec, inv) != HIPFFT_SUCCESS)
            return fail(ctx, ARTIA_ERR_FFT, "hipfftExecC2R (batch %d) failed", b);
        if (direct) {
            // the slab images are written
        } else if (forFP)
            hipLaunchKernelGGL(finish_sum_kernel, grid2d(W, H), dim3(64, 4), 0, ctx->stream, fb.square, b, N, bx, by, W,
                               H, out, pitch, (float)(N * N), crop, cp);
        else
            hipLaunchKernelGGL(finish_kernel, grid2d(W, H), dim3(64, 4), 0, ctx->stream, fb.square, b, N, bx, by, W, H,
                               out + (size_t)k0 * out_stride, pitch, (float)(N * N), 0, crop, cp, out_stride);
        st = check_hip(ctx, hipGetLastError(), "stacked slab filter");
        if (st != ARTIA_OK) return st;
        k0 += b;
    }
    return ARTIA_OK;
}

artia_status wbp_filter(artia_ctx* ctx, const float* img, int W, int H, size_t pitch, float* out, size_t out_pitch,
                        float psi_deg, int method) {
    const int N = W > H ? W : H;
    artia_status st = ctx->fft.ensure(ctx, N);
    if (st != ARTIA_OK) return st;
    launch_make_square(ctx, W, H, pitch, img, N, ctx->fft.square, (N - W) / 2, (N - H) / 2, 0, 0);
    if (hipfftSetStream(ctx->fft.r2c, ctx->stream) != HIPFFT_SUCCESS ||
        hipfftSetStream(ctx->fft.c2r, ctx->stream) != HIPFFT_SUCCESS)
        return fail(ctx, ARTIA_ERR_FFT, "hipfftSetStream failed");
    if (hipfftExecR2C(ctx->fft.r2c, ctx->fft.square, (hipfftComplex*)ctx->fft.spec) != HIPFFT_SUCCESS)
        return fail(ctx, ARTIA_ERR_FFT, "hipfftExecR2C failed");
    st = run_wbp_weight(ctx, ctx->fft.spec, N, (size_t)(N / 2 + 1), psi_deg, method);
    if (st != ARTIA_OK) return st;
    if (hipfftExecC2R(ctx->fft.c2r, (hipfftComplex*)ctx->fft.spec, ctx->fft.square) != HIPFFT_SUCCESS)
        return fail(ctx, ARTIA_ERR_FFT, "hipfftExecC2R failed");
    const CropParams cp{};
    hipLaunchKernelGGL(finish_kernel, grid2d(W, H), dim3(64, 4), 0, ctx->stream, ctx->fft.square, 1, N, (N - W) / 2,
                       (N - H) / 2, W, H, out, out_pitch, (float)(N * N), 0, 0, cp, (size_t)0);
    return check_hip(ctx, hipGetLastError(), "wbp filter");
}

}  // namespace artia

using namespace artia;

namespace {
bool valid_ctf(const artia_ctf_params* p) {
    return p && p->voltage_kV > 0.0f && p->pixelsize_nm > 0.0f && p->wiener >= 0.0f;
}
}  // namespace

extern "C" {

artia_status artia_ctf_apply(artia_ctx* ctx, void* spectrum, int n, size_t pitch, const artia_ctf_params* p,
                             float defocus_min_nm, float defocus_max_nm, float angle_deg, int for_fp, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!spectrum || n <= 1 || pitch < (size_t)(n / 2 + 1) || !valid_ctf(p))
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_ctf_apply: invalid arguments");
    Timing t(ctx, ms);
    artia_status st = run_ctf(ctx, (float2*)spectrum, n, pitch, *p, defocus_min_nm, defocus_max_nm, angle_deg, for_fp);
    if (st != ARTIA_OK) return st;
    return t.finish();
}

artia_status artia_wbp_weight(artia_ctx* ctx, void* spectrum, int n, size_t pitch, float psi_deg, int method,
                              float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!spectrum || n <= 1 || pitch < (size_t)(n / 2 + 1) || method < 0 || method > 4)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_wbp_weight: invalid arguments");
    Timing t(ctx, ms);
    artia_status st = run_wbp_weight(ctx, (float2*)spectrum, n, pitch, psi_deg, method);
    if (st != ARTIA_OK) return st;
    return t.finish();
}

artia_status artia_ctf_slab_filter(artia_ctx* ctx, float* img, int width, int height, size_t pitch, float* out,
                                   size_t out_pitch, int accumulate, const int corners[8], const float cut_length[2],
                                   const float dim_length[2], const artia_ctf_params* p, float defocus_min_nm,
                                   float defocus_max_nm, float angle_deg, int for_fp, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!img || !out || (corners && (!cut_length || !dim_length)) || !valid_ctf(p) || width <= 0 || height <= 0 ||
        pitch < (size_t)width || out_pitch < (size_t)width)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_ctf_slab_filter: invalid arguments");
    Timing t(ctx, ms);
    artia_status st = ctf_slab_filter(ctx, img, width, height, pitch, out, out_pitch, accumulate, corners, cut_length,
                                      dim_length, *p, defocus_min_nm, defocus_max_nm, angle_deg, for_fp);
    if (st != ARTIA_OK) return st;
    return t.finish();
}

artia_status artia_wbp_filter(artia_ctx* ctx, const float* img, int width, int height, size_t pitch, float* out,
                              size_t out_pitch, float psi_deg, int method, float* ms) {
    if (!ctx) return ARTIA_ERR_INVALID_ARGUMENT;
    if (!img || !out || width <= 0 || height <= 0 || pitch < (size_t)width || out_pitch < (size_t)width ||
        method < 0 || method > 4)
        return fail(ctx, ARTIA_ERR_INVALID_ARGUMENT, "artia_wbp_filter: invalid arguments");
    Timing t(ctx, ms);
    artia_status st = wbp_filter(ctx, img, width, height, pitch, out, out_pitch, psi_deg, method);
    if (st != ARTIA_OK) return st;
    return t.finish();
}

}  // extern "C"
