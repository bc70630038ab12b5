// A C++ host program that reaches libartia only through include/artia_kernels.hpp, the functor
// adapter shaped like the reference's Kernels.h, the way Reconstructor-style host code would.
// Inputs come from files the calling test writes; outputs go back as raw files for a bit-for-bit
// comparison with the same calls made through the C ABI from Python.
//
//   client <dir>
//     reads  vg.bin (artia_vol_geom), tg.bin (artia_tilt_geom), params.txt ("W H os lambda"),
//            vol_f32.bin (nx*ny*nz floats), vol_f16.bin (nx*ny*nz IEEE half bits), img.bin (W*H floats)
//     writes proj_f32.bin, dist_f32.bin, proj_f16.bin (FPKernel on both volumes),
//            bp_f32.bin, bp_f16.bin (BPKernel, POINT, ungated, on both volumes)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "artia_kernels.hpp"

namespace {

template <typename T>
std::vector<T> read_file(const std::string& path, size_t count) {
    std::vector<T> v(count);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(T), count, f) != count) throw std::runtime_error("cannot read " + path);
    std::fclose(f);
    return v;
}

template <typename T>
void write_file(const std::string& path, const std::vector<T>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) throw std::runtime_error("cannot write " + path);
    std::fclose(f);
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
T* to_device(const std::vector<T>& h) {
    T* d = nullptr;
    hip_check(hipMalloc(&d, h.size() * sizeof(T)), "hipMalloc");
    hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return d;
}

template <typename T>
std::vector<T> to_host(const T* d, size_t count) {
    std::vector<T> h(count);
    hip_check(hipMemcpy(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return h;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    try {
        const artia_vol_geom vg = read_file<artia_vol_geom>(dir + "/vg.bin", 1)[0];
        const artia_tilt_geom tg = read_file<artia_tilt_geom>(dir + "/tg.bin", 1)[0];
        int W = 0, H = 0, os = 0;
        float lambda = 0.0f;
        FILE* pf = std::fopen((dir + "/params.txt").c_str(), "r");
        if (!pf || std::fscanf(pf, "%d %d %d %f", &W, &H, &os, &lambda) != 4) throw std::runtime_error("params.txt");
        std::fclose(pf);
        const size_t nvox = (size_t)vg.nx * (size_t)vg.ny * (size_t)vg.nz, npix = (size_t)W * (size_t)H;
        float* vol = to_device(read_file<float>(dir + "/vol_f32.bin", nvox));
        uint16_t* volh = to_device(read_file<uint16_t>(dir + "/vol_f16.bin", nvox));
        float* img = to_device(read_file<float>(dir + "/img.bin", npix));
        float* proj = to_device(std::vector<float>(npix, 0.0f));
        float* dist = to_device(std::vector<float>(npix, 0.0f));

        artia::Context ctx(0);
        artia::FPKernel fp(ctx);
        artia::BPKernel bp(ctx);
        float ms = fp(vg, tg, vol, W, H, (size_t)W, proj, dist);
        write_file(dir + "/proj_f32.bin", to_host(proj, npix));
        write_file(dir + "/dist_f32.bin", to_host(dist, npix));
        hip_check(hipMemset(proj, 0, npix * sizeof(float)), "hipMemset");
        ms += fp(vg, tg, volh, W, H, (size_t)W, proj, dist);
        write_file(dir + "/proj_f16.bin", to_host(proj, npix));
        ms += bp(vg, tg, vol, img, W, H, (size_t)W, lambda, os, 0.0f, 9999999999999.0f, ARTIA_FILTER_POINT);
        write_file(dir + "/bp_f32.bin", to_host(vol, nvox));
        ms += bp(vg, tg, volh, img, W, H, (size_t)W, lambda, os, 0.0f, 9999999999999.0f, ARTIA_FILTER_POINT);
        write_file(dir + "/bp_f16.bin", to_host(volh, nvox));
        // an error comes back as artia::Error with the library's message
        bool threw = false;
        try {
            bp(vg, tg, vol, img, W, H, (size_t)W, lambda, 0, 0.0f, 1.0f, ARTIA_FILTER_POINT);
        } catch (const artia::Error& e) {
            threw = e.status == ARTIA_ERR_INVALID_ARGUMENT;
        }
        for (void* p : {(void*)vol, (void*)volh, (void*)img, (void*)proj, (void*)dist}) (void)hipFree(p);
        std::printf("client: 4 functor calls in %.3f ms; bad oversampling %s\n", ms,
                    threw ? "raised artia::Error" : "NOT RAISED");
        return threw ? 0 : 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "client: %s\n", e.what());
        return 1;
    }
}
