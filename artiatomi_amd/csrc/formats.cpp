// Headers of EmSART's on-disk stacks and volumes: the EM format (EmTools/FileIO/EmHeader.h,
// EmFile.cpp) and MRC (mrcHeader.h, MRCFile.cpp).  Parsing reads the fields at their byte offsets
// (no struct overlay), so the result does not depend on this compiler's struct packing; the fields
// are little-endian, and an MRC header whose NX reads above 32000 is taken as big-endian and
// byte-swapped, as MRCFile::OpenAndRead does (MRCFile.cpp:84-88).
#include <cstdint>
#include <cstring>

#include "../../include/artia.h"

namespace {

constexpr size_t kEmHeaderBytes = 512;
constexpr size_t kMrcHeaderBytes = 1024;

uint32_t rd32(const unsigned char* p, size_t off, bool swap) {
    uint32_t v;
    std::memcpy(&v, p + off, 4);
    return swap ? __builtin_bswap32(v) : v;
}
void wr32(unsigned char* p, size_t off, uint32_t v) { std::memcpy(p + off, &v, 4); }
void wrf(unsigned char* p, size_t off, float v) { std::memcpy(p + off, &v, 4); }
float as_float(uint32_t v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

// EM DataType (EmHeader.h EmDataType_Enum) -> element bytes and the artia_dtype of
// PrepareProjection's conversions (Reconstructor.cpp:1165-1190), -1 where it has none
bool em_type(int t, int& bytes, int& dtype) {
    switch (t) {
    case 1: bytes = 1; dtype = -1; return true;                 // byte (read as unsigned char)
    case 2: bytes = 2; dtype = ARTIA_DT_SHORT; return true;
    case 4: bytes = 4; dtype = ARTIA_DT_INT; return true;
    case 5: bytes = 4; dtype = ARTIA_DT_FLOAT; return true;
    case 8: bytes = 8; dtype = -1; return true;                 // complex
    case 9: bytes = 8; dtype = -1; return true;                 // double
    case 10: bytes = 2; dtype = -1; return true;                // half (FP16 volumes)
    default: return false;
    }
}

// MRC MODE (mrcHeader.h MrcMode_Enum; 12 is MRC2014's float16)
bool mrc_type(uint32_t m, int& bytes, int& dtype) {
    switch (m) {
    case 0: bytes = 1; dtype = -1; return true;
    case 1: bytes = 2; dtype = ARTIA_DT_SHORT; return true;
    case 2: bytes = 4; dtype = ARTIA_DT_FLOAT; return true;
    case 3: bytes = 4; dtype = -1; return true;                 // complex int16
    case 4: bytes = 8; dtype = -1; return true;                 // complex float
    case 6: bytes = 2; dtype = ARTIA_DT_USHORT; return true;
    case 12: bytes = 2; dtype = -1; return true;
    case 16: bytes = 3; dtype = -1; return true;                // RGB
    case 64: bytes = 2; dtype = -1; return true;                // half, as EmSART writes FP16 volumes
    default: return false;
    }
}

}  // namespace

extern "C" {

artia_status artia_file_parse_header(int format, const void* header, size_t size, artia_file_info* out) {
    if (!header || !out) return ARTIA_ERR_INVALID_ARGUMENT;
    const unsigned char* h = static_cast<const unsigned char*>(header);
    artia_file_info info{};
    if (format == ARTIA_FILE_EM) {
        if (size < kEmHeaderBytes) return ARTIA_ERR_INVALID_ARGUMENT;
        // EmFile::OpenAndRead (EmFile.cpp:75-78): PC machine coding only
        if (h[0] != 6) return ARTIA_ERR_UNSUPPORTED;
        info.raw_type = h[3];
        if (!em_type(info.raw_type, info.elem_bytes, info.dtype)) return ARTIA_ERR_UNSUPPORTED;
        info.nx = (int)rd32(h, 4, false);
        info.ny = (int)rd32(h, 8, false);
        info.nz = (int)rd32(h, 12, false);
        info.data_offset = kEmHeaderBytes;
        info.pixel_size_nm = 1.0f;   // TiltSeries / Volume::GetPixelSize for EM files
    } else if (format == ARTIA_FILE_MRC) {
        if (size < kMrcHeaderBytes) return ARTIA_ERR_INVALID_ARGUMENT;
        const bool swap = rd32(h, 0, false) > 32000u;
        info.byte_swapped = swap ? 1 : 0;
        info.raw_type = (int)rd32(h, 12, swap);
        if (!mrc_type((uint32_t)info.raw_type, info.elem_bytes, info.dtype)) return ARTIA_ERR_UNSUPPORTED;
        const uint32_t nx = rd32(h, 0, swap);
        info.nx = (int)nx;
        info.ny = (int)rd32(h, 4, swap);
        info.nz = (int)rd32(h, 8, swap);
        info.data_offset = kMrcHeaderBytes + (size_t)rd32(h, 92, swap);   // + NEXT extended bytes
        const float xlen = as_float(rd32(h, 40, swap));
        // MRCFile::GetPixelsize (MRCFile.cpp:585-591): Xlen / NX in Angstrom, as nm
        info.pixel_size_nm = (nx == 0 || xlen == 0.0f) ? 0.0f : xlen / (float)nx / 10.0f;
    } else {
        return ARTIA_ERR_INVALID_ARGUMENT;
    }
    if (info.nx <= 0 || info.ny <= 0 || info.nz <= 0) return ARTIA_ERR_INVALID_ARGUMENT;
    *out = info;
    return ARTIA_OK;
}

artia_status artia_file_volume_header(int format, int half, int nx, int ny, int nz, float pixel_size_nm,
                                      const float voxel_size[3], void* out, size_t capacity, size_t* header_bytes) {
    if (!out || !voxel_size || nx <= 0 || ny <= 0 || nz <= 0) return ARTIA_ERR_INVALID_ARGUMENT;
    unsigned char* h = static_cast<unsigned char*>(out);
    if (format == ARTIA_FILE_MRC) {
        // EmSart.cpp:717-746: zeroed header, sizes, MODE, Xlen = pixel size * dim * voxel size * 10
        if (capacity < kMrcHeaderBytes) return ARTIA_ERR_INVALID_ARGUMENT;
        std::memset(h, 0, kMrcHeaderBytes);
        wr32(h, 0, (uint32_t)nx);
        wr32(h, 4, (uint32_t)ny);
        wr32(h, 8, (uint32_t)nz);
        wr32(h, 12, half ? 64u : 2u);
        wr32(h, 28, (uint32_t)nx);   // MX, MY, MZ
        wr32(h, 32, (uint32_t)ny);
        wr32(h, 36, (uint32_t)nz);
        wrf(h, 40, pixel_size_nm * (float)nx * voxel_size[0] * 10.0f);
        wrf(h, 44, pixel_size_nm * (float)ny * voxel_size[1] * 10.0f);
        wrf(h, 48, pixel_size_nm * (float)nz * voxel_size[2] * 10.0f);
        wr32(h, 64, 1u);   // MAPC, MAPR, MAPS = x, y, z
        wr32(h, 68, 2u);
        wr32(h, 72, 3u);
        if (header_bytes) *header_bytes = kMrcHeaderBytes;
        return ARTIA_OK;
    }
    if (format == ARTIA_FILE_EM) {
        // EmSart.cpp:748-759: PC machine coding, float (or half) data, ObjectPixelSize in pm
        if (capacity < kEmHeaderBytes) return ARTIA_ERR_INVALID_ARGUMENT;
        std::memset(h, 0, kEmHeaderBytes);
        h[0] = 6;
        h[3] = half ? 10 : 5;
        wr32(h, 4, (uint32_t)nx);
        wr32(h, 8, (uint32_t)ny);
        wr32(h, 12, (uint32_t)nz);
        wr32(h, 96 + 6 * 4, (uint32_t)(int)(pixel_size_nm * voxel_size[0] * 1000.0f));
        if (header_bytes) *header_bytes = kEmHeaderBytes;
        return ARTIA_OK;
    }
    return ARTIA_ERR_INVALID_ARGUMENT;
}

}  // extern "C"
