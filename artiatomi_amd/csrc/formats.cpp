// Headers of EmSART's on-disk stacks and volumes: the EM format (EmTools/FileIO/EmHeader.h,
// EmFile.cpp) and MRC (mrcHeader.h, MRCFile.cpp).  Parsing reads the fields at their byte offsets
// (no struct overlay), so the result does not depend on this compiler's struct packing; the fields
// are little-endian, and an MRC header whose NX reads above 32000 is taken as big-endian and
// byte-swapped, as MRCFile::OpenAndRead does (MRCFile.cpp:84-88).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

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

// struct motive (MotiveListe.h:81-106): 20 floats per particle row
constexpr int kMotiveColumns = 20;

// MotiveList::GetAt (MotiveList.cpp:76-90): the row, positions and shifts scaled by the binnings
artia_motive motive_at(const float* data, int index, float click, float shift) {
    artia_motive m;
    std::memcpy(&m, data + (size_t)index * kMotiveColumns, sizeof(m));
    m.x_coord_2d *= click;
    m.y_coord_2d *= click;
    m.x *= click;
    m.y *= click;
    m.z *= click;
    m.x_shift *= shift;
    m.y_shift *= shift;
    m.z_shift *= shift;
    return m;
}

// MotiveList::GetDistance (MotiveList.cpp:126-135), in float
float motive_distance(const artia_motive& a, const artia_motive& b) {
    const float x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
    return std::sqrt(x * x + y * y + z * z);
}

// The distinct class numbers in file order, kept as int (MotiveList.cpp:57-73: an int list
// compared against the float class, appended with the float truncated)
std::vector<int> class_numbers(const float* data, int rows) {
    std::vector<int> cls;
    for (int i = 0; i < rows; i++) {
        const float c = data[(size_t)i * kMotiveColumns + 19];
        bool seen = false;
        for (int k : cls) seen = seen || (float)k == c;
        if (!seen) cls.push_back((int)c);
    }
    return cls;
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

// ---------------------------------------------------------------- ShiftFile

artia_status artia_shift_file_get(const float* data, int dim_x, int dim_y, int dim_z, int proj, int motive,
                                  float out[2]) {
    if (!data || !out || dim_x <= 0 || dim_y <= 0 || dim_z < 2 || proj < 0 || proj >= dim_y || motive < 0 ||
        motive >= dim_x)
        return ARTIA_ERR_INVALID_ARGUMENT;
    const size_t plane = (size_t)dim_x * (size_t)dim_y, at = (size_t)proj * (size_t)dim_x + (size_t)motive;
    out[0] = -data[at];
    out[1] = -data[plane + at];
    return ARTIA_OK;
}

artia_status artia_shift_file_set(float* data, int dim_x, int dim_y, int dim_z, int proj, int motive,
                                  const float value[2]) {
    if (!data || !value || dim_x <= 0 || dim_y <= 0 || dim_z < 2 || proj < 0 || proj >= dim_y || motive < 0 ||
        motive >= dim_x)
        return ARTIA_ERR_INVALID_ARGUMENT;
    const size_t plane = (size_t)dim_x * (size_t)dim_y, at = (size_t)proj * (size_t)dim_x + (size_t)motive;
    data[at] = value[0];
    data[plane + at] = value[1];
    return ARTIA_OK;
}

// ---------------------------------------------------------------- MotiveList

artia_status artia_motive_get(const float* data, int dim_x, int dim_y, int index, float binning_click,
                              float binning_shift, artia_motive* out) {
    static_assert(sizeof(artia_motive) == kMotiveColumns * sizeof(float), "artia_motive is 20 floats");
    if (!data || !out || dim_x != kMotiveColumns || dim_y <= 0 || index < 0 || index >= dim_y)
        return ARTIA_ERR_INVALID_ARGUMENT;
    *out = motive_at(data, index, binning_click, binning_shift);
    return ARTIA_OK;
}

artia_status artia_motive_group_count(const float* data, int dim_x, int dim_y, int mode, int* count) {
    if (!data || !count || dim_x != kMotiveColumns || dim_y <= 0) return ARTIA_ERR_INVALID_ARGUMENT;
    if (mode == ARTIA_GROUP_BY_CLASS) {
        *count = (int)class_numbers(data, dim_y).size();
        return ARTIA_OK;
    }
    if (mode == ARTIA_GROUP_MAX_DIST || mode == ARTIA_GROUP_MAX_COUNT) {
        *count = dim_y;
        return ARTIA_OK;
    }
    return ARTIA_ERR_INVALID_ARGUMENT;
}

artia_status artia_motive_group(const float* data, int dim_x, int dim_y, int mode, int group, float max_distance,
                                int group_size, float binning_click, float binning_shift, int* out, int capacity,
                                int* count) {
    if (!data || !count || dim_x != kMotiveColumns || dim_y <= 0 || capacity < 0 || (capacity > 0 && !out))
        return ARTIA_ERR_INVALID_ARGUMENT;
    std::vector<int> members;
    if (mode == ARTIA_GROUP_BY_CLASS) {
        // GetGroupIdx: -1 (no members) outside the class list; a negative class number also
        // yields no members (MotiveList.cpp:206-207)
        const std::vector<int> classes = class_numbers(data, dim_y);
        if (group >= 0 && group < (int)classes.size() && classes[(size_t)group] >= 0) {
            const float cls = (float)classes[(size_t)group];
            for (int i = 0; i < dim_y; i++)
                if (data[(size_t)i * kMotiveColumns + 19] == cls) members.push_back(i);
        }
    } else if (mode == ARTIA_GROUP_MAX_DIST || mode == ARTIA_GROUP_MAX_COUNT) {
        if (group < 0 || group >= dim_y) return ARTIA_ERR_INVALID_ARGUMENT;
        const artia_motive c = motive_at(data, group, binning_click, binning_shift);
        if (mode == ARTIA_GROUP_MAX_DIST) {
            members.push_back(group);
            for (int i = 0; i < dim_y; i++)
                if (i != group && motive_distance(motive_at(data, i, binning_click, binning_shift), c) <= max_distance)
                    members.push_back(i);
        } else {
            std::vector<std::pair<float, int>> d((size_t)dim_y);
            for (int i = 0; i < dim_y; i++)
                d[(size_t)i] = {motive_distance(motive_at(data, i, binning_click, binning_shift), c), i};
            std::sort(d.begin(), d.end());
            const int n = std::min(group_size, dim_y);
            for (int i = 0; i < n; i++) members.push_back(d[(size_t)i].second);
        }
    } else {
        return ARTIA_ERR_INVALID_ARGUMENT;
    }
    *count = (int)members.size();
    if ((int)members.size() > capacity) return ARTIA_ERR_INVALID_ARGUMENT;
    std::copy(members.begin(), members.end(), out);
    return ARTIA_OK;
}

}  // extern "C"
