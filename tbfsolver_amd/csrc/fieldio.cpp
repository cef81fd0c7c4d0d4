// fieldio.cpp -- the reference's field files and restart folders, host side.
//
// A field file (field.f90:419-476, readInternalField / writeField) is an
// unformatted stream: one default integer (int32) with the number of values,
// then the values (FP64) of the field's index range in column-major order,
// x fastest.  Index ranges (setIndexesBounds, field.f90:1746-1789): cell
// centres 'cl' 1..n on every axis; staggered 'sx', 'sy', 'sz' 0..n on their own
// axis.  The engine keeps the staggered x and z faces 1..n (face 0 is the
// periodic image of face n) and the y faces 0..n (walls at 0 and n).
// A time folder (time.f90:309-343) is named by the output number (format I6,
// left-adjusted) and holds one file per field plus info_restart, the time
// (ES25.15E3) and the bubble count (I6) as text lines.
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "tbfgpu.h"

namespace {

bool extents(int type, const int32_t n[3], int64_t ext[3])
{
    if (!n || n[0] < 1 || n[1] < 1 || n[2] < 1 || type < TBFG_FIELD_CL || type > TBFG_FIELD_SZ) return false;
    for (int d = 0; d < 3; ++d) ext[d] = n[d];
    if (type != TBFG_FIELD_CL) ext[type - 1] += 1;  // the range 0..n of the staggered axis
    return true;
}

// the file value at (i, j, k) of the file's range, from the engine array
double engine_value(int type, const int32_t n[3], const double* f, int64_t i, int64_t j, int64_t k)
{
    const int64_t nx = n[0], ny = n[1], nz = n[2];
    switch (type) {
    case TBFG_FIELD_SX:  // file i = 0..nx, engine faces 1..nx at 0..nx-1
        return f[(i == 0 ? nx - 1 : i - 1) + nx * (j + ny * k)];
    case TBFG_FIELD_SY:  // file j = 0..ny = engine j
        return f[i + nx * (j + (ny + 1) * k)];
    case TBFG_FIELD_SZ:
        return f[i + nx * (j + ny * (k == 0 ? nz - 1 : k - 1))];
    default:
        return f[i + nx * (j + ny * k)];
    }
}

}  // namespace

extern "C" int tbfg_field_count(int type, const int32_t n[3], int64_t* count)
{
    int64_t e[3];
    if (!count || !extents(type, n, e)) {
        tbfg::set_error("tbfg_field_count: invalid field type or size");
        return TBFG_EINVAL;
    }
    *count = e[0] * e[1] * e[2];
    return TBFG_OK;
}

extern "C" int tbfg_field_write(const char* path, int type, const int32_t n[3], const double* f)
{
    int64_t e[3];
    if (!path || !f || !extents(type, n, e)) {
        tbfg::set_error("tbfg_field_write: invalid argument");
        return TBFG_EINVAL;
    }
    const int64_t count = e[0] * e[1] * e[2];
    if (count > INT32_MAX) {
        tbfg::set_error("tbfg_field_write: %lld values exceed the int32 count of the format", (long long)count);
        return TBFG_EINVAL;
    }
    std::vector<double> buf((size_t)count);
    size_t t = 0;
    for (int64_t k = 0; k < e[2]; ++k)
        for (int64_t j = 0; j < e[1]; ++j)
            for (int64_t i = 0; i < e[0]; ++i) buf[t++] = engine_value(type, n, f, i, j, k);
    FILE* fp = std::fopen(path, "wb");
    if (!fp) {
        tbfg::set_error("tbfg_field_write: cannot open %s for writing", path);
        return TBFG_EIO;
    }
    const int32_t c32 = (int32_t)count;
    const bool ok = std::fwrite(&c32, sizeof c32, 1, fp) == 1 && std::fwrite(buf.data(), 8, buf.size(), fp) == buf.size();
    if (std::fclose(fp) != 0 || !ok) {
        tbfg::set_error("tbfg_field_write: write to %s failed", path);
        return TBFG_EIO;
    }
    return TBFG_OK;
}

extern "C" int tbfg_field_read(const char* path, int type, const int32_t n[3], double* f)
{
    int64_t e[3];
    if (!path || !f || !extents(type, n, e)) {
        tbfg::set_error("tbfg_field_read: invalid argument");
        return TBFG_EINVAL;
    }
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        tbfg::set_error("tbfg_field_read: cannot open %s", path);
        return TBFG_EIO;
    }
    const int64_t count = e[0] * e[1] * e[2];
    int32_t c32 = 0;
    if (std::fread(&c32, sizeof c32, 1, fp) != 1) {
        std::fclose(fp);
        tbfg::set_error("tbfg_field_read: %s: no value count", path);
        return TBFG_EIO;
    }
    if ((int64_t)c32 != count) {  // readInternalField (field.f90:438-441)
        std::fclose(fp);
        tbfg::set_error("tbfg_field_read: dimension of field %s (%d) not equal to mesh size (%lld)", path, c32,
                        (long long)count);
        return TBFG_EINVAL;
    }
    std::vector<double> buf((size_t)count);
    const size_t got = std::fread(buf.data(), 8, buf.size(), fp);
    std::fclose(fp);
    if (got != buf.size()) {
        tbfg::set_error("tbfg_field_read: %s: %zu of %lld values", path, got, (long long)count);
        return TBFG_EIO;
    }
    // the engine keeps faces 1..n of sx / sz (file entry 0 is their periodic image)
    const int64_t nx = n[0], ny = n[1];
    size_t t = 0;
    for (int64_t k = 0; k < e[2]; ++k)
        for (int64_t j = 0; j < e[1]; ++j)
            for (int64_t i = 0; i < e[0]; ++i, ++t) {
                int64_t ii = i, kk = k;
                if (type == TBFG_FIELD_SX) {
                    if (i == 0) continue;
                    ii = i - 1;
                }
                if (type == TBFG_FIELD_SZ) {
                    if (k == 0) continue;
                    kk = k - 1;
                }
                const int64_t nyr = type == TBFG_FIELD_SY ? ny + 1 : ny;
                f[ii + nx * (j + nyr * kk)] = buf[t];
            }
    return TBFG_OK;
}

extern "C" int tbfg_time_folder_name(int32_t n, char* out, int32_t len)
{
    char tmp[32];
    std::snprintf(tmp, sizeof tmp, "%6d", (int)n);  // write(dirName, '(I6)'), then adjustl / trim
    const char* p = tmp;
    while (*p == ' ') ++p;
    if (!out || len < (int32_t)std::strlen(p) + 1 || std::strlen(tmp) > 6) {
        tbfg::set_error("tbfg_time_folder_name: invalid argument");
        return TBFG_EINVAL;
    }
    std::strcpy(out, p);
    return TBFG_OK;
}
