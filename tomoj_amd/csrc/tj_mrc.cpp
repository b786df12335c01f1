// tj_mrc.cpp -- MRC files, the container TomoJ reads tilt series from and
// writes reconstructions to (InputOutput/MRC_Reader.java, MRC_Writer.java).
// Host code only: the data go to and from the device through tj_load_projections
// / tj_get_volume.
//
// Reading follows MRC_Reader.getMRCFileInfo (MRC_Reader.java:176-233): the mode
// word is read big-endian first; mode 0 switches to little-endian when the
// big-endian sizes are negative or >= 65536; a mode other than 1, 2, 6 is
// re-read little-endian and must then be 1, 2 or 6.  Data start at
// 1024 + the extended-header size (word 23); an extended header of 131072
// bytes is FEI's, whose per-image tilt angle is the float at 1024 + 128 i
// (readTiltAngles :235-256).  Pixel types: mode 0 unsigned 8-bit, 1 signed
// 16-bit, 2 float, 6 unsigned 16-bit.
//
// Writing follows MRC_Writer.createHeader (MRC_Writer.java:104-160) for float
// images: little-endian, mode 2, sizes in words 0-2, map axes 1, 2, 3 in words
// 16-18, and the minimum, maximum and (float)(double sum / count) mean in
// words 19-21, the other header bytes zero.
#include "../../include/tomoj_hip.h"

#include <cmath>
#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

void tj_record_error(const char* msg);

namespace {

int mrc_fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    tj_record_error(buf);
    return code;
}

int32_t word(const unsigned char* h, int index, bool little)
{
    const uint32_t b1 = h[index], b2 = h[index + 1], b3 = h[index + 2], b4 = h[index + 3];
    return (int32_t)(little ? (b4 << 24 | b3 << 16 | b2 << 8 | b1) : (b1 << 24 | b2 << 16 | b3 << 8 | b4));
}

void put_word(unsigned char* h, int index, uint32_t v)  // little-endian
{
    h[index] = (unsigned char)v;
    h[index + 1] = (unsigned char)(v >> 8);
    h[index + 2] = (unsigned char)(v >> 16);
    h[index + 3] = (unsigned char)(v >> 24);
}

int parse(FILE* f, const char* path, tj_mrc_info* info)
{
    unsigned char h[1024];
    if (fread(h, 1, 1024, f) != 1024) return mrc_fail(TJ_EIO, "%s: shorter than an MRC header", path);
    bool little = false;
    int mode = word(h, 12, false);
    if (mode == 0) {
        const int w = word(h, 0, false), ht = word(h, 4, false), n = word(h, 8, false);
        if (w < 0 || ht < 0 || n < 0 || w >= 65536 || ht >= 65536 || n >= 65536) little = true;
    } else if (!(mode == 1 || mode == 2 || mode == 6)) {
        little = true;
        mode = word(h, 12, true);
        if (!(mode == 1 || mode == 2 || mode == 6))
            return mrc_fail(TJ_EUNSUP, "%s: unimplemented MRC data mode %d", path, mode);
    }
    memset(info, 0, sizeof *info);
    info->nx = word(h, 0, little);
    info->ny = word(h, 4, little);
    info->nz = word(h, 8, little);
    info->mode = mode;
    info->little_endian = little ? 1 : 0;
    info->extra_bytes = word(h, 92, little);
    info->fei = info->extra_bytes == 131072;
    int32_t a = word(h, 76, little), b = word(h, 80, little), c = word(h, 84, little);
    memcpy(&info->amin, &a, 4);
    memcpy(&info->amax, &b, 4);
    memcpy(&info->amean, &c, 4);
    if (info->nx <= 0 || info->ny <= 0 || info->nz <= 0 || info->extra_bytes < 0)
        return mrc_fail(TJ_EINVAL, "%s: bad MRC sizes", path);
    return TJ_OK;
}

}  // namespace

extern "C" {

int tj_mrc_read_info(const char* path, tj_mrc_info* info)
{
    if (!path || !info) return mrc_fail(TJ_EINVAL, "tj_mrc_read_info: null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return mrc_fail(TJ_EIO, "%s: cannot open", path);
    const int rc = parse(f, path, info);
    fclose(f);
    return rc;
}

int tj_mrc_read(const char* path, float* out, float* tilts)
{
    if (!path || !out) return mrc_fail(TJ_EINVAL, "tj_mrc_read: null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return mrc_fail(TJ_EIO, "%s: cannot open", path);
    tj_mrc_info in;
    int rc = parse(f, path, &in);
    if (rc) {
        fclose(f);
        return rc;
    }
    const bool little = in.little_endian != 0;
    if (tilts) {
        if (in.fei) {
            std::vector<unsigned char> ext((size_t)128 * in.nz);
            if (fseek(f, 1024, SEEK_SET) != 0 || fread(ext.data(), 1, ext.size(), f) != ext.size()) {
                fclose(f);
                return mrc_fail(TJ_EIO, "%s: truncated FEI extended header", path);
            }
            for (int i = 0; i < in.nz; ++i) {
                const int32_t v = word(ext.data(), 128 * i, little);
                memcpy(&tilts[i], &v, 4);
            }
        } else {
            for (int i = 0; i < in.nz; ++i) tilts[i] = NAN;
        }
    }
    const int bpp = in.mode == 0 ? 1 : (in.mode == 2 ? 4 : 2);
    const size_t n = (size_t)in.nx * in.ny * in.nz;
    std::vector<unsigned char> raw(n * bpp);
    if (fseek(f, 1024L + in.extra_bytes, SEEK_SET) != 0 || fread(raw.data(), 1, raw.size(), f) != raw.size()) {
        fclose(f);
        return mrc_fail(TJ_EIO, "%s: truncated MRC data", path);
    }
    fclose(f);
    for (size_t k = 0; k < n; ++k) {
        const unsigned char* p = raw.data() + k * bpp;
        if (in.mode == 0) {
            out[k] = (float)p[0];
        } else if (in.mode == 2) {
            const int32_t v = word(p, 0, little);
            memcpy(&out[k], &v, 4);
        } else {
            const uint32_t u = little ? (uint32_t)(p[1] << 8 | p[0]) : (uint32_t)(p[0] << 8 | p[1]);
            out[k] = in.mode == 1 ? (float)(int16_t)(uint16_t)u : (float)(uint16_t)u;
        }
    }
    return TJ_OK;
}

int tj_mrc_write(const char* path, const float* data, int nx, int ny, int nz)
{
    if (!path || !data || nx <= 0 || ny <= 0 || nz <= 0) return mrc_fail(TJ_EINVAL, "tj_mrc_write: bad argument");
    const size_t n = (size_t)nx * ny * nz;
    // createHeader's statistics: a value above the running maximum is not
    // compared with the minimum (the else-if of MRC_Writer.java:137-141)
    double sum = 0.0;
    float mn = NAN, mx = NAN;
    for (size_t k = 0; k < n; ++k) {
        const float v = data[k];
        sum += v;
        if (std::isnan(mn)) mn = v;
        if (std::isnan(mx)) mx = v;
        if (mx < v) mx = v;
        else if (mn > v) mn = v;
    }
    const float avg = (float)(sum / (double)n);
    unsigned char h[1024];
    memset(h, 0, sizeof h);
    uint32_t u;
    put_word(h, 0, (uint32_t)nx);
    put_word(h, 4, (uint32_t)ny);
    put_word(h, 8, (uint32_t)nz);
    put_word(h, 12, 2u);
    put_word(h, 64, 1u);
    put_word(h, 68, 2u);
    put_word(h, 72, 3u);
    memcpy(&u, &mn, 4);
    put_word(h, 76, u);
    memcpy(&u, &mx, 4);
    put_word(h, 80, u);
    memcpy(&u, &avg, 4);
    put_word(h, 84, u);
    FILE* f = fopen(path, "wb");
    if (!f) return mrc_fail(TJ_EIO, "%s: cannot create", path);
    bool ok = fwrite(h, 1, 1024, f) == 1024;
    std::vector<unsigned char> buf;
    const size_t chunk = (size_t)1 << 20;
    for (size_t k0 = 0; ok && k0 < n; k0 += chunk) {
        const size_t m = std::min(chunk, n - k0);
        buf.resize(m * 4);
        for (size_t k = 0; k < m; ++k) {
            memcpy(&u, &data[k0 + k], 4);
            put_word(buf.data(), (int)(4 * k), u);
        }
        ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) return mrc_fail(TJ_EIO, "%s: write failed", path);
    return TJ_OK;
}

}  // extern "C"
