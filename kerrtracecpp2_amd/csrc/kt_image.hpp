// ============================================================================
//  kt_image.hpp -- host-side file formats of the renderer:
//    * .kgeo  GeoPixel container, byte-compatible with the reference's
//             save_kgeo / load_kgeo (main.cpp:96-137): "KGEO", u32 version 1,
//             KGeoMeta (88 bytes), W*H 28-byte GeoPixel records.
//    * PNG    8-bit RGB writer (stored rows, filter 0, zlib deflate) and a
//             reader for non-interlaced 8-bit grey / grey+alpha / RGB / RGBA /
//             palette images (the equirect backgrounds).
//    * PPM    binary P6 reader (maxval 255).
//  The reference decodes backgrounds with stb_image (main.cpp:145-152), which
//  also reads JPEG; JPEG is not provided here and is reported as a load failure.
// ============================================================================
#pragma once
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/kt_hip.h"

namespace kthip {

static_assert(sizeof(kt_geopixel) == 28, "GeoPixel is 28 bytes (main.cpp:94)");

struct KGeoMeta {
    uint32_t W, H;
    double M_bh, a_bh, Q_bh, Lam;
    double r_isco, r_disk_in, r_disk_out;
    double theta_obs, phi_obs, r_obs;  // degrees, degrees, M
};
static_assert(sizeof(KGeoMeta) == 88, "KGeoMeta layout (main.cpp:99-104)");

inline bool save_kgeo(const std::string& path, const std::vector<kt_geopixel>& geo, const KGeoMeta& meta) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    const uint32_t version = 1;
    f.write("KGEO", 4);
    f.write(reinterpret_cast<const char*>(&version), 4);
    f.write(reinterpret_cast<const char*>(&meta), sizeof(meta));
    f.write(reinterpret_cast<const char*>(geo.data()), (std::streamsize)(geo.size() * sizeof(kt_geopixel)));
    return f.good();
}

inline bool load_kgeo(const std::string& path, std::vector<kt_geopixel>& geo, KGeoMeta& meta, std::string* err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { if (err) *err = "cannot open " + path; return false; }
    char magic[4] = {0, 0, 0, 0};
    uint32_t version = 0;
    f.read(magic, 4);
    if (!f || std::memcmp(magic, "KGEO", 4) != 0) { if (err) *err = "bad magic in " + path; return false; }
    f.read(reinterpret_cast<char*>(&version), 4);
    if (!f || version != 1) { if (err) *err = "unsupported version " + std::to_string(version); return false; }
    f.read(reinterpret_cast<char*>(&meta), sizeof(meta));
    if (!f || meta.W == 0 || meta.H == 0 || (uint64_t)meta.W * meta.H > (1ull << 31)) {
        if (err) *err = "bad header in " + path;
        return false;
    }
    geo.resize((size_t)meta.W * meta.H);
    f.read(reinterpret_cast<char*>(geo.data()), (std::streamsize)(geo.size() * sizeof(kt_geopixel)));
    if (!f) { if (err) *err = "truncated pixel data in " + path; return false; }
    return true;
}

// ---------------------------------------------------------------- PNG ------
namespace png_detail {
inline void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24)); v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8)); v.push_back((uint8_t)x);
}
inline uint32_t be32(const uint8_t* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}
inline void chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* data, size_t n) {
    put_be32(out, (uint32_t)n);
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    if (n) out.insert(out.end(), data, data + n);
    const uLong crc = crc32(0L, out.data() + start, (uInt)(n + 4));
    put_be32(out, (uint32_t)crc);
}
inline int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}
}  // namespace png_detail

// rgb: W*H*3 bytes, row-major, top row first.
inline bool write_png(const std::string& path, const uint8_t* rgb, int W, int H) {
    using namespace png_detail;
    if (W <= 0 || H <= 0) return false;
    std::vector<uint8_t> raw((size_t)H * (1 + 3 * (size_t)W));
    for (int y = 0; y < H; ++y) {
        uint8_t* row = &raw[(size_t)y * (1 + 3 * (size_t)W)];
        row[0] = 0;
        std::memcpy(row + 1, rgb + (size_t)y * 3 * W, 3 * (size_t)W);
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK) return false;
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)W);
    put_be32(ihdr, (uint32_t)H);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8 bit, RGB, deflate, filter 0, no interlace
    chunk(out, "IHDR", ihdr.data(), ihdr.size());
    chunk(out, "IDAT", z.data(), zn);
    chunk(out, "IEND", nullptr, 0);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0) && ok;
}

inline bool read_file(const std::string& path, std::vector<uint8_t>& bytes) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    bytes.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

// Decodes to RGB8.  Returns false with a message for anything outside the supported subset.
inline bool decode_png(const std::vector<uint8_t>& b, std::vector<uint8_t>& rgb, int& W, int& H, std::string* err) {
    using namespace png_detail;
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (b.size() < 8 || std::memcmp(b.data(), sig, 8) != 0) { if (err) *err = "not a PNG"; return false; }
    size_t p = 8;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte;
    W = H = 0;
    while (p + 12 <= b.size()) {
        const uint32_t n = be32(&b[p]);
        if (p + 12 + (size_t)n > b.size()) break;
        const char* t = reinterpret_cast<const char*>(&b[p + 4]);
        const uint8_t* d = &b[p + 8];
        if (!std::memcmp(t, "IHDR", 4) && n >= 13) {
            W = (int)be32(d); H = (int)be32(d + 4);
            depth = d[8]; ctype = d[9]; interlace = d[12];
        } else if (!std::memcmp(t, "PLTE", 4)) {
            plte.assign(d, d + n);
        } else if (!std::memcmp(t, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + n);
        } else if (!std::memcmp(t, "IEND", 4)) {
            break;
        }
        p += 12 + (size_t)n;
    }
    int ch = 0;
    switch (ctype) {
        case 0: ch = 1; break;
        case 2: ch = 3; break;
        case 3: ch = 1; break;
        case 4: ch = 2; break;
        case 6: ch = 4; break;
        default: break;
    }
    if (W <= 0 || H <= 0 || depth != 8 || ch == 0 || interlace != 0 || (ctype == 3 && plte.size() < 3) ||
        (uint64_t)W * H > (1ull << 28)) {
        if (err) *err = "unsupported PNG (need 8-bit, non-interlaced)";
        return false;
    }
    const size_t stride = (size_t)W * ch;
    std::vector<uint8_t> raw((stride + 1) * H);
    uLongf rn = (uLongf)raw.size();
    if (uncompress(raw.data(), &rn, idat.data(), (uLong)idat.size()) != Z_OK || rn != raw.size()) {
        if (err) *err = "corrupt PNG image data";
        return false;
    }
    std::vector<uint8_t> px(stride * H), prev(stride, 0);
    for (int y = 0; y < H; ++y) {
        const uint8_t f = raw[y * (stride + 1)];
        const uint8_t* s = &raw[y * (stride + 1) + 1];
        uint8_t* o = &px[y * stride];
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)ch ? o[i - ch] : 0, up = prev[i], c = i >= (size_t)ch ? prev[i - ch] : 0;
            int v = s[i];
            switch (f) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += up; break;
                case 3: v += (a + up) >> 1; break;
                case 4: v += paeth(a, up, c); break;
                default: if (err) *err = "bad PNG filter"; return false;
            }
            o[i] = (uint8_t)v;
        }
        std::memcpy(prev.data(), o, stride);
    }
    rgb.resize((size_t)W * H * 3);
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const uint8_t* s = &px[i * ch];
        uint8_t* o = &rgb[i * 3];
        if (ctype == 3) {
            const size_t k = (size_t)s[0] * 3;
            if (k + 3 > plte.size()) { o[0] = o[1] = o[2] = 0; continue; }
            o[0] = plte[k]; o[1] = plte[k + 1]; o[2] = plte[k + 2];
        } else if (ch >= 3) {
            o[0] = s[0]; o[1] = s[1]; o[2] = s[2];
        } else {
            o[0] = o[1] = o[2] = s[0];
        }
    }
    return true;
}

inline bool decode_ppm(const std::vector<uint8_t>& b, std::vector<uint8_t>& rgb, int& W, int& H, std::string* err) {
    size_t p = 2;
    auto token = [&](long& v) -> bool {
        while (p < b.size()) {
            if (b[p] == '#') { while (p < b.size() && b[p] != '\n') ++p; }
            else if (b[p] == ' ' || b[p] == '\n' || b[p] == '\r' || b[p] == '\t') ++p;
            else break;
        }
        if (p >= b.size() || b[p] < '0' || b[p] > '9') return false;
        v = 0;
        while (p < b.size() && b[p] >= '0' && b[p] <= '9' && v < (1L << 30)) v = v * 10 + (b[p++] - '0');
        return true;
    };
    long w = 0, h = 0, mx = 0;
    // width and height are each below 2^30 after token(); capped at 65536 before the product
    if (b.size() < 2 || b[0] != 'P' || b[1] != '6' || !token(w) || !token(h) || !token(mx) || mx != 255 || w <= 0 ||
        h <= 0 || w > 65536 || h > 65536 || w * h > (1L << 28)) {
        if (err) *err = "unsupported PPM (need binary P6, maxval 255)";
        return false;
    }
    ++p;  // single whitespace after maxval
    if (b.size() < p + (size_t)w * h * 3) { if (err) *err = "truncated PPM"; return false; }
    W = (int)w; H = (int)h;
    rgb.assign(b.begin() + p, b.begin() + p + (size_t)w * h * 3);
    return true;
}

// Background loader: PNG or binary PPM, chosen by the file's magic bytes.
inline bool load_rgb_image(const std::string& path, std::vector<uint8_t>& rgb, int& W, int& H, std::string* err) {
    std::vector<uint8_t> b;
    if (!read_file(path, b)) { if (err) *err = "cannot open " + path; return false; }
    if (b.size() >= 8 && b[0] == 0x89 && b[1] == 'P') return decode_png(b, rgb, W, H, err);
    if (b.size() >= 2 && b[0] == 'P' && b[1] == '6') return decode_ppm(b, rgb, W, H, err);
    if (err) *err = "unsupported image format (PNG or binary PPM)";
    return false;
}

// Packed ABGR (R in bits 0-7) -> RGB8.
inline void abgr_to_rgb(const uint32_t* px, size_t n, uint8_t* rgb) {
    for (size_t i = 0; i < n; ++i) {
        rgb[3 * i + 0] = (uint8_t)(px[i] & 0xFF);
        rgb[3 * i + 1] = (uint8_t)((px[i] >> 8) & 0xFF);
        rgb[3 * i + 2] = (uint8_t)((px[i] >> 16) & 0xFF);
    }
}

}  // namespace kthip
