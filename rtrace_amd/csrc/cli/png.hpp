// png.hpp — minimal PNG writer (8-bit RGB, one IDAT, zlib) for the CLI and the Node addon.
#pragma once

#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace rtpng {

inline void put32(std::vector<unsigned char>& v, uint32_t x) {
    v.push_back(x >> 24);
    v.push_back(x >> 16);
    v.push_back(x >> 8);
    v.push_back(x);
}
inline void chunk(std::vector<unsigned char>& png, const char* type, const std::vector<unsigned char>& data) {
    put32(png, (uint32_t)data.size());
    size_t start = png.size();
    png.insert(png.end(), type, type + 4);
    png.insert(png.end(), data.begin(), data.end());
    uint32_t crc = crc32(0, png.data() + start, (uInt)(png.size() - start));
    put32(png, crc);
}
inline bool write_png(const char* path, const uint8_t* rgb, int w, int h) {
    std::vector<unsigned char> raw((size_t)(w * 3 + 1) * h);
    for (int y = 0; y < h; ++y) {
        raw[(size_t)y * (w * 3 + 1)] = 0;
        memcpy(&raw[(size_t)y * (w * 3 + 1) + 1], rgb + (size_t)y * w * 3, (size_t)w * 3);
    }
    uLongf zlen = compressBound(raw.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), raw.size(), 6) != Z_OK) return false;
    z.resize(zlen);
    std::vector<unsigned char> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    bool ok = fwrite(png.data(), 1, png.size(), f) == png.size();
    return fclose(f) == 0 && ok;
}
}  // namespace rtpng
