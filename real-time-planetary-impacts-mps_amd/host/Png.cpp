/*
 * Png.cpp -- the snapshot image of Compute::drawSnapshot (Compute.cpp:416-446):
 * material ids coloured (402 grey, 400 orange, anything else white) and
 * written as an 8-bit RGBA PNG (the reference uses lodepng; this writes the
 * chunks itself and deflates with zlib).
 */
#include <zlib.h>

#include <cstdint>
#include <cstring>
#include <exception>
#include <fstream>
#include <string>
#include <vector>

#include "Compute.hpp"

namespace {

void put32(std::vector<uint8_t>& b, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
}

void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put32(out, (uint32_t)data.size());
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    const uint32_t crc = (uint32_t)crc32(0L, out.data() + at, (uInt)(4 + data.size()));
    put32(out, crc);
}

}  // namespace

bool writePNG(const std::string& path, const uint8_t* rgba, int w, int h, std::string& err) {
    if (w < 1 || h < 1) {
        err = "empty image";
        return false;
    }
    /* scanlines, each with filter type 0 */
    const size_t row = 4 * (size_t)w;
    std::vector<uint8_t> raw((row + 1) * (size_t)h);
    for (int y = 0; y < h; ++y) {
        raw[(row + 1) * y] = 0;
        std::memcpy(&raw[(row + 1) * y + 1], rgba + row * y, row);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) {
        err = "zlib compress2 failed";
        return false;
    }
    z.resize(zlen);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0}); /* 8-bit RGBA, deflate, no filter choice, no interlace */
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f || !f.write((const char*)out.data(), (std::streamsize)out.size())) {
        err = "cannot write " + path;
        return false;
    }
    return true;
}

bool writeSnapshotPNG(const std::string& path, const int32_t* materials, int res, std::string& err) {
    std::vector<uint8_t> rgba(4 * (size_t)res * res);
    for (size_t i = 0; i < (size_t)res * res; ++i) {
        uint8_t* p = &rgba[4 * i];
        if (materials[i] == 402) {
            p[0] = 160, p[1] = 160, p[2] = 160;
        } else if (materials[i] == 400) {
            p[0] = 255, p[1] = 100, p[2] = 0;
        } else {
            p[0] = 255, p[1] = 255, p[2] = 255;
        }
        p[3] = 255;
    }
    return writePNG(path, rgba.data(), res, res, err);
}

extern "C" {

int sph_write_snapshot_png(const int32_t* materials, int res, const char* path) {
    if (!materials || !path || res < 1) return SPH_E_ARG;
    try {
        std::string err;
        if (!writeSnapshotPNG(path, materials, res, err)) {
            setHdfError(err);
            return SPH_E_ARG;
        }
        return SPH_OK;
    } catch (const std::exception& e) {
        setHdfError(e.what());
        return SPH_E_CAPACITY;
    }
}

}  // extern "C"
