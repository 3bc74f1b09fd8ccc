// Film output on the host: the resolve to 8-bit sRGB, the PNG writer and an
// FP32 PFM dump (cav_film_to_srgb8, cav_write_png, cav_write_pfm).
//
// Reference: Film::readCurrentResult + apply (Core/Film.cu:20-46),
// clampBetween0And1 / gammaCorrect / writeColorAt (Core/Color.h:53-81),
// RenderResult::saveToPNG (Core/RenderResult.cu:1-12, svpng). The reference
// writes only 8-bit PNG; the PFM dump keeps the film's float values.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/cavalry.h"

namespace {

// Core/Color.h:57-61
float gamma_correct(float v) {
    if (v <= 0.0031308f) return 12.92f * v;
    return 1.055f * std::pow(v, (float)(1.f / 2.4f)) - 0.055f;
}

// CUDA min/max on floats: a NaN operand is ignored
float clamp01(float v) { return std::fmax(0.f, std::fmin(1.f, v)); }

// CRC-32 (ISO 3309 polynomial, reflected) as PNG chunks use it
struct CrcTable {
    uint32_t t[256];
    CrcTable() {
        for (uint32_t n = 0; n < 256; ++n) {
            uint32_t c = n;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
            t[n] = c;
        }
    }
};

uint32_t crc_update(uint32_t c, const uint8_t* p, size_t n) {
    static const CrcTable tab;  // initialised once, thread-safe
    for (size_t i = 0; i < n; ++i) c = tab.t[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put_be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    const uint32_t crc = crc_update(0xffffffffu, out.data() + start, out.size() - start) ^ 0xffffffffu;
    put_be32(out, crc);
}

}  // namespace

extern "C" {

int cav_film_to_srgb8(const float* rgb, const float* weight, int width, int height, uint8_t* out) {
    if (!rgb || !weight || !out || width < 1 || height < 1) return CAV_ERR_INVALID;
    const int64_t n = (int64_t)width * height;
    for (int64_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            const float v = clamp01(rgb[i * 3 + c] / weight[i]);
            out[i * 3 + c] = (uint8_t)(gamma_correct(v) * 255);
        }
    }
    return CAV_OK;
}

int cav_write_png(const char* path, const uint8_t* rgb8, int width, int height) {
    if (!path || !rgb8 || width < 1 || height < 1) return CAV_ERR_INVALID;
    // zlib stream of stored (uncompressed) deflate blocks; every row starts
    // with filter type 0
    const size_t row = (size_t)width * 3 + 1;
    std::vector<uint8_t> raw(row * height);
    for (int y = 0; y < height; ++y) {
        raw[y * row] = 0;
        for (size_t k = 0; k < (size_t)width * 3; ++k) raw[y * row + 1 + k] = rgb8[(size_t)y * width * 3 + k];
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    size_t pos = 0;
    do {
        const size_t len = raw.size() - pos < 65535 ? raw.size() - pos : 65535;
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(len & 0xff));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xff));
        z.push_back((uint8_t)((~len >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, no filter / interlace
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    FILE* f = std::fopen(path, "wb");
    if (!f) return CAV_ERR_INVALID;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? CAV_OK : CAV_ERR_INVALID;
}

int cav_write_pfm(const char* path, const float* rgb, const float* weight, int width, int height) {
    if (!path || !rgb || width < 1 || height < 1) return CAV_ERR_INVALID;
    FILE* f = std::fopen(path, "wb");
    if (!f) return CAV_ERR_INVALID;
    // little-endian colour PFM; rows run bottom to top
    std::fprintf(f, "PF\n%d %d\n-1.0\n", width, height);
    std::vector<float> line((size_t)width * 3);
    bool ok = true;
    for (int y = height - 1; y >= 0 && ok; --y) {
        for (int x = 0; x < width; ++x) {
            const int64_t i = (int64_t)y * width + x;
            const float w = weight ? weight[i] : 1.f;
            for (int c = 0; c < 3; ++c) line[(size_t)x * 3 + c] = rgb[i * 3 + c] / w;
        }
        ok = std::fwrite(line.data(), sizeof(float), line.size(), f) == line.size();
    }
    return (std::fclose(f) == 0 && ok) ? CAV_OK : CAV_ERR_INVALID;
}

}  // extern "C"
