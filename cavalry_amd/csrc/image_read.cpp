// Texture image files for the scene front end: PNG and TGA decoded to RGBA8
// the way Texture2D::createTextureFromFile (Core/Texture.cu:30-85) receives
// them from its image loader: four channels (grey replicated, alpha 255 when
// the file has none), 16-bit samples reduced to their high byte, rows flipped
// so that row 0 is the bottom of the image (texture coordinate v = 0), and
// for .png / .tga files the sRGB curve removed from r, g, b:
// (uint8)(255 * inverseGammaCorrect(x / 255)) (Core/Color.h:61-65).
//
// PNG: 8- and 16-bit grey, grey + alpha, RGB, RGBA and 8-bit palette images
// (with tRNS transparency), non-interlaced; the zlib stream is inflated with
// the system zlib. TGA: uncompressed and run-length true-colour (24 / 32 bit)
// and grey (8 bit) images, either row origin.
#include "image_read.h"

#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace cav {
namespace {

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// Paeth predictor of the PNG specification (filter type 4)
int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

bool decode_png(const std::string& data, Image& img, std::string& err) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    const uint8_t* d = (const uint8_t*)data.data();
    const size_t n = data.size();
    if (n < 8 || std::memcmp(d, sig, 8) != 0) return err = "not a PNG file", false;
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    bool ended = false;
    while (pos + 12 <= n && !ended) {
        const uint32_t len = be32(d + pos);
        if (len > n - pos - 12) return err = "truncated PNG chunk", false;
        const std::string type((const char*)d + pos + 4, 4);
        const uint8_t* body = d + pos + 8;
        if (type == "IHDR") {
            if (len < 13) return err = "short IHDR", false;
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
        } else if (type == "PLTE") {
            plte.assign(body, body + len);
        } else if (type == "tRNS") {
            trns.assign(body, body + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), body, body + len);
        } else if (type == "IEND") {
            ended = true;
        }
        pos += 12 + (size_t)len;
    }
    if (ctype < 0) return err = "PNG without IHDR", false;
    if (w < 1 || h < 1 || w > 65536 || h > 65536 || (uint64_t)w * h > (1ull << 28))
        return err = "PNG size out of range", false;
    if (interlace != 0) return err = "interlaced PNG is not supported", false;
    int chans;
    switch (ctype) {
        case 0: chans = 1; break;
        case 2: chans = 3; break;
        case 3: chans = 1; break;
        case 4: chans = 2; break;
        case 6: chans = 4; break;
        default: return err = "unknown PNG colour type", false;
    }
    if (!(depth == 8 || (depth == 16 && ctype != 3))) return err = "PNG bit depth " + std::to_string(depth) + " is not supported", false;
    if (ctype == 3 && plte.size() < 3) return err = "palette PNG without PLTE", false;
    const int bpp = chans * depth / 8;  // bytes per pixel
    const size_t stride = (size_t)w * bpp;
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf out_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &out_len, idat.data(), (uLong)idat.size()) != Z_OK || out_len != raw.size())
        return err = "corrupt PNG image data", false;
    // undo the per-row filters in place
    std::vector<uint8_t> px(stride * h);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t* src = raw.data() + y * (stride + 1) + 1;
        uint8_t* row = px.data() + y * stride;
        const uint8_t* up = y ? row - stride : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)bpp ? row[i - bpp] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= (size_t)bpp) ? up[i - bpp] : 0;
            int v;
            switch (ft) {
                case 0: v = src[i]; break;
                case 1: v = src[i] + a; break;
                case 2: v = src[i] + b; break;
                case 3: v = src[i] + ((a + b) >> 1); break;
                case 4: v = src[i] + paeth(a, b, c); break;
                default: return err = "unknown PNG filter type", false;
            }
            row[i] = (uint8_t)v;
        }
    }
    img.w = (int)w;
    img.h = (int)h;
    img.rgba.assign((size_t)w * h * 4, 0);
    const int step = depth / 8;  // the high byte of a 16-bit sample comes first
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const uint8_t* s = px.data() + y * stride + (size_t)x * bpp;
            uint8_t* o = img.rgba.data() + ((size_t)y * w + x) * 4;
            switch (ctype) {
                case 0: o[0] = o[1] = o[2] = s[0]; o[3] = 255; break;
                case 4: o[0] = o[1] = o[2] = s[0]; o[3] = s[step]; break;
                case 2: o[0] = s[0]; o[1] = s[step]; o[2] = s[2 * step]; o[3] = 255; break;
                case 6: o[0] = s[0]; o[1] = s[step]; o[2] = s[2 * step]; o[3] = s[3 * step]; break;
                case 3: {
                    const size_t k = s[0];
                    if (3 * k + 2 >= plte.size()) return err = "PNG palette index out of range", false;
                    o[0] = plte[3 * k]; o[1] = plte[3 * k + 1]; o[2] = plte[3 * k + 2];
                    o[3] = k < trns.size() ? trns[k] : 255;
                    break;
                }
            }
        }
    return true;
}

bool decode_tga(const std::string& data, Image& img, std::string& err) {
    const uint8_t* d = (const uint8_t*)data.data();
    const size_t n = data.size();
    if (n < 18) return err = "truncated TGA header", false;
    const int idlen = d[0], cmap = d[1], type = d[2];
    const int w = d[12] | d[13] << 8, h = d[14] | d[15] << 8, bits = d[16], desc = d[17];
    if (cmap != 0) return err = "colour-mapped TGA is not supported", false;
    const bool rle = type == 10 || type == 11;
    const bool grey = type == 3 || type == 11;
    if (!(type == 2 || type == 3 || type == 10 || type == 11)) return err = "TGA image type " + std::to_string(type) + " is not supported", false;
    if (grey ? bits != 8 : (bits != 24 && bits != 32)) return err = "TGA pixel depth " + std::to_string(bits) + " is not supported", false;
    if (w < 1 || h < 1) return err = "TGA size out of range", false;
    const int bpp = bits / 8;
    size_t pos = 18 + (size_t)idlen;
    std::vector<uint8_t> px((size_t)w * h * bpp);
    if (!rle) {
        if (pos > n || px.size() > n - pos) return err = "truncated TGA data", false;
        std::memcpy(px.data(), d + pos, px.size());
    } else {
        size_t k = 0;
        while (k < px.size()) {
            if (pos >= n) return err = "truncated TGA data", false;
            const uint8_t hdr = d[pos++];
            const size_t bytes = (size_t)((hdr & 127) + 1) * bpp;
            if (bytes > px.size() - k) return err = "TGA run past the image", false;
            if (hdr & 128) {  // one pixel repeated
                if (pos + bpp > n) return err = "truncated TGA data", false;
                for (size_t r = 0; r < bytes; r += bpp) std::memcpy(px.data() + k + r, d + pos, bpp);
                pos += bpp;
            } else {          // literal pixels
                if (pos + bytes > n) return err = "truncated TGA data", false;
                std::memcpy(px.data() + k, d + pos, bytes);
                pos += bytes;
            }
            k += bytes;
        }
    }
    // rows in file order run bottom-up unless bit 5 of the descriptor is set
    const bool top_first = (desc & 32) != 0;
    img.w = w;
    img.h = h;
    img.rgba.assign((size_t)w * h * 4, 0);
    for (int fy = 0; fy < h; ++fy) {
        const int y = top_first ? fy : h - 1 - fy;  // row from the top
        for (int x = 0; x < w; ++x) {
            const uint8_t* s = px.data() + ((size_t)fy * w + x) * bpp;
            uint8_t* o = img.rgba.data() + ((size_t)y * w + x) * 4;
            if (grey) {
                o[0] = o[1] = o[2] = s[0];
                o[3] = 255;
            } else {  // stored as b, g, r (, a)
                o[0] = s[2]; o[1] = s[1]; o[2] = s[0];
                o[3] = bpp == 4 ? s[3] : 255;
            }
        }
    }
    return true;
}

std::string lower_ext(const std::string& path) {
    const size_t dot = path.find_last_of('.');
    const size_t slash = path.find_last_of('/');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    std::string e = path.substr(dot + 1);
    for (char& c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

// Core/Color.h:61-65
float inverse_gamma(float value) {
    if (value <= 0.04045f) return value * 1.f / 12.92f;
    return std::pow((value + 0.055f) * 1.f / 1.055f, 2.4f);
}

}  // namespace

bool read_texture_file(const std::string& path, Image& img, std::string& err) {
    const std::string ext = lower_ext(path);
    if (ext != "png" && ext != "tga")
        return err = "texture format ." + ext + " is not supported (PNG and TGA are): " + path, false;
    std::ifstream f(path, std::ios::binary);
    if (!f) return err = "cannot open texture " + path, false;
    std::ostringstream ss;
    ss << f.rdbuf();
    const std::string data = ss.str();
    const bool ok = ext == "png" ? decode_png(data, img, err) : decode_tga(data, img, err);
    if (!ok) {
        err += ": " + path;
        return false;
    }
    // rows bottom-up: row 0 at v = 0
    const size_t row = (size_t)img.w * 4;
    for (int y = 0; y < img.h / 2; ++y)
        std::swap_ranges(img.rgba.begin() + (size_t)y * row, img.rgba.begin() + (size_t)(y + 1) * row,
                         img.rgba.begin() + (size_t)(img.h - 1 - y) * row);
    uint8_t lut[256];
    for (int v = 0; v < 256; ++v) lut[v] = (uint8_t)(255.f * inverse_gamma((float)v / 255.f));
    for (size_t i = 0; i < img.rgba.size(); i += 4)
        for (int c = 0; c < 3; ++c) img.rgba[i + c] = lut[img.rgba[i + c]];
    return true;
}

}  // namespace cav
