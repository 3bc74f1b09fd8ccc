// Texture image files (PNG, TGA) decoded to bottom-up RGBA8; see image_read.cpp.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace cav {

struct Image {
    int w = 0, h = 0;
    std::vector<uint8_t> rgba;  // w * h * 4, row 0 = bottom of the image
};

// false with a message in err when the file cannot be read or decoded
bool read_texture_file(const std::string& path, Image& img, std::string& err);

}  // namespace cav
