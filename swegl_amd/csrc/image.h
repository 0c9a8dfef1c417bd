// image.h -- texture decoders of the asset loader (reference:
// src/misc/image.cpp).  Every decoder returns row-major width*height
// pixel_colors words (b | g << 8 | r << 16 | a << 24), top row first.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace sw_img {

// PNG as libpng reads it with the reference's transformations
// (image.cpp:93-173): 16-bit samples stripped to their high byte, palettes
// expanded, gray below 8 bits scaled to 8, tRNS turned into alpha, a 0xFF
// alpha filler for formats without alpha, gray copied to r, g and b; then r
// and b swapped into pixel_colors order.  Interlaced (Adam7) images are
// de-interlaced as png_read_image does.
bool decode_png(const uint8_t *data, size_t n, std::vector<uint32_t> &out, int &w, int &h,
                std::string &err);

// 24-bit BMP as the reference reads it (image.cpp:30-83): the header fields
// at their fixed offsets, pixel data right after the 54-byte header (the
// header's data offset and bit depth are not consulted), rows stored bottom
// up, (width * 3) % 4 bytes skipped after each row, alpha 255.
bool decode_bmp(const uint8_t *data, size_t n, std::vector<uint32_t> &out, int &w, int &h,
                std::string &err);

}  // namespace sw_img
