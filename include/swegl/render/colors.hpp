// swegl/render/colors.hpp -- pixel_colors (reference: swegl/render/colors.hpp):
// b, g, r, a bytes, one 32-bit word in memory.
#pragma once

namespace swegl
{

struct pixel_colors {
    union {
        struct {
            unsigned char b, g, r, a;
        } o;
        int i;
    };
    pixel_colors() = default;
    pixel_colors(unsigned char b, unsigned char g, unsigned char r, unsigned char a) : o{b, g, r, a} {}
    pixel_colors(int v) : i(v) {}
    int to_int() const { return i; }
};

}  // namespace swegl
