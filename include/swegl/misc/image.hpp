// swegl/misc/image.hpp -- read_image_file (reference: swegl/misc/image.hpp,
// src/misc/image.cpp:20-28) for .png and .bmp files over the library's
// decoders; JPEG files raise (no JPEG decoder in this build).
#pragma once

#include <stdexcept>
#include <string>

#include "../../swegl_hip.h"
#include "../data/model.hpp"

namespace swegl
{

inline texture_t read_image_file(const std::string &filename)
{
    uint32_t *px = nullptr, w = 0, h = 0;
    if (swegl_image_read(filename.c_str(), &px, &w, &h) != SWEGL_OK)
        throw std::runtime_error("swegl::read_image_file(" + filename + "): " + swegl_last_error());
    texture_t t(px, (int)w, (int)h);
    swegl_image_free(px);
    return t;
}

}  // namespace swegl
