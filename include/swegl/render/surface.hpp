// swegl/render/surface.hpp -- the render target.  With SDL2 available this
// is SDL's own SDL_Surface; otherwise a layout-compatible stand-in with the
// fields the renderer reads (w, h, pitch, pixels, format->BytesPerPixel).
#pragma once

#if defined(SWEGL_USE_SDL) && __has_include(<SDL.h>)
#include <SDL.h>
#else
#include <cstdint>
#include <vector>

struct SDL_PixelFormat {
    std::uint8_t BytesPerPixel = 4;
};

struct SDL_Surface {
    int w = 0, h = 0;
    int pitch = 0;
    void *pixels = nullptr;
    SDL_PixelFormat *format = nullptr;
};

namespace swegl
{
// An owning 32-bit surface (what sdl_t::make_surface provides in the demo).
struct owned_surface {
    SDL_PixelFormat fmt;
    std::vector<std::uint32_t> storage;
    SDL_Surface surface;
    owned_surface(int w, int h) : storage((size_t)w * h, 0u)
    {
        surface.w = w;
        surface.h = h;
        surface.pitch = 4 * w;
        surface.pixels = storage.data();
        surface.format = &fmt;
    }
};
}  // namespace swegl
#endif
