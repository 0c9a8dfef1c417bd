// swegl/render/viewport.hpp -- viewport_t and shader plugin types of the
// swegl API (reference: swegl/render/viewport.hpp, pixel_shaders.hpp,
// post_shaders.hpp).  Shader objects are type tags naming the GPU kernel
// variant that computes them: every pixel shader type of the reference and
// every pixel_shader_light_and_texture<L, T> of its light and texture types,
// post_shader_t (no-op without transparency) and the depth-of-field post
// shaders.
#pragma once

#include <memory>
#include <vector>

#include "../data/model.hpp"
#include "../projection/camera.hpp"
#include "surface.hpp"

namespace swegl
{

// pixel_shader_t shades with the material colour (pixel_shaders.hpp:15-29)
struct pixel_shader_t {
    virtual ~pixel_shader_t() = default;
    virtual int gpu_kind() const { return SWEGL_SHADER_DEFAULT; }
};
struct pixel_shader_lights_phong : pixel_shader_t {
    int gpu_kind() const override { return SWEGL_SHADER_LIGHTS_PHONG; }
};
struct pixel_shader_lights_flat : pixel_shader_t {
    int gpu_kind() const override { return SWEGL_SHADER_LIGHTS_FLAT; }
};
struct pixel_shader_texture_bilinear : pixel_shader_t {
    int gpu_kind() const override { return SWEGL_SHADER_TEXTURE_BILINEAR; }
};
struct pixel_shader_texture : pixel_shader_t {
    int gpu_kind() const override { return SWEGL_SHADER_TEXTURE_NEAREST; }
};

namespace detail
{
template <typename L, typename T>
struct combined_kind;   // only light x texture combinations exist
template <>
struct combined_kind<pixel_shader_lights_phong, pixel_shader_texture_bilinear> {
    static constexpr int value = SWEGL_SHADER_PHONG_BILINEAR;
};
template <>
struct combined_kind<pixel_shader_lights_phong, pixel_shader_texture> {
    static constexpr int value = SWEGL_SHADER_PHONG_NEAREST;
};
template <>
struct combined_kind<pixel_shader_lights_flat, pixel_shader_texture_bilinear> {
    static constexpr int value = SWEGL_SHADER_FLAT_BILINEAR;
};
template <>
struct combined_kind<pixel_shader_lights_flat, pixel_shader_texture> {
    static constexpr int value = SWEGL_SHADER_FLAT_NEAREST;
};
}  // namespace detail

template <typename L, typename T>
struct pixel_shader_light_and_texture : pixel_shader_t {
    int gpu_kind() const override { return detail::combined_kind<L, T>::value; }
};

struct viewport_t;

struct post_shader_t {
    virtual ~post_shader_t() = default;
    virtual int gpu_kind() const { return SWEGL_POST_NULL; }
    float focal_distance = 0, focal_depth = 0;
    int max_radius = 0;
};

// post_shader_depth_box(dist, depth, vp) (post_shaders.hpp:51-61): blur
// radius up to 5, box taps.
struct post_shader_depth_box : post_shader_t {
    post_shader_depth_box(float dist, float depth, viewport_t &)
    {
        focal_distance = dist;
        focal_depth = depth;
        max_radius = 5;
    }
    int gpu_kind() const override { return SWEGL_POST_DEPTH_BOX; }
};

// Same blur factor, 3x3 taps at {-r, 0, r} and a configurable maximum radius.
struct post_shader_depth_tap9 : post_shader_t {
    post_shader_depth_tap9(float dist, float depth, viewport_t &, int r_max = 8)
    {
        focal_distance = dist;
        focal_depth = depth;
        max_radius = r_max;
    }
    int gpu_kind() const override { return SWEGL_POST_DEPTH_TAP9; }
};

struct viewport_t {
    int m_x, m_y;
    int m_w, m_h;
    SDL_Surface *m_screen;
    std::unique_ptr<float[]> m_zbuffer;
    camera_t m_camera;
    std::shared_ptr<pixel_shader_t> m_pixel_shader;
    post_shader_t *m_post_shader = nullptr;
    int m_transparency_layer_count;
    bool m_got_transparency;
    swegl_target *m_device_depth = nullptr;   // device mirror of m_zbuffer

    viewport_t(int x, int y, int w, int h, SDL_Surface *screen, std::shared_ptr<pixel_shader_t> &pixel_shader,
               int transparency_layer_count)
        : m_x(x), m_y(y), m_w(w), m_h(h), m_screen(screen), m_zbuffer(new float[(size_t)w * h]),
          m_camera((float)(1.0 * w / h)), m_pixel_shader(pixel_shader),
          m_transparency_layer_count(transparency_layer_count),
          m_got_transparency(transparency_layer_count > 0)
    {
    }
    ~viewport_t();
    viewport_t(const viewport_t &) = delete;
    viewport_t &operator=(const viewport_t &) = delete;

    void set_post_shader(post_shader_t &p) { m_post_shader = &p; }
    float *zbuffer() { return m_zbuffer.get(); }
    const float *zbuffer() const { return m_zbuffer.get(); }
    camera_t &camera() { return m_camera; }
    const camera_t &camera() const { return m_camera; }
};

}  // namespace swegl
