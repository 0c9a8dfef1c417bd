// swegl/data/model.hpp -- scene description of the swegl API (reference:
// swegl/data/model.hpp, swegl/data/texture.hpp), kept in the reference's
// array-of-structures form.  The GPU render does not write the per-vertex
// scratch fields (v_world, v_viewport, normal_world, yes): they stay for
// layout compatibility only.
#pragma once

#include <atomic>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../swegl_hip.h"

#include "../projection/points.hpp"
#include "../render/colors.hpp"

namespace swegl
{

using vertex_idx = std::uint32_t;

struct mesh_vertex_t {
    vertex_t v;
    vertex_t v_world;
    vertex_t v_viewport;
    vec2f_t tex_coords;
    normal_t normal;
    normal_t normal_world;
    bool yes = false;
};

struct primitive_t {
    enum index_mode_t {
        POINTS = 0,
        LINES = 1,
        LINE_LOOP = 2,
        LINE_STRIP = 3,
        TRIANGLES = 4,
        TRIANGLE_STRIP = 5,
        TRIANGLE_FAN = 6,
    };
    std::vector<mesh_vertex_t> vertices;
    std::vector<vertex_idx> indices;
    index_mode_t mode;
    int material_id;
};

struct node_t {
    vertex_t scale = vertex_t(1, 1, 1);
    matrix44_t rotation = matrix44_t::identity();
    vertex_t translation = vertex_t(0, 0, 0);
    std::vector<primitive_t> primitives;
    std::vector<int> children_idx;
    bool root = true;
};

struct point_source_light {
    vertex_t position;
    float intensity;
};

struct material_t {
    pixel_colors color = pixel_colors(255, 255, 255, 255);
    float metallic = 1;
    float roughness = 1;
    int texture_idx = -1;
    bool double_sided = false;
};

// mip 0 of a texture_t (texture.hpp:12-38)
struct mipmap_t {
    std::vector<unsigned int> m_bitmap;
    unsigned int m_width = 0, m_height = 0;
};

class texture_t
{
public:
    std::vector<std::shared_ptr<mipmap_t>> m_mipmaps;
    texture_t(const unsigned int *data, int w, int h)
    {
        auto m = std::make_shared<mipmap_t>();
        m->m_bitmap.assign(data, data + (size_t)w * h);
        m->m_width = (unsigned)w;
        m->m_height = (unsigned)h;
        m_mipmaps.push_back(m);
    }
};

// animation_step_t, animation_channel_t, animation_t (model.hpp:89-127)
struct animation_step_t {
    float time;
    vec4f_t value;
};
struct animation_channel_t {
    enum path_t {
        SCALE = SWEGL_PATH_SCALE,
        ROTATION = SWEGL_PATH_ROTATION,
        TRANSLATION = SWEGL_PATH_TRANSLATION,
        WEIGHTS = SWEGL_PATH_NONE,
        NONE = SWEGL_PATH_NONE,
    };
    int node_idx;
    path_t path;
    std::vector<animation_step_t> steps;   // sorted by time
};
struct animation_t {
    float end_time = 0.0f;
    std::vector<animation_channel_t> channels;
};

namespace detail
{
// Each scene_t object gets an id of its own (constructed, copied or moved
// into), so that device copies are never matched to another scene that
// happens to reuse an address; the renderer installs the hook that frees a
// scene's device copy when the scene object goes away.
inline std::uint64_t next_scene_id()
{
    static std::atomic<std::uint64_t> n{0};
    return ++n;
}
struct scene_hook {
    void (*fn)(void *, std::uint64_t) = nullptr;   // trivially destructible: safe at exit
    void *arg = nullptr;
};
inline scene_hook &scene_gone_hook()
{
    static scene_hook h;
    return h;
}
}  // namespace detail

struct scene_t {
    scene_t() = default;
    scene_t(const scene_t &o) = default;
    scene_t(scene_t &&o) = default;
    scene_t &operator=(const scene_t &o) = default;
    scene_t &operator=(scene_t &&o) = default;
    ~scene_t()
    {
        const detail::scene_hook h = detail::scene_gone_hook();
        if (h.fn)
            h.fn(h.arg, device_id.v);
    }

    std::vector<node_t> nodes;
    std::vector<int> root_nodes;
    float ambient_light_intensity = 0;
    normal_t sun_direction = normal_t(0, -1, 0);
    float sun_intensity = 0;
    material_t default_material;
    std::vector<point_source_light> point_source_lights;
    std::vector<material_t> materials;
    std::vector<texture_t> images;
    std::vector<animation_t> animations;
    // Bump after editing vertex, index, texture or material values in
    // place: render() re-uploads meshes and textures when this changes, or
    // when an array was resized, replaced or re-allocated (noticed by its
    // address and size); node poses and lights are sent on every call.
    std::uint64_t device_generation = 0;
    // Identity of this scene object for the renderer's device copy: a fresh
    // id on construction and on every copy or move into it.
    struct object_id {
        std::uint64_t v = detail::next_scene_id();
        object_id() = default;
        object_id(const object_id &) : v(detail::next_scene_id()) {}
        object_id &operator=(const object_id &)
        {
            v = detail::next_scene_id();
            return *this;
        }
    } device_id;

    // scene_t::animate (model.hpp:146-177): poses the animated nodes; the
    // next render picks the poses up.
    void animate(const float elapsed_seconds)
    {
        if (animations.empty())
            return;
        std::vector<swegl_node_desc> nd(nodes.size());
        for (size_t i = 0; i < nodes.size(); i++) {
            const node_t &n = nodes[i];
            nd[i].scale[0] = n.scale.x();
            nd[i].scale[1] = n.scale.y();
            nd[i].scale[2] = n.scale.z();
            std::memcpy(nd[i].rotation, &n.rotation.m[0][0], sizeof nd[i].rotation);
            nd[i].translation[0] = n.translation.x();
            nd[i].translation[1] = n.translation.y();
            nd[i].translation[2] = n.translation.z();
        }
        std::vector<std::vector<swegl_anim_step>> steps;
        std::vector<std::vector<swegl_anim_channel>> chans(animations.size());
        std::vector<swegl_animation> anims(animations.size());
        for (size_t a = 0; a < animations.size(); a++) {
            for (const animation_channel_t &c : animations[a].channels) {
                std::vector<swegl_anim_step> st;
                for (const animation_step_t &s : c.steps)
                    st.push_back({s.time, {s.value.x(), s.value.y(), s.value.z(), s.value.w()}});
                steps.push_back(std::move(st));
                chans[a].push_back({c.node_idx, (int32_t)c.path, (uint32_t)steps.back().size(),
                                    steps.back().data()});
            }
            anims[a] = {animations[a].end_time, (uint32_t)chans[a].size(), chans[a].data()};
        }
        if (swegl_animate(anims.data(), (uint32_t)anims.size(), elapsed_seconds, nd.data(),
                          (uint32_t)nd.size()) != SWEGL_OK)
            throw std::runtime_error(std::string("swegl::scene_t::animate: ") + swegl_last_error());
        for (size_t i = 0; i < nodes.size(); i++) {
            node_t &n = nodes[i];
            n.scale = vertex_t(nd[i].scale[0], nd[i].scale[1], nd[i].scale[2]);
            std::memcpy(&n.rotation.m[0][0], nd[i].rotation, sizeof nd[i].rotation);
            n.translation = vertex_t(nd[i].translation[0], nd[i].translation[1], nd[i].translation[2]);
        }
    }
};

// make_cube (model.hpp:254-325): six 4-vertex triangle fans.
inline node_t make_cube(float size, int material_idx)
{
    struct face {
        float p[4][3];
        float t[4][2];
        float n[3];
    };
    const float h = size / 2.0f;
    const face faces[6] = {
        {{{-h, h, h}, {h, h, h}, {h, h, -h}, {-h, h, -h}}, {{0.0, 0.0}, {0.5, 0.0}, {0.5, 0.333}, {0.0, 0.333}}, {0, 1, 0}},
        {{{-h, -h, h}, {h, -h, h}, {h, h, h}, {-h, h, h}}, {{0.5, 0.0}, {1.0, 0.0}, {1.0, 0.333}, {0.5, 0.333}}, {0, 0, 1}},
        {{{h, -h, h}, {h, -h, -h}, {h, h, -h}, {h, h, h}}, {{0.0, 0.333}, {0.5, 0.333}, {0.5, 0.667}, {0.0, 0.667}}, {1, 0, 0}},
        {{{h, -h, -h}, {-h, -h, -h}, {-h, h, -h}, {h, h, -h}}, {{0.5, 0.667}, {1.0, 0.667}, {1.0, 1.0}, {0.5, 1.0}}, {0, 0, -1}},
        {{{-h, -h, -h}, {-h, -h, h}, {-h, h, h}, {-h, h, -h}}, {{0.5, 0.333}, {1.0, 0.333}, {1.0, 0.667}, {0.5, 0.667}}, {-1, 0, 0}},
        {{{-h, -h, -h}, {h, -h, -h}, {h, -h, h}, {-h, -h, h}}, {{0.0, 0.667}, {0.5, 0.667}, {0.5, 1.0}, {0.0, 1.0}}, {0, -1, 0}},
    };
    node_t node;
    for (const face &f : faces) {
        primitive_t p;
        for (int i = 0; i < 4; i++) {
            mesh_vertex_t mv;
            mv.v = vertex_t(f.p[i][0], f.p[i][1], f.p[i][2]);
            mv.tex_coords = vec2f_t(f.t[i][0], f.t[i][1]);
            mv.normal = normal_t(f.n[0], f.n[1], f.n[2]);
            p.vertices.push_back(mv);
        }
        p.indices = {0, 1, 2, 3};
        p.mode = primitive_t::TRIANGLE_FAN;
        p.material_id = material_idx;
        node.primitives.push_back(p);
    }
    return node;
}

// make_tri (model.hpp:233-252): one triangle in the z = 0 plane.
inline node_t make_tri(float size, int material_idx)
{
    node_t node;
    primitive_t p;
    const float pos[3][3] = {{0.0f, 0.0f, 0.0f}, {size, 0.0f, 0.0f}, {0.0f, size, 0.0f}};
    const float tex[3][2] = {{0.0f, 0.0f}, {0.0f, 1.0f}, {1.0f, 0.0f}};
    for (int i = 0; i < 3; i++) {
        mesh_vertex_t mv;
        mv.v = vertex_t(pos[i][0], pos[i][1], pos[i][2]);
        mv.tex_coords = vec2f_t(tex[i][0], tex[i][1]);
        mv.normal = normal_t(0, 0, 1);
        p.vertices.push_back(mv);
    }
    p.indices = {0, 1, 2};
    p.mode = primitive_t::TRIANGLES;
    p.material_id = material_idx;
    node.primitives.push_back(p);
    return node;
}

namespace detail {
// precision + 1 strips from swegl_make_tore / swegl_make_sphere
template <typename F>
inline node_t strips(unsigned int precision, int material_idx, F fill)
{
    const size_t n = precision + 1, V = n * 2 * n;
    std::vector<float> pos(V * 3), nor(V * 3), tex(V * 2);
    std::vector<uint32_t> idx(V);
    if (fill(pos.data(), nor.data(), tex.data(), idx.data()) != SWEGL_OK)
        throw std::runtime_error(std::string("swegl mesh generator: ") + swegl_last_error());
    node_t node;
    for (size_t k = 0; k < n; k++) {
        primitive_t p;
        for (size_t i = 2 * n * k; i < 2 * n * (k + 1); i++) {
            mesh_vertex_t mv;
            mv.v = vertex_t(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
            mv.normal.x() = nor[3 * i];   // already normal_t values
            mv.normal.y() = nor[3 * i + 1];
            mv.normal.z() = nor[3 * i + 2];
            mv.tex_coords = vec2f_t(tex[2 * i], tex[2 * i + 1]);
            p.vertices.push_back(mv);
        }
        p.indices.assign(idx.begin() + 2 * n * k, idx.begin() + 2 * n * (k + 1));
        p.mode = primitive_t::TRIANGLE_STRIP;
        p.material_id = material_idx;
        node.primitives.push_back(std::move(p));
    }
    return node;
}
}  // namespace detail

// make_tore (model.hpp:327-391)
inline node_t make_tore(unsigned int precision, int material_idx)
{
    return detail::strips(precision, material_idx, [&](float *p, float *n, float *t, uint32_t *i) {
        return swegl_make_tore(precision, p, n, t, i);
    });
}

// make_sphere (model.hpp:393-464)
inline node_t make_sphere(unsigned int precision, float radius, int material_idx)
{
    return detail::strips(precision, material_idx, [&](float *p, float *n, float *t, uint32_t *i) {
        return swegl_make_sphere(precision, radius, p, n, t, i);
    });
}

// calculate_face_normals (model.hpp:181-231)
inline void calculate_face_normals(primitive_t &primitive)
{
    std::vector<float> pos, nor;
    for (const mesh_vertex_t &v : primitive.vertices) {
        pos.insert(pos.end(), {v.v.x(), v.v.y(), v.v.z()});
        nor.insert(nor.end(), {v.normal.x(), v.normal.y(), v.normal.z()});
    }
    if (swegl_calculate_face_normals((int32_t)primitive.mode, pos.data(), nor.data(),
                                     (uint32_t)primitive.vertices.size(), primitive.indices.data(),
                                     (uint32_t)primitive.indices.size()) != SWEGL_OK)
        throw std::runtime_error(std::string("swegl::calculate_face_normals: ") + swegl_last_error());
    for (size_t i = 0; i < primitive.vertices.size(); i++) {
        primitive.vertices[i].normal.x() = nor[3 * i];
        primitive.vertices[i].normal.y() = nor[3 * i + 1];
        primitive.vertices[i].normal.z() = nor[3 * i + 2];
    }
}

}  // namespace swegl
