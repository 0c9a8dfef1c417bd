// swegl/render/renderer.hpp -- swegl::render(scene, viewports...) on the GPU.
//
// Same signature and meaning as the reference's CPU renderer
// (swegl/render/renderer.hpp:27-34): world transform once, then each
// viewport in order, including clear and post shader; results land in each
// viewport's SDL_Surface rectangle and m_zbuffer.  Underneath it calls the C
// ABI of libswegl_hip.so (swegl_hip.h).  Meshes and textures are uploaded on
// first use and again when scene_t::device_generation changes; node poses
// and lights are refreshed on every call.  Link with -lswegl_hip.
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../swegl_hip.h"
#include "../data/model.hpp"
#include "viewport.hpp"

namespace swegl
{
namespace detail
{

inline void check(int rc, const char *what)
{
    if (rc != SWEGL_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                                 "): " + swegl_last_error());
}

struct device_state {
    swegl_ctx *ctx = nullptr;
    struct scene_entry {
        swegl_scene *ds = nullptr;
        std::uint64_t gen = 0;
        std::uint64_t shape = 0;   // scene_shape() at upload
    };
    std::unordered_map<std::uint64_t, scene_entry> scenes;   // by scene_t::device_id
    struct surface_entry {
        swegl_target *t = nullptr;
        int w = 0, h = 0;
    };
    std::unordered_map<const SDL_Surface *, surface_entry> surfaces;
    device_state()
    {
        check(swegl_ctx_create(0, &ctx), "swegl_ctx_create");
        // a scene object that goes away takes its device copy with it
        swegl::detail::scene_gone_hook() = {&device_state::scene_gone, this};
    }
    static void scene_gone(void *self, std::uint64_t id)
    {
        device_state &st = *(device_state *)self;
        auto it = st.scenes.find(id);
        if (it != st.scenes.end()) {
            swegl_scene_destroy(st.ctx, it->second.ds);
            st.scenes.erase(it);
        }
    }
    ~device_state()
    {
        swegl::detail::scene_gone_hook() = {};
        for (auto &kv : scenes)
            swegl_scene_destroy(ctx, kv.second.ds);
        for (auto &kv : surfaces)
            swegl_target_destroy(ctx, kv.second.t);
        swegl_ctx_destroy(ctx);
    }
};

// A cheap signature of what an upload copies: node and primitive counts,
// each primitive's vertex and index array addresses and sizes, each
// texture's bitmap address and size.  A change (arrays grown, replaced or
// re-allocated, textures added) triggers a re-upload by itself; edits of
// values in place keep the signature and need device_generation bumped.
inline std::uint64_t scene_shape(const scene_t &s)
{
    std::uint64_t h = 1469598103934665603ull;
    auto mix = [&h](std::uint64_t v) {
        h ^= v;
        h *= 1099511628211ull;
    };
    mix(s.nodes.size());
    for (const node_t &n : s.nodes) {
        mix(n.primitives.size());
        for (const primitive_t &p : n.primitives) {
            mix((std::uint64_t)(std::uintptr_t)p.vertices.data());
            mix(p.vertices.size());
            mix((std::uint64_t)(std::uintptr_t)p.indices.data());
            mix(p.indices.size());
            mix((std::uint64_t)p.mode);
            mix((std::uint64_t)(std::int64_t)p.material_id);
        }
    }
    mix(s.images.size());
    for (const texture_t &t : s.images)
        for (const auto &m : t.m_mipmaps) {
            mix((std::uint64_t)(std::uintptr_t)m->m_bitmap.data());
            mix(m->m_bitmap.size());
        }
    mix(s.materials.size());
    return h;
}

inline device_state &state()
{
    static device_state s;
    return s;
}

// scene_t (array of structures) -> swegl_scene_desc (structure of arrays)
struct scene_arrays {
    std::vector<std::vector<float>> pos, nor, tex;
    std::vector<swegl_primitive_desc> prims;
    std::vector<swegl_node_desc> nodes;
    std::vector<int32_t> roots, children;
    std::vector<swegl_material_desc> mats;
    std::vector<swegl_texture_desc> texs;
    std::vector<swegl_point_light> lights;
    swegl_scene_desc d{};

    static swegl_material_desc mat(const material_t &m)
    {
        swegl_material_desc r;
        r.color = (uint32_t)m.color.to_int();
        r.metallic = m.metallic;
        r.roughness = m.roughness;
        r.texture_idx = m.texture_idx;
        r.double_sided = m.double_sided ? 1 : 0;
        return r;
    }

    // with_meshes false: nodes, lights, ambient and sun only (a re-pose)
    explicit scene_arrays(const scene_t &s, bool with_meshes = true)
    {
        uint32_t n_prims = 0;
        for (const node_t &n : s.nodes) {
            swegl_node_desc nd{};
            nd.scale[0] = n.scale.x();
            nd.scale[1] = n.scale.y();
            nd.scale[2] = n.scale.z();
            std::memcpy(nd.rotation, &n.rotation.m[0][0], sizeof nd.rotation);
            nd.translation[0] = n.translation.x();
            nd.translation[1] = n.translation.y();
            nd.translation[2] = n.translation.z();
            nd.first_primitive = n_prims;
            nd.n_primitives = (uint32_t)n.primitives.size();
            n_prims += nd.n_primitives;
            nd.first_child = (uint32_t)children.size();
            nd.n_children = (uint32_t)n.children_idx.size();
            children.insert(children.end(), n.children_idx.begin(), n.children_idx.end());
            for (const primitive_t &p : n.primitives) {
                if (!with_meshes)
                    break;
                std::vector<float> P, N, T;
                for (const mesh_vertex_t &v : p.vertices) {
                    P.insert(P.end(), {v.v.x(), v.v.y(), v.v.z()});
                    N.insert(N.end(), {v.normal.x(), v.normal.y(), v.normal.z()});
                    T.insert(T.end(), {v.tex_coords.x(), v.tex_coords.y()});
                }
                pos.push_back(std::move(P));
                nor.push_back(std::move(N));
                tex.push_back(std::move(T));
                swegl_primitive_desc pd{};
                pd.n_vertices = (uint32_t)p.vertices.size();
                pd.positions = pos.back().data();
                pd.normals = nor.back().data();
                pd.tex_coords = tex.back().data();
                pd.n_indices = (uint32_t)p.indices.size();
                pd.indices = p.indices.data();
                pd.mode = (int32_t)p.mode;
                pd.material_id = p.material_id;
                prims.push_back(pd);
            }
            nodes.push_back(nd);
        }
        roots.assign(s.root_nodes.begin(), s.root_nodes.end());
        for (const material_t &m : s.materials)
            mats.push_back(mat(m));
        for (const texture_t &t : s.images) {
            if (!with_meshes)
                break;
            const mipmap_t &m = *t.m_mipmaps.at(0);
            texs.push_back({m.m_bitmap.data(), m.m_width, m.m_height});
        }
        for (const point_source_light &l : s.point_source_lights)
            lights.push_back({{l.position.x(), l.position.y(), l.position.z()}, l.intensity});
        d.n_nodes = (uint32_t)nodes.size();
        d.nodes = nodes.data();
        d.n_root_nodes = (uint32_t)roots.size();
        d.root_nodes = roots.data();
        d.n_child_indices = (uint32_t)children.size();
        d.child_indices = children.data();
        d.n_primitives = (uint32_t)prims.size();
        d.primitives = prims.data();
        d.n_materials = (uint32_t)mats.size();
        d.materials = mats.data();
        d.default_material = mat(s.default_material);
        d.n_textures = (uint32_t)texs.size();
        d.textures = texs.data();
        d.n_point_lights = (uint32_t)lights.size();
        d.point_lights = lights.data();
        d.ambient_light_intensity = s.ambient_light_intensity;
        d.sun_direction[0] = s.sun_direction.x();
        d.sun_direction[1] = s.sun_direction.y();
        d.sun_direction[2] = s.sun_direction.z();
        d.sun_intensity = s.sun_intensity;
    }
};

inline swegl_target *surface_target(device_state &st, SDL_Surface *s)
{
    auto &e = st.surfaces[s];
    if (e.t && (e.w != s->w || e.h != s->h)) {
        swegl_target_destroy(st.ctx, e.t);
        e.t = nullptr;
    }
    if (!e.t) {
        check(swegl_target_create(st.ctx, SWEGL_TARGET_COLOR, s->w, s->h, &e.t), "swegl_target_create");
        e.w = s->w;
        e.h = s->h;
    }
    return e.t;
}

inline void render_impl(scene_t &scene, viewport_t *const *vps, size_t n)
{
    device_state &st = state();
    auto &se = st.scenes[scene.device_id.v];
    const std::uint64_t shape = scene_shape(scene);
    if (!se.ds || se.gen != scene.device_generation || se.shape != shape) {
        if (se.ds)
            swegl_scene_destroy(st.ctx, se.ds);
        se.ds = nullptr;
        scene_arrays arrays(scene);
        check(swegl_scene_upload(st.ctx, &arrays.d, &se.ds), "swegl_scene_upload");
        se.gen = scene.device_generation;
        se.shape = shape;
    } else {
        // meshes and textures are resident: send the poses and lights only
        scene_arrays arrays(scene, false);
        check(swegl_scene_update_pose(st.ctx, se.ds, &arrays.d), "swegl_scene_update_pose");
    }
    std::vector<swegl_viewport_desc> descs(n);
    for (size_t i = 0; i < n; i++) {
        viewport_t &vp = *vps[i];
        swegl_viewport_desc &d = descs[i];
        std::memset(&d, 0, sizeof d);
        if (vp.m_screen->format && vp.m_screen->format->BytesPerPixel != 4)
            throw std::runtime_error("swegl::render: the surface must be 32 bits per pixel");
        d.x = vp.m_x;
        d.y = vp.m_y;
        d.w = vp.m_w;
        d.h = vp.m_h;
        d.color = surface_target(st, vp.m_screen);
        if (!vp.m_device_depth)
            check(swegl_target_create(st.ctx, SWEGL_TARGET_FLOAT, vp.m_w, vp.m_h, &vp.m_device_depth),
                  "swegl_target_create");
        d.depth = vp.m_device_depth;
        std::memcpy(d.view, &vp.m_camera.m_viewmatrix.m[0][0], sizeof d.view);
        std::memcpy(d.projection, &vp.m_camera.m_projectionmatrix.m[0][0], sizeof d.projection);
        d.camera_position[0] = vp.m_camera.m_center.x();
        d.camera_position[1] = vp.m_camera.m_center.y();
        d.camera_position[2] = vp.m_camera.m_center.z();
        d.pixel_shader = vp.m_pixel_shader ? vp.m_pixel_shader->gpu_kind() : 0;
        // the reference dereferences m_post_shader unconditionally
        if (!vp.m_post_shader)
            throw std::runtime_error("swegl::render: viewport has no post shader (set_post_shader)");
        d.post_shader = vp.m_post_shader->gpu_kind();
        d.focal_distance = vp.m_post_shader->focal_distance;
        d.focal_depth = vp.m_post_shader->focal_depth;
        d.dof_max_radius = vp.m_post_shader->max_radius;
        d.transparency_layer_count = vp.m_transparency_layer_count;
    }
    check(swegl_render(st.ctx, se.ds, descs.data(), (uint32_t)n, nullptr), "swegl_render");
    // the viewport rectangles and z-buffers back into host memory; with
    // transparency layers the copy writes columns from 0 (post_shaders.hpp:24-28)
    for (size_t i = 0; i < n; i++) {
        viewport_t &vp = *vps[i];
        const int x0 = vp.m_transparency_layer_count > 0 ? 0 : vp.m_x;
        char *dst = (char *)vp.m_screen->pixels + (size_t)vp.m_y * vp.m_screen->pitch + (size_t)x0 * 4;
        check(swegl_target_read_rect(st.ctx, descs[i].color, x0, vp.m_y, vp.m_x + vp.m_w - x0, vp.m_h, dst,
                                     vp.m_screen->pitch),
              "swegl_target_read_rect");
        check(swegl_target_read(st.ctx, vp.m_device_depth, vp.m_zbuffer.get(), vp.m_w * 4), "swegl_target_read");
    }
}

}  // namespace detail

inline viewport_t::~viewport_t()
{
    if (m_device_depth)
        swegl_target_destroy(detail::state().ctx, m_device_depth);
}

template <typename... T>
void render(scene_t &scene, T &...viewports)
{
    viewport_t *vps[] = {&viewports...};
    detail::render_impl(scene, vps, sizeof...(T));
}

}  // namespace swegl
