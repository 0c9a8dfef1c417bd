// swegl/data/gltf.hpp -- swegl::load_scene (reference: swegl/data/gltf.hpp:97,
// src/data/gltf.cpp:420-464) over the library's asset loader.  The file is
// read as the reference reads it (see swegl_asset_load in swegl_hip.h); the
// result is an ordinary scene_t with its animations, ready for render().
#pragma once

#include <stdexcept>
#include <string>

#include "../../swegl_hip.h"
#include "model.hpp"

namespace swegl
{

inline scene_t load_scene(const std::string &filename)
{
    swegl_asset *a = nullptr;
    if (swegl_asset_load(filename.c_str(), &a) != SWEGL_OK)
        throw std::runtime_error("swegl::load_scene(" + filename + "): " + swegl_last_error());
    struct holder {
        swegl_asset *a;
        ~holder() { swegl_asset_destroy(a); }
    } h{a};
    swegl_scene_desc d;
    swegl_asset_scene(a, &d);
    scene_t s;
    for (uint32_t i = 0; i < d.n_textures; i++)
        s.images.emplace_back(d.textures[i].bitmap, (int)d.textures[i].width, (int)d.textures[i].height);
    for (uint32_t i = 0; i < d.n_materials; i++) {
        const swegl_material_desc &m = d.materials[i];
        s.materials.push_back(material_t{pixel_colors((int)m.color), m.metallic, m.roughness, m.texture_idx,
                                         m.double_sided != 0});
    }
    for (uint32_t i = 0; i < d.n_nodes; i++) {
        const swegl_node_desc &nd = d.nodes[i];
        node_t n;
        n.scale = vertex_t(nd.scale[0], nd.scale[1], nd.scale[2]);
        std::memcpy(&n.rotation.m[0][0], nd.rotation, sizeof nd.rotation);
        n.translation = vertex_t(nd.translation[0], nd.translation[1], nd.translation[2]);
        for (uint32_t k = 0; k < nd.n_primitives; k++) {
            const swegl_primitive_desc &p = d.primitives[nd.first_primitive + k];
            primitive_t q;
            q.vertices.resize(p.n_vertices);
            for (uint32_t v = 0; v < p.n_vertices; v++) {
                mesh_vertex_t &mv = q.vertices[v];
                mv.v = vertex_t(p.positions[3 * v], p.positions[3 * v + 1], p.positions[3 * v + 2]);
                mv.normal.x() = p.normals[3 * v];          // as read: not renormalised
                mv.normal.y() = p.normals[3 * v + 1];
                mv.normal.z() = p.normals[3 * v + 2];
                mv.tex_coords = vec2f_t(p.tex_coords[2 * v], p.tex_coords[2 * v + 1]);
            }
            q.indices.assign(p.indices, p.indices + p.n_indices);
            q.mode = (primitive_t::index_mode_t)p.mode;
            q.material_id = p.material_id;
            n.primitives.push_back(std::move(q));
        }
        n.children_idx.assign(d.child_indices + nd.first_child, d.child_indices + nd.first_child + nd.n_children);
        s.nodes.push_back(std::move(n));
    }
    for (uint32_t i = 0; i < d.n_root_nodes; i++)
        s.root_nodes.push_back(d.root_nodes[i]);
    for (int c : std::vector<int>(d.child_indices, d.child_indices + d.n_child_indices))
        s.nodes[c].root = false;
    const swegl_animation *an = nullptr;
    uint32_t na = 0;
    swegl_asset_animations(a, &an, &na);
    for (uint32_t i = 0; i < na; i++) {
        animation_t A;
        A.end_time = an[i].end_time;
        for (uint32_t k = 0; k < an[i].n_channels; k++) {
            const swegl_anim_channel &c = an[i].channels[k];
            animation_channel_t C{c.node, (animation_channel_t::path_t)c.path, {}};
            for (uint32_t t = 0; t < c.n_steps; t++) {
                const swegl_anim_step &st = c.steps[t];
                C.steps.push_back({st.time, vec4f_t(st.value[0], st.value[1], st.value[2], st.value[3])});
            }
            A.channels.push_back(std::move(C));
        }
        s.animations.push_back(std::move(A));
    }
    return s;
}

}  // namespace swegl
