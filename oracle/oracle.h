/*
 * oracle.h -- CPU restatement of swegl's render path, used ONLY as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
 * It is never linked into or called by the product library.
 *
 * The scene input is the boundary's data format (include/swegl_hip.h);
 * viewports carry host buffers instead of device targets.
 */
#ifndef SWEGL_ORACLE_H
#define SWEGL_ORACLE_H

#include <stdint.h>
#include "../include/swegl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_viewport {
    int32_t   x, y, w, h;
    uint32_t *surface;                 /* SDL_Surface::pixels (host)          */
    int32_t   surface_w, surface_h;    /* SDL_Surface::w, h                   */
    int32_t   pitch;                   /* bytes per surface row               */
    float    *zbuffer;                 /* w*h                                 */
    uint32_t *triangle_id;             /* w*h or NULL: winning draw sequence  */
    float     view[16];
    float     projection[16];
    float     camera_position[3];
    int32_t   post_shader;             /* SWEGL_POST_*                        */
    float     focal_distance, focal_depth;
    int32_t   dof_max_radius;
    uint32_t *color64;                 /* w*h or NULL: FP64 re-shade of each  */
                                       /* pixel's final winner, before the    */
                                       /* post shader; 0 where nothing drew   */
                                       /* (SWEGL_SHADER_PHONG_BILINEAR only)  */
    int32_t   pixel_shader;            /* SWEGL_SHADER_*                      */
    int32_t   transparency_layer_count;/* 0..SWEGL_MAX_TRANSPARENCY_LAYERS;    */
                                       /* > 0 only with SWEGL_POST_NULL       */
} oracle_viewport;

typedef struct oracle_stats {
    uint64_t fragments;
    uint64_t nan_fragments;
    uint64_t drawn_subtriangles;
    uint64_t texel_oob;
    uint64_t clipped_triangles;        /* fill_triangle calls taking a clip branch */
} oracle_stats;

/* swegl::render(scene, vp...) restated; stats[n] may be NULL. */
int oracle_render(const swegl_scene_desc *scene, oracle_viewport *vps, uint32_t n,
                  oracle_stats *stats);
/* v_world, normal_world, v_viewport (xyz per vertex) and yes after rendering
 * one viewport, vertices in node, primitive, vertex order */
int oracle_vertex_state(const swegl_scene_desc *scene, oracle_viewport *vp, float *v_world,
                        float *normal_world, float *v_viewport, uint8_t *yes);

/* ---- single pieces, exposed so tests can pin them in isolation ---------- */
/* transform(v, m) (src/projection/points.cpp:8-13) */
void oracle_transform(const float v[3], const float m[16], float out[3]);
/* vertex stage for one vertex: v_world, normal_world and frustum coords
 * (vertex_shaders.hpp:16-71) for a node with world matrix M and local
 * rotation/scale RS = scale(rotation, scale). */
void oracle_vertex(const float v[3], const float n[3], const float M[16],
                   const float RS[16], const float view[16], const float proj[16],
                   float v_world[3], float n_world[3], float v_frustum[3]);
/* make_clip_vertex (renderer.cpp:290-297) between a and b = {v_world xyz,
 * normal xyz, tex uv} at `cut`, then world_to_viewport with node RS, view,
 * projection and viewport matrix vpm = {m00, m03, m11, m13}.
 * out = {v_world xyz, tex uv, normal_world xyz, v_viewport xyz}. */
void oracle_clip_vertex(const float a[8], const float b[8], float cut, const float RS[16],
                        const float view[16], const float proj[16], const float vpm[4],
                        float out[11]);
/* interpolator_g<1>: InitSelf(d,z1,z2); DisplaceStartingPoint(m); Step()
 * `steps` times.  out = {progress, value}. (interpolator.hpp:52-104) */
void oracle_interp(float d, float z1, float z2, float m, int steps, float out[2]);
/* Rasterise one screen-space triangle (fill_triangle_2 + fill_half_triangle,
 * renderer.cpp:361-558) into a w*h viewport at (0,0): writes per-pixel
 * fragment count (coverage, no depth test) and the last z. */
int oracle_raster_triangle(const float v[9], int w, int h, uint32_t *cover, float *zlast);
/* pixel_shader_texture_bilinear::shade at texel coords t (already scaled by
 * width/height), pixel_shaders.cpp:348-384. */
uint32_t oracle_bilinear(const uint32_t *bitmap, int w, int h, float tx, float ty);
/* pixel_shader_lights_phong::shade at one point (pixel_shaders.cpp:159-205). */
int32_t oracle_phong(const float center[3], const float normal[3], const float cam[3],
                     float ambient, const float sun_dir[3], float sun_intensity,
                     const swegl_point_light *lights, int n_lights);
/* pixel_shader_light_and_texture::shade combine step (pixel_shaders.hpp:159-178). */
uint32_t oracle_combine(uint32_t color, int32_t light);
/* depth-of-field post pass over a w*h image (build spec of
 * post_shader_depth_box, post_shaders.hpp:51-131). */
void oracle_dof(const uint32_t *src, const float *z, int w, int h, float focal_distance,
                float focal_depth, int max_radius, int taps9, uint32_t *dst);
/* x86-64 cvttss2si semantics: out of range / NaN -> INT32_MIN. */
int32_t oracle_cvtt(float f);
/* pixel_shader_texture::shade texel fetch (pixel_shaders.cpp:275-281) */
uint32_t oracle_nearest(const uint32_t *bitmap, int w, int h, float tx, float ty);

#ifdef __cplusplus
}
#endif
#endif
