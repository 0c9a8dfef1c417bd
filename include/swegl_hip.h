/*
 * swegl_hip.h -- C ABI of the MI355X renderer (libswegl_hip.so).
 *
 * This is the drop-in boundary for swegl's CPU render call
 *   template<typename...T> void swegl::render(scene_t&, T&... viewports)
 *   (swegl/render/renderer.hpp:27-34, src/render/renderer.cpp:77-235).
 * The reference call is a header-only C++ template with no ABI; the entry
 * points below are what an FFI for that path binds: plain structs, pointers
 * and sizes, integer status codes, no exceptions and no C++ types.
 *
 * Conventions
 *  - Every function returns 0 (SWEGL_OK) or a negative errno-style code.
 *  - Work is enqueued on the context's HIP stream; results are valid after
 *    swegl_ctx_synchronize() (or a blocking read-back call).
 *  - Matrices are row-major float[16] with the reference's meaning:
 *    m[r*4+c] == matrix44_t[r][c] (swegl/projection/matrix44.hpp:13-44).
 *  - Colours are swegl::pixel_colors stored as a 32-bit word: byte 0 = b,
 *    1 = g, 2 = r, 3 = a (swegl/render/colors.hpp:9-33).
 *  - One context serves one host thread at a time (the reference render is
 *    not re-entrant either: stateful shaders, mutated scene).
 */
#ifndef SWEGL_HIP_H
#define SWEGL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SWEGL_OK          0
#define SWEGL_EIO        (-5)   /* HIP runtime error (message via swegl_last_error) */
#define SWEGL_EAGAIN     (-11)  /* admission: not enough budget / slots right now */
#define SWEGL_ENOMEM     (-12)  /* device allocation failed / over HBM budget */
#define SWEGL_ENODEV     (-19)  /* no usable gfx950 device */
#define SWEGL_EINVAL     (-22)  /* malformed input (indices out of range, bad rect, ...) */
#define SWEGL_ERANGE     (-34)  /* a fixed capacity was exceeded (bins, queue) */
#define SWEGL_ENOTSUP    (-95)  /* shader/post-shader/transparency mode not on the GPU path */

/* ---- opaque handles ---------------------------------------------------- */
typedef struct swegl_ctx    swegl_ctx;
typedef struct swegl_scene  swegl_scene;
typedef struct swegl_target swegl_target;
typedef struct swegl_server swegl_server;

/* ---- scene description (swegl/data/model.hpp) --------------------------- */

/* primitive_t::index_mode_t (model.hpp:33-42); only triangle modes render,
 * points and lines are accepted and draw nothing, as in the reference */
#define SWEGL_POINTS          0
#define SWEGL_LINES           1
#define SWEGL_LINE_LOOP       2
#define SWEGL_LINE_STRIP      3
#define SWEGL_TRIANGLES       4
#define SWEGL_TRIANGLE_STRIP  5
#define SWEGL_TRIANGLE_FAN    6

/* primitive_t (model.hpp:31-48) with mesh_vertex_t inputs (model.hpp:20-29)
 * given as separate arrays: v (xyz), normal (xyz), tex_coords (xy). */
typedef struct swegl_primitive_desc {
    uint32_t        n_vertices;
    const float    *positions;   /* n_vertices*3 : mesh_vertex_t::v          */
    const float    *normals;     /* n_vertices*3 : mesh_vertex_t::normal     */
    const float    *tex_coords;  /* n_vertices*2 : mesh_vertex_t::tex_coords */
    uint32_t        n_indices;
    const uint32_t *indices;     /* primitive_t::indices                     */
    int32_t         mode;        /* SWEGL_TRIANGLES / _STRIP / _FAN          */
    int32_t         material_id; /* index into materials, -1 = none          */
} swegl_primitive_desc;

/* node_t (model.hpp:50-71).  Its primitives are
 * primitives[first_primitive .. first_primitive+n_primitives) and its
 * children are child_indices[first_child .. first_child+n_children). */
typedef struct swegl_node_desc {
    float    scale[3];           /* node_t::scale        */
    float    rotation[16];       /* node_t::rotation     */
    float    translation[3];     /* node_t::translation  */
    uint32_t first_primitive, n_primitives;
    uint32_t first_child, n_children;
} swegl_node_desc;

/* material_t (model.hpp:80-87) */
typedef struct swegl_material_desc {
    uint32_t color;              /* pixel_colors word */
    float    metallic, roughness;
    int32_t  texture_idx;        /* -1 = none */
    int32_t  double_sided;
} swegl_material_desc;

/* texture_t mip 0 (swegl/data/texture.hpp:12-38): row-major width*height
 * pixel_colors words. */
typedef struct swegl_texture_desc {
    const uint32_t *bitmap;
    uint32_t        width, height;
} swegl_texture_desc;

/* point_source_light (model.hpp:74-78) */
typedef struct swegl_point_light {
    float position[3];
    float intensity;
} swegl_point_light;

/* scene_t (model.hpp:128-178).  Draw order is nodes[] order, as
 * _render iterates scene.nodes (renderer.cpp:83). */
typedef struct swegl_scene_desc {
    uint32_t                    n_nodes;
    const swegl_node_desc      *nodes;
    uint32_t                    n_root_nodes;
    const int32_t              *root_nodes;
    uint32_t                    n_child_indices;
    const int32_t              *child_indices;
    uint32_t                    n_primitives;
    const swegl_primitive_desc *primitives;
    uint32_t                    n_materials;
    const swegl_material_desc  *materials;
    swegl_material_desc         default_material;
    uint32_t                    n_textures;
    const swegl_texture_desc   *textures;
    uint32_t                    n_point_lights;
    const swegl_point_light    *point_lights;
    float                       ambient_light_intensity;
    float                       sun_direction[3];   /* normal_t */
    float                       sun_intensity;
} swegl_scene_desc;

/* ---- viewport (swegl/render/viewport.hpp:27-62) ------------------------- */

/* pixel shader plugins (swegl/render/pixel_shaders.hpp): every concrete
 * shader type of the reference and every pixel_shader_light_and_texture<L, T>
 * (pixel_shaders.hpp:121-179) combination of its light and texture types.
 * The shaders that sample a texture index materials[material_id] and so
 * reject primitives with material_id == -1 (UB in the reference,
 * pixel_shaders.cpp:213-215, 288-290); the others use default_material. */
#define SWEGL_SHADER_PHONG_BILINEAR   1   /* light_and_texture<lights_phong, texture_bilinear> */
#define SWEGL_SHADER_PHONG_NEAREST    2   /* light_and_texture<lights_phong, texture>          */
#define SWEGL_SHADER_FLAT_BILINEAR    3   /* light_and_texture<lights_flat, texture_bilinear>  */
#define SWEGL_SHADER_FLAT_NEAREST     4   /* light_and_texture<lights_flat, texture>           */
#define SWEGL_SHADER_DEFAULT          5   /* pixel_shader_t: the material colour (:28)         */
#define SWEGL_SHADER_TEXTURE_NEAREST  6   /* pixel_shader_texture alone: the texel             */
#define SWEGL_SHADER_TEXTURE_BILINEAR 7   /* pixel_shader_texture_bilinear alone               */
#define SWEGL_SHADER_LIGHTS_FLAT      8   /* pixel_shader_lights_flat alone: the light word    */
#define SWEGL_SHADER_LIGHTS_PHONG     9   /* pixel_shader_lights_phong alone                   */

/* post shaders (swegl/render/post_shaders.hpp) */
/* transparency layers per viewport (viewport_t's transparency_layer_count);
 * with layers the post shader must be post_shader_t */
#define SWEGL_MAX_TRANSPARENCY_LAYERS 4
#define SWEGL_POST_NULL        0   /* post_shader_t, no transparency: no-op (:33-47)   */
#define SWEGL_POST_DEPTH_BOX   1   /* post_shader_depth_box: box taps (:51-131)          */
#define SWEGL_POST_DEPTH_TAP9  2   /* same blur radius, 3x3 taps at {-r,0,r}^2           */

typedef struct swegl_viewport_desc {
    int32_t       x, y, w, h;        /* viewport_t::m_x, m_y, m_w, m_h           */
    swegl_target *color;             /* the SDL_Surface (viewport_t::m_screen)    */
    swegl_target *depth;             /* viewport_t::m_zbuffer, w*h floats         */
    swegl_target *triangle_id;       /* optional (may be NULL): per-pixel winning
                                        draw sequence, w*h uint32, 0xFFFFFFFF = none */
    float         view[16];          /* camera_t::m_viewmatrix                    */
    float         projection[16];    /* camera_t::m_projectionmatrix              */
    float         camera_position[3];/* camera_t::m_center (camera_t::position()) */
    int32_t       pixel_shader;      /* SWEGL_SHADER_*                            */
    int32_t       post_shader;       /* SWEGL_POST_*                              */
    float         focal_distance;    /* post_shader_depth_box::focal_distance     */
    float         focal_depth;       /* post_shader_depth_box::focal_depth        */
    int32_t       dof_max_radius;    /* 5 in the reference (post_shaders.hpp:69)  */
    int32_t       transparency_layer_count; /* 0..SWEGL_MAX_TRANSPARENCY_LAYERS   */
} swegl_viewport_desc;

/* Per-render counters, filled when the render's work completes. */
typedef struct swegl_frame_stats {
    uint64_t fragments;          /* (sub-triangle, pixel) samples with z > 0.001 */
    uint64_t nan_fragments;      /* samples whose z is NaN (order-dependent in
                                    the reference; must be 0 for exact parity) */
    uint64_t drawn_subtriangles; /* fill_triangle_2 calls that reach the DDA     */
    uint64_t bin_entries;        /* (triangle, tile) pairs                        */
    uint64_t texel_oob;          /* bilinear taps outside the bitmap (UB in the
                                    reference; clamped here)                      */
    uint64_t big_triangles;      /* triangles over the bin limit (tested by every
                                    tile of their viewport)                       */
} swegl_frame_stats;

/* ---- context ----------------------------------------------------------- */
int  swegl_ctx_create(int device, swegl_ctx **out);
int  swegl_ctx_destroy(swegl_ctx *ctx);
int  swegl_ctx_synchronize(swegl_ctx *ctx);
const char *swegl_last_error(void);
/* Device facts: number of CUs, gcn arch name (e.g. "gfx950:sramecc+:xnack-") */
int  swegl_device_info(swegl_ctx *ctx, int *n_cus, char *arch, int arch_cap,
                       uint64_t *hbm_bytes);

/* Raster tuning: tile size in pixels (8x8, 16x16, 32x16, 32x32, 64x32 or
 * 64x64); the autotuner writes it. */
int  swegl_ctx_set_tile(swegl_ctx *ctx, int tile_w, int tile_h);
int  swegl_ctx_get_tile(swegl_ctx *ctx, int *tile_w, int *tile_h);
/* With auto tiles on (the default), a render whose viewports give fewer
 * tiles than there are persistent raster workgroups uses the next smaller
 * tile sizes (64x64 > 64x32 > 32x32 > 32x16 > 16x16) until they do, so a
 * small viewport still spreads over the device; off: always the set tile. */
int  swegl_ctx_set_tile_auto(swegl_ctx *ctx, int on);
/* Bin limit: a triangle whose tile rectangle holds more than `tiles` tiles
 * is not binned per tile but listed once per viewport and tested by every
 * tile (default 16; 0 sends every triangle through that list). */
int  swegl_ctx_set_bin_limit(swegl_ctx *ctx, uint32_t tiles);
/* Persistent tile-raster workgroups per compute unit (1..16, default 4). */
int  swegl_ctx_set_raster_groups(swegl_ctx *ctx, uint32_t per_cu);

/* Per-kernel HIP-event timing on the context stream (0 = off). */
int  swegl_ctx_set_profiling(swegl_ctx *ctx, int on);
/* Average milliseconds per launch of each kernel since profiling was
 * switched on: names[i] (NUL-separated into name_buf), ms[i], counts[i]. */
int  swegl_ctx_kernel_times(swegl_ctx *ctx, char *name_buf, int name_cap,
                            double *ms, uint64_t *counts, int cap, int *n);

/* ---- device-resident surfaces ----------------------------------------- */
/* A colour surface (width*height pixel_colors words, pitch = width*4) or a
 * float/uint32 plane of width*height elements (depth, triangle id). */
#define SWEGL_TARGET_COLOR  0
#define SWEGL_TARGET_FLOAT  1
#define SWEGL_TARGET_UINT   2
int  swegl_target_create(swegl_ctx *ctx, int kind, int32_t width, int32_t height,
                         swegl_target **out);
int  swegl_target_destroy(swegl_ctx *ctx, swegl_target *t);
/* Blocking copies between a target and host memory with the given row pitch
 * in bytes (SDL_Surface::pitch). */
int  swegl_target_read(swegl_ctx *ctx, swegl_target *t, void *dst, int32_t dst_pitch);
int  swegl_target_write(swegl_ctx *ctx, swegl_target *t, const void *src, int32_t src_pitch);
/* Blocking copy of the rectangle [x, x+w) x [y, y+h) of a target to host
 * memory (dst points at the rectangle's first element, dst_pitch in bytes). */
int  swegl_target_read_rect(swegl_ctx *ctx, swegl_target *t, int32_t x, int32_t y, int32_t w,
                            int32_t h, void *dst, int32_t dst_pitch);
/* Enqueue a copy of one target into another of the same size and kind. */
int  swegl_target_copy(swegl_ctx *ctx, swegl_target *dst, swegl_target *src);
/* Enqueue a fill of every element with a 32-bit pattern. */
int  swegl_target_fill(swegl_ctx *ctx, swegl_target *t, uint32_t word);

/* ---- scenes ------------------------------------------------------------ */
/* Copy meshes, textures, materials, lights and poses to the device.  The
 * scene stays resident until swegl_scene_destroy. */
int  swegl_scene_upload(swegl_ctx *ctx, const swegl_scene_desc *desc, swegl_scene **out);
/* Re-pose a resident scene: node scale/rotation/translation, lights,
 * ambient and sun from desc (same node/primitive topology; meshes and
 * textures are not re-sent). */
int  swegl_scene_update_pose(swegl_ctx *ctx, swegl_scene *scene, const swegl_scene_desc *desc);
int  swegl_scene_destroy(swegl_ctx *ctx, swegl_scene *scene);
/* Device bytes held by a scene (meshes, textures, per-frame scratch). */
int  swegl_scene_bytes(swegl_ctx *ctx, swegl_scene *scene, uint64_t *bytes);

/* ---- render: swegl::render(scene, vp...) ------------------------------- */
/* Replaces renderer.hpp:27-34: world transform once, then each viewport in
 * order (renderer.cpp:77-235), including clear (viewport.cpp:88-113) and the
 * post shader.  stats may be NULL.  Asynchronous: results are complete after
 * swegl_ctx_synchronize or a target read; renders queued back to back
 * overlap (one render's binning runs while the previous one's raster
 * finishes) and write their targets in call order. */
int  swegl_render(swegl_ctx *ctx, swegl_scene *scene,
                  const swegl_viewport_desc *viewports, uint32_t n_viewports,
                  swegl_frame_stats *stats);

/* Parity aid: render one viewport (as swegl_render) and return the scene
 * fields the reference's render writes into scene_t (model.hpp:20-29):
 * v_world, normal_world and v_viewport (xyz per vertex) and yes, for every
 * vertex in node, primitive, vertex order.  v_viewport is the frustum
 * position, moved to screen coordinates for marked vertices as
 * frustum_to_viewport does (vertex_shaders.hpp:72-84). */
int  swegl_debug_vertex_state(swegl_ctx *ctx, swegl_scene *scene, const swegl_viewport_desc *viewport,
                              float *v_world, float *normal_world, float *v_viewport, uint8_t *yes);

/* Allocation-free variant of swegl_render.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels run on a fixed per-device scratch array of
 * SWEGL_NOALLOC_SCRATCH_BYTES (held by every device the library runs on
 * until the process ends) instead of buffers the library allocates; the
 * scene and the viewports' targets are the caller's, as in swegl_render.
 * The call allocates no device or pinned memory.  Renders whose scratch
 * does not fit (per viewport: 17 B per vertex, 12 B per triangle, 5 B per
 * pixel with DoF, 4 B per pixel with layers; per call: the bins, 4 B per
 * triangle and bin-limit tile, and the tile counters) return SWEGL_ERANGE
 * before anything is enqueued.  Every call on a device shares its array:
 * calls from contexts that overlap on one device race; calls on one context
 * are ordered by its stream. */
#define SWEGL_NOALLOC_SCRATCH_BYTES ((uint64_t)64 << 20)
int  swegl_render_noalloc(swegl_ctx *ctx, swegl_scene *scene,
                          const swegl_viewport_desc *viewports, uint32_t n_viewports,
                          swegl_frame_stats *stats);

/* Plain path as one replayed hipGraph: the first call for a given
 * (scene, viewport set) captures, later calls replay.  A viewport set that
 * combines a depth-of-field post shader with a pixel shader other than
 * Phong + bilinear renders eagerly instead (same results, no capture). */
int  swegl_render_graph(swegl_ctx *ctx, swegl_scene *scene,
                        const swegl_viewport_desc *viewports, uint32_t n_viewports,
                        swegl_frame_stats *stats);

/* ---- batched: many independent jobs in one launch sequence ------------- */
typedef struct swegl_job {
    swegl_scene               *scene;
    const swegl_viewport_desc *viewports;
    uint32_t                   n_viewports;
} swegl_job;
int  swegl_render_batch(swegl_ctx *ctx, const swegl_job *jobs, uint32_t n_jobs,
                        swegl_frame_stats *stats);
/* The same batch as a replayed hipGraph (captured on the first call with a
 * given launch shape: job count, scene sizes, viewport sizes, post shaders). */
int  swegl_render_batch_graph(swegl_ctx *ctx, const swegl_job *jobs, uint32_t n_jobs,
                              swegl_frame_stats *stats);

/* ---- served: continuous batching ----------------------------------------- */
/* Requests (a scene plus its viewports) are queued by swegl_server_submit.
 * swegl_server_poll retires finished batches and admits queued requests into
 * every free lane as one batched launch sequence, FIFO, as long as the
 * device bytes the server holds stay within the HBM budget: each lane's
 * scratch region (kept between batches; idle lanes give theirs back when
 * the queue's head would not fit otherwise) and the target planes of the
 * requests in flight.  A request's scratch is bounded by the formulas the
 * launch sizes its buffers with.  Batches on different lanes run
 * concurrently, so requests that arrive while others are on the GPU start
 * as soon as a lane and memory free up.  Each result is bit-identical to
 * rendering the request alone. */
typedef struct swegl_server_config {
    uint64_t hbm_budget_bytes;   /* admission budget for requests in flight     */
    uint32_t max_batch;          /* requests per launch sequence (0 = 256)       */
    uint32_t lanes;              /* batches in flight at once, 1..8 (0 = 2)      */
} swegl_server_config;

typedef struct swegl_server_stats {
    uint64_t submitted, completed, batches;
    uint64_t bytes_in_flight;        /* scratch bounds + target planes of requests in flight */
    uint64_t peak_bytes_in_flight;   /* most device bytes held at once (lane regions + targets) */
    uint32_t largest_batch;
} swegl_server_stats;

int  swegl_server_create(swegl_ctx *ctx, const swegl_server_config *cfg, swegl_server **out);
int  swegl_server_destroy(swegl_server *srv);
/* Queue one request; its viewport array is copied.  EAGAIN when the request
 * alone exceeds the budget.  Targets of requests in flight must not be
 * reused before those requests finish. */
int  swegl_server_submit(swegl_server *srv, const swegl_job *job, uint64_t *ticket);
/* Admit and retire; returns up to `cap` finished tickets. */
int  swegl_server_poll(swegl_server *srv, uint64_t *done, uint32_t cap, uint32_t *n_done);
/* Block until every submitted request has finished. */
int  swegl_server_drain(swegl_server *srv);
int  swegl_server_get_stats(swegl_server *srv, swegl_server_stats *st);
/* Replay an arrival trace natively: request i is submitted arrival_s[i]
 * seconds after the call starts (later if a target it writes is still in
 * flight) and latency_s[i] = completion time - arrival time. */
int  swegl_server_run_trace(swegl_server *srv, const swegl_job *jobs, const double *arrival_s,
                            uint32_t n, double *latency_s);

/* ---- assets: glTF 2.0 / GLB files (src/data/gltf.cpp) ------------------ */
/* A loaded file in host memory: the scene arrays and the animations.  The
 * reading follows the reference's loader, including where it departs from
 * the glTF specification (transposed TEXCOORD_0, 16-bit index reads, the
 * quaternion's x taken as its scalar part, the row-length scale of a node
 * matrix); PNG and BMP textures are decoded, JPEG returns SWEGL_ENOTSUP. */
typedef struct swegl_asset swegl_asset;

/* animation_step_t (model.hpp:90-94): value is (x, y, z) for scale and
 * translation, (x, y, z, w) for rotation */
typedef struct swegl_anim_step {
    float time;
    float value[4];
} swegl_anim_step;
/* animation_channel_t::path_t (model.hpp:97-105) */
#define SWEGL_PATH_SCALE       0
#define SWEGL_PATH_ROTATION    1
#define SWEGL_PATH_TRANSLATION 2
#define SWEGL_PATH_NONE        3   /* "weights" and unknown paths: not applied */
/* animation_channel_t (model.hpp:95-122): steps sorted by time */
typedef struct swegl_anim_channel {
    int32_t                node;
    int32_t                path;
    uint32_t               n_steps;
    const swegl_anim_step *steps;
} swegl_anim_channel;
/* animation_t (model.hpp:123-127) */
typedef struct swegl_animation {
    float                     end_time;
    uint32_t                  n_channels;
    const swegl_anim_channel *channels;
} swegl_animation;

/* swegl::load_scene(filename) (src/data/gltf.cpp:455-464, .glb or .gltf). */
int  swegl_asset_load(const char *filename, swegl_asset **out);
/* The asset as a scene description for swegl_scene_upload; its arrays
 * belong to the asset.  Lights, ambient and sun are left at zero (sun
 * direction (0, -1, 0)), as the reference leaves them to the caller. */
int  swegl_asset_scene(const swegl_asset *asset, swegl_scene_desc *desc);
/* scene_t::animations (model.hpp:141), owned by the asset */
int  swegl_asset_animations(const swegl_asset *asset, const swegl_animation **animations,
                            uint32_t *n_animations);
void swegl_asset_destroy(swegl_asset *asset);
/* scene_t::animate(elapsed_seconds) (model.hpp:146-177): sets the targeted
 * nodes' rotation, translation and scale; pass the result to
 * swegl_scene_update_pose. */
int  swegl_animate(const swegl_animation *animations, uint32_t n_animations, float elapsed_seconds,
                   swegl_node_desc *nodes, uint32_t n_nodes);
/* read_image_file (src/misc/image.cpp:20-28) for .png and .bmp: row-major
 * pixel_colors words; free with swegl_image_free. */
int  swegl_image_read(const char *filename, uint32_t **pixels, uint32_t *width, uint32_t *height);
void swegl_image_free(uint32_t *pixels);

/* ---- procedural meshes (swegl/data/model.hpp:181-466) ------------------ */
/* make_tore(precision, material) (model.hpp:327-391) and
 * make_sphere(precision, radius, material) (:393-464): precision + 1 triangle
 * strips of 2 * (precision + 1) vertices each.  The caller's arrays hold
 * V = (precision + 1) * 2 * (precision + 1) vertices: positions and normals
 * V * 3 floats, tex_coords V * 2 floats, strip after strip; indices
 * V uint32, 2 * (precision + 1) per strip, local to the strip. */
int  swegl_make_tore(uint32_t precision, float *positions, float *normals, float *tex_coords,
                     uint32_t *indices);
int  swegl_make_sphere(uint32_t precision, float radius, float *positions, float *normals,
                       float *tex_coords, uint32_t *indices);
/* calculate_face_normals (model.hpp:181-231) on one primitive, in place,
 * as the reference computes them (strips and fans; its last loop runs for
 * fans, so triangle lists keep their normals). */
int  swegl_calculate_face_normals(int32_t mode, const float *positions, float *normals,
                                  uint32_t n_vertices, const uint32_t *indices, uint32_t n_indices);

/* ---- host math helpers (same float semantics as the reference) --------- */
/* matrix44_t::rotate_x/y/z (src/projection/matrix44.cpp:12-67), translate
 * (:87-92), swegl::scale (src/projection/points.cpp:14-30). */
void swegl_matrix_rotate_x(float m[16], float a);
void swegl_matrix_rotate_y(float m[16], float a);
void swegl_matrix_rotate_z(float m[16], float a);
void swegl_matrix_translate(float m[16], float x, float y, float z);
/* matrix44_t::from_quaternion(q0, q1, q2, q3) (swegl/projection/matrix44.hpp:34-43) */
void swegl_matrix_from_quaternion(float m[16], float q0, float q1, float q2, float q3);
/* camera_t(aspect) (src/projection/camera.cpp:8-35): view, projection,
 * center. */
void swegl_camera_init(float aspect_ratio, float view[16], float projection[16],
                       float center[3]);
/* camera_t::translate (camera.cpp:50-58) and camera_t::rotate_x/y/z
 * (camera.cpp:37-48). axis: 0=x, 1=y, 2=z. */
void swegl_camera_translate(float view[16], float center[3], float x, float y, float z);
void swegl_camera_rotate(float view[16], int axis, float a);

#ifdef __cplusplus
}
#endif
#endif /* SWEGL_HIP_H */
