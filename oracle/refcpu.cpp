/*
 * refcpu.cpp -- ORACLE (test infrastructure only).
 *
 * A sequential CPU restatement of swegl's render path, written to follow the
 * reference step by step so that every float operation happens in the same
 * order with the same operands.  Only tests/, __graft_entry__.smoke() and
 * bench.py (cpu_baseline) may load it; the product library never does.
 *
 * Build: g++ -O3 -msse4 -ffp-contract=off (the reference's release flags,
 * Makefile:17, plus an explicit contraction-off).
 *
 * Reference files followed (paths relative to the reference root):
 *   swegl/render/renderer.hpp:20-34         render / _render recursion
 *   src/render/renderer.cpp:58-235          frustum test, facing, marking, draw loops
 *   src/render/renderer.cpp:240-359         fill_triangle (z sort, near clip)
 *   src/render/renderer.cpp:361-460         fill_triangle_2 (y sort, DDA setup)
 *   src/render/renderer.cpp:462-558         fill_half_triangle (spans, z test)
 *   swegl/render/interpolator.hpp:52-104    interpolator_g<1>
 *   swegl/render/vertex_shaders.hpp:16-88   vertex stage
 *   src/render/pixel_shaders.cpp:14-31,88-205,284-384  phong + bilinear
 *   swegl/render/pixel_shaders.hpp:121-179  light_and_texture combiner
 *   src/render/colors.cpp:11-30             pixel_colors_f ops
 *   src/render/viewport.cpp:14-40,88-129    viewport matrix, clear, transform
 *   swegl/render/post_shaders.hpp:51-131, swegl/misc/lerp.hpp:24-43  DoF
 *   src/projection/points.cpp:8-49, swegl/projection/points.hpp:71-151  vector math
 *
 * Parity status of each piece (how it is pinned) is listed in the test
 * files tests/test_oracle_*.py.  The one piece with no pin is the matrix
 * product used for multi-level node hierarchies (freon::Matrix operator*,
 * not present in the reference tree): parity unpinned for scenes whose
 * nodes have children.  For root-only scenes the product is Identity*M,
 * which is exact in any summation order.
 */
#include "oracle.h"

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

/* ---------------- scalar semantics of the x86-64 reference build -------- */

/* (int)f compiled to cvttss2si: NaN and out-of-range give INT32_MIN. */
inline int32_t cvtt(float f)
{
    if (f >= -2147483648.0f && f < 2147483648.0f)
        return (int32_t)f;
    return INT32_MIN;
}
inline int32_t cvtt_d(double f)
{
    if (f >= -2147483648.0 && f < 2147483648.0)
        return (int32_t)f;
    return INT32_MIN;
}
/* (unsigned char)f: cvttss2si then keep the low byte. */
inline uint8_t to_u8(float f) { return (uint8_t)(uint32_t)cvtt(f); }
inline uint8_t to_u8_d(double f) { return (uint8_t)(uint32_t)cvtt_d(f); }
/* 32-bit wrapping add (signed overflow wraps in the x86 build). */
inline int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

/* ---------------- vertex_t / vector_t / normal_t ------------------------ */
struct vec3 { float x, y, z; };
struct vec2 { float x, y; };

inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 neg(vec3 a) { return {-a.x, -a.y, -a.z}; }
inline bool same(vec3 a, vec3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline vec2 operator+(vec2 a, vec2 b) { return {a.x + b.x, a.y + b.y}; }
inline vec2 operator-(vec2 a, vec2 b) { return {a.x - b.x, a.y - b.y}; }
inline vec2 operator*(vec2 a, float s) { return {a.x * s, a.y * s}; }

/* vector_t::dot (points.hpp:93): (xx + yy) + zz */
inline float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* vector_t::len_squared (points.hpp:75-78) */
inline float len2(vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
/* vector_t::normalize (points.hpp:80-90) */
inline vec3 normalize(vec3 a)
{
    float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    if (l != 0) {
        a.x /= l;
        a.y /= l;
        a.z /= l;
    }
    return a;
}
/* cross() returns a normal_t, whose constructor normalises (points.cpp:32-37) */
inline vec3 cross_n(vec3 l, vec3 r)
{
    return normalize({l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x});
}
/* transform(v, m) (points.cpp:8-13) */
inline vec3 transform(vec3 v, const float *m)
{
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3],
            m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7],
            m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11]};
}
/* rotate(n, m) (points.cpp:44-49): the normal_t result is normalised */
inline vec3 rotate_n(vec3 v, const float *m)
{
    return normalize({m[0] * v.x + m[1] * v.y + m[2] * v.z,
                      m[4] * v.x + m[5] * v.y + m[6] * v.z,
                      m[8] * v.x + m[9] * v.y + m[10] * v.z});
}

/* ---------------- scene mirror (AoS, as model.hpp:20-71) ---------------- */
struct mesh_vertex {
    vec3 v, v_world, v_viewport;
    vec2 tex;
    vec3 normal, normal_world;
    bool yes;
};
struct primitive {
    std::vector<mesh_vertex> vertices;
    std::vector<uint32_t> indices;
    int mode;
    int material_id;
};
struct node {
    float local[16];   /* get_local_world_matrix() (model.hpp:56-61)   */
    float rs[16];      /* scale(rotation, scale) used for normals       */
    float o2w[16];     /* original_to_world_matrix                      */
    std::vector<primitive> prims;
    std::vector<int> children;
};
struct scene_m {
    std::vector<node> nodes;
    std::vector<int> roots;
    std::vector<swegl_material_desc> materials;
    swegl_material_desc default_material;
    std::vector<swegl_texture_desc> textures;
    std::vector<swegl_point_light> lights;
    float ambient;
    vec3 sun_dir;
    float sun_intensity;
};

/* freon::Matrix product, textbook order: parity unpinned (see header). */
void matmul(const float *a, const float *b, float *r)
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r[i * 4 + j] = a[i * 4 + 0] * b[0 * 4 + j] + a[i * 4 + 1] * b[1 * 4 + j] +
                           a[i * 4 + 2] * b[2 * 4 + j] + a[i * 4 + 3] * b[3 * 4 + j];
}

/* swegl::scale(m, v) (points.cpp:14-30): columns 0..2 of every row */
void scale_cols(const float *m, const float *s, float *r)
{
    std::memcpy(r, m, 16 * sizeof(float));
    for (int row = 0; row < 4; row++) {
        r[row * 4 + 0] *= s[0];
        r[row * 4 + 1] *= s[1];
        r[row * 4 + 2] *= s[2];
    }
}

const float kIdentity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

/* max_z (renderer.cpp:15-19): bytes 0x7F7F7F7F */
float max_z()
{
    uint32_t i = 0x7F7F7F7Fu;
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}

/* ---------------- interpolator_g<1> (interpolator.hpp:52-104) ----------- */
struct interp {
    float topalpha, topstep, ualpha, bottomalpha, bottomstep;
    float v0, v1;
    void init_self(float dist, float z1, float z2)
    {
        v0 = z1;
        v1 = z2 - z1;
        float alphastep = 1.0f / dist;
        bottomalpha = 1.0f / z1;
        float invz2 = 1.0f / z2;
        ualpha = 0.0f;
        topalpha = ualpha;
        topalpha *= bottomalpha;
        topstep = (1.0f * invz2 - topalpha) * alphastep;
        bottomstep = (invz2 - bottomalpha) * alphastep;
    }
    void displace(float move)
    {
        topalpha += topstep * move;
        bottomalpha += bottomstep * move;
        ualpha = topalpha / bottomalpha;
    }
    void step()
    {
        topalpha += topstep;
        bottomalpha += bottomstep;
        ualpha = topalpha / bottomalpha;
    }
    float progress() const { return ualpha; }
    float value() const { return v0 + v1 * progress(); }
};

struct line_side {
    interp it;
    float ratio;
    float x;
};

/* blend(back, front) (colors.cpp:39-57): alpha from the front, integer
 * alpha union, channels mixed in double and truncated to 8 bits */
uint32_t blend(uint32_t back, uint32_t front)
{
    const int a = (int)(front >> 24), ba = (int)(back >> 24);
    const uint32_t na = (uint32_t)(255 - ((255 - a) * (255 - ba) / 255));
    if (a == 255)
        return (front & 0xFFFFFFu) | (na << 24);
    if (a == 0)
        return (back & 0xFFFFFFu) | (na << 24);
    uint32_t out = na << 24;
    for (int c = 0; c < 3; c++) {
        const double v = ((back >> (8 * c)) & 0xFF) * ((256 - a) / 256.0) + ((front >> (8 * c)) & 0xFF) * (a / 256.0);
        out |= (uint32_t)to_u8_d(v) << (8 * c);
    }
    return out;
}

/* ---------------- viewport ---------------------------------------------- */
struct viewport_m {
    int x, y, w, h;
    uint32_t *pixels;
    int pitch_words;
    float *zb;
    uint32_t *tid;
    uint32_t *color64 = nullptr;   /* w*h FP64 re-shade, or NULL */
    /* transparency layers (viewport.hpp:27-62): K of them, each a z-buffer
     * and a colour plane of w*h; layer 0 is the farthest */
    int K = 0;
    std::vector<float> lz;
    std::vector<uint32_t> lc;
    const float *view, *proj;
    vec3 cam;
    float m00, m03, m11, m13;   /* viewport matrix (viewport.cpp:30-35) */
};

/* ---------------- pixel shader: light_and_texture<phong, bilinear> ------ */
struct shader {
    const scene_m *sc;
    const viewport_m *vp;
    const primitive *prim;
    /* phong (pixel_shaders.hpp:42-63) */
    vec3 v0, v1, v2, vleft, vright, vleftdir, vrightdir, v, vdir;
    vec3 n0, n1, n2, nleft, nright, nleftdir, nrightdir, n, ndir;
    /* bilinear (pixel_shaders.hpp:93-118) */
    vec2 t0, t1, t2, side_long_t_dir, side_short_t, side_short_t_dir, t_left, t_dir;
    bool long_line_on_right;
    uint32_t default_bitmap;
    const uint32_t *tbitmap;
    int twidth, theight;
    uint64_t texel_oob = 0;
    /* kind of pixel shader (SWEGL_SHADER_*), split into its light part
     * (0 none, 1 flat, 2 phong) and texture part (0 the material colour of
     * pixel_shader_t, 1 nearest, 2 bilinear) */
    int light_kind = 2, tex_kind = 2;
    bool combined = true;          /* pixel_shader_light_and_texture<L, T> */
    uint32_t material_color;       /* pixel_shader_t::color (pixel_shaders.cpp:14-31) */
    int32_t flat_light;            /* pixel_shader_lights_flat::light, as returned by shade */

    void set_kind(int kind)
    {
        static const int L[10] = {0, 2, 2, 1, 1, 0, 0, 0, 1, 2};
        static const int T[10] = {0, 2, 1, 2, 1, 0, 1, 2, 0, 0};
        light_kind = L[kind];
        tex_kind = T[kind];
        combined = kind >= 1 && kind <= 4;
    }

    /* pixel_shader_t::prepare_for_primitive (pixel_shaders.cpp:14-31), then the
     * texture shaders' (:209-228, :284-302).  A texture shader indexes
     * materials[material_id], UB for -1: rejected before rendering. */
    void prepare_for_primitive(const primitive &p)
    {
        prim = &p;
        material_color = p.material_id != -1 ? sc->materials[p.material_id].color : sc->default_material.color;
        if (tex_kind == 0) {
            tbitmap = &default_bitmap;
            twidth = theight = 1;
            return;
        }
        const swegl_material_desc &m = sc->materials[p.material_id];
        if (m.texture_idx == -1) {
            default_bitmap = m.color;
            tbitmap = &default_bitmap;
            twidth = 1;
            theight = 1;
        } else {
            const swegl_texture_desc &t = sc->textures[m.texture_idx];
            tbitmap = t.bitmap;
            twidth = (int)t.width;
            theight = (int)t.height;
        }
    }
    void prepare_for_triangle(uint32_t i0, uint32_t i1, uint32_t i2, bool inverted)
    {
        const std::vector<mesh_vertex> &vs = prim->vertices;
        /* phong (pixel_shaders.cpp:88-105) */
        v0 = vs[i0].v_world;
        v1 = vs[i1].v_world;
        v2 = vs[i2].v_world;
        if (!inverted) {
            n0 = vs[i0].normal_world;
            n1 = vs[i1].normal_world;
            n2 = vs[i2].normal_world;
        } else {
            n0 = neg(vs[i0].normal_world);
            n1 = neg(vs[i1].normal_world);
            n2 = neg(vs[i2].normal_world);
        }
        /* bilinear (pixel_shaders.cpp:304-318) */
        t0 = vs[i0].tex;
        t1 = vs[i1].tex;
        t2 = vs[i2].tex;
        t0.x *= twidth;
        t0.y *= theight;
        t1.x *= twidth;
        t1.y *= theight;
        t2.x *= twidth;
        t2.y *= theight;
        side_long_t_dir = t2 - t0;
        if (light_kind == 1)
            flat_light = flat(vs[i0].v_world, vs[i1].v_world, vs[i2].v_world, inverted, vp->cam, *sc);
    }
    void prepare_for_upper_triangle(bool lr)
    {
        vleft = v0;
        vright = v0;
        if (lr) {
            vleftdir = v1 - v0;
            vrightdir = v2 - v0;
        } else {
            vleftdir = v2 - v0;
            vrightdir = v1 - v0;
        }
        nleft = n0;
        nright = n0;
        if (lr) {
            nleftdir = n1 - n0;
            nrightdir = n2 - n0;
        } else {
            nleftdir = n2 - n0;
            nrightdir = n1 - n0;
        }
        long_line_on_right = lr;
        side_short_t = t0;
        side_short_t_dir = t1 - t0;
    }
    void prepare_for_lower_triangle(bool lr)
    {
        if (lr) {
            vright = v0;
            vrightdir = v2 - v0;
            vleft = v1;
            vleftdir = v2 - v1;
            nright = n0;
            nrightdir = n2 - n0;
            nleft = n1;
            nleftdir = n2 - n1;
        } else {
            vleft = v0;
            vleftdir = v2 - v0;
            vright = v1;
            vrightdir = v2 - v1;
            nleft = n0;
            nleftdir = n2 - n0;
            nright = n1;
            nrightdir = n2 - n1;
        }
        long_line_on_right = lr;
        side_short_t = t1;
        side_short_t_dir = t2 - t1;
    }
    void prepare_for_scanline(float pl, float pr)
    {
        v = vleft + vleftdir * pl;
        vdir = vright + vrightdir * pr - v;
        n = nleft + nleftdir * pl;
        ndir = nright + nrightdir * pr - n;
        if (long_line_on_right) {
            t_left = side_short_t + side_short_t_dir * pl;
            t_dir = t0 + side_long_t_dir * pr - t_left;
        } else {
            t_left = t0 + side_long_t_dir * pl;
            t_dir = side_short_t + side_short_t_dir * pr - t_left;
        }
    }
    uint32_t shade(float p)
    {
        uint32_t color = material_color;
        if (tex_kind == 1)
            color = nearest(tbitmap, (uint32_t)twidth, (uint32_t)theight, t_left + t_dir * p);
        else if (tex_kind == 2)
            color = bilinear(tbitmap, twidth, theight, t_left + t_dir * p, &texel_oob);
        if (light_kind == 0)
            return color;
        int32_t light;
        if (light_kind == 1) {
            light = flat_light;
        } else {
            vec3 center = v + vdir * p;
            vec3 normal = normalize(n + ndir * p);   /* normal_t(vector_t&&) normalises */
            light = phong(center, normal, vp->cam, *sc);
        }
        return combined ? combine(color, light) : (uint32_t)light;
    }

    /* pixel_shader_texture::shade (pixel_shaders.cpp:275-281): twidth and
     * theight are unsigned there, so (int)t % twidth is an unsigned
     * remainder of the int's bits; column from t.x, row from t.y. */
    static uint32_t nearest(const uint32_t *pc, uint32_t w, uint32_t h, vec2 t)
    {
        uint32_t u = (uint32_t)cvtt(t.x) % w;
        uint32_t v = (uint32_t)cvtt(t.y) % h;
        return pc[v * w + u];
    }

    /* pixel_shader_lights_flat::prepare_for_triangle (pixel_shaders.cpp:33-84):
     * the face normal normal_t(cross(...)) is normalised once, and once more
     * when negated (unary minus builds a normal_t, points.hpp:171-174); the
     * centroid ((v0 + v1) + v2) / 3; then the Phong light sum at the centroid,
     * light = (ambient + sun + lights) * 65536 returned as int. */
    static int32_t flat(vec3 w0, vec3 w1, vec3 w2, bool inverted, vec3 cam, const scene_m &sc)
    {
        vec3 nrm = cross_n(w1 - w0, w2 - w0);
        if (inverted)
            nrm = normalize(neg(nrm));
        vec3 s01 = w0 + w1;
        vec3 s = s01 + w2;
        vec3 center = {s.x / 3, s.y / 3, s.z / 3};
        return phong(center, nrm, cam, sc);
    }

    /* pixel_shader_texture_bilinear::shade (pixel_shaders.cpp:348-384) */
    static uint32_t bilinear(const uint32_t *pc, int twidth, int theight, vec2 t,
                             uint64_t *oob)
    {
        float v = t.x;   /* row coordinate comes from t.x */
        float u = t.y;
        float u1 = (float)((double)u - 0.5);
        float u2 = (float)((double)u + 0.5);
        float v1 = (float)((double)v - 0.5);
        float v2 = (float)((double)v + 0.5);
        u = std::floor(u2);
        v = std::floor(v2);
        int v1m = wrap_add(cvtt(v1), theight) % theight;
        int v2m = v1m + 1;
        if (v2m == theight)
            v2m = 0;
        v1m *= twidth;
        v2m *= twidth;
        int u1m = wrap_add(cvtt(u1), twidth) % twidth;
        int u2m = u1m + 1;
        if (u2m == twidth)
            u2m = 0;
        int64_t lim = (int64_t)twidth * theight;
        auto fetch = [&](int64_t idx) -> uint32_t {
            if (idx < 0 || idx >= lim) {   /* out-of-bounds read in the reference */
                ++*oob;
                idx = idx < 0 ? 0 : lim - 1;
            }
            return pc[idx];
        };
        uint32_t p11 = fetch((int64_t)v1m + u1m), p21 = fetch((int64_t)v2m + u1m);
        uint32_t p12 = fetch((int64_t)v1m + u2m), p22 = fetch((int64_t)v2m + u2m);
        float w11 = (u - u1) * (v - v1);
        float w21 = (u - u1) * (v2 - v);
        float w12 = (u2 - u) * (v - v1);
        float w22 = (u2 - u) * (v2 - v);
        uint32_t out = 0;
        for (int c = 0; c < 4; c++) {
            /* pixel_colors * float, then _mm_add_ps left to right (colors.cpp:11-30) */
            float a = (float)((p11 >> (8 * c)) & 0xFF) * w11;
            float b = (float)((p21 >> (8 * c)) & 0xFF) * w21;
            float d = (float)((p12 >> (8 * c)) & 0xFF) * w12;
            float e = (float)((p22 >> (8 * c)) & 0xFF) * w22;
            float s = ((a + b) + d) + e;
            out |= (uint32_t)to_u8(std::round(s)) << (8 * c);
        }
        return out;
    }

    /* pixel_shader_lights_phong::shade (pixel_shaders.cpp:159-205) */
    static int32_t phong(vec3 center_vertex, vec3 normal, vec3 cam, const scene_m &sc)
    {
        vec3 camera_vector = normalize(cam - center_vertex);
        float face_sun_intensity = -dot(normal, sc.sun_dir);
        if (face_sun_intensity < 0.0f)
            face_sun_intensity = 0.0f;
        else
            face_sun_intensity *= sc.sun_intensity;
        float dynamic_lights_intensity = 0.0f;
        for (const swegl_point_light &psl : sc.lights) {
            vec3 pos = {psl.position[0], psl.position[1], psl.position[2]};
            vec3 light_direction = center_vertex - pos;
            float light_distance_squared = len2(light_direction);
            float diffuse = psl.intensity / light_distance_squared;
            if (diffuse < 0.05)
                continue;
            light_direction = normalize(light_direction);
            float alignment = -dot(normal, light_direction);
            if (alignment < 0.0f)
                continue;
            diffuse *= alignment;
            vec3 reflection = light_direction + normal * (alignment * 2);
            float specular = dot(reflection, camera_vector);
            if (specular > 0) {
                const int p = 32;
                specular = (float)std::pow((double)specular, (double)p);
                specular = specular * p / 2;
                dynamic_lights_intensity += diffuse + specular / light_distance_squared;
            } else {
                dynamic_lights_intensity += diffuse;
            }
        }
        return cvtt(65536 * (sc.ambient + face_sun_intensity + dynamic_lights_intensity));
    }

    /* pixel_shader_light_and_texture::shade (pixel_shaders.hpp:159-178) */
    static uint32_t combine(uint32_t color, int32_t light_i)
    {
        float light = (float)((double)light_i / 65536.0);
        uint8_t c[4] = {(uint8_t)color, (uint8_t)(color >> 8), (uint8_t)(color >> 16),
                        (uint8_t)(color >> 24)};
        if (light < 1) {
            for (int k = 0; k < 3; k++)
                c[k] = to_u8((float)c[k] * light);
        } else {
            light = std::sqrt(light);
            light = std::sqrt(light);
            for (int k = 0; k < 3; k++)
                c[k] = (uint8_t)(255 - to_u8((float)(255 - c[k]) / light));
        }
        return (uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16) |
               ((uint32_t)c[3] << 24);
    }
};

/* ---------------- FP64 re-shade ------------------------------------------
 * The same shading model as `shader` (phong + bilinear + combine) with every
 * shading quantity carried in double: vertex attributes are the FP32 v_world,
 * normal_world and tex coords widened to double, the interpolation
 * parameters (pl, pr, p) are the FP32 raster's, and the model's own integer
 * steps (texel choice, 65536 fixed-point light, 8-bit truncation) are kept.
 * Comparing its output with the FP32 colours measures the colour error of
 * FP32 shading alone, for the same winning triangle of each pixel. */
struct dvec3 { double x, y, z; };
struct dvec2 { double x, y; };
inline dvec3 D(vec3 a) { return {a.x, a.y, a.z}; }
inline dvec3 operator+(dvec3 a, dvec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline dvec3 operator-(dvec3 a, dvec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline dvec3 operator*(dvec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline dvec2 operator+(dvec2 a, dvec2 b) { return {a.x + b.x, a.y + b.y}; }
inline dvec2 operator-(dvec2 a, dvec2 b) { return {a.x - b.x, a.y - b.y}; }
inline dvec2 operator*(dvec2 a, double s) { return {a.x * s, a.y * s}; }
inline double ddot(dvec3 a, dvec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline dvec3 dnormalize(dvec3 a)
{
    double l = std::sqrt(ddot(a, a));
    return l != 0 ? dvec3{a.x / l, a.y / l, a.z / l} : a;
}

struct shader64 {
    const scene_m *sc;
    const viewport_m *vp;
    const shader *sh;   /* texture binding of the FP32 shader */
    dvec3 v0, v1, v2, vleft, vright, vleftdir, vrightdir, v, vdir;
    dvec3 n0, n1, n2, nleft, nright, nleftdir, nrightdir, n, ndir;
    dvec2 t0, t1, t2, side_long_t_dir, side_short_t, side_short_t_dir, t_left, t_dir;
    bool long_line_on_right;

    void prepare_for_triangle(uint32_t i0, uint32_t i1, uint32_t i2, bool inverted)
    {
        const std::vector<mesh_vertex> &vs = sh->prim->vertices;
        const double sgn = inverted ? -1.0 : 1.0;
        v0 = D(vs[i0].v_world);
        v1 = D(vs[i1].v_world);
        v2 = D(vs[i2].v_world);
        n0 = D(vs[i0].normal_world) * sgn;
        n1 = D(vs[i1].normal_world) * sgn;
        n2 = D(vs[i2].normal_world) * sgn;
        const double tw = sh->twidth, th = sh->theight;
        t0 = {vs[i0].tex.x * tw, vs[i0].tex.y * th};
        t1 = {vs[i1].tex.x * tw, vs[i1].tex.y * th};
        t2 = {vs[i2].tex.x * tw, vs[i2].tex.y * th};
        side_long_t_dir = t2 - t0;
    }
    void prepare_for_upper_triangle(bool lr)
    {
        vleft = vright = v0;
        nleft = nright = n0;
        vleftdir = lr ? v1 - v0 : v2 - v0;
        vrightdir = lr ? v2 - v0 : v1 - v0;
        nleftdir = lr ? n1 - n0 : n2 - n0;
        nrightdir = lr ? n2 - n0 : n1 - n0;
        long_line_on_right = lr;
        side_short_t = t0;
        side_short_t_dir = t1 - t0;
    }
    void prepare_for_lower_triangle(bool lr)
    {
        if (lr) {
            vright = v0; vrightdir = v2 - v0; vleft = v1; vleftdir = v2 - v1;
            nright = n0; nrightdir = n2 - n0; nleft = n1; nleftdir = n2 - n1;
        } else {
            vleft = v0; vleftdir = v2 - v0; vright = v1; vrightdir = v2 - v1;
            nleft = n0; nleftdir = n2 - n0; nright = n1; nrightdir = n2 - n1;
        }
        long_line_on_right = lr;
        side_short_t = t1;
        side_short_t_dir = t2 - t1;
    }
    void prepare_for_scanline(double pl, double pr)
    {
        v = vleft + vleftdir * pl;
        vdir = vright + vrightdir * pr - v;
        n = nleft + nleftdir * pl;
        ndir = nright + nrightdir * pr - n;
        if (long_line_on_right) {
            t_left = side_short_t + side_short_t_dir * pl;
            t_dir = t0 + side_long_t_dir * pr - t_left;
        } else {
            t_left = t0 + side_long_t_dir * pl;
            t_dir = side_short_t + side_short_t_dir * pr - t_left;
        }
    }
    uint32_t bilinear(dvec2 t) const
    {
        const uint32_t *pc = sh->tbitmap;
        const int W = sh->twidth, H = sh->theight;
        double v = t.x, u = t.y;   /* row coordinate comes from t.x */
        double u1 = u - 0.5, u2 = u + 0.5, v1 = v - 0.5, v2 = v + 0.5;
        u = std::floor(u2);
        v = std::floor(v2);
        int r1 = wrap_add(cvtt_d(v1), H) % H, r2 = r1 + 1 == H ? 0 : r1 + 1;
        int c1 = wrap_add(cvtt_d(u1), W) % W, c2 = c1 + 1 == W ? 0 : c1 + 1;
        const int64_t lim = (int64_t)W * H;
        auto fetch = [&](int64_t r, int64_t c) {
            int64_t idx = r * W + c;
            idx = idx < 0 ? 0 : idx >= lim ? lim - 1 : idx;
            return pc[idx];
        };
        uint32_t p11 = fetch(r1, c1), p21 = fetch(r2, c1), p12 = fetch(r1, c2), p22 = fetch(r2, c2);
        double w11 = (u - u1) * (v - v1), w21 = (u - u1) * (v2 - v);
        double w12 = (u2 - u) * (v - v1), w22 = (u2 - u) * (v2 - v);
        uint32_t out = 0;
        for (int c = 0; c < 4; c++) {
            double s = (double)((p11 >> (8 * c)) & 0xFF) * w11 + (double)((p21 >> (8 * c)) & 0xFF) * w21 +
                       (double)((p12 >> (8 * c)) & 0xFF) * w12 + (double)((p22 >> (8 * c)) & 0xFF) * w22;
            out |= (uint32_t)to_u8_d(std::round(s)) << (8 * c);
        }
        return out;
    }
    int32_t phong(dvec3 P, dvec3 N) const
    {
        dvec3 C = dnormalize(D(vp->cam) - P);
        double sun = -ddot(N, D(sc->sun_dir));
        sun = sun < 0 ? 0 : sun * sc->sun_intensity;
        double acc = 0;
        for (const swegl_point_light &l : sc->lights) {
            dvec3 L = P - dvec3{l.position[0], l.position[1], l.position[2]};
            double d2 = ddot(L, L);
            double diffuse = l.intensity / d2;
            if (diffuse < 0.05)
                continue;
            L = dnormalize(L);
            double a = -ddot(N, L);
            if (a < 0)
                continue;
            diffuse *= a;
            double s = ddot(L + N * (a * 2), C);
            acc += s > 0 ? diffuse + std::pow(s, 32.0) * 16.0 / d2 : diffuse;
        }
        return cvtt_d(65536.0 * (sc->ambient + sun + acc));
    }
    uint32_t shade(double p) const
    {
        uint32_t color = bilinear(t_left + t_dir * p);
        int32_t li = phong(v + vdir * p, dnormalize(n + ndir * p));
        double light = li / 65536.0;
        uint32_t out = color & 0xFF000000u;
        for (int k = 0; k < 3; k++) {
            double c = (color >> (8 * k)) & 0xFF;
            uint32_t o = light < 1 ? to_u8_d(c * light)
                                   : (uint8_t)(255 - to_u8_d((255 - c) / std::sqrt(std::sqrt(light))));
            out |= o << (8 * k);
        }
        return out;
    }
};

/* ---------------- fragment sinks ---------------------------------------- */
/* The full renderer: z test, shade, write (renderer.cpp:486-499). */
struct render_sink {
    viewport_m *vp;
    shader *sh;
    shader64 *sh64 = nullptr;   /* set with viewport_m::color64 */
    uint32_t seq;
    uint64_t fragments = 0, nan_fragments = 0;
    void span(int y, int x1, int x2, interp &q)
    {
        uint32_t *video = &vp->pixels[(int64_t)y * vp->pitch_words + x1];
        int64_t off = (int64_t)(y - vp->y) * vp->w + (x1 - vp->x);
        for (; x1 < x2; x1++, video++, off++, q.step()) {
            float z = q.value();
            if (z <= 0.001)   /* near rejection, double literal as in :489 */
                continue;
            fragments++;
            if (std::isnan(z))
                nan_fragments++;
            if (z >= vp->zb[off])
                continue;
            uint32_t color = sh->shade(q.progress());
            if (vp->K > 0 && (color >> 24) != 255) {
                insert_layer(off, z, color);   /* transparent: no base write */
                continue;
            }
            *video = color;
            if (sh64)
                vp->color64[off] = sh64->shade(q.progress());
            vp->zb[off] = z;
            if (vp->tid)
                vp->tid[off] = seq;
            if (vp->K > 0)
                drop_layers_behind(off, z);
        }
    }
    /* renderer.cpp:500-515: the first layer that is empty or nearer than z */
    int layer_at(int64_t off, float z) const
    {
        const int64_t n = (int64_t)vp->w * vp->h;
        int k = 0;
        for (; k < vp->K; k++) {
            float lz = vp->lz[k * n + off];
            if (lz == max_z() || lz < z)
                break;
        }
        return k;
    }
    /* renderer.cpp:516-532: an opaque fragment hides the layers behind it */
    void drop_layers_behind(int64_t off, float z)
    {
        const int64_t n = (int64_t)vp->w * vp->h;
        int i = 0, k = layer_at(off, z);
        for (; k < vp->K; i++, k++) {
            vp->lz[i * n + off] = vp->lz[k * n + off];
            vp->lc[i * n + off] = vp->lc[k * n + off];
        }
        for (; i < vp->K; i++) {
            vp->lz[i * n + off] = max_z();
            vp->lc[i * n + off] = 0;
        }
    }
    /* renderer.cpp:533-550: insert a transparent fragment, dropping the
     * farthest layer when all are used */
    void insert_layer(int64_t off, float z, uint32_t color)
    {
        const int64_t n = (int64_t)vp->w * vp->h;
        int k = layer_at(off, z);
        if (vp->lz[(int64_t)(vp->K - 1) * n + off] != max_z()) {
            while (k-- > 0) {
                std::swap(vp->lz[k * n + off], z);
                std::swap(vp->lc[k * n + off], color);
            }
        } else {
            for (; k < vp->K; k++) {
                std::swap(vp->lz[k * n + off], z);
                std::swap(vp->lc[k * n + off], color);
                if (z == max_z())
                    break;
            }
        }
    }
    void prepare_for_triangle(uint32_t i0, uint32_t i1, uint32_t i2, bool inv)
    {
        sh->prepare_for_triangle(i0, i1, i2, inv);
        if (sh64)
            sh64->prepare_for_triangle(i0, i1, i2, inv);
    }
    void prepare_for_upper_triangle(bool lr)
    {
        sh->prepare_for_upper_triangle(lr);
        if (sh64)
            sh64->prepare_for_upper_triangle(lr);
    }
    void prepare_for_lower_triangle(bool lr)
    {
        sh->prepare_for_lower_triangle(lr);
        if (sh64)
            sh64->prepare_for_lower_triangle(lr);
    }
    void prepare_for_scanline(float pl, float pr)
    {
        sh->prepare_for_scanline(pl, pr);
        if (sh64)
            sh64->prepare_for_scanline(pl, pr);
    }
};

/* Coverage-only sink for the single-triangle probe. */
struct cover_sink {
    int w;
    uint32_t *cover;
    float *zlast;
    void span(int y, int x1, int x2, interp &q)
    {
        for (; x1 < x2; x1++, q.step()) {
            int64_t off = (int64_t)y * w + x1;
            cover[off]++;
            zlast[off] = q.value();
        }
    }
    void prepare_for_triangle(uint32_t, uint32_t, uint32_t, bool) {}
    void prepare_for_upper_triangle(bool) {}
    void prepare_for_lower_triangle(bool) {}
    void prepare_for_scanline(float, float) {}
};

/* ---------------- fill_half_triangle (renderer.cpp:462-558) ------------- */
template <class Sink>
void fill_half_triangle(int y, int y_end, line_side &L, line_side &R, int vx, int vw, Sink &sink)
{
    for (; y < y_end; y++) {
        int x1 = std::max(cvtt(std::ceil(L.x)), vx);
        int x2 = std::min(cvtt(std::ceil(R.x)), vx + vw);
        if (x1 < x2) {
            sink.prepare_for_scanline(L.it.progress(), R.it.progress());
            interp q;
            q.init_self(R.x - L.x, L.it.value(), R.it.value());
            q.displace((float)x1 - L.x);
            sink.span(y, x1, x2, q);
        }
        L.x += L.ratio;
        R.x += R.ratio;
        L.it.step();
        R.it.step();
    }
}

/* ---------------- fill_triangle_2 (renderer.cpp:361-460) --------------- */
template <class Sink>
bool fill_triangle_2(const std::vector<mesh_vertex> &vs, uint32_t i0, uint32_t i1, uint32_t i2,
                     int vx, int vy, int vw, int vh, Sink &sink, bool front_face_visible)
{
    const vec3 *v0 = &vs[i0].v_viewport;
    const vec3 *v1 = &vs[i1].v_viewport;
    const vec3 *v2 = &vs[i2].v_viewport;
    bool inverted = !front_face_visible;
    if (v1->y < v0->y) { std::swap(v0, v1); std::swap(i0, i1); }
    if (v2->y < v1->y) { std::swap(v1, v2); std::swap(i1, i2); }
    if (v1->y < v0->y) { std::swap(v0, v1); std::swap(i0, i1); }

    int y0 = cvtt(std::ceil(v0->y));
    int y1 = cvtt(std::ceil(v1->y));
    int y2 = cvtt(std::ceil(v2->y));
    if (y0 == y2)
        return false;

    line_side side_long, side_short;
    side_long.ratio = (v2->x - v0->x) / (v2->y - v0->y);
    side_long.it.init_self(v2->y - v0->y, v0->z, v2->z);
    if (y0 < vy) {
        side_long.it.displace((float)vy - v0->y);
        side_long.x = v0->x + side_long.ratio * ((float)vy - v0->y);
    } else {
        side_long.it.displace((float)y0 - v0->y);
        side_long.x = v0->x + side_long.ratio * ((float)y0 - v0->y);
    }
    sink.prepare_for_triangle(i0, i1, i2, inverted);

    if (y1 >= vy) {
        side_short.ratio = (v1->x - v0->x) / (v1->y - v0->y);
        side_short.it.init_self(v1->y - v0->y, v0->z, v1->z);
        int y = std::max(y0, vy);
        int y_end = std::min(y1, vy + vh);
        side_short.it.displace((float)y - v0->y);
        side_short.x = v0->x + side_short.ratio * ((float)y - v0->y);
        bool lr = side_long.ratio > side_short.ratio;
        sink.prepare_for_upper_triangle(lr);
        if (lr)
            fill_half_triangle(y, y_end, side_short, side_long, vx, vw, sink);
        else
            fill_half_triangle(y, y_end, side_long, side_short, vx, vw, sink);
    }
    if (y1 < vy + vh) {
        side_short.ratio = (v2->x - v1->x) / (v2->y - v1->y);
        side_short.it.init_self(v2->y - v1->y, v1->z, v2->z);
        int y = std::max(y1, vy);
        int y_end = std::min(y2, vy + vh);
        side_short.it.displace((float)y - v1->y);
        side_short.x = v1->x + side_short.ratio * ((float)y - v1->y);
        bool lr = side_long.ratio < side_short.ratio;
        sink.prepare_for_lower_triangle(lr);
        if (lr)
            fill_half_triangle(y, y_end, side_short, side_long, vx, vw, sink);
        else
            fill_half_triangle(y, y_end, side_long, side_short, vx, vw, sink);
    }
    return true;
}

/* ---------------- vertex stage (vertex_shaders.hpp) --------------------- */
/* camera_to_frustum (:61-71).  normal_world is normalised three times:
 * rotate() builds a normal_t, .normalize(), then normal_t::operator=. */
void camera_to_frustum(mesh_vertex &mv, const node &nd, const viewport_m &vp)
{
    mv.normal_world = normalize(normalize(rotate_n(mv.normal, nd.rs)));
    mv.v_viewport = transform(mv.v_viewport, vp.proj);
    if (mv.v_viewport.z != 0) {
        mv.v_viewport.x /= std::fabs(mv.v_viewport.z);
        mv.v_viewport.y /= std::fabs(mv.v_viewport.z);
    }
}
/* viewport_t::transform(mesh_vertex_t&) (viewport.cpp:123-129) */
inline void viewport_transform(mesh_vertex &mv, const viewport_m &vp)
{
    mv.v_viewport.x = vp.m00 * mv.v_viewport.x + vp.m03;
    mv.v_viewport.y = vp.m11 * mv.v_viewport.y + vp.m13;
}
/* world_to_viewport (:54-59) */
void world_to_viewport(mesh_vertex &mv, const node &nd, const viewport_m &vp)
{
    mv.v_viewport = transform(mv.v_world, vp.view);
    camera_to_frustum(mv, nd, vp);
    viewport_transform(mv, vp);
}

/* ---------------- visibility marking (renderer.cpp:58-75) --------------- */
bool inside_camera_frustum(const mesh_vertex &a, const mesh_vertex &b, const mesh_vertex &c)
{
    const vec3 &p = a.v_viewport, &q = b.v_viewport, &r = c.v_viewport;
    return (p.x >= -1 || q.x >= -1 || r.x >= -1) && (p.y >= -1 || q.y >= -1 || r.y >= -1) &&
           (p.x < 1 || q.x < 1 || r.x < 1) && (p.y < 1 || q.y < 1 || r.y < 1) &&
           (p.z >= 0.001 || q.z >= 0.001 || r.z >= 0.001) && (p.x != q.x || p.x != r.x) &&
           (p.y != q.y || p.y != r.y);
}
bool front_face_visible(const mesh_vertex &a, const mesh_vertex &b, const mesh_vertex &c)
{
    return cross_n(b.v_viewport - a.v_viewport, c.v_viewport - a.v_viewport).z > 0;
}

/* build a new clip vertex between a and b at parameter cut (renderer.cpp:290-297) */
void make_clip_vertex(std::vector<mesh_vertex> &vs, uint32_t ia, uint32_t ib, float cut,
                      const node &nd, const viewport_m &vp)
{
    mesh_vertex nv{};
    const mesh_vertex &a = vs[ia], &b = vs[ib];
    nv.v_world = a.v_world + (b.v_world - a.v_world) * cut;
    nv.tex = a.tex + (b.tex - a.tex) * cut;
    if (same(b.normal, a.normal))
        nv.normal = a.normal;
    else
        nv.normal = normalize(a.normal + (b.normal - a.normal) * cut);
    nv.yes = false;
    world_to_viewport(nv, nd, vp);
    vs.push_back(nv);
}

/* ---------------- fill_triangle (renderer.cpp:240-359) ------------------ */
struct draw_ctx {
    viewport_m *vp;
    render_sink *sink;
    uint64_t drawn = 0;
    uint64_t clipped = 0;
};

void fill_triangle(uint32_t i0, uint32_t i1, uint32_t i2, const node &nd, primitive &prim,
                   draw_ctx &dc, uint32_t tri_index)
{
    std::vector<mesh_vertex> &vs = prim.vertices;
    if (!vs[i0].yes || !vs[i1].yes || !vs[i2].yes)
        return;
    viewport_m &vp = *dc.vp;
    vec3 v0 = vs[i0].v_viewport, v1 = vs[i1].v_viewport, v2 = vs[i2].v_viewport;
    bool front = cross_n(v1 - v0, v2 - v0).z < 0;
    bool inverted_order = false;
    if (v1.z > v0.z) { std::swap(v0, v1); std::swap(i0, i1); inverted_order = !inverted_order; }
    if (v2.z > v1.z) { std::swap(v1, v2); std::swap(i1, i2); inverted_order = !inverted_order; }
    if (v1.z > v0.z) { std::swap(v0, v1); std::swap(i0, i1); inverted_order = !inverted_order; }

    const uint32_t seq0 = 2 * tri_index;
    if (v2.z >= 0.001) {
        dc.sink->seq = seq0;
        dc.drawn += fill_triangle_2(vs, i0, i1, i2, vp.x, vp.y, vp.w, vp.h, *dc.sink, front);
    } else if (v1.z < 0.001) {
        /* only v0 in front of the camera (:286-317) */
        dc.clipped++;
        float cut_1 = (v0.z - 0.001f) / (v0.z - v1.z);
        make_clip_vertex(vs, i0, i1, cut_1, nd, vp);
        float cut_2 = (v0.z - 0.001f) / (v0.z - v2.z);
        make_clip_vertex(vs, i0, i2, cut_2, nd, vp);
        uint32_t n1 = (uint32_t)vs.size() - 2, n2 = (uint32_t)vs.size() - 1;
        front = cross_n(vs[n1].v_viewport - v0, vs[n2].v_viewport - v0).z < 0;
        if (inverted_order)
            front = !front;
        dc.sink->seq = seq0;
        dc.drawn += fill_triangle_2(vs, i0, n1, n2, vp.x, vp.y, vp.w, vp.h, *dc.sink, front);
        vs.pop_back();
        vs.pop_back();
    } else if (v2.z < 0.001) {
        /* only v2 behind the camera (:318-356) */
        dc.clipped++;
        float cut_0 = (v0.z - 0.001f) / (v0.z - v2.z);
        make_clip_vertex(vs, i0, i2, cut_0, nd, vp);
        float cut_1 = (v1.z - 0.001f) / (v1.z - v2.z);
        make_clip_vertex(vs, i1, i2, cut_1, nd, vp);
        uint32_t n1 = (uint32_t)vs.size() - 2, n2 = (uint32_t)vs.size() - 1;
        front = cross_n(v1 - v0, vs[n2].v_viewport - v0).z < 0;
        if (inverted_order)
            front = !front;
        dc.sink->seq = seq0;
        dc.drawn += fill_triangle_2(vs, i0, i1, n2, vp.x, vp.y, vp.w, vp.h, *dc.sink, front);
        front = cross_n(vs[n2].v_viewport - v0, vs[n1].v_viewport - v0).z < 0;
        if (inverted_order)
            front = !front;
        dc.sink->seq = seq0 + 1;
        dc.drawn += fill_triangle_2(vs, i0, n2, n1, vp.x, vp.y, vp.w, vp.h, *dc.sink, front);
        vs.pop_back();
        vs.pop_back();
    }
}

/* ---------------- DoF (build spec of post_shader_depth_box) ------------- */
/* remap_clipped (lerp.hpp:24-43) */
float inv_lerp_clipped(float a, float b, float x)
{
    if (a == b)
        return 0.5;
    if (x <= a)
        return 0;
    if (x >= b)
        return 1;
    return (x - a) / (b - a);
}
float lerp_clipped(float a, float b, float x)
{
    if (x <= 0)
        return a;
    if (x >= b)
        return b;
    return a + b * x;
}
float remap_clipped(float a, float b, float u, float v, float t)
{
    return lerp_clipped(u, v, inv_lerp_clipped(a, b, t));
}

void dof(const uint32_t *src, int src_pitch, const float *z, int w, int h, float fdist,
         float fdepth, int rmax, bool taps9, uint32_t *dst, int dst_pitch)
{
    /* translate_z_to_blur_factor (post_shaders.hpp:63-70): the float factor
     * is kept; the radius is its truncation and a tap counts as unfocused
     * when its float factor is non-zero (:84, :95). */
    std::vector<float> bf((size_t)w * h);
    for (int i = 0; i < w * h; i++)
        bf[i] = remap_clipped(1.0f, fdepth, 0.0f, (float)rmax, std::fabs(fdist - z[i]));
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            uint32_t s = src[(int64_t)y * src_pitch + x];
            uint32_t &d = dst[(int64_t)y * dst_pitch + x];
            int radius = cvtt(bf[(size_t)y * w + x]);
            if (radius == 0) {
                d = s;
                continue;
            }
            int b = 0, g = 0, r = 0, count = 0;
            auto tap = [&](int i, int j) {
                if (bf[(size_t)j * w + i] != 0) {
                    count++;
                    uint32_t p = src[(int64_t)j * src_pitch + i];
                    b += p & 0xFF;
                    g += (p >> 8) & 0xFF;
                    r += (p >> 16) & 0xFF;
                }
            };
            if (!taps9) {
                for (int j = std::max(0, y - radius); j < std::min(h, y + radius); j++)
                    for (int i = std::max(0, x - radius); i < std::min(w, x + radius); i++)
                        tap(i, j);
            } else {
                const int off[3] = {-radius, 0, radius};
                for (int a = 0; a < 3; a++) {
                    int j = y + off[a];
                    if (j < 0 || j >= h)
                        continue;
                    for (int c = 0; c < 3; c++) {
                        int i = x + off[c];
                        if (i < 0 || i >= w)
                            continue;
                        tap(i, j);
                    }
                }
            }
            if (count)
                d = (uint32_t)(uint8_t)(b / count) | ((uint32_t)(uint8_t)(g / count) << 8) |
                    ((uint32_t)(uint8_t)(r / count) << 16) | (255u << 24);
            else
                d = s;
        }
    }
}

/* ---------------- scene import / validation ----------------------------- */
int import_scene(const swegl_scene_desc *d, scene_m &sc)
{
    sc.nodes.resize(d->n_nodes);
    for (uint32_t i = 0; i < d->n_nodes; i++) {
        const swegl_node_desc &nd = d->nodes[i];
        node &n = sc.nodes[i];
        scale_cols(nd.rotation, nd.scale, n.rs);
        std::memcpy(n.local, n.rs, sizeof n.local);
        n.local[3] += nd.translation[0];
        n.local[7] += nd.translation[1];
        n.local[11] += nd.translation[2];
        std::memcpy(n.o2w, kIdentity, sizeof n.o2w);
        if ((uint64_t)nd.first_child + nd.n_children > d->n_child_indices)
            return SWEGL_EINVAL;
        for (uint32_t c = 0; c < nd.n_children; c++) {
            int ci = d->child_indices[nd.first_child + c];
            if (ci < 0 || (uint32_t)ci >= d->n_nodes)
                return SWEGL_EINVAL;
            n.children.push_back(ci);
        }
        if ((uint64_t)nd.first_primitive + nd.n_primitives > d->n_primitives)
            return SWEGL_EINVAL;
        for (uint32_t p = 0; p < nd.n_primitives; p++) {
            const swegl_primitive_desc &pd = d->primitives[nd.first_primitive + p];
            primitive pr;
            pr.mode = pd.mode;
            pr.material_id = pd.material_id;
            /* points and lines are read but never drawn (renderer.cpp:83-185,
             * 197-229 only handle the triangle modes) */
            if (pd.mode < SWEGL_POINTS || pd.mode > SWEGL_TRIANGLE_FAN)
                return SWEGL_EINVAL;
            if (pd.material_id < -1 || pd.material_id >= (int)d->n_materials)
                return SWEGL_EINVAL;
            int ti = pd.material_id == -1 ? -1 : d->materials[pd.material_id].texture_idx;
            if (ti < -1 || ti >= (int)d->n_textures)
                return SWEGL_EINVAL;
            pr.vertices.resize(pd.n_vertices);
            pr.vertices.reserve(pd.n_vertices + 2);
            for (uint32_t v = 0; v < pd.n_vertices; v++) {
                mesh_vertex &mv = pr.vertices[v];
                mv.v = {pd.positions[3 * v], pd.positions[3 * v + 1], pd.positions[3 * v + 2]};
                mv.normal = {pd.normals[3 * v], pd.normals[3 * v + 1], pd.normals[3 * v + 2]};
                mv.tex = {pd.tex_coords[2 * v], pd.tex_coords[2 * v + 1]};
                mv.yes = false;
            }
            pr.indices.assign(pd.indices, pd.indices + pd.n_indices);
            for (uint32_t idx : pr.indices)
                if (idx >= pd.n_vertices)
                    return SWEGL_EINVAL;
            n.prims.push_back(std::move(pr));
        }
    }
    std::vector<char> reach(d->n_nodes, 0);
    std::vector<int> stack;
    for (uint32_t r = 0; r < d->n_root_nodes; r++) {
        int ri = d->root_nodes[r];
        if (ri < 0 || (uint32_t)ri >= d->n_nodes)
            return SWEGL_EINVAL;
        sc.roots.push_back(ri);
        stack.push_back(ri);
    }
    size_t guard = 0;
    while (!stack.empty()) {
        int i = stack.back();
        stack.pop_back();
        reach[i] = 1;
        for (int c : sc.nodes[i].children)
            stack.push_back(c);
        if (++guard > (size_t)64 * (d->n_nodes + 1))
            return SWEGL_EINVAL;   /* cycle */
    }
    for (uint32_t i = 0; i < d->n_nodes; i++)
        if (!reach[i])
            return SWEGL_EINVAL;   /* v_world would never be computed */
    sc.materials.assign(d->materials, d->materials + d->n_materials);
    sc.default_material = d->default_material;
    sc.textures.assign(d->textures, d->textures + d->n_textures);
    for (const swegl_texture_desc &t : sc.textures)
        if (t.width == 0 || t.height == 0 || !t.bitmap)
            return SWEGL_EINVAL;
    sc.lights.assign(d->point_lights, d->point_lights + d->n_point_lights);
    sc.ambient = d->ambient_light_intensity;
    sc.sun_dir = {d->sun_direction[0], d->sun_direction[1], d->sun_direction[2]};
    sc.sun_intensity = d->sun_intensity;
    return SWEGL_OK;
}

/* original_to_world (vertex_shaders.hpp:16-33) */
void original_to_world(scene_m &sc, int ni, const float *parent)
{
    node &n = sc.nodes[ni];
    matmul(parent, n.local, n.o2w);
    for (primitive &p : n.prims)
        for (mesh_vertex &mv : p.vertices)
            mv.v_world = transform(mv.v, n.o2w);
    for (int c : n.children)
        original_to_world(sc, c, n.o2w);
}

/* _render (renderer.cpp:77-235) */
void render_viewport(scene_m &sc, viewport_m &vp, const oracle_viewport &ov, oracle_stats *st)
{
    /* world_to_camera_or_frustum (vertex_shaders.hpp:35-52) */
    for (node &n : sc.nodes)
        for (primitive &p : n.prims)
            for (mesh_vertex &mv : p.vertices) {
                mv.yes = false;
                mv.v_viewport = transform(mv.v_world, vp.view);
                camera_to_frustum(mv, n, vp);
            }
    /* clear (viewport.cpp:88-113) */
    for (int j = vp.y; j < vp.y + vp.h; j++)
        for (int i = vp.x; i < vp.x + vp.w; i++)
            vp.pixels[(int64_t)j * vp.pitch_words + i] = 0;
    const float mz = max_z();
    for (int i = 0; i < vp.w * vp.h; i++) {
        vp.zb[i] = mz;
        if (vp.tid)
            vp.tid[i] = 0xFFFFFFFFu;
        if (vp.color64)
            vp.color64[i] = 0;
    }
    vp.lz.assign((size_t)vp.K * vp.w * vp.h, max_z());   /* viewport.cpp:106-110 */
    vp.lc.assign((size_t)vp.K * vp.w * vp.h, 0u);

    shader sh{};
    sh.sc = &sc;
    sh.vp = &vp;
    sh.set_kind(ov.pixel_shader);
    shader64 sh64{};
    sh64.sc = &sc;
    sh64.vp = &vp;
    sh64.sh = &sh;
    render_sink sink;
    sink.vp = &vp;
    sink.sh = &sh;
    if (vp.color64)
        sink.sh64 = &sh64;
    draw_ctx dc;
    dc.vp = &vp;
    dc.sink = &sink;
    uint32_t tri_index = 0;

    for (node &n : sc.nodes) {
        for (primitive &p : n.prims) {
            std::vector<mesh_vertex> &V = p.vertices;
            const std::vector<uint32_t> &I = p.indices;
            const bool ds = p.material_id != -1 && sc.materials[p.material_id].double_sided;
            const size_t ni = I.size();
            auto mark = [&](uint32_t a, uint32_t b, uint32_t c, uint32_t fa, uint32_t fb,
                            uint32_t fc) {
                if (inside_camera_frustum(V[a], V[b], V[c]) &&
                    (ds || front_face_visible(V[fa], V[fb], V[fc]))) {
                    V[a].yes = true;
                    V[b].yes = true;
                    V[c].yes = true;
                }
            };
            if (p.mode == SWEGL_TRIANGLE_STRIP)
                for (size_t i = 2; i < ni; i++)
                    mark(I[i - 2], I[i - 1], I[i], I[i - 2], I[i - 1 + (i & 1)], I[i - (i & 1)]);
            if (p.mode == SWEGL_TRIANGLE_FAN)
                for (size_t i = 2; i < ni; i++)
                    mark(I[0], I[i - 1], I[i], I[0], I[i - 1], I[i]);
            if (p.mode == SWEGL_TRIANGLES)
                for (size_t i = 2; i < ni; i += 3)
                    mark(I[i - 2], I[i - 1], I[i], I[i - 2], I[i - 1], I[i]);
        }
        /* frustum_to_viewport(node) (vertex_shaders.hpp:72-84) */
        for (primitive &p : n.prims)
            for (mesh_vertex &mv : p.vertices)
                if (mv.yes)
                    viewport_transform(mv, vp);
        for (primitive &p : n.prims) {
            sh.prepare_for_primitive(p);
            const std::vector<uint32_t> &I = p.indices;
            const size_t ni = I.size();
            if (p.mode == SWEGL_TRIANGLE_STRIP)
                for (size_t i = 2; i < ni; i++)
                    fill_triangle(I[i - 2], I[i - 1 + (i & 1)], I[i - (i & 1)], n, p, dc, tri_index++);
            if (p.mode == SWEGL_TRIANGLE_FAN)
                for (size_t i = 2; i < ni; i++)
                    fill_triangle(I[0], I[i - 1], I[i], n, p, dc, tri_index++);
            if (p.mode == SWEGL_TRIANGLES)
                for (size_t i = 2; i < ni; i += 3)
                    fill_triangle(I[i - 2], I[i - 1], I[i], n, p, dc, tri_index++);
        }
    }

    if (vp.K > 0) {
        /* viewport_t::flatten (viewport.cpp:43-86): layers 1..K-1 blended into
         * layer 0, then layer 0 over the screen, which is read from the
         * surface's origin, not the viewport's (:61); post_shader_t::shade
         * copies layer 0 back to rows y + j but columns from 0 (post_shaders.hpp:22-30) */
        const int64_t n = (int64_t)vp.w * vp.h;
        for (int k = 1; k < vp.K; k++)
            for (int64_t p = 0; p < n; p++)
                if (vp.lc[k * n + p] >> 24)
                    vp.lc[p] = blend(vp.lc[p], vp.lc[k * n + p]);
        for (int j = 0; j < vp.h; j++)
            for (int i = 0; i < vp.w; i++) {
                uint32_t &front = vp.lc[(int64_t)j * vp.w + i];
                const uint32_t back = vp.pixels[(int64_t)j * vp.pitch_words + i];
                front = (front >> 24) ? blend(back, front) : back;
            }
        for (int j = 0; j < vp.h; j++)
            for (int i = 0; i < vp.w; i++)
                vp.pixels[(int64_t)(j + vp.y) * vp.pitch_words + i] = vp.lc[(int64_t)j * vp.w + i];
    }
    if (ov.post_shader == SWEGL_POST_DEPTH_BOX || ov.post_shader == SWEGL_POST_DEPTH_TAP9) {
        std::vector<uint32_t> src((size_t)vp.w * vp.h);
        for (int j = 0; j < vp.h; j++)
            for (int i = 0; i < vp.w; i++)
                src[(size_t)j * vp.w + i] = vp.pixels[(int64_t)(j + vp.y) * vp.pitch_words + vp.x + i];
        dof(src.data(), vp.w, vp.zb, vp.w, vp.h, ov.focal_distance, ov.focal_depth,
            ov.dof_max_radius, ov.post_shader == SWEGL_POST_DEPTH_TAP9,
            vp.pixels + (int64_t)vp.y * vp.pitch_words + vp.x, vp.pitch_words);
    }
    if (st) {
        st->fragments = sink.fragments;
        st->nan_fragments = sink.nan_fragments;
        st->drawn_subtriangles = dc.drawn;
        st->texel_oob = sh.texel_oob;
        st->clipped_triangles = dc.clipped;
    }
}

}  // namespace

/* ======================== exported entry points ========================= */
extern "C" {

static int render_scene(const swegl_scene_desc *scene, oracle_viewport *vps, uint32_t n, oracle_stats *stats,
                        scene_m &sc)
{
    int rc = import_scene(scene, sc);
    if (rc)
        return rc;
    for (uint32_t i = 0; i < n; i++) {
        const oracle_viewport &o = vps[i];
        if (o.w <= 0 || o.h <= 0 || o.x < 0 || o.y < 0 || o.x + o.w > o.surface_w ||
            o.y + o.h > o.surface_h || o.pitch < 4 * o.surface_w || (o.pitch & 3))
            return SWEGL_EINVAL;
        if (o.post_shader < SWEGL_POST_NULL || o.post_shader > SWEGL_POST_DEPTH_TAP9)
            return SWEGL_ENOTSUP;
        if (o.pixel_shader < SWEGL_SHADER_PHONG_BILINEAR || o.pixel_shader > SWEGL_SHADER_LIGHTS_PHONG)
            return SWEGL_ENOTSUP;
        if (o.color64 && (o.pixel_shader != SWEGL_SHADER_PHONG_BILINEAR || o.transparency_layer_count))
            return SWEGL_ENOTSUP;
        if (o.transparency_layer_count < 0 || o.transparency_layer_count > SWEGL_MAX_TRANSPARENCY_LAYERS ||
            (o.transparency_layer_count > 0 && o.post_shader != SWEGL_POST_NULL))
            return SWEGL_ENOTSUP;
        const bool textured = o.pixel_shader <= SWEGL_SHADER_FLAT_NEAREST ||
                              o.pixel_shader == SWEGL_SHADER_TEXTURE_NEAREST ||
                              o.pixel_shader == SWEGL_SHADER_TEXTURE_BILINEAR;
        if (textured)
            for (const node &nd : sc.nodes)
                for (const primitive &pr : nd.prims)
                    if (pr.material_id == -1)
                        return SWEGL_EINVAL;   /* materials[-1] in the texture shader */
    }
    for (int r : sc.roots)
        original_to_world(sc, r, kIdentity);
    for (uint32_t i = 0; i < n; i++) {
        oracle_viewport &o = vps[i];
        viewport_m vp;
        vp.x = o.x;
        vp.y = o.y;
        vp.w = o.w;
        vp.h = o.h;
        vp.pixels = o.surface;
        vp.pitch_words = o.pitch / 4;
        vp.zb = o.zbuffer;
        vp.tid = o.triangle_id;
        vp.color64 = o.color64;
        vp.K = o.transparency_layer_count;
        vp.view = o.view;
        vp.proj = o.projection;
        vp.cam = {o.camera_position[0], o.camera_position[1], o.camera_position[2]};
        /* viewport_t ctor (viewport.cpp:30-35) */
        vp.m03 = o.x + o.w / 2.0f;
        vp.m13 = o.y + o.h / 2.0f;
        vp.m00 = o.w / 2.0f;
        vp.m11 = -o.h / 2.0f;
        render_viewport(sc, vp, o, stats ? &stats[i] : nullptr);
    }
    return SWEGL_OK;
}

int oracle_render(const swegl_scene_desc *scene, oracle_viewport *vps, uint32_t n, oracle_stats *stats)
{
    scene_m sc;
    return render_scene(scene, vps, n, stats, sc);
}

/* The scene fields render() writes (model.hpp:20-29) after rendering one
 * viewport: v_world, normal_world, v_viewport and yes of every vertex, in
 * node, primitive, vertex order (n_vertices entries; xyz triples). */
int oracle_vertex_state(const swegl_scene_desc *scene, oracle_viewport *vp, float *v_world,
                        float *normal_world, float *v_viewport, uint8_t *yes)
{
    scene_m sc;
    int rc = render_scene(scene, vp, 1, nullptr, sc);
    if (rc)
        return rc;
    size_t k = 0;
    for (const node &nd : sc.nodes)
        for (const primitive &pr : nd.prims)
            for (const mesh_vertex &mv : pr.vertices) {
                const vec3 *src[3] = {&mv.v_world, &mv.normal_world, &mv.v_viewport};
                float *dst[3] = {v_world, normal_world, v_viewport};
                for (int a = 0; a < 3; a++) {
                    dst[a][3 * k] = src[a]->x;
                    dst[a][3 * k + 1] = src[a]->y;
                    dst[a][3 * k + 2] = src[a]->z;
                }
                yes[k] = mv.yes ? 1 : 0;
                k++;
            }
    return SWEGL_OK;
}

void oracle_transform(const float v[3], const float m[16], float out[3])
{
    vec3 r = transform({v[0], v[1], v[2]}, m);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void oracle_vertex(const float v[3], const float n[3], const float M[16], const float RS[16],
                   const float view[16], const float proj[16], float v_world[3],
                   float n_world[3], float v_frustum[3])
{
    node nd;
    std::memcpy(nd.rs, RS, sizeof nd.rs);
    viewport_m vp{};
    vp.view = view;
    vp.proj = proj;
    mesh_vertex mv{};
    mv.v = {v[0], v[1], v[2]};
    mv.normal = {n[0], n[1], n[2]};
    mv.v_world = transform(mv.v, M);
    mv.v_viewport = transform(mv.v_world, view);
    camera_to_frustum(mv, nd, vp);
    v_world[0] = mv.v_world.x;
    v_world[1] = mv.v_world.y;
    v_world[2] = mv.v_world.z;
    n_world[0] = mv.normal_world.x;
    n_world[1] = mv.normal_world.y;
    n_world[2] = mv.normal_world.z;
    v_frustum[0] = mv.v_viewport.x;
    v_frustum[1] = mv.v_viewport.y;
    v_frustum[2] = mv.v_viewport.z;
}

/* make_clip_vertex between a and b at `cut` (renderer.cpp:290-297): a, b =
 * {v_world xyz, normal xyz, tex uv}; node rotation/scale RS, view and
 * projection, viewport matrix {m00, m03, m11, m13}.  out = {v_world xyz,
 * tex uv, normal_world xyz, v_viewport xyz}. */
void oracle_clip_vertex(const float a[8], const float b[8], float cut, const float RS[16],
                        const float view[16], const float proj[16], const float vpm[4],
                        float out[11])
{
    node nd;
    std::memcpy(nd.rs, RS, sizeof nd.rs);
    viewport_m vp{};
    vp.view = view;
    vp.proj = proj;
    vp.m00 = vpm[0];
    vp.m03 = vpm[1];
    vp.m11 = vpm[2];
    vp.m13 = vpm[3];
    std::vector<mesh_vertex> vs(2);
    const float *src[2] = {a, b};
    for (int i = 0; i < 2; i++) {
        vs[i].v_world = {src[i][0], src[i][1], src[i][2]};
        vs[i].normal = {src[i][3], src[i][4], src[i][5]};
        vs[i].tex = {src[i][6], src[i][7]};
    }
    make_clip_vertex(vs, 0, 1, cut, nd, vp);
    const mesh_vertex &nv = vs[2];
    const float o[11] = {nv.v_world.x, nv.v_world.y, nv.v_world.z, nv.tex.x, nv.tex.y,
                         nv.normal_world.x, nv.normal_world.y, nv.normal_world.z,
                         nv.v_viewport.x, nv.v_viewport.y, nv.v_viewport.z};
    std::memcpy(out, o, sizeof o);
}

void oracle_interp(float d, float z1, float z2, float m, int steps, float out[2])
{
    interp it;
    it.init_self(d, z1, z2);
    it.displace(m);
    for (int i = 0; i < steps; i++)
        it.step();
    out[0] = it.progress();
    out[1] = it.value();
}

int oracle_raster_triangle(const float v[9], int w, int h, uint32_t *cover, float *zlast)
{
    std::vector<mesh_vertex> vs(3);
    for (int i = 0; i < 3; i++)
        vs[i].v_viewport = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    cover_sink sink{w, cover, zlast};
    return fill_triangle_2(vs, 0, 1, 2, 0, 0, w, h, sink, true) ? 1 : 0;
}

uint32_t oracle_bilinear(const uint32_t *bitmap, int w, int h, float tx, float ty)
{
    uint64_t oob = 0;
    return shader::bilinear(bitmap, w, h, {tx, ty}, &oob);
}

int32_t oracle_phong(const float center[3], const float normal[3], const float cam[3],
                     float ambient, const float sun_dir[3], float sun_intensity,
                     const swegl_point_light *lights, int n_lights)
{
    scene_m sc;
    sc.ambient = ambient;
    sc.sun_dir = {sun_dir[0], sun_dir[1], sun_dir[2]};
    sc.sun_intensity = sun_intensity;
    sc.lights.assign(lights, lights + n_lights);
    return shader::phong({center[0], center[1], center[2]}, {normal[0], normal[1], normal[2]},
                         {cam[0], cam[1], cam[2]}, sc);
}

uint32_t oracle_combine(uint32_t color, int32_t light) { return shader::combine(color, light); }

void oracle_dof(const uint32_t *src, const float *z, int w, int h, float focal_distance,
                float focal_depth, int max_radius, int taps9, uint32_t *dst)
{
    dof(src, w, z, w, h, focal_distance, focal_depth, max_radius, taps9 != 0, dst, w);
}

int32_t oracle_cvtt(float f) { return cvtt(f); }

uint32_t oracle_nearest(const uint32_t *bitmap, int w, int h, float tx, float ty)
{
    return shader::nearest(bitmap, (uint32_t)w, (uint32_t)h, {tx, ty});
}

}  // extern "C"
