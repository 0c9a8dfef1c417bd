// shade.h -- pixel_shader_light_and_texture<pixel_shader_lights_phong,
// pixel_shader_texture_bilinear> for gfx950 (swegl/render/pixel_shaders.hpp,
// src/render/pixel_shaders.cpp), same float operation order as the CPU.
#pragma once
#include "core.h"

namespace sw {

// (float)pow((double)s, 32) (pixel_shaders.cpp:193): glibc's pow rounds
// s^32 to double, the store rounds that to float.  Five plain double
// squarings give s^32 within 31 double-rounding errors (relative 2^-48);
// when every value within 2^-46 of it rounds to the same float, that float
// is the answer.  Otherwise (values near a float rounding boundary) s^32 is
// recomputed in double-double by exact squarings (two-product through fma)
// and rounded once to double and then to float.
SW_HD float pow32(float s)
{
    double p = (double)s;
    p *= p;
    p *= p;
    p *= p;
    p *= p;
    p *= p;
    const float lo = (float)(p * (1.0 - 0x1p-46)), hi = (float)(p * (1.0 + 0x1p-46));
    if (lo == hi)
        return lo;
    double h = (double)s, l = 0.0;
    for (int i = 0; i < 5; i++) {
        double q = h * h;
        double e = fma(h, h, -q);
        e += 2.0 * h * l;
        double h2 = q + e;
        l = e - (h2 - q);
        h = h2;
    }
    return (float)(h + l);
}

struct Light4 { float x, y, z, I; };

// pixel_shader_lights_phong::shade (pixel_shaders.cpp:159-205)
SW_D int32_t phong(f3 center, f3 normal, f3 cam, float ambient, f3 sun, float sun_I,
                   const Light4 *lights, int n_lights)
{
    f3 camv = normalize(cam - center);
    float fs = -dot(normal, sun);
    if (fs < 0.0f)
        fs = 0.0f;
    else
        fs *= sun_I;
    float acc = 0.0f;
    for (int i = 0; i < n_lights; i++) {
        const Light4 l = ro(lights)[i];
        f3 L = center - f3{l.x, l.y, l.z};
        float d2 = L.x * L.x + L.y * L.y + L.z * L.z;
        float diffuse = l.I / d2;
        if (diffuse < SW_DIFFUSE_CUT)
            continue;
        L = normalize(L);
        float a = -dot(normal, L);
        if (a < 0.0f)
            continue;
        diffuse *= a;
        f3 R = L + normal * (a * 2.0f);
        float s = dot(R, camv);
        if (s > 0) {
            s = pow32(s);
            s = s * 32.0f / 2.0f;
            acc += diffuse + s / d2;
        } else {
            acc += diffuse;
        }
    }
    return cvtt(65536.0f * (ambient + fs + acc));
}

// pixel_shader_texture_bilinear::shade (pixel_shaders.cpp:348-384).  The
// reference reads outside the bitmap for coordinates that wrap to a negative
// index (UB); here such a tap is clamped and counted.
// x % m (C remainder); a mask when m is a power of two and x >= 0.
SW_D int rem(int x, int m)
{
    if (x >= 0 && (m & (m - 1)) == 0)
        return x & (m - 1);
    return x % m;
}

// A texture as the shaders read it: w x h texels at p, or, with p null, the
// 1x1 texture holding the material colour (pixel_shaders.cpp:288-302).
// Keeping that colour by value (not behind a pointer to a local) keeps the
// kernels free of private memory.
struct TexView {
    const uint32_t *p;
    uint32_t one;
    int w, h;
};
SW_D uint32_t texel(const TexView &X, int64_t i) { return X.p ? ro(X.p)[i] : X.one; }

// The four texels and weights of one bilinear sample.
struct BilTaps {
    uint32_t p[4];
    float w[4];
};

SW_D BilTaps bilinear_taps(const TexView &X, f2 t, uint32_t &oob)
{
    const int tw = X.w, th = X.h;
    float v = t.x;
    float u = t.y;
    // (float)((double)u - 0.5), float minus double literal stored to float,
    // equals the float operation u - 0.5f: the double difference is exact
    // whenever |u| >= 2^-30, and below that both round to -0.5 (or +0.5).
    float u1 = u - 0.5f;
    float u2 = u + 0.5f;
    float v1 = v - 0.5f;
    float v2 = v + 0.5f;
    u = floorf(u2);
    v = floorf(v2);
    int v1m = rem(wrap_add(cvtt(v1), th), th);
    int v2m = v1m + 1;
    if (v2m == th)
        v2m = 0;
    v1m *= tw;
    v2m *= tw;
    int u1m = rem(wrap_add(cvtt(u1), tw), tw);
    int u2m = u1m + 1;
    if (u2m == tw)
        u2m = 0;
    const int64_t lim = (int64_t)tw * th;
    int64_t idx[4] = {(int64_t)v1m + u1m, (int64_t)v2m + u1m, (int64_t)v1m + u2m,
                      (int64_t)v2m + u2m};
    BilTaps T;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int64_t i = idx[k];
        if (i < 0 || i >= lim) {
            oob++;
            i = i < 0 ? 0 : lim - 1;
        }
        T.p[k] = texel(X, i);
    }
    T.w[0] = (u - u1) * (v - v1);
    T.w[1] = (u - u1) * (v2 - v);
    T.w[2] = (u2 - u) * (v - v1);
    T.w[3] = (u2 - u) * (v2 - v);
    return T;
}

// Channel c (0 b, 1 g, 2 r, 3 alpha) of a bilinear sample.
SW_D uint32_t bilinear_channel(const BilTaps &T, int c)
{
    float a = (float)((T.p[0] >> (8 * c)) & 0xFF) * T.w[0];
    float b = (float)((T.p[1] >> (8 * c)) & 0xFF) * T.w[1];
    float d = (float)((T.p[2] >> (8 * c)) & 0xFF) * T.w[2];
    float e = (float)((T.p[3] >> (8 * c)) & 0xFF) * T.w[3];
    float s = ((a + b) + d) + e;
    return (uint32_t)to_u8(roundf(s));
}

SW_D uint32_t bilinear(const TexView &X, f2 t, uint32_t &oob)
{
    const BilTaps T = bilinear_taps(X, t, oob);
    uint32_t out = 0;
#pragma unroll
    for (int c = 0; c < 4; c++)
        out |= bilinear_channel(T, c) << (8 * c);
    return out;
}

// pixel_shader_texture::shade (pixel_shaders.cpp:275-281): twidth and
// theight are unsigned there, so (int)t % twidth is the unsigned remainder
// of the int's bits, always inside the bitmap; column from t.x, row from t.y.
SW_D uint32_t nearest(const TexView &X, f2 t)
{
    const uint32_t tw = (uint32_t)X.w, th = (uint32_t)X.h;
    const uint32_t x = (uint32_t)cvtt(t.x), y = (uint32_t)cvtt(t.y);
    const uint32_t u = (tw & (tw - 1)) == 0 ? x & (tw - 1) : x % tw;
    const uint32_t v = (th & (th - 1)) == 0 ? y & (th - 1) : y % th;
    return texel(X, v * tw + u);
}

// pixel_shader_lights_flat::prepare_for_triangle (pixel_shaders.cpp:33-84),
// returned by shade as int: the face normal normal_t(cross(e1, e2)) is
// normalised once, and again when negated for an inverted triangle (unary
// minus builds a normal_t, points.hpp:171-174); the light is evaluated at
// the centroid ((v0 + v1) + v2) / 3 exactly as Phong evaluates a pixel.
SW_D int32_t flat_light(f3 w0, f3 w1, f3 w2, bool inverted, f3 cam, float ambient, f3 sun, float sun_I,
                        const Light4 *lights, int n_lights)
{
    const f3 e1 = w1 - w0, e2 = w2 - w0;
    f3 nrm = normalize({e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x});
    if (inverted)
        nrm = normalize(neg(nrm));
    const f3 s = (w0 + w1) + w2;
    const f3 center = {s.x / 3.0f, s.y / 3.0f, s.z / 3.0f};
    return phong(center, nrm, cam, ambient, sun, sun_I, lights, n_lights);
}

// pixel_shader_light_and_texture::shade combine (pixel_shaders.hpp:159-178)
SW_D uint32_t combine(uint32_t color, int32_t light_i)
{
    // (float)((double)light_i / 65536.0): the double quotient is exact (a
    // power-of-two scale of an int), so this is the int rounded to float and
    // scaled by 2^-16, which the float operations below reproduce exactly.
    float light = (float)light_i * 0x1p-16f;
    uint32_t out = color & 0xFF000000u;
    if (light < 1) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            uint32_t c = (color >> (8 * k)) & 0xFF;
            out |= (uint32_t)to_u8((float)c * light) << (8 * k);
        }
    } else {
        light = sqrtf(light);
        light = sqrtf(light);
        const FDiv by_light(light);   // 1 <= light <= 32768^(1/4); 0 <= 255 - c <= 255
#pragma unroll
        for (int k = 0; k < 3; k++) {
            uint32_t c = (color >> (8 * k)) & 0xFF;
            uint8_t r = (uint8_t)(255 - to_u8(by_light((float)(255 - (int)c))));
            out |= (uint32_t)r << (8 * k);
        }
    }
    return out;
}

// Shader state of one triangle half and one scanline
// (prepare_for_triangle / _upper / _lower / _scanline, pixel_shaders.cpp:88-158
// and :304-346).
struct TriAttr {
    f3 w[3];    // v_world of the y-sorted vertices
    f3 n[3];    // normal_world, negated when the triangle is inverted
    f2 t[3];    // tex_coords * (width, height)
};

struct ScanShade {
    f3 v, vdir, n, ndir;
    f2 t_left, t_dir;
};

SW_D ScanShade prepare_scanline(const TriAttr &A, bool lower, bool lr, float pl, float pr)
{
    f3 vleft, vright, vld, vrd, nleft, nright, nld, nrd;
    f2 sst, sstd;
    const f2 sltd = A.t[2] - A.t[0];
    if (!lower) {
        vleft = A.w[0];
        vright = A.w[0];
        nleft = A.n[0];
        nright = A.n[0];
        if (lr) {
            vld = A.w[1] - A.w[0];
            vrd = A.w[2] - A.w[0];
            nld = A.n[1] - A.n[0];
            nrd = A.n[2] - A.n[0];
        } else {
            vld = A.w[2] - A.w[0];
            vrd = A.w[1] - A.w[0];
            nld = A.n[2] - A.n[0];
            nrd = A.n[1] - A.n[0];
        }
        sst = A.t[0];
        sstd = A.t[1] - A.t[0];
    } else {
        if (lr) {
            vright = A.w[0];
            vrd = A.w[2] - A.w[0];
            vleft = A.w[1];
            vld = A.w[2] - A.w[1];
            nright = A.n[0];
            nrd = A.n[2] - A.n[0];
            nleft = A.n[1];
            nld = A.n[2] - A.n[1];
        } else {
            vleft = A.w[0];
            vld = A.w[2] - A.w[0];
            vright = A.w[1];
            vrd = A.w[2] - A.w[1];
            nleft = A.n[0];
            nld = A.n[2] - A.n[0];
            nright = A.n[1];
            nrd = A.n[2] - A.n[1];
        }
        sst = A.t[1];
        sstd = A.t[2] - A.t[1];
    }
    ScanShade S;
    S.v = vleft + vld * pl;
    S.vdir = vright + vrd * pr - S.v;
    S.n = nleft + nld * pl;
    S.ndir = nright + nrd * pr - S.n;
    if (lr) {
        S.t_left = sst + sstd * pl;
        S.t_dir = A.t[0] + sltd * pr - S.t_left;
    } else {
        S.t_left = A.t[0] + sltd * pl;
        S.t_dir = sst + sstd * pr - S.t_left;
    }
    return S;
}

}  // namespace sw
