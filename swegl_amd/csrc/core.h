// core.h -- device-side arithmetic of the render path for gfx950.
//
// Every function here reproduces one step of swegl's CPU renderer with the
// same float operations in the same order (the library is compiled with
// -ffp-contract=off; HIP's default IEEE division and square root).  The
// x86-64 conversion semantics of the reference build (cvttss2si: NaN and
// out-of-range -> INT32_MIN, float -> uint8 through a 32-bit int) are
// emulated explicitly because v_cvt_i32_f32 saturates instead.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SW_HD __host__ __device__ __forceinline__
#define SW_D __device__ __forceinline__

namespace sw {

// Arrays a kernel only reads (the scene, the frustum coordinates, marks, bins
// and rectangles written by earlier kernels) are read through the constant
// address space: uniform addresses then take scalar loads, divergent ones
// global loads with a scalar base.  Never for an array the same kernel
// writes.
#if defined(__HIP_DEVICE_COMPILE__)
#define SW_RO __attribute__((address_space(4)))
#else
#define SW_RO
#endif
template <class T>
SW_HD const SW_RO T *ro(const T *p)
{
    return (const SW_RO T *)p;
}
// HIP's vector types cannot be copied out of another address space.
SW_HD float4 ro4(const float4 *p, uint32_t i)
{
    const SW_RO float *q = ro((const float *)(p + i));
    return make_float4(q[0], q[1], q[2], q[3]);
}
SW_HD uint2 ro2(const uint2 *p, uint32_t i)
{
    const SW_RO uint32_t *q = ro((const uint32_t *)(p + i));
    return make_uint2(q[0], q[1]);
}

// ------------------------------------------------------------------ scalars
SW_HD int32_t cvtt(float f)
{
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int32_t)f : INT32_MIN;
}
SW_HD uint8_t to_u8(float f) { return (uint8_t)(uint32_t)cvtt(f); }
SW_HD int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

// `z <= 0.001` with z a float and 0.001 a double literal (renderer.cpp:489)
// is exactly `z < 0.001f`: 0.001f is the least float above 0.001.  The same
// holds for `>= 0.001` (renderer.cpp:65, 281, 286, 318) and `< 0.05`
// (pixel_shaders.cpp:181).  A CPU test pins both facts.
#define SW_NEAR 0.001f
#define SW_DIFFUSE_CUT 0.05f

struct f3 { float x, y, z; };
struct f2 { float x, y; };
SW_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
SW_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
SW_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
SW_HD f2 operator+(f2 a, f2 b) { return {a.x + b.x, a.y + b.y}; }
SW_HD f2 operator-(f2 a, f2 b) { return {a.x - b.x, a.y - b.y}; }
SW_HD f2 operator*(f2 a, float s) { return {a.x * s, a.y * s}; }
SW_HD f3 neg(f3 a) { return {-a.x, -a.y, -a.z}; }
// vector_t::dot (points.hpp:93): (xx + yy) + zz
SW_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Correctly rounded a / b for several numerators over one denominator b,
// for callers that know 1 <= b <= 2^20 and a == 0 or 2^-60 <= |a| <= 2^60.
// The compiler expands each IEEE '/' into: v_div_scale of a and b, r =
// rcp(b) refined by one fma step, q = a*r, two fma residual corrections
// (the last through v_div_fmas), then v_div_fixup.  In that range the
// scaling and the fixup are identities, so the refined reciprocal depends
// on b alone and is computed once; each quotient then takes five
// instructions and equals '/' bit for bit (a zero numerator keeps its
// signed zero a*r).
struct FDiv {
    float b, r;
    SW_D explicit FDiv(float den) : b(den)
    {
        const float r0 = __builtin_amdgcn_rcpf(den);
        r = __builtin_fmaf(__builtin_fmaf(-den, r0, 1.0f), r0, r0);
    }
    SW_D float operator()(float a) const { return a == 0.0f ? a * r : nonzero(a); }
    // the quotient of a nonzero numerator in the range above
    SW_D float nonzero(float a) const
    {
        const float q0 = a * r;
        float q = __builtin_fmaf(__builtin_fmaf(-b, q0, a), r, q0);
        return __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
    }
};

// vector_t::normalize (points.hpp:80-90).  When every component is at
// least 2^-60 in magnitude and the length l at most 2^40, the compiler's
// expansion of each x / l scales nothing (l and 1/l normal, the numerator's
// exponent within 96 of l's and above 2^-103, the quotient at least 2^-100),
// so the three quotients share l's refined reciprocal (FDiv's steps, here
// without its zero case); otherwise the plain divisions.
SW_D f3 normalize(f3 a)
{
    float l = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    const float m = fminf(fabsf(a.x), fminf(fabsf(a.y), fabsf(a.z)));
    if (m >= 0x1p-60f && l <= 0x1p40f) {
        const FDiv by(l);
        return {by.nonzero(a.x), by.nonzero(a.y), by.nonzero(a.z)};
    }
    if (l != 0) {
        a.x /= l;
        a.y /= l;
        a.z /= l;
    }
    return a;
}
// cross() builds a normal_t, which normalises (points.cpp:32-37); only the
// sign of the normalised z component is ever read on the path:
// sign(normalize(cross(l, r)).z) compared with 0 (dir = +1: > 0, -1: < 0)
// without the square root and division where they cannot change it: with
// s = x^2 + y^2 + z^2 (the normalisation's own sum) zero, normalize leaves the
// vector unchanged; with s <= 2^120 and |z| >= 2^-60, z / sqrt(s) is a
// normal float of z's sign.  Otherwise (overflow, NaN, tiny z) it divides.
SW_HD bool cross_z_sign(f3 l, f3 r, int dir)
{
    const float x = l.y * r.z - l.z * r.y, y = l.z * r.x - l.x * r.z, z = l.x * r.y - l.y * r.x;
    const float s = x * x + y * y + z * z;
    float v = z;
    if (!(s == 0.0f || (s <= 0x1p120f && fabsf(z) >= 0x1p-60f)))
        v = z / sqrtf(s);
    return dir > 0 ? v > 0.0f : v < 0.0f;
}

// 3x4 affine part of a matrix44_t, row-major.
struct m34 { float m[12]; };
// transform(v, m) (points.cpp:8-13): ((a+b)+c)+d per row
SW_HD f3 xform(const m34 &M, f3 v)
{
    const float *m = M.m;
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3],
            m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7],
            m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11]};
}
// rotate(n, m) then .normalize() then normal_t::operator= : three
// normalisations (vertex_shaders.hpp:63, points.cpp:44-49, points.hpp:143-150)
SW_D f3 normal_world(const m34 &RS, f3 n)
{
    const float *m = RS.m;
    f3 r = {m[0] * n.x + m[1] * n.y + m[2] * n.z, m[4] * n.x + m[5] * n.y + m[6] * n.z,
            m[8] * n.x + m[9] * n.y + m[10] * n.z};
    return normalize(normalize(normalize(r)));
}

// ---------------------------------------------------------------- per view
struct ViewXf {
    m34 view, proj;
    float m00, m03, m11, m13;   // viewport matrix (viewport.cpp:30-35)
    int vx, vy, vw, vh;
    f3 cam;                     // camera_t::m_center
};

// world -> frustum coordinates (vertex_shaders.hpp:35-47, 61-71)
SW_D f3 world_to_frustum(const ViewXf &V, f3 w)
{
    f3 c = xform(V.view, w);
    f3 p = xform(V.proj, c);
    if (p.z != 0) {
        p.x /= fabsf(p.z);
        p.y /= fabsf(p.z);
    }
    return p;
}
// viewport_t::transform (viewport.cpp:123-129)
SW_HD f3 frustum_to_screen(const ViewXf &V, f3 p) { return {V.m00 * p.x + V.m03, V.m11 * p.y + V.m13, p.z}; }

// inside_camera_frustum && facing (renderer.cpp:58-75) on frustum coords
SW_HD bool inside_frustum(f3 a, f3 b, f3 c)
{
    return (a.x >= -1 || b.x >= -1 || c.x >= -1) && (a.y >= -1 || b.y >= -1 || c.y >= -1) &&
           (a.x < 1 || b.x < 1 || c.x < 1) && (a.y < 1 || b.y < 1 || c.y < 1) &&
           (a.z >= SW_NEAR || b.z >= SW_NEAR || c.z >= SW_NEAR) && (a.x != b.x || a.x != c.x) &&
           (a.y != b.y || a.y != c.y);
}

// --------------------------------------------------- interpolator_g<1>
// (interpolator.hpp:52-104).  The state only accumulates by repeated adds,
// so any sample is a replay of the same add sequence.
struct Interp {
    float ta, ts, ba, bs, v0, v1;
    SW_HD void init(float dist, float z1, float z2)
    {
        v0 = z1;
        v1 = z2 - z1;
        float as = 1.0f / dist;
        ba = 1.0f / z1;
        float iz2 = 1.0f / z2;
        ta = 0.0f;
        ta *= ba;                  // 0*inf = NaN exactly as the reference
        ts = (1.0f * iz2 - ta) * as;
        bs = (iz2 - ba) * as;
    }
    SW_HD void displace(float m)
    {
        ta += ts * m;
        ba += bs * m;
    }
    SW_HD void step()
    {
        ta += ts;
        ba += bs;
    }
    // n single steps (the same adds in the same order), four per loop trip:
    // a long replay then spends its time on the adds, not on loop control
    SW_HD void steps(int n)
    {
        int i = 0;
        for (; i + 4 <= n; i += 4) {
            ta += ts;
            ba += bs;
            ta += ts;
            ba += bs;
            ta += ts;
            ba += bs;
            ta += ts;
            ba += bs;
        }
        for (; i < n; i++) {
            ta += ts;
            ba += bs;
        }
    }
    SW_HD float u() const { return ta / ba; }
    SW_HD float value(float u) const { return v0 + v1 * u; }
};

struct Side {
    Interp it;
    float ratio, x;
    SW_HD void steps(int n)
    {
        int i = 0;
        for (; i + 4 <= n; i += 4) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                x += ratio;
                it.ta += it.ts;
                it.ba += it.bs;
            }
        }
        for (; i < n; i++) {
            x += ratio;
            it.ta += it.ts;
            it.ba += it.bs;
        }
    }
};

// ------------------------------------------------------ vertex records
// A vertex of a drawn (sub-)triangle: screen position plus where its
// shading attributes come from: an original vertex, or a near-clip vertex
// interpolated between vertices a and b at `cut` (renderer.cpp:286-356).
struct SVert {
    f3 s;          // v_viewport after the viewport transform
    uint32_t a, b; // original vertex ids (b == a for an original vertex)
    float cut;     // clip parameter; only meaningful when b != a
    int clip;      // 1 when this is a clip vertex
};

// Sub-triangle after fill_triangle's z-sort and clip, in the vertex order
// passed to fill_triangle_2, with its facing flag.
struct SubTri {
    SVert v[3];
    bool front;
};

// Sub-triangle after fill_triangle_2's y-sort, with the DDA setup.
struct Setup {
    SVert v[3];        // sorted by screen y
    bool inverted;
    int iy0, iy1, iy2;
    Side L;            // long side at row ysL = max(iy0, vy)
    int ysL;
    bool has_up, has_lo;
    int up_y0, up_y1;  // upper half rows [up_y0, up_y1)
    int lo_y0, lo_y1;  // lower half rows [lo_y0, lo_y1)
    Side Su, Sl;       // short sides at their first rows
    bool lr_up, lr_lo; // long_line_on_right per half
    int up_rows;       // rows the upper half actually processes (long-side steps)
};

// fill_triangle_2 setup (renderer.cpp:361-460).  Returns false when the
// sub-triangle spans no scanline (y0 == y2).  With `row` given, only the
// short side of the half that processes that row is set up (the other
// half's row range is still computed); row_state then works for that row.
// fill_triangle_2's y-sort (renderer.cpp:375-387): S.v and S.inverted only.
SW_D void sort_subtri(const SubTri &t, Setup &S)
{
    SVert a = t.v[0], b = t.v[1], c = t.v[2];
    S.inverted = !t.front;
    if (b.s.y < a.s.y) { SVert tmp = a; a = b; b = tmp; }
    if (c.s.y < b.s.y) { SVert tmp = b; b = c; c = tmp; }
    if (b.s.y < a.s.y) { SVert tmp = a; a = b; b = tmp; }
    S.v[0] = a;
    S.v[1] = b;
    S.v[2] = c;
}

SW_D bool setup_subtri(const SubTri &t, const ViewXf &V, Setup &S, int row = INT32_MIN)
{
    sort_subtri(t, S);
    const f3 p0 = S.v[0].s, p1 = S.v[1].s, p2 = S.v[2].s;
    S.iy0 = cvtt(ceilf(p0.y));
    S.iy1 = cvtt(ceilf(p1.y));
    S.iy2 = cvtt(ceilf(p2.y));
    if (S.iy0 == S.iy2)
        return false;
    const int vy = V.vy, vh = V.vh;
    S.L.ratio = (p2.x - p0.x) / (p2.y - p0.y);
    S.L.it.init(p2.y - p0.y, p0.z, p2.z);
    S.ysL = S.iy0 < vy ? vy : S.iy0;
    S.L.it.displace((float)S.ysL - p0.y);
    S.L.x = p0.x + S.L.ratio * ((float)S.ysL - p0.y);

    S.has_up = S.iy1 >= vy;
    S.up_rows = 0;
    if (S.has_up) {
        int y = S.iy0 > vy ? S.iy0 : vy;
        int ye = S.iy1 < vy + vh ? S.iy1 : vy + vh;
        S.up_y0 = y;
        S.up_y1 = ye;
        S.up_rows = ye > y ? ye - y : 0;
        if (row == INT32_MIN || (row >= y && row < ye)) {
            S.Su.ratio = (p1.x - p0.x) / (p1.y - p0.y);
            S.Su.it.init(p1.y - p0.y, p0.z, p1.z);
            S.Su.it.displace((float)y - p0.y);
            S.Su.x = p0.x + S.Su.ratio * ((float)y - p0.y);
            S.lr_up = S.L.ratio > S.Su.ratio;
        }
    }
    S.has_lo = S.iy1 < vy + vh;
    if (S.has_lo) {
        int y = S.iy1 > vy ? S.iy1 : vy;
        int ye = S.iy2 < vy + vh ? S.iy2 : vy + vh;
        S.lo_y0 = y;
        S.lo_y1 = ye;
        if (row == INT32_MIN || (row >= y && row < ye)) {
            S.Sl.ratio = (p2.x - p1.x) / (p2.y - p1.y);
            S.Sl.it.init(p2.y - p1.y, p1.z, p2.z);
            S.Sl.it.displace((float)y - p1.y);
            S.Sl.x = p1.x + S.Sl.ratio * ((float)y - p1.y);
            S.lr_lo = S.L.ratio < S.Sl.ratio;
        }
    }
    return true;
}

// The two sides of scanline y, replayed exactly (fill_half_triangle's
// per-row `x += ratio; Step()` at renderer.cpp:553-556).  half: 0 upper,
// 1 lower.  Returns false if no half processes row y.
struct Row {
    Side left, right;
    bool lower;
    bool lr;
};
SW_HD bool row_state(const Setup &S, int y, Row &R)
{
    Side lng = S.L, sh;
    if (S.has_up && y >= S.up_y0 && y < S.up_y1) {
        lng.steps(y - S.ysL);
        sh = S.Su;
        sh.steps(y - S.up_y0);
        R.lower = false;
        R.lr = S.lr_up;
    } else if (S.has_lo && y >= S.lo_y0 && y < S.lo_y1) {
        lng.steps(S.up_rows + (y - S.lo_y0));
        sh = S.Sl;
        sh.steps(y - S.lo_y0);
        R.lower = true;
        R.lr = S.lr_lo;
    } else {
        return false;
    }
    if (R.lr) {
        R.left = sh;
        R.right = lng;
    } else {
        R.left = lng;
        R.right = sh;
    }
    return true;
}

// Scanline span and its pixel interpolator (renderer.cpp:469-480).
struct Span {
    int x1, x2;
    Interp q;       // displaced to x1
    float pl, pr;   // side progresses passed to prepare_for_scanline
};
SW_D bool span_of(const Row &R, const ViewXf &V, Span &sp)
{
    int x1 = cvtt(ceilf(R.left.x));
    int x2 = cvtt(ceilf(R.right.x));
    x1 = x1 > V.vx ? x1 : V.vx;
    x2 = x2 < V.vx + V.vw ? x2 : V.vx + V.vw;
    sp.x1 = x1;
    sp.x2 = x2;
    if (!(x1 < x2))
        return false;
    sp.pl = R.left.it.u();
    sp.pr = R.right.it.u();
    sp.q.init(R.right.x - R.left.x, R.left.it.value(sp.pl), R.right.it.value(sp.pr));
    sp.q.displace((float)x1 - R.left.x);
    return true;
}

// ---------------------------------------------------------------- packing
// Visibility key: depth bits above draw sequence.  z > 0.001 > 0 so the bit
// pattern orders like the value; the smaller sequence wins ties, which is
// the reference's first-minimum-wins rule (renderer.cpp:491).
SW_HD uint64_t vis_key(float z, uint32_t seq)
{
    return ((uint64_t)__builtin_bit_cast(uint32_t, z) << 32) | seq;
}
#define SW_KEY_EMPTY 0x7F7F7F7FFFFFFFFFull   // max_z (renderer.cpp:15-19), no sequence
#define SW_MAX_Z_BITS 0x7F7F7F7Fu

}  // namespace sw
