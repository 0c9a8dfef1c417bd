// hostmath.h -- host-side matrix and camera arithmetic with the reference's
// float semantics (row-major m[r*4+c] == matrix44_t[r][c]).
#pragma once
#include <cmath>
#include <cstring>

namespace hm {

static const float kIdentity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

// matrix44_t::rotate_x (src/projection/matrix44.cpp:12-29).  `cos(a)` with a
// float argument resolves to std::cos(float) through <math.h>.
inline void rotate_x(float *m, float a)
{
    float cosa = cosf(a), sina = sinf(a);
    float d4 = m[4], d5 = m[5], d6 = m[6], d7 = m[7];
    m[4] = cosa * d4 + sina * m[8];
    m[5] = cosa * d5 + sina * m[9];
    m[6] = cosa * d6 + sina * m[10];
    m[7] = cosa * d7 + sina * m[11];
    m[8] = -sina * d4 + cosa * m[8];
    m[9] = -sina * d5 + cosa * m[9];
    m[10] = -sina * d6 + cosa * m[10];
    m[11] = -sina * d7 + cosa * m[11];
}
// matrix44_t::rotate_y (matrix44.cpp:31-48)
inline void rotate_y(float *m, float a)
{
    float cosa = cosf(a), sina = sinf(a);
    float d0 = m[0], d1 = m[1], d2 = m[2], d3 = m[3];
    m[0] = cosa * d0 + sina * m[8];
    m[1] = cosa * d1 + sina * m[9];
    m[2] = cosa * d2 + sina * m[10];
    m[3] = cosa * d3 + sina * m[11];
    m[8] = -sina * d0 + cosa * m[8];
    m[9] = -sina * d1 + cosa * m[9];
    m[10] = -sina * d2 + cosa * m[10];
    m[11] = -sina * d3 + cosa * m[11];
}
// matrix44_t::rotate_z (matrix44.cpp:50-67)
inline void rotate_z(float *m, float a)
{
    float cosa = cosf(a), sina = sinf(a);
    float d0 = m[0], d1 = m[1], d2 = m[2], d3 = m[3];
    m[0] = cosa * d0 + sina * m[4];
    m[1] = cosa * d1 + sina * m[5];
    m[2] = cosa * d2 + sina * m[6];
    m[3] = cosa * d3 + sina * m[7];
    m[4] = -sina * d0 + cosa * m[4];
    m[5] = -sina * d1 + cosa * m[5];
    m[6] = -sina * d2 + cosa * m[6];
    m[7] = -sina * d3 + cosa * m[7];
}
// matrix44_t::translate (matrix44.cpp:87-92)
inline void translate(float *m, float x, float y, float z)
{
    m[3] += x;
    m[7] += y;
    m[11] += z;
}
// node_t::get_local_world_matrix (model.hpp:56-61) = translate(scale(R, S));
// rs = scale(R, S) (src/projection/points.cpp:14-30)
inline void local_matrix(const float *R, const float *S, const float *T, float *local, float *rs)
{
    std::memcpy(rs, R, 16 * sizeof(float));
    for (int r = 0; r < 4; r++) {
        rs[r * 4 + 0] *= S[0];
        rs[r * 4 + 1] *= S[1];
        rs[r * 4 + 2] *= S[2];
    }
    std::memcpy(local, rs, 16 * sizeof(float));
    translate(local, T[0], T[1], T[2]);
}
// Matrix product for node hierarchies (freon::Matrix operator*, not in the
// reference tree): textbook order.  For root nodes the parent is the
// identity and the product is exact in any order.
inline void matmul(const float *a, const float *b, float *r)
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r[i * 4 + j] = a[i * 4 + 0] * b[0 * 4 + j] + a[i * 4 + 1] * b[1 * 4 + j] +
                           a[i * 4 + 2] * b[2 * 4 + j] + a[i * 4 + 3] * b[3 * 4 + j];
}
// viewport_t ctor (viewport.cpp:30-35)
inline void viewport_matrix(int x, int y, int w, int h, float *m00, float *m03, float *m11, float *m13)
{
    *m03 = x + w / 2.0f;
    *m13 = y + h / 2.0f;
    *m00 = w / 2.0f;
    *m11 = -h / 2.0f;
}
// camera_t::camera_t (src/projection/camera.cpp:8-35)
inline void camera_init(float aspect_ratio, float *view, float *proj, float *center)
{
    std::memcpy(view, kIdentity, sizeof kIdentity);
    std::memcpy(proj, kIdentity, sizeof kIdentity);
    view[10] = -1;
    const float n = 0.5f, f = 10.0f, w = 1.0f, h = 1.0f;
    if (aspect_ratio > 1) {
        proj[0] = 2 * n / w;
        proj[5] = aspect_ratio * 2 * n / h;
    } else {
        proj[0] = (1.0f / aspect_ratio) * 2 * n / w;
        proj[5] = 2 * n / h;
    }
    proj[10] = f / (f - n);
    proj[11] = -f * n / (f - n);
    proj[14] = 1;
    center[0] = center[1] = center[2] = 0;
}
// camera_t::translate (camera.cpp:50-58)
inline void camera_translate(float *view, float *center, float x, float y, float z)
{
    translate(view, -x, -y, -z);
    const float *m = view;
    center[0] += x * m[0] + y * m[4] + z * m[8];
    center[1] += x * m[1] + y * m[5] + z * m[9];
    center[2] += x * m[2] + y * m[6] + z * m[10];
}

}  // namespace hm
