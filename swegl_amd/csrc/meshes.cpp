// meshes.cpp -- the reference's procedural meshes and face normals
// (swegl/data/model.hpp:181-466) with its float arithmetic: make_tore and
// make_sphere build their vertices by transforming points through matrices
// turned step by step with matrix44_t::rotate_y/z, normals are normal_t
// values (normalised on construction, points.hpp:131-148), texture
// coordinates are double quotients stored to float.
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/swegl_hip.h"
#include "hostmath.h"

int sw_fail_msg(int rc, const std::string &msg);

namespace {

struct V3 {
    float x, y, z;
};

// vector_t::normalize (points.hpp:76-86) through vector_t::len (:71-74)
V3 normalized(V3 v)
{
    const float l = (float)std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (l != 0) {
        v.x /= l;
        v.y /= l;
        v.z /= l;
    }
    return v;
}

// transform(vertex_t, matrix44_t) (src/projection/points.cpp:8-13)
V3 transform(V3 v, const float *m)
{
    return V3{m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3], m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7],
              m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11]};
}

// cross (points.cpp:32-37): a normal_t, so normalised
V3 cross(V3 l, V3 r)
{
    return normalized(V3{l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x});
}

V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }

void put3(float *a, size_t i, V3 v)
{
    a[3 * i] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
}
V3 get3(const float *a, size_t i) { return V3{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }

}  // namespace

// make_tore (model.hpp:327-391)
extern "C" int swegl_make_tore(uint32_t precision, float *positions, float *normals, float *tex_coords,
                               uint32_t *indices)
{
    if (precision == 0 || precision > 4096 || !positions || !normals || !tex_coords || !indices)
        return sw_fail_msg(SWEGL_EINVAL, "make_tore: precision in [1, 4096] and output arrays required");
    const uint32_t n = precision + 1;
    const float angle = (2 * 3.141592653589f) / precision;
    float big[16], big_normal[16];
    std::memcpy(big, hm::kIdentity, sizeof big);
    std::memcpy(big_normal, hm::kIdentity, sizeof big_normal);
    hm::translate(big, 2.0f, 0.0f, 0.0f);
    size_t v = 0;
    for (uint32_t bg = 0; bg <= precision; bg++) {
        for (int row = 0; row < 2; row++) {
            if (row == 1) {
                hm::rotate_y(big, angle);
                hm::rotate_y(big_normal, angle);
            }
            float small[16], small_normal[16];
            std::memcpy(small, hm::kIdentity, sizeof small);
            std::memcpy(small_normal, hm::kIdentity, sizeof small_normal);
            hm::translate(small, 0.8f, 0.0f, 0.0f);
            for (uint32_t sm = 0; sm <= precision; sm++, v++) {
                put3(positions, v, transform(transform(V3{0, 0, 0}, small), big));
                const V3 nrm = normalized(transform(normalized(transform(V3{1, 0, 0}, small_normal)), big_normal));
                put3(normals, v, nrm);
                tex_coords[2 * v] = (float)(1.0 * (bg + row) / precision);
                tex_coords[2 * v + 1] = (float)(1.0 * sm / precision);
                hm::rotate_z(small, angle);
                hm::rotate_z(small_normal, angle);
            }
        }
        for (uint32_t sm = 0; sm <= precision; sm++) {
            indices[(size_t)bg * 2 * n + 2 * sm] = n + sm;
            indices[(size_t)bg * 2 * n + 2 * sm + 1] = sm;
        }
    }
    return SWEGL_OK;
}

// make_sphere (model.hpp:393-464)
extern "C" int swegl_make_sphere(uint32_t precision, float radius, float *positions, float *normals,
                                 float *tex_coords, uint32_t *indices)
{
    if (precision == 0 || precision > 4096 || !positions || !normals || !tex_coords || !indices)
        return sw_fail_msg(SWEGL_EINVAL, "make_sphere: precision in [1, 4096] and output arrays required");
    const uint32_t n = precision + 1;
    const float angle = (2 * 3.141592653589f) / precision;
    float big[16];
    std::memcpy(big, hm::kIdentity, sizeof big);
    size_t v = 0;
    for (uint32_t bg = 0; bg <= precision; bg++) {
        for (int row = 0; row < 2; row++) {
            if (row == 1)
                hm::rotate_y(big, angle);
            float small[16];
            std::memcpy(small, hm::kIdentity, sizeof small);
            for (uint32_t sm = 0; sm <= precision; sm++, v++) {
                const V3 p = transform(transform(V3{0, radius, 0}, small), big);
                put3(positions, v, p);
                put3(normals, v, normalized(p));
                tex_coords[2 * v] = (float)(1.0 * sm / precision);
                tex_coords[2 * v + 1] = (float)(1.0 * (bg + row) / precision);
                hm::rotate_z(small, angle / 2);
            }
        }
        for (uint32_t sm = 0; sm <= precision; sm++) {
            indices[(size_t)bg * 2 * n + 2 * sm] = n + sm;
            indices[(size_t)bg * 2 * n + 2 * sm + 1] = sm;
        }
    }
    return SWEGL_OK;
}

// calculate_face_normals (model.hpp:181-231), including its last loop,
// which the reference runs for fans again (not for triangle lists) and
// which renormalises through normal_t::operator=(const vector_t &)
extern "C" int swegl_calculate_face_normals(int32_t mode, const float *positions, float *normals,
                                            uint32_t n_vertices, const uint32_t *indices, uint32_t n_indices)
{
    if ((n_vertices && (!positions || !normals)) || (n_indices && !indices))
        return sw_fail_msg(SWEGL_EINVAL, "calculate_face_normals: null array");
    for (uint32_t i = 0; i < n_indices; i++)
        if (indices[i] >= n_vertices)
            return sw_fail_msg(SWEGL_EINVAL, "calculate_face_normals: index out of range");
    if ((mode == SWEGL_TRIANGLE_STRIP || mode == SWEGL_TRIANGLE_FAN) && n_indices < 3)
        return sw_fail_msg(SWEGL_EINVAL, "calculate_face_normals: a strip or fan needs 3 indices");
    auto P = [&](uint32_t i) { return get3(positions, i); };
    if (mode == SWEGL_TRIANGLE_STRIP) {
        uint32_t i0 = indices[0], i1 = indices[1], i2 = indices[2];
        const V3 nrm = cross(sub(P(i1), P(i0)), sub(P(i2), P(i0)));
        put3(normals, i0, nrm);
        put3(normals, i1, nrm);
        put3(normals, i2, nrm);
        for (uint32_t i = 3; i < n_indices; i++, i0 = i1, i1 = i2) {
            i2 = indices[i];
            put3(normals, i2, (i & 1) == 0 ? cross(sub(P(i1), P(i0)), sub(P(i2), P(i0)))
                                           : cross(sub(P(i2), P(i0)), sub(P(i1), P(i0))));
        }
    }
    if (mode == SWEGL_TRIANGLE_FAN) {
        const uint32_t i0 = indices[0];
        uint32_t i1 = indices[1], i2 = indices[2];
        const V3 nrm = cross(sub(P(i1), P(i0)), sub(P(i2), P(i0)));
        put3(normals, i0, nrm);
        put3(normals, i1, nrm);
        put3(normals, i2, nrm);
        for (uint32_t i = 3; i < n_indices; i++, i1 = i2) {
            i2 = indices[i];
            put3(normals, i2, cross(sub(P(i1), P(i0)), sub(P(i2), P(i0))));
        }
        for (uint32_t i = 2; i < n_indices; i += 3) {
            const uint32_t a = indices[i - 2], b = indices[i - 1], c = indices[i];
            const V3 nrm3 = normalized(cross(sub(P(b), P(a)), sub(P(c), P(a))));
            put3(normals, a, nrm3);
            put3(normals, b, nrm3);
            put3(normals, c, nrm3);
        }
    }
    return SWEGL_OK;
}
