// swegl/projection/points.hpp -- point, vector and matrix types of the
// swegl API (reference: swegl/projection/points.hpp, matrix44.hpp,
// vec2f.hpp), with the same float semantics, over libswegl_hip.so.
#pragma once

#include <cmath>
#include <cstring>

#include "../../swegl_hip.h"

namespace swegl
{

class vertex_t
{
    float _x, _y, _z;

public:
    vertex_t() = default;
    vertex_t(float x, float y, float z) : _x(x), _y(y), _z(z) {}
    float &x() { return _x; }
    float x() const { return _x; }
    float &y() { return _y; }
    float y() const { return _y; }
    float &z() { return _z; }
    float z() const { return _z; }
};

class vector_t
{
    float _x, _y, _z;

public:
    vector_t() = default;
    vector_t(float x, float y, float z) : _x(x), _y(y), _z(z) {}
    float &x() { return _x; }
    float x() const { return _x; }
    float &y() { return _y; }
    float y() const { return _y; }
    float &z() { return _z; }
    float z() const { return _z; }
    float len() const { return (float)std::sqrt(x() * x() + y() * y() + z() * z()); }
    float len_squared() const { return x() * x() + y() * y() + z() * z(); }
    vector_t &normalize()
    {
        float l = len();
        if (l != 0) {
            _x /= l;
            _y /= l;
            _z /= l;
        }
        return *this;
    }
    float dot(const vector_t &o) const { return x() * o.x() + y() * o.y() + z() * o.z(); }
};

// A vector normalised on construction and assignment (points.hpp:131-151).
class normal_t : public vector_t
{
public:
    normal_t() : vector_t() {}
    normal_t(const vector_t &v) : vector_t(v) { normalize(); }
    normal_t(float x, float y, float z) : vector_t(x, y, z) { normalize(); }
    normal_t &operator=(const vector_t &v)
    {
        x() = v.x();
        y() = v.y();
        z() = v.z();
        normalize();
        return *this;
    }
};

class vec2f_t
{
    float _x, _y;

public:
    vec2f_t() = default;
    vec2f_t(float x, float y) : _x(x), _y(y) {}
    float &x() { return _x; }
    float x() const { return _x; }
    float &y() { return _y; }
    float y() const { return _y; }
};

// vec4f_t (swegl/projection/vec2f.hpp:45-93): keyframe values of animations
class vec4f_t
{
    float _x = 0, _y = 0, _z = 0, _w = 0;

public:
    vec4f_t() = default;
    vec4f_t(float x, float y, float z, float w) : _x(x), _y(y), _z(z), _w(w) {}
    float &x() { return _x; }
    float x() const { return _x; }
    float &y() { return _y; }
    float y() const { return _y; }
    float &z() { return _z; }
    float z() const { return _z; }
    float &w() { return _w; }
    float w() const { return _w; }
};

// 4x4 row-major float matrix (matrix44_t, freon::Matrix<float,4,4> storage).
class matrix44_t
{
public:
    float m[4][4];
    matrix44_t() { std::memset(m, 0, sizeof m); }
    float *operator[](int r) { return m[r]; }
    const float *operator[](int r) const { return m[r]; }
    static matrix44_t identity()
    {
        matrix44_t r;
        for (int i = 0; i < 4; i++)
            r.m[i][i] = 1.0f;
        return r;
    }
    void rotate_x(float a) { swegl_matrix_rotate_x(&m[0][0], a); }
    void rotate_y(float a) { swegl_matrix_rotate_y(&m[0][0], a); }
    void rotate_z(float a) { swegl_matrix_rotate_z(&m[0][0], a); }
    void translate(float x, float y, float z) { swegl_matrix_translate(&m[0][0], x, y, z); }
    // matrix44.hpp:34-43: q0 is the scalar part
    static matrix44_t from_quaternion(float q0, float q1, float q2, float q3)
    {
        matrix44_t r;
        swegl_matrix_from_quaternion(&r.m[0][0], q0, q1, q2, q3);
        return r;
    }
};

}  // namespace swegl
