// swegl/projection/camera.hpp -- camera_t (reference:
// swegl/projection/camera.hpp, src/projection/camera.cpp).
#pragma once

#include "points.hpp"

namespace swegl
{

class camera_t
{
public:
    float m_xfov = 0, m_yfov = 0;   // unused, as in the reference
    vertex_t m_center;
    matrix44_t m_viewmatrix;
    matrix44_t m_projectionmatrix;

    explicit camera_t(float aspect_ratio)
    {
        float c[3];
        swegl_camera_init(aspect_ratio, &m_viewmatrix.m[0][0], &m_projectionmatrix.m[0][0], c);
        m_center = vertex_t(c[0], c[1], c[2]);
    }
    void rotate_x(float a) { swegl_camera_rotate(&m_viewmatrix.m[0][0], 0, a); }
    void rotate_y(float a) { swegl_camera_rotate(&m_viewmatrix.m[0][0], 1, a); }
    void rotate_z(float a) { swegl_camera_rotate(&m_viewmatrix.m[0][0], 2, a); }
    void translate(float x, float y, float z)
    {
        float c[3] = {m_center.x(), m_center.y(), m_center.z()};
        swegl_camera_translate(&m_viewmatrix.m[0][0], c, x, y, z);
        m_center = vertex_t(c[0], c[1], c[2]);
    }
    vertex_t position() const { return m_center; }
};

}  // namespace swegl
