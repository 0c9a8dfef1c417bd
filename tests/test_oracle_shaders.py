"""Pins of the oracle's other pixel shaders (pixel_shaders.hpp/.cpp): the
material colour of pixel_shader_t, pixel_shader_texture (nearest texel with
an unsigned remainder), pixel_shader_lights_flat (one light per triangle at
its centroid) and the standalone light shaders, against closed forms and
against each other."""
import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A
from swegl_amd import host, scenes
from swegl_amd.host import pixel_colors

from test_oracle_scene import A_, B_, C_, flat_scene, view


def render(scene, shader, vp=None):
    vp = vp or view()
    vp.m_pixel_shader = shader
    fr = oracle.render(scene, [vp])[0]
    return fr, fr.pixels.reshape(vp.m_screen.h, vp.m_screen.w), fr.tid != 0xFFFFFFFF


def test_default_shader_is_material_colour():
    s = flat_scene([[A_, B_, C_]])
    s.materials[0].color = pixel_colors(10, 20, 30, 40)
    fr, px, drawn = render(s, host.pixel_shader_t())
    assert drawn.sum() > 50
    assert (px[drawn] == pixel_colors(10, 20, 30, 40)).all()


def test_default_shader_takes_default_material_for_minus_one():
    s = flat_scene([[A_, B_, C_]])
    s.nodes[0].primitives[0].material_id = -1
    s.default_material = host.material_t(pixel_colors(7, 8, 9, 10))
    fr, px, drawn = render(s, host.pixel_shader_t())
    assert (px[drawn] == pixel_colors(7, 8, 9, 10)).all()


@pytest.mark.parametrize("shader", [host.pixel_shader_texture(), host.pixel_shader_texture_bilinear(),
                                    host.pixel_shader_light_and_texture(host.pixel_shader_lights_flat,
                                                                        host.pixel_shader_texture)])
def test_texture_shaders_refuse_material_minus_one(shader):
    s = flat_scene([[A_, B_, C_]])
    s.nodes[0].primitives[0].material_id = -1
    with pytest.raises(RuntimeError, match="EINVAL"):
        render(s, shader)


def test_nearest_unsigned_remainder():
    """(int)t % twidth with twidth unsigned: the int's bits as unsigned."""
    assert (2**32 - 1) % 3 == 0 and (2**32 - 2) % 3 == 2
    w, h = 3, 4
    bm = np.arange(w * h, dtype=np.uint32).reshape(h, w) + 1000
    L = oracle.lib()
    import ctypes as C
    L.oracle_nearest.restype = C.c_uint32
    L.oracle_nearest.argtypes = [A.U32P, C.c_int, C.c_int, C.c_float, C.c_float]
    def near(tx, ty):
        return L.oracle_nearest(A.u32ptr(bm), w, h, tx, ty)
    assert near(0.5, 0.5) == bm[0, 0]
    assert near(2.9, 3.9) == bm[3, 2]          # column from t.x, row from t.y
    assert near(5.7, 1.0) == bm[1, 5 % 3]
    assert near(-1.0, 0.0) == bm[0, 0]        # (2^32 - 1) % 3 == 0, not a wrap to column 2
    assert near(-2.0, 0.0) == bm[0, 2]        # (2^32 - 2) % 3 == 2
    assert near(0.0, -1.0) == bm[(2**32 - 1) % 4, 0]
    assert near(-0.5, 0.0) == bm[0, 0]        # truncation toward zero
    assert near(float("nan"), 0.0) == bm[0, 2**31 % 3]   # cvttss2si gives INT32_MIN


def test_nearest_constant_texture_and_texel_centres():
    s = flat_scene([[A_, B_, C_]])
    fr, px, drawn = render(s, host.pixel_shader_texture())
    assert (px[drawn] == pixel_colors(100, 150, 200, 255)).all()


def test_flat_light_closed_form_sun_only():
    """Face normal (0, 0, 1), sun (0, 0, -1) at 0.5, ambient 0.5: light word
    trunc(65536 * (0.5 + 0.5)) = 65536 on every pixel of the triangle."""
    s = flat_scene([[A_, B_, C_]])
    fr, px, drawn = render(s, host.pixel_shader_lights_flat())
    assert (px[drawn] == 65536).all()
    s.ambient_light_intensity = 0.1
    s.sun_intensity = 0.25
    fr, px, drawn = render(s, host.pixel_shader_lights_flat())
    want = int(np.float32(np.float32(0.1) + np.float32(0.25)) * np.float32(65536))
    assert (px[drawn] == want).all()


def test_flat_light_constant_per_triangle_and_negated_when_inverted():
    """A back-facing triangle drawn through a double-sided material uses the
    negated face normal.  Both windings sort to (C, A, B) by screen y, whose
    cross product is +z; inverted, the normal is -z, so a sun travelling
    toward +z lights it and one travelling toward -z does not."""
    s = flat_scene([[A_, C_, B_]], double_sided=True)     # clockwise on screen: inverted
    s.sun_intensity = 0.5
    s.ambient_light_intensity = 0.25
    for sun, want in (([0, 0, 1], 0.75), ([0, 0, -1], 0.25)):
        s.sun_direction = np.array(sun, np.float32)
        fr, px, drawn = render(s, host.pixel_shader_lights_flat())
        assert drawn.sum() > 50
        vals = np.unique(px[drawn])
        assert len(vals) == 1
        assert vals[0] == int(np.float32(want) * np.float32(65536))


def test_flat_light_matches_phong_with_a_point_light_over_the_centroid():
    """Flat light = the Phong sum at the centroid with the face normal: for a
    triangle whose vertex normals equal its face normal, the Phong light of
    a pixel at the centroid equals the flat light (oracle_phong at that
    point vs the flat shader's word)."""
    tri = [[-0.6, -0.5, 0.0], [0.6, -0.5, 0.0], [0.0, 0.7, 0.0]]
    s = flat_scene([tri])
    s.point_source_lights = [host.point_source_light(np.array([0.0, 0.0, 1.5], np.float32), np.float32(1.0))]
    s.ambient_light_intensity = 0.1
    s.sun_intensity = 0.0
    vp = view(96, 96)
    fr, px, drawn = render(s, host.pixel_shader_lights_flat(), vp)
    vals = np.unique(px[drawn])
    assert len(vals) == 1
    w = np.array(tri, np.float32)
    c = ((w[0] + w[1]) + w[2]) / np.float32(3)
    L = oracle.lib()
    lights = (A.PointLight * 1)()
    lights[0] = A.PointLight((0.0, 0.0, 1.5), 1.0)
    got = L.oracle_phong(A.fptr(c.astype(np.float32)), A.fptr(np.array([0, 0, 1], np.float32)),
                         A.fptr(vp.m_camera.m_center.astype(np.float32)), np.float32(0.1),
                         A.fptr(np.array(s.sun_direction, np.float32)), np.float32(0.0), lights, 1)
    assert vals[0] == got


@pytest.mark.parametrize("seed", [0, 3])
def test_standalone_phong_light_word_drives_the_combined_shader(seed):
    """pixel_shader_lights_phong alone returns the light word that
    light_and_texture<phong, T> combines: with the nearest texture, each
    pixel equals combine(texel, word)."""
    sc, vps = scenes.cube_scene(seed=seed)
    vp = vps[0]
    f_light, light, drawn = render(sc, host.pixel_shader_lights_phong(), vp)
    f_tex, tex, _ = render(sc, host.pixel_shader_texture(), vp)
    f_comb, comb, _ = render(sc, host.pixel_shader_light_and_texture(host.pixel_shader_lights_phong,
                                                                     host.pixel_shader_texture), vp)
    assert np.array_equal(f_light.tid, f_comb.tid) and np.array_equal(f_tex.tid, f_comb.tid)
    L = oracle.lib()
    ys, xs = np.nonzero(drawn)
    for y, x in list(zip(ys, xs))[::97]:
        word = int(np.array([light[y, x]], np.uint32).view(np.int32)[0])
        assert L.oracle_combine(int(tex[y, x]), word) == comb[y, x]


def test_every_shader_kind_renders_the_same_coverage():
    sc, vps = scenes.cube_scene(seed=2)
    vp = vps[0]
    base = None
    for sh in [host.pixel_shader_t(), host.pixel_shader_lights_flat(), host.pixel_shader_lights_phong(),
               host.pixel_shader_texture(), host.pixel_shader_texture_bilinear()] + \
              [host.pixel_shader_light_and_texture(Lt, Tt)
               for Lt in (host.pixel_shader_lights_flat, host.pixel_shader_lights_phong)
               for Tt in (host.pixel_shader_texture, host.pixel_shader_texture_bilinear)]:
        fr, px, drawn = render(sc, sh, vp)
        if base is None:
            base = fr
        assert np.array_equal(fr.tid, base.tid)
        assert np.array_equal(fr.zbuffer.view(np.uint32), base.zbuffer.view(np.uint32))
