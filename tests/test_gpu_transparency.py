"""GPU parity of transparency layers (renderer.cpp:486-550, viewport.cpp:43-86,
colors.cpp:39-57, post_shaders.hpp:22-47) against the oracle: layered quads
of known alpha, offset viewports (whose flatten reads the surface origin),
and textured scenes whose random texel alpha makes most fragments
transparent, with 1 to 4 layers."""
import numpy as np
import pytest

import oracle
from swegl_amd import host, scenes
from swegl_amd.device import Context
from swegl_amd.host import pixel_colors

from _helpers import compare, gpu_render
from test_oracle_transparency import BG, quads_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def check(ctx, scene, vps):
    for vp in vps:
        vp.m_screen.pixels[:] = BG
    want = oracle.render(scene, vps)
    got = gpu_render(ctx, scene, vps)
    for g, w in zip(got, want):
        compare(g, w, colour_tol=1)
    return got, want


@pytest.mark.parametrize("layers", [1, 2, 3, 4])
def test_layered_quads(ctx, layers):
    quads = [(0.5, 0.5, pixel_colors(0, 0, 200, 100)), (-0.5, 0.7, pixel_colors(200, 0, 0, 90)),
             (0.0, 0.3, pixel_colors(0, 200, 0, 255)), (0.2, 0.9, pixel_colors(90, 90, 90, 60)),
             (-0.2, 0.4, pixel_colors(10, 200, 30, 140))]
    scene = quads_scene(quads)
    surf = host.Surface(64, 48)
    vp = host.viewport_t(0, 0, 64, 48, surf, host.pixel_shader_t(), layers)
    vp.m_camera.translate(0, 0, -3)
    vp.set_post_shader(host.post_shader_t())
    got, want = check(ctx, scene, [vp])
    assert (want[0].pixels != BG).sum() > 500


def test_offset_viewport(ctx):
    scene = quads_scene([(0.0, 2.0, pixel_colors(200, 100, 50, 128)), (0.3, 0.5, pixel_colors(9, 9, 200, 255))])
    surf = host.Surface(48, 30)
    vp = host.viewport_t(20, 10, 16, 12, surf, host.pixel_shader_t(), 2)
    vp.m_camera.translate(0, 0, -3)
    vp.set_post_shader(host.post_shader_t())
    check(ctx, scene, [vp])


@pytest.mark.parametrize("seed,layers", [(0, 3), (4, 1), (6, 4)])
def test_textured_cube_with_layers(ctx, seed, layers):
    """Random texel alpha: almost every fragment is transparent."""
    scene, vps = scenes.cube_scene(seed)
    vps[0].transparency_layer_count = layers
    vps[0].m_got_transparency = True
    check(ctx, scene, vps)


def test_height_field_with_layers_and_other_shaders(ctx):
    scene, vps = scenes.heightfield_scene(3, n=40, width=160, height=120, tex=64, dof=False)
    vps[0].transparency_layer_count = 2
    for sh in (host.pixel_shader_light_and_texture(), host.pixel_shader_texture(),
               host.pixel_shader_light_and_texture(host.pixel_shader_lights_flat, host.pixel_shader_texture)):
        vps[0].m_pixel_shader = sh
        check(ctx, scene, vps)


def test_layers_next_to_plain_viewports(ctx):
    """A viewport with layers between two plain ones on the same surface keeps
    the reference's order of effects."""
    scene, _ = scenes.cube_scene(2)
    surf = host.Surface(200, 120)
    vps = []
    for k, (x, layers) in enumerate([(0, 0), (60, 2), (120, 0)]):
        vp = host.viewport_t(x, 20, 80, 80, surf, host.pixel_shader_light_and_texture(), layers)
        vp.m_camera.translate(0, 0, -3)
        scenes.set_fov(vp.m_camera, np.float32(1.0))
        vp.set_post_shader(host.post_shader_t())
        vps.append(vp)
    check(ctx, scene, vps)
