"""Edge cases of the render path on the GPU against the oracle: empty scenes
and primitives, viewports of 1x1 and ragged sizes at the surface's edges,
degenerate and enormous triangles, triangles crossing the near plane and
behind the camera, and non-finite vertices."""
import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A
from swegl_amd import host, scenes
from swegl_amd.device import Context

from _helpers import compare, gpu_render
from test_oracle_scene import A_, B_, C_, flat_scene, view

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def check(ctx, scene, vps):
    want = oracle.render(scene, vps)
    got = gpu_render(ctx, scene, vps)
    for g, w in zip(got, want):
        compare(g, w, colour_tol=1)
    return got, want


def test_empty_scenes(ctx):
    s = flat_scene([[A_, B_, C_]])
    s.nodes[0].primitives = []                       # a node without primitives
    got, want = check(ctx, s, [view()])
    assert (want[0].tid == 0xFFFFFFFF).all() and (want[0].pixels == 0).all()
    s = flat_scene([[A_, B_, C_]])
    p = s.nodes[0].primitives[0]
    s.nodes[0].primitives = [host.primitive_t(p.positions, p.normals, p.tex_coords, [], A.TRIANGLES, 0)]
    check(ctx, s, [view()])                          # a primitive without indices
    s.nodes = []
    s.root_nodes = []
    check(ctx, s, [view()])                          # no nodes at all


@pytest.mark.parametrize("x,y,w,h", [(0, 0, 1, 1), (639, 479, 1, 1), (5, 7, 3, 5), (600, 3, 40, 17),
                                     (0, 460, 65, 20), (17, 0, 33, 480), (1, 1, 638, 478)])
def test_small_and_ragged_viewports(ctx, x, y, w, h):
    s, vps = scenes.cube_scene(4)
    surf = host.Surface(640, 480)
    surf.pixels[:] = host.pixel_colors(9, 8, 7, 6)
    vp = host.viewport_t(x, y, w, h, surf, host.pixel_shader_light_and_texture(), 0)
    vp.m_camera.translate(0, 0, -3)
    vp.set_post_shader(host.post_shader_t())
    check(ctx, s, [vp])


def test_degenerate_triangles(ctx):
    """Zero area, collinear, repeated vertices, and a triangle whose rows
    collapse to nothing, mixed with ordinary ones."""
    tris = [[A_, B_, C_],
            [A_, A_, A_],
            [A_, B_, B_],
            [[-0.5, 0.0, 0.0], [0.0, 0.0, 0.0], [0.5, 0.0, 0.0]],       # collinear, one row
            [[-0.5, -0.5, 0.0], [0.0, 0.0, 0.0], [0.5, 0.5, 0.0]],      # collinear diagonal
            [[0.0, 0.1, 0.1], [0.0004, 0.1, 0.1], [0.0002, 0.1003, 0.1]]]  # sub-pixel
    check(ctx, flat_scene(tris, double_sided=True), [view(96, 80)])


def test_enormous_triangles(ctx):
    """Triangles far larger than the viewport (many tiles, the big list) and
    with coordinates near the float range edges of the projection."""
    big = [[-1e4, -1e4, -1.0], [1e4, -1e4, -1.0], [0.0, 1e4, -1.0]]
    huge = [[-1e20, -1.0, -2.0], [1e20, -1.0, -2.0], [0.0, 1e20, -2.0]]
    check(ctx, flat_scene([big, huge, [A_, B_, C_]], double_sided=True), [view(120, 90)])


def test_near_plane_and_behind_the_camera(ctx):
    """Triangles cut by the near plane (one and two vertices behind it) and
    wholly behind the camera (renderer.cpp:262-356)."""
    one_behind = [[-0.4, -0.3, 0.0], [0.4, -0.3, 0.0], [0.0, -0.3, 2.9]]
    two_behind = [[-0.4, 0.3, 2.8], [0.4, 0.3, 2.9], [0.0, 0.1, 0.0]]
    behind = [[-0.4, 0.3, 3.5], [0.4, 0.3, 3.5], [0.0, 0.5, 3.6]]
    s = flat_scene([one_behind, two_behind, behind, [A_, B_, C_]], double_sided=True)
    got, want = check(ctx, s, [view(80, 60)])
    assert want[0].stats["clipped_triangles"] >= 2


def test_non_finite_vertices(ctx):
    """A vertex at NaN or infinity: the triangles using it are not drawn in
    either renderer (their comparisons fail) and the rest are untouched."""
    nan, inf = float("nan"), float("inf")
    tris = [[A_, B_, C_], [[nan, 0.0, 0.0], [0.5, 0.5, 0.0], [0.0, 0.5, 0.0]],
            [[inf, 0.0, 0.0], [0.5, -0.5, 0.0], [0.0, -0.5, 0.0]]]
    got, want = check(ctx, flat_scene(tris, double_sided=True), [view()])
    assert want[0].stats["nan_fragments"] == 0
