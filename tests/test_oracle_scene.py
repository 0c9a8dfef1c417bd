"""Scene-level pins of the oracle: draw-set semantics of renderer.cpp:77-359
(visibility-mark quirk, double-sided materials, near clip, first minimum
wins) and invariants of full frames."""
import math

import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A
from swegl_amd import scenes
from swegl_amd.host import (Surface, material_t, node_t, pixel_colors,
                            pixel_shader_light_and_texture, post_shader_t,
                            primitive_t, scene_t, texture_t, viewport_t)

MAXZ_BITS = 0x7F7F7F7F


def flat_scene(tris, double_sided=False, mode=A.TRIANGLES):
    """Triangles (lists of 3 world points, camera looks down -z from z=+3)."""
    pts = np.array([p for t in tris for p in t], np.float32)
    n = len(pts)
    prim = primitive_t(pts, np.tile([0, 0, 1], (n, 1)), np.full((n, 2), 0.5), np.arange(n), mode, 0)
    s = scene_t()
    s.images = [texture_t(np.full((4, 4), pixel_colors(100, 150, 200, 255), np.uint32))]
    s.materials = [material_t(pixel_colors(255, 255, 255, 255), 1, 1, 0, double_sided)]
    s.nodes = [node_t([prim])]
    s.root_nodes = [0]
    s.ambient_light_intensity = 0.5
    s.sun_direction = np.array([0, 0, -1], np.float32)
    s.sun_intensity = 0.5
    return s


def view(w=64, h=64, z=3.0):
    vp = viewport_t(0, 0, w, h, Surface(w, h), pixel_shader_light_and_texture(), 0)
    vp.m_camera.translate(0, 0, -z)
    vp.set_post_shader(post_shader_t())
    return vp


A_, B_, C_ = [-0.5, -0.5, 0], [0.5, -0.5, 0], [0.0, 0.5, 0]


def test_back_face_culled_alone():
    s = flat_scene([[A_, C_, B_]])       # clockwise on screen: back facing
    fr = oracle.render(s, [view()])[0]
    assert fr.stats["drawn_subtriangles"] == 0
    assert (fr.tid == 0xFFFFFFFF).all()


def test_marking_quirk_draws_back_face_with_shared_vertices():
    """A back-facing triangle whose three vertices are marked by a front
    triangle is rasterised (renderer.cpp:83-185 marks vertices, :248-253
    tests only the marks)."""
    pts = np.array([A_, B_, C_], np.float32)
    prim = primitive_t(pts, np.tile([0, 0, 1], (3, 1)), np.full((3, 2), 0.5), [0, 1, 2, 0, 2, 1],
                       A.TRIANGLES, 0)
    s = flat_scene([[A_, B_, C_]])
    s.nodes = [node_t([prim])]
    fr = oracle.render(s, [view()])[0]
    assert fr.stats["drawn_subtriangles"] == 2


def test_double_sided_draws_back_face():
    s = flat_scene([[A_, C_, B_]], double_sided=True)
    fr = oracle.render(s, [view()])[0]
    assert fr.stats["drawn_subtriangles"] == 1
    assert (fr.tid != 0xFFFFFFFF).sum() > 50


def test_first_minimum_wins_and_order():
    """Equal depth: the earlier draw keeps the pixel (z >= zb rejects,
    renderer.cpp:491)."""
    s = flat_scene([[A_, B_, C_], [A_, B_, C_]])
    fr = oracle.render(s, [view()])[0]
    ids = fr.tid[fr.tid != 0xFFFFFFFF]
    assert ids.size > 50 and (ids == 0).all()
    # a nearer second triangle wins everywhere it covers
    near = [[p[0], p[1], 0.25] for p in (A_, B_, C_)]
    s = flat_scene([[A_, B_, C_], near])
    fr = oracle.render(s, [view()])[0]
    ids = fr.tid[fr.tid != 0xFFFFFFFF]
    assert (ids == 2).mean() > 0.9


def test_near_clip_fragments_stay_in_front():
    """A triangle crossing the near plane z' = 0.001 is cut (renderer.cpp:286-356);
    all its written depths exceed 0.001 and it reaches the viewport edge."""
    # camera at z=3 looking down -z; near distance 0.5 => z' = 0 at world z = 2.5
    tri = [[-0.4, -0.3, 0.0], [0.4, -0.3, 0.0], [0.0, -0.3, 2.9]]
    s = flat_scene([tri], double_sided=True)
    fr = oracle.render(s, [view(80, 60)])[0]
    assert fr.stats["clipped_triangles"] == 1
    zz = fr.zbuffer[fr.tid != 0xFFFFFFFF]
    assert zz.size > 0 and (zz > np.float32(0.001)).all()


def test_strip_and_fan_winding():
    """Strips flip winding on odd triangles (renderer.cpp:147,199-202); a
    front-facing strip quad draws both triangles, as does a fan."""
    quad = [[-0.5, -0.5, 0], [0.5, -0.5, 0], [-0.5, 0.5, 0], [0.5, 0.5, 0]]
    pts = np.array(quad, np.float32)
    for mode, idx in ((A.TRIANGLE_STRIP, [0, 1, 2, 3]), (A.TRIANGLE_FAN, [0, 1, 3, 2])):
        prim = primitive_t(pts, np.tile([0, 0, 1], (4, 1)), np.full((4, 2), 0.5), idx, mode, 0)
        s = flat_scene([[A_, B_, C_]])
        s.nodes = [node_t([prim])]
        fr = oracle.render(s, [view()])[0]
        assert fr.stats["drawn_subtriangles"] == 2, mode


@pytest.mark.parametrize("seed", range(4))
def test_cube_frame_invariants(seed):
    scene, vps = scenes.cube_scene(seed)
    fr = oracle.render(scene, vps)[0]
    cov = fr.tid != 0xFFFFFFFF
    zb = fr.zbuffer.view(np.uint32)
    assert (zb[~cov] == MAXZ_BITS).all()
    assert (fr.pixels[~cov] == 0).all()
    z = fr.zbuffer[cov]
    # corners lie within sqrt(3)/2 of the centre at distance 3: z' = (10/9.5)(z_c - 0.5)
    assert (z > (10 / 9.5) * (2.5 - 0.867)).all() and (z < (10 / 9.5) * (2.5 + 0.867)).all()
    # a convex object: each covered pixel is covered by exactly one front face,
    # so fragments == covered pixels
    assert fr.stats["fragments"] == cov.sum()
    # only whole faces (both triangles of a fan) are drawn
    faces = np.unique(fr.tid[cov] // 4)
    assert 1 <= len(faces) <= 3


def test_heightfield_frame_invariants():
    scene, vps = scenes.heightfield_scene(5, n=96, width=320, height=200, tex=64, dof=False)
    fr = oracle.render(scene, vps)[0]
    cov = fr.tid != 0xFFFFFFFF
    assert cov.mean() > 0.3
    assert fr.stats["fragments"] >= cov.sum()
    assert fr.stats["nan_fragments"] == 0
    assert (fr.tid[cov] // 2 < scene.n_triangles()).all()


@pytest.mark.parametrize("shift", [(-4, 0), (4, 0), (0, -4), (0, 4)])
def test_frustum_rejects_triangles_outside_each_side(shift):
    """inside_camera_frustum (renderer.cpp:58-70): a triangle wholly beyond
    one side of the frustum marks nothing and is never filled."""
    dx, dy = shift
    tri = [[p[0] + dx, p[1] + dy, p[2]] for p in (A_, B_, C_)]
    fr = oracle.render(flat_scene([tri], double_sided=True), [view()])[0]
    assert fr.stats["drawn_subtriangles"] == 0


def test_strip_odd_triangle_keeps_front_facing_shading():
    """The odd triangle of a strip is drawn with the flipped order
    (renderer.cpp:199-202), so a flat front-facing strip is lit uniformly."""
    quad = [[-0.5, -0.5, 0], [0.5, -0.5, 0], [-0.5, 0.5, 0], [0.5, 0.5, 0]]
    prim = primitive_t(np.array(quad, np.float32), np.tile([0, 0, 1], (4, 1)), np.full((4, 2), 0.5),
                       [0, 1, 2, 3], A.TRIANGLE_STRIP, 0)
    s = flat_scene([[A_, B_, C_]])
    s.nodes = [node_t([prim])]
    fr = oracle.render(s, [view()])[0]
    cov = fr.tid != 0xFFFFFFFF
    c0 = np.unique(fr.pixels[cov & (fr.tid == 0)])
    c1 = np.unique(fr.pixels[cov & (fr.tid == 2)])
    assert len(c0) == 1 and len(c1) == 1 and c0[0] == c1[0]


def test_inverted_triangle_negates_normals():
    """A back-facing double-sided triangle is shaded with negated normals
    (pixel_shaders.cpp:97-104): normals pointing away from the camera and a
    sun shining from the camera side light it fully."""
    pts = np.array([A_, C_, B_], np.float32)
    prim = primitive_t(pts, np.tile([0, 0, -1], (3, 1)), np.full((3, 2), 0.5), [0, 1, 2], A.TRIANGLES, 0)
    s = flat_scene([[A_, C_, B_]], double_sided=True)
    s.nodes = [node_t([prim])]
    s.ambient_light_intensity = 0.25
    s.sun_direction = np.array([0, 0, -1], np.float32)
    s.sun_intensity = 0.25
    fr = oracle.render(s, [view()])[0]
    px = np.unique(fr.pixels[fr.tid != 0xFFFFFFFF])
    # light 0.5: texel (100,150,200) halves
    assert len(px) == 1 and px[0] == pixel_colors(50, 75, 100, 255)


def test_frustum_side_met_by_any_vertex():
    """inside_camera_frustum (renderer.cpp:58-70) asks, per side, that at
    least one vertex be inside it, not that every vertex be: a triangle with
    one vertex left of x' = -1 is drawn, and a triangle around the whole
    view, no vertex of which lies inside, covers every pixel."""
    # world z = 0 is at z' = (10/9.5)(3 - 0.5): x' = x / 2.63, so x = -4 is left of -1
    fr = oracle.render(flat_scene([[[-4, -0.5, 0], B_, C_]]), [view()])[0]
    assert fr.stats["drawn_subtriangles"] == 1
    assert (fr.tid != 0xFFFFFFFF).sum() > 300
    big = [[-20, -20, 0], [20, -20, 0], [0, 20, 0]]
    fr = oracle.render(flat_scene([big]), [view()])[0]
    assert fr.stats["drawn_subtriangles"] == 1
    assert (fr.tid != 0xFFFFFFFF).all()


def _identity_view(w, h):
    """A viewport whose view and projection are identities: z' is the world z
    exactly and x' = x / |z|, y' = y / |z|."""
    vp = viewport_t(0, 0, w, h, Surface(w, h), pixel_shader_light_and_texture(), 0)
    vp.m_camera.m_viewmatrix[:] = np.eye(4, dtype=np.float32)
    vp.m_camera.m_projectionmatrix[:] = np.eye(4, dtype=np.float32)
    vp.set_post_shader(post_shader_t())
    return vp


def test_near_rejection_compares_in_double():
    """The near tests compare the float z' with the double literal 0.001
    (renderer.cpp:489, :58-70, :281): 0.001f is above 0.001, so a triangle of
    constant depth 0.001f is marked, set up and every fragment kept (a float
    literal 0.001f would reject them all); at the float below, nothing is
    marked."""
    z = np.float32(0.001)
    for depth, drawn in ((z, True), (np.nextafter(z, np.float32(0)), False)):
        h = np.float32(0.5) * depth
        tri = [[-h, -h, depth], [h, -h, depth], [0, h, depth]]
        fr = oracle.render(flat_scene([tri], double_sided=True), [_identity_view(64, 64)])[0]
        cov = fr.tid != 0xFFFFFFFF
        if drawn:
            assert fr.stats["fragments"] > 200 and cov.sum() == fr.stats["fragments"]
            assert (fr.zbuffer[cov].view(np.uint32) == np.float32(depth).view(np.uint32)).all()
        else:
            assert fr.stats["drawn_subtriangles"] == 0 and not cov.any()


def test_near_clip_cuts_at_the_near_plane():
    """Only v0 in front (renderer.cpp:286-317): the edges are cut where
    z' = 0.001, cut = (z0 - 0.001f) / (z0 - z1).  With identity view and
    projection the vertex (0, 0, 1) and the two behind it at (0.001, 0, -1)
    and (0, 0.001, -1) are cut at 0.4995, whose x' = 0.0004995 / 0.001 ~ 0.4995:
    on a 400x400 viewport the clipped triangle is the right triangle with
    corners (200, 200), (299.9, 200), (200, 100.1), about 4,990 pixels.  A cut
    at z' = 0 would leave a sub-pixel sliver (no divide at z' = 0)."""
    tri = [[0, 0, 1], [0.001, 0, -1], [0, 0.001, -1]]
    fr = oracle.render(flat_scene([tri], double_sided=True), [_identity_view(400, 400)])[0]
    assert fr.stats["clipped_triangles"] == 1 and fr.stats["drawn_subtriangles"] == 1
    cov = fr.tid != 0xFFFFFFFF
    assert 4800 < cov.sum() < 5200
    ys, xs = np.nonzero(cov)
    assert xs.min() == 200 and 298 <= xs.max() <= 300
    assert 100 <= ys.min() <= 102 and ys.max() == 199
    # depths run from 1 at the front vertex down to the plane
    z = fr.zbuffer[cov]
    assert z.max() <= 1 and z.min() < np.float32(0.0011) and (z > np.float32(0.001)).all()
