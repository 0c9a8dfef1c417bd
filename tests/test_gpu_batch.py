"""Batched and graph-replayed renders: every job's result is bit-identical to
rendering it alone (plain path) and matches the oracle."""
import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A
from swegl_amd import scenes
from swegl_amd.device import Context, check, lib

from _helpers import Batch, compare, gpu_render, same_frame

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def small_batch(seed=0, count=12):
    return scenes.batch_scenes(seed, count=count, tmin=500, tmax=20_000, sides=(96, 128, 160, 200))


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_plain_and_oracle(ctx, graph):
    pairs = small_batch(1)
    b = Batch(ctx, pairs)
    b.render(graph=graph)
    got = b.frames()
    for k, (s, vps) in enumerate(pairs):
        alone = gpu_render(ctx, s, vps)[0]
        assert same_frame(got[k][0], alone), k
        if k < 4:
            want = oracle.render(s, vps)[0]
            compare(got[k][0], want, colour_tol=2)


def test_heavy_batch_equals_plain(ctx):
    """Multi-viewport jobs (4 viewports per scene) batched in one graph equal
    each scene rendered alone."""
    pairs = scenes.heavy_scenes(4, count=3, n_clusters=30, width=320, height=180, tex=256)
    b = Batch(ctx, pairs)
    b.render(graph=True)
    got = b.frames()
    for k, (s, vps) in enumerate(pairs):
        alone = gpu_render(ctx, s, vps)
        for i in range(len(vps)):
            assert same_frame(got[k][i], alone[i]), (k, i)
    want = oracle.render(*pairs[0])
    for i in range(len(pairs[0][1])):
        compare(got[0][i], want[i], colour_tol=2)


def test_graph_replays_follow_camera_and_stay_exact(ctx):
    """Capture once, replay with new camera poses and eager work in between;
    every replay equals the eager render of the same pose."""
    scene, vps = scenes.heightfield_scene(7, n=120, width=256, height=192, tex=128, dof=True)
    b = Batch(ctx, [(scene, vps)])
    for step in range(4):
        vps[0].m_camera.rotate_y(0.05 * (step + 1))
        b.refresh_cameras()
        b.render(graph=True)
        eager = gpu_render(ctx, scene, vps)[0]       # eager work between replays
        got = b.frames()[0][0]
        assert same_frame(got, eager), step
        if step == 3:
            compare(got, oracle.render(scene, vps)[0], colour_tol=2)


def test_overlapping_viewports_keep_order(ctx):
    """Two viewports on one surface overlap: the second one's pixels win
    where they overlap, as in the reference's sequential _render calls."""
    scene, vps = scenes.heightfield_scene(8, n=80, width=200, height=150, tex=64, dof=False)
    from swegl_amd.host import pixel_shader_light_and_texture, post_shader_t, viewport_t
    surf = vps[0].m_screen
    v2 = viewport_t(60, 40, 120, 100, surf, pixel_shader_light_and_texture(), 0)
    v2.m_camera.translate(0.2, 1.0, -1.5)
    v2.m_camera.rotate_x(-0.5)
    v2.set_post_shader(post_shader_t())
    both = [vps[0], v2]
    want = oracle.render(scene, both)
    got = gpu_render(ctx, scene, both)
    assert np.array_equal(got[1].pixels, want[1].pixels)
    for g, w in zip(got, want):
        compare(g, w)


def test_multi_wave_graph_replays(ctx):
    """Overlapping viewports on one surface make several launch waves in one
    graph; every replay places each wave's job table where the capture put
    it, so each replay equals the eager render."""
    from swegl_amd.device import DeviceScene, Target
    from swegl_amd.host import SceneDescHolder, pixel_shader_light_and_texture, post_shader_t, viewport_t
    scene, vps = scenes.heightfield_scene(9, n=80, width=200, height=150, tex=64, dof=False)
    surf = vps[0].m_screen
    views = [vps[0]]
    for k in range(2):
        v = viewport_t(30 + 40 * k, 20 + 30 * k, 120, 100, surf, pixel_shader_light_and_texture(), 0)
        v.m_camera.translate(0.2 * k, 1.0, -1.5)
        v.m_camera.rotate_x(-0.5)
        v.set_post_shader(post_shader_t())
        views.append(v)
    eager = gpu_render(ctx, scene, views)
    ds = DeviceScene(ctx, SceneDescHolder(scene))
    col = Target(ctx, A.TARGET_COLOR, surf.w, surf.h)
    planes = [(Target(ctx, A.TARGET_FLOAT, v.m_w, v.m_h), Target(ctx, A.TARGET_UINT, v.m_w, v.m_h)) for v in views]
    d = (A.ViewportDesc * len(views))()
    for i, v in enumerate(views):
        d[i] = v.desc(col, planes[i][0], planes[i][1])
    for rep in range(4):
        col.write(surf.pixels)
        check(lib().swegl_render_graph(ctx.h, ds.h, d, len(views), None), "swegl_render_graph")
        assert np.array_equal(col.read(), eager[0].pixels), rep
        for i in range(len(views)):
            assert np.array_equal(planes[i][1].read(), eager[i].tid), (rep, i)
            assert np.array_equal(planes[i][0].read().view(np.uint32), eager[i].zbuffer.view(np.uint32)), (rep, i)
    ds.close()


@pytest.mark.parametrize("graph", [False, True])
def test_one_job_overflows_its_tiles(ctx, graph):
    """A batch of a crowded height field (19,602 triangles seen from 30 units
    back: its busiest tiles hold more entries than its share of k_front's
    fixed layout, so that job alone takes the counted, scanned and filled
    layout) between mixed-set scenes that keep their fixed layouts: every
    job equals its render alone, and the crowded one and a neighbour match
    the oracle."""
    crowd, vps = scenes.heightfield_scene(2, n=100, width=1920, height=1080, dof=False)
    vps[0].m_camera.translate(0.0, 0.0, -30.0)
    pairs = small_batch(4, count=4)
    pairs = pairs[:2] + [(crowd, vps)] + pairs[2:]
    b = Batch(ctx, pairs)
    st = b.render(graph=graph)
    got = b.frames()
    for k, (s, v) in enumerate(pairs):
        alone = gpu_render(ctx, s, v)[0]
        assert same_frame(got[k][0], alone), k
    for k in (1, 2):
        s, v = pairs[k]
        compare(got[k][0], oracle.render(s, v)[0], colour_tol=1)
    if not graph:
        assert st["fragments"] > 0
