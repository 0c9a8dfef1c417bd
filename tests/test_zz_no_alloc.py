"""The allocation-free entry point swegl_render_noalloc: the same kernels on
the device's fixed scratch array give the same frames as swegl_render, on two
calls in a row (the second smaller, so it runs on what the first left in the
array), and sizes beyond the array are refused before anything runs."""
import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A
from swegl_amd import scenes
from swegl_amd.device import Context, DeviceScene, SweglError, check, lib
from swegl_amd.host import SceneDescHolder, pixel_shader_texture, post_shader_t, viewport_t

from _helpers import compare, gpu_render, same_frame

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _both(ctx, scene, vps):
    print(f"noalloc: {scene.n_triangles()} triangles, viewports "
          f"{[(v.m_w, v.m_h) for v in vps]}", flush=True)
    got = gpu_render(ctx, scene, vps, entry="swegl_render_noalloc")
    want = gpu_render(ctx, scene, vps, entry="swegl_render")
    for g, w in zip(got, want):
        assert same_frame(g, w)
        assert g.stats == w.stats
    return got


def test_noalloc_equals_render_on_two_calls_second_smaller(ctx):
    big = scenes.heightfield_scene(3, n=200, width=512, height=384, tex=256, dof=True)
    small = scenes.cube_scene(5, width=96, height=72)
    got = _both(ctx, *big)
    want = oracle.render(*big)
    compare(got[0], want[0], colour_tol=2)
    got = _both(ctx, *small)
    compare(got[0], oracle.render(*small)[0], colour_tol=1)


def test_noalloc_two_viewports_and_generic_shader(ctx):
    """Two viewports of one scene (two waves when they overlap) and a shader
    other than Phong + bilinear."""
    scene, vps = scenes.cube_scene(2, width=160, height=120)
    vp2 = viewport_t(40, 30, 100, 80, vps[0].m_screen, pixel_shader_texture(), 0)
    vp2.m_camera = vps[0].m_camera
    vp2.set_post_shader(post_shader_t())
    _both(ctx, scene, [vps[0], vp2])


def test_noalloc_refuses_sizes_beyond_the_array(ctx):
    """Bins of 4096 tiles per triangle for a 20k-triangle scene need far more
    than the array: SWEGL_ERANGE before anything is enqueued; the next call
    that fits still works."""
    scene, vps = scenes.heightfield_scene(4, n=100, width=256, height=192, tex=64, dof=False)
    check(lib().swegl_ctx_set_bin_limit(ctx.h, 4096), "set_bin_limit")
    try:
        with pytest.raises(SweglError) as e:
            gpu_render(ctx, scene, vps, entry="swegl_render_noalloc")
        assert e.value.rc == A.ERANGE
    finally:
        check(lib().swegl_ctx_set_bin_limit(ctx.h, 16), "set_bin_limit")
    _both(ctx, scene, vps)


def test_noalloc_no_viewports_zero_stats(ctx):
    scene, _ = scenes.cube_scene(1, width=32, height=32)
    ds = DeviceScene(ctx, SceneDescHolder(scene))
    st = (A.FrameStats * 1)()
    st[0].fragments = 12345
    for fn in ("swegl_render_noalloc", "swegl_render", "swegl_render_graph"):
        check(getattr(lib(), fn)(ctx.h, ds.h, None, 0, st), fn)
        assert st[0].fragments == 0
        st[0].fragments = 777
    ds.close()
