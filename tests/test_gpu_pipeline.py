"""Renders queued back to back without synchronisation.  swegl_render runs a
render's front end (k_vertex, k_front, k_bin_fill) on a second stream, with
one of three sets of front-end buffers taken in turn, so that it overlaps
the previous renders' rasters and depth of field.  Every frame of such a queue,
interleaved with pose updates, graph replays and renders with statistics,
must equal the same render done alone, bit for bit."""
import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A
from swegl_amd import scenes
from swegl_amd.device import Context, DeviceScene, Target, check, lib
from swegl_amd.host import SceneDescHolder

from _helpers import compare, gpu_render

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


class Job:
    """One render: a scene, a viewport posed by a camera turn, own targets."""

    def __init__(self, ctx, ds, vp, turn):
        vp.m_camera.rotate_y(turn)
        self.ds = ds
        self.col = Target(ctx, A.TARGET_COLOR, vp.m_screen.w, vp.m_screen.h)
        self.z = Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)
        self.tid = Target(ctx, A.TARGET_UINT, vp.m_w, vp.m_h)
        self.desc = (A.ViewportDesc * 1)()
        self.desc[0] = vp.desc(self.col, self.z, self.tid)
        vp.m_camera.rotate_y(-turn)

    def render(self, ctx, entry="swegl_render", stats=None):
        check(getattr(lib(), entry)(ctx.h, self.ds.h, self.desc, 1, stats), entry)

    def read(self):
        return tuple(t.read().view(np.uint32).copy() for t in (self.col, self.z, self.tid))

    def clear(self):
        for t in (self.col, self.z, self.tid):
            t.fill(0)


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_queued_renders_equal_alone(ctx):
    """Twelve renders of three scenes (a cube, a 20k-triangle mixed-set
    scene, the 999,698-triangle 1080p height field with DoF) at several
    camera turns, queued without a sync, with a pose update, a graph replay
    and a render with statistics among them; then each alone."""
    pairs = [scenes.cube_scene(0), scenes.batch_scenes(3, count=4, tmax=20_000)[1],
             scenes.heightfield_scene(0)]
    ds = [DeviceScene(ctx, SceneDescHolder(s)) for s, _ in pairs]
    plan = [(0, 0.0), (2, 0.0), (1, 0.1), (2, 0.05), (0, -0.2), (1, -0.1), (2, -0.05), (0, 0.3),
            (2, 0.1), (1, 0.2), (0, 0.1), (2, -0.1)]
    jobs = [Job(ctx, ds[k], pairs[k][1][0], turn) for k, turn in plan]
    print(f"queue of {len(jobs)} renders", flush=True)
    st = (A.FrameStats * 1)()
    for i, j in enumerate(jobs):
        if i == 4:
            ds[0].update_pose(SceneDescHolder(pairs[0][0]))   # same pose: ordering only
        if i == 6:
            j.render(ctx, "swegl_render_graph")
        elif i == 9:
            j.render(ctx, stats=st)
        else:
            j.render(ctx)
    ctx.sync()
    queued = [j.read() for j in jobs]
    for i, j in enumerate(jobs):
        j.clear()
        ctx.sync()
        j.render(ctx)
        ctx.sync()
        assert same(j.read(), queued[i]), i
    for d in ds:
        d.close()


def test_one_target_rendered_twice_keeps_the_last(ctx):
    """Two renders into the same targets, queued back to back with different
    camera turns: the targets hold the second render."""
    s, vps = scenes.heightfield_scene(1, n=300, width=640, height=360)
    d = DeviceScene(ctx, SceneDescHolder(s))
    a = Job(ctx, d, vps[0], 0.0)
    b = Job(ctx, d, vps[0], 0.2)
    b.col, b.z, b.tid = a.col, a.z, a.tid
    vps[0].m_camera.rotate_y(0.2)
    b.desc[0] = vps[0].desc(a.col, a.z, a.tid)
    vps[0].m_camera.rotate_y(-0.2)
    for _ in range(3):
        a.render(ctx)
        b.render(ctx)
    ctx.sync()
    got = a.read()
    a.clear()
    ctx.sync()
    b.render(ctx)
    ctx.sync()
    assert same(a.read(), got)
    d.close()


@pytest.mark.parametrize("back", [10.0, 30.0])
def test_crowded_tiles_match_the_oracle(ctx, back):
    """A 19,602-triangle height field seen from afar: its triangles crowd
    into ~60 (10 units back) or ~12 (30 back) of the 2,040 32x32 tiles.  At
    30 the busiest tiles hold more entries than k_front's fixed bin layout
    has room for (512 per tile), so the frame takes the counted, scanned and
    filled layout; at 10 they fit.  Both match the oracle, through the
    plain render and through queued renders without statistics."""
    s, vps = scenes.heightfield_scene(2, n=100, width=1920, height=1080, dof=False)
    vps[0].m_camera.translate(0.0, 0.0, -back)
    want = oracle.render(s, vps)[0]
    got = gpu_render(ctx, s, vps)[0]
    compare(got, want)
    d = DeviceScene(ctx, SceneDescHolder(s))
    jobs = [Job(ctx, d, vps[0], 0.0) for _ in range(3)]
    for j in jobs:
        j.render(ctx)
    ctx.sync()
    for j in jobs:
        col, z, tid = j.read()
        assert np.array_equal(z, got.zbuffer.view(np.uint32).reshape(z.shape))
        assert np.array_equal(tid, got.tid.reshape(tid.shape))
    d.close()


def test_dof_and_plain_renders_into_one_surface(ctx):
    """A viewport with depth of field and one without, on one scene, writing
    the same colour surface and depth plane, queued in turns without a sync:
    one writes the surface from its depth-of-field pass, the other from its
    raster, and the targets hold the last render."""
    s, vd = scenes.heightfield_scene(3, n=300, width=640, height=360, dof=True)
    _, vp = scenes.heightfield_scene(3, n=300, width=640, height=360, dof=False)
    d = DeviceScene(ctx, SceneDescHolder(s))
    a = Job(ctx, d, vd[0], 0.0)
    b = Job(ctx, d, vp[0], 0.0)
    b.col, b.z, b.tid = a.col, a.z, a.tid
    b.desc[0] = vp[0].desc(a.col, a.z, a.tid)
    want = {}
    for name, j in (("dof", a), ("plain", b)):
        a.clear()
        ctx.sync()
        j.render(ctx)
        ctx.sync()
        want[name] = a.read()
    assert not np.array_equal(want["dof"][0], want["plain"][0])
    for order in (("dof", "plain"), ("plain", "dof"), ("dof", "dof", "plain"), ("plain", "plain", "dof")):
        a.clear()
        ctx.sync()
        for _ in range(3):
            for name in order:
                (a if name == "dof" else b).render(ctx)
        ctx.sync()
        assert same(a.read(), want[order[-1]]), order
    d.close()
