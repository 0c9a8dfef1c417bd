"""The batched and served forms at the sizes of the measured workloads:
the full 256-scene mixed batch, full-size heavy scenes (500,000 triangles,
four 1080p viewports each) and a served stream with 5 % 1M-triangle 1080p
requests.  Every batched or served frame must equal the same job rendered
alone, bit for bit, and a few of them must match the oracle."""
import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A
from swegl_amd import autotune, scenes
from swegl_amd.device import Context, DeviceScene, Server, Target, check, lib
from swegl_amd.host import SceneDescHolder

from _helpers import Batch, GpuFrame, compare, same_frame

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


class Alone:
    """Targets for rendering one job at a time with swegl_render, reused
    across jobs of the same viewport sizes."""

    def __init__(self, ctx):
        self.ctx, self.pool = ctx, {}

    def render(self, ds, vps):
        descs = (A.ViewportDesc * len(vps))()
        tg = []
        for i, vp in enumerate(vps):
            key = (i, vp.m_screen.w, vp.m_screen.h, vp.m_w, vp.m_h)
            if key not in self.pool:
                self.pool[key] = (Target(self.ctx, A.TARGET_COLOR, vp.m_screen.w, vp.m_screen.h),
                                  Target(self.ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h),
                                  Target(self.ctx, A.TARGET_UINT, vp.m_w, vp.m_h))
            col, z, tid = self.pool[key]
            col.write(vp.m_screen.pixels)
            descs[i] = vp.desc(col, z, tid)
            tg.append((col, z, tid))
        check(lib().swegl_render(self.ctx.h, ds.h, descs, len(vps), None), "swegl_render")
        self.ctx.sync()
        return [GpuFrame(c.read(), z.read(), t.read(), None) for c, z, t in tg]


def test_full_mixed_batch_equals_plain(ctx):
    """All 256 scenes of the mixed batch (1,000-200,000 triangles, sides
    256-1024, DoF on half) in one batched graph, with the settings tuned for
    batches (tile, bin limit, workgroups per CU), equal each scene alone
    with the settings tuned for single renders."""
    tuning = autotune.load()
    pairs = scenes.batch_scenes(1)
    assert len(pairs) == 256
    print(f"batch of {len(pairs)} scenes, {sum(s.n_triangles() for s, _ in pairs)} triangles", flush=True)
    if "batch" in tuning:
        ctx.apply_tuning(tuning["batch"])
    b = Batch(ctx, pairs)
    b.render(graph=True, stats=False)
    b.render(graph=True, stats=False)   # a replay of the captured graph
    if "heightfield" in tuning:
        ctx.apply_tuning(tuning["heightfield"])
    alone = Alone(ctx)
    for k, (s, vps) in enumerate(pairs):
        got = [GpuFrame(c.read(), z.read(), t.read(), None) for c, z, t in b.targets[k]]
        want = alone.render(b.dscenes[k], vps)
        for i in range(len(vps)):
            assert same_frame(got[i], want[i]), (k, i)
        if k in (0, 1, 2):
            ref = oracle.render(s, vps)
            for i in range(len(vps)):
                compare(got[i], ref[i], colour_tol=2)
    for d in b.dscenes:
        d.close()


def test_full_heavy_scenes_batched_equal_plain(ctx):
    """Four heavy scenes at full size (500 clusters of 1,000 small triangles
    stacked 8-20 deep, four 1920x1080 viewports with DoF, 4096^2 texture,
    sun + 8 point lights) batched in one graph equal each scene alone; one
    viewport matches the oracle."""
    pairs = scenes.heavy_scenes(3, count=4)
    print(f"heavy batch of {len(pairs)} scenes x {len(pairs[0][1])} viewports", flush=True)
    b = Batch(ctx, pairs)
    b.render(graph=True, stats=False)
    alone = Alone(ctx)
    for k, (s, vps) in enumerate(pairs):
        got = [GpuFrame(c.read(), z.read(), t.read(), None) for c, z, t in b.targets[k]]
        want = alone.render(b.dscenes[k], vps)
        for i in range(len(vps)):
            assert same_frame(got[i], want[i]), (k, i)
        if k == 0:
            ref = oracle.render(s, vps[:1])
            compare(got[0], ref[0], colour_tol=2)
    for d in b.dscenes:
        d.close()


def test_served_stream_with_large_requests(ctx):
    """200 requests through 3 lanes: 95 % drawn from the 256-scene mixed set,
    5 % of the 999,698-triangle 1920x1080 DoF shape, each with its own camera
    pose.  Every result equals the request rendered alone, and three of them
    (one of the large shape) match the oracle."""
    rng = np.random.default_rng(11)
    pairs = scenes.batch_scenes(2)
    big = scenes.heightfield_scene(0)
    n_req = 200
    which = np.where(rng.random(n_req) < 0.05, -1, rng.integers(0, len(pairs), n_req))
    which[7] = -1                       # at least one large request, early in the stream
    print(f"served stream: {n_req} requests, {(which == -1).sum()} large", flush=True)
    ds = {}
    reqs = []
    for i in range(n_req):
        k = int(which[i])
        s, vps = big if k < 0 else pairs[k]
        if k not in ds:
            ds[k] = DeviceScene(ctx, SceneDescHolder(s))
        vp = vps[0]
        vp.m_camera.rotate_y(float(rng.uniform(-0.2, 0.2)))
        col = Target(ctx, A.TARGET_COLOR, vp.m_screen.w, vp.m_screen.h)
        z = Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)
        tid = Target(ctx, A.TARGET_UINT, vp.m_w, vp.m_h)
        arr = (A.ViewportDesc * 1)()
        arr[0] = vp.desc(col, z, tid)
        reqs.append(dict(k=k, desc=arr, targets=(col, z, tid), view=vp.m_camera.m_viewmatrix.copy(),
                         job=A.Job(ds[k].h, arr, 1)))
    alone = []
    for r in reqs:
        check(lib().swegl_render(ctx.h, ds[r["k"]].h, r["desc"], 1, None), "render")
        ctx.sync()
        alone.append(tuple(t.read() for t in r["targets"]))
        for t in r["targets"]:
            t.fill(0)
    srv = Server(ctx, budget_bytes=200 * 10**9, max_batch=64, lanes=3)
    for r in reqs:
        srv.submit(r["job"])
    srv.drain()
    st = srv.stats()
    assert st["completed"] == n_req and st["batches"] >= 3
    assert st["peak_bytes_in_flight"] <= 200 * 10**9
    for i, (r, a) in enumerate(zip(reqs, alone)):
        got = tuple(t.read() for t in r["targets"])
        for g, w in zip(got, a):
            assert np.array_equal(g.view(np.uint32), w.view(np.uint32)), i
    checked = [7] + [i for i in range(n_req) if which[i] >= 0][:2]
    for i in checked:
        r = reqs[i]
        s, vps = big if r["k"] < 0 else pairs[r["k"]]
        vps[0].m_camera.m_viewmatrix[...] = r["view"]
        want = oracle.render(s, vps[:1])[0]
        col, z, tid = (t.read() for t in r["targets"])
        compare(GpuFrame(col, z, tid, None), want, colour_tol=2)
    srv.close()
    for d in ds.values():
        d.close()
