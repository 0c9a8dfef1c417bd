"""Shared helpers for the GPU parity tests: render a scene on the GPU through
the C ABI with per-pixel triangle ids, and compare against the oracle."""
import ctypes as C

import numpy as np

from swegl_amd import _abi as A
from swegl_amd.device import DeviceScene, Target, check, lib
from swegl_amd.host import SceneDescHolder


class GpuFrame:
    def __init__(self, pixels, zbuffer, tid, stats):
        self.pixels, self.zbuffer, self.tid, self.stats = pixels, zbuffer, tid, stats


def gpu_render(ctx, scene, viewports, dscene=None, with_tid=True, entry="swegl_render"):
    """Render on the GPU through the C entry point `entry` (swegl_render or
    swegl_render_noalloc); each viewport's surface starts from its host pixels."""
    holder = SceneDescHolder(scene)
    own = dscene is None
    if own:
        dscene = DeviceScene(ctx, holder)
    surfaces = {}
    descs = (A.ViewportDesc * len(viewports))()
    planes = []
    for i, vp in enumerate(viewports):
        s = vp.m_screen
        if id(s) not in surfaces:
            t = Target(ctx, A.TARGET_COLOR, s.w, s.h)
            t.write(s.pixels)
            surfaces[id(s)] = t
        z = Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)
        tid = Target(ctx, A.TARGET_UINT, vp.m_w, vp.m_h) if with_tid else None
        planes.append((surfaces[id(s)], z, tid))
        descs[i] = vp.desc(surfaces[id(s)], z, tid)
    st = (A.FrameStats * 1)()
    check(getattr(lib(), entry)(ctx.h, dscene.h, descs, len(viewports), st), entry)
    frames = []
    stats = {k: getattr(st[0], k) for k, _ in A.FrameStats._fields_}
    for (t, z, tid) in planes:
        frames.append(GpuFrame(t.read(), z.read(), tid.read() if tid is not None else None, stats))
    if own:
        dscene.close()
    return frames


def compare(got, want, colour_tol=1, check_tid=True):
    """Assert bit-exact depth and triangle ids, colour within colour_tol LSB.
    Returns (max colour diff, number of differing pixels)."""
    zg = got.zbuffer.view(np.uint32)
    zw = want.zbuffer.view(np.uint32)
    bad = np.argwhere(zg != zw)
    assert bad.size == 0, f"depth differs at {len(bad)} pixels, first {bad[:5].tolist()}"
    if check_tid and got.tid is not None:
        bad = np.argwhere(got.tid != want.tid)
        assert bad.size == 0, f"triangle id differs at {len(bad)} pixels, first {bad[:5].tolist()}"
    pg = got.pixels.view(np.uint8).astype(np.int16)
    pw = want.pixels.view(np.uint8).astype(np.int16)
    d = np.abs(pg - pw)
    mx = int(d.max()) if d.size else 0
    ndiff = int((d.reshape(got.pixels.shape + (4,)).max(-1) > 0).sum())
    assert mx <= colour_tol, f"colour differs by up to {mx} LSB at {ndiff} pixels"
    return mx, ndiff


class Batch:
    """Device state for a list of (scene, viewports) jobs."""

    def __init__(self, ctx, pairs, with_tid=True):
        self.ctx = ctx
        self.pairs = pairs
        self.dscenes = [DeviceScene(ctx, SceneDescHolder(s)) for s, _ in pairs]
        self.targets = []
        self.descs = []
        for s, vps in pairs:
            arr = (A.ViewportDesc * len(vps))()
            tl = []
            for i, vp in enumerate(vps):
                col = Target(ctx, A.TARGET_COLOR, vp.m_screen.w, vp.m_screen.h)
                col.write(vp.m_screen.pixels)
                z = Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)
                tid = Target(ctx, A.TARGET_UINT, vp.m_w, vp.m_h) if with_tid else None
                arr[i] = vp.desc(col, z, tid)
                tl.append((col, z, tid))
            self.descs.append(arr)
            self.targets.append(tl)
        self.jobs = (A.Job * len(pairs))()
        for k, (s, vps) in enumerate(pairs):
            self.jobs[k] = A.Job(self.dscenes[k].h, self.descs[k], len(vps))

    def refresh_cameras(self):
        for k, (s, vps) in enumerate(self.pairs):
            for i, vp in enumerate(vps):
                col, z, tid = self.targets[k][i]
                self.descs[k][i] = vp.desc(col, z, tid)

    def render(self, graph=False, stats=True):
        st = (A.FrameStats * 1)()
        fn = lib().swegl_render_batch_graph if graph else lib().swegl_render_batch
        check(fn(self.ctx.h, self.jobs, len(self.pairs), st if stats else None), "render_batch")
        if not stats:
            self.ctx.sync()
        return {k: getattr(st[0], k) for k, _ in A.FrameStats._fields_}

    def frames(self):
        return [[GpuFrame(c.read(), z.read(), t.read() if t is not None else None, None)
                 for (c, z, t) in tl] for tl in self.targets]


def same_frame(a, b):
    return (np.array_equal(a.pixels, b.pixels) and np.array_equal(a.zbuffer.view(np.uint32),
                                                                 b.zbuffer.view(np.uint32))
            and (a.tid is None or b.tid is None or np.array_equal(a.tid, b.tid)))
