#!/usr/bin/env python3
"""Benchmark of the MI355X render path.

Default workload (one GPU): the 999,698-triangle height field (708x708 grid,
2048^2 texture, sun + 4 point lights) rendered to a 1920x1080 viewport with
the 9-tap depth of field, scene resident in HBM.  One step = one frame of
the plain path (vertex, marking, binning, tile raster + shading, DoF).

Prints ONE JSON line.  Extra fields: Gfragments/s, p50/p99 frame latency,
per-kernel averages, the roofline of the dominant kernel (HIP events on the
library stream) and the CPU oracle timed on this host on a bounded sample.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# 256 CUs x 4 SIMDs, one wave64 FP32 VALU instruction per SIMD every 4 cycles, 2.4 GHz
VALU_PEAK_G = 256 * 4 * 2.4 / 4
METRIC = "frames/s + Gfragments/s, batched vs plain on GPU; p50/p99 frame latency"


def algorithmic_bytes(scene, vp, f_vis):
    """Bytes a frame must move at minimum: vertex inputs (pos+normal+uv,
    32 B), indices (4 B), texels touched (bounded by 16 B per shaded pixel),
    colour+depth written (8 B/px) and the DoF pass (12 B/px)."""
    V = scene.n_vertices()
    I = sum(len(p.indices) for n in scene.nodes for p in n.primitives)
    tex = sum(t.width * t.height * 4 for t in scene.images)
    P = vp.m_w * vp.m_h
    dof = vp.m_post_shader is not None and vp.m_post_shader.kind != 0
    frame = V * 32 + I * 4 + min(tex, 16 * f_vis) + P * 8 + (P * 12 if dof else 0)
    raster = min(tex, 16 * f_vis) + P * 8
    return frame, raster


def job_bytes(scene, vps, f_vis):
    """Algorithmic bytes of one job (a scene and its viewports): the scene's
    vertex inputs and indices once, texels bounded by 16 B per shaded pixel
    over all its viewports (f_vis: shaded pixels per viewport), colour and
    depth written per viewport pixel and the DoF pass where it is on."""
    V = scene.n_vertices()
    I = sum(len(p.indices) for n in scene.nodes for p in n.primitives)
    tex = sum(t.width * t.height * 4 for t in scene.images)
    b = V * 32 + I * 4 + min(tex, 16 * int(sum(f_vis)))
    for vp in vps:
        P = vp.m_w * vp.m_h
        dof = vp.m_post_shader is not None and vp.m_post_shader.kind != 0
        b += P * 8 + (P * 12 if dof else 0)
    return b


def roofline_of(nbytes, seconds):
    gbs = nbytes / seconds / 1e9
    return {"bound": "hbm", "algorithmic_bytes": int(nbytes), "achieved": gbs, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}


def shaded_pixels(z_target):
    import numpy as np
    return int((z_target.read().view(np.uint32) != 0x7F7F7F7F).sum())


def _cpu_model():
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cpu


def cpu_baseline(scene, vps, budget_s, threads):
    """The CPU oracle on this host: one frame per thread on `threads` threads
    (throughput, the figure reported), and single-threaded frames (the
    reference renders a frame on one thread)."""
    import concurrent.futures as cf

    import oracle
    oracle.render(scene, vps)          # warm (page in, build caches)
    n1, t0 = 0, time.perf_counter()
    while True:
        oracle.render(scene, vps)
        n1 += 1
        el1 = time.perf_counter() - t0
        if el1 >= budget_s / 2 or n1 >= 200:
            break
    # oracle_render runs without the GIL (ctypes), so threads render in parallel
    stop = time.perf_counter() + budget_s / 2

    def worker():
        k = 0
        while time.perf_counter() < stop:
            oracle.render(scene, vps)
            k += 1
        return k

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        nt = sum(f.result() for f in [ex.submit(worker) for _ in range(threads)])
    elt = time.perf_counter() - t0
    cpu = _cpu_model()
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"value": nt / elt, "unit": "frames/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(), "usable_cpus": usable,
            "threads_note": "one frame per thread on the host threads allotted to one GPU",
            "sample": f"{nt} frames of the same workload, one frame per thread on {threads} threads, "
                      f"{elt:.1f} s ({cpu})",
            "single_thread": {"frames_per_s": n1 / el1, "ms_per_frame": el1 * 1e3 / n1,
                              "sample": f"{n1} frames, {el1:.1f} s"}}


def fp64_reshade(ctx, A, check, L, ds, scene, vp):
    """The frame's FP32 colours (post shader off) against the oracle's FP64
    re-shade of the same winners; depth and triangle ids must agree."""
    import numpy as np

    import oracle
    from swegl_amd import host
    from swegl_amd.device import Target

    keep = vp.m_post_shader
    vp.set_post_shader(host.post_shader_t())
    try:
        col = Target(ctx, A.TARGET_COLOR, vp.m_screen.w, vp.m_screen.h)
        z = Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)
        tid = Target(ctx, A.TARGET_UINT, vp.m_w, vp.m_h)
        d = (A.ViewportDesc * 1)()
        d[0] = vp.desc(col, z, tid)
        check(L.swegl_render(ctx.h, ds.h, d, 1, None), "swegl_render")
        ctx.sync()
        want = oracle.render(scene, [vp], reshade64=True)[0]
        g_tid, g_z = tid.read(), z.read()
        same = bool(np.array_equal(g_tid, want.tid) and
                    np.array_equal(g_z.view(np.uint32), want.zbuffer.view(np.uint32)))
        px = col.read().reshape(vp.m_screen.h, vp.m_screen.w)[vp.m_y:vp.m_y + vp.m_h, vp.m_x:vp.m_x + vp.m_w]
        e = oracle.reshade_error(px, g_tid, want.color64)
        e["depth_and_ids_bit_exact"] = same
        return e
    finally:
        vp.set_post_shader(keep)


def fp64_reshade_jobs(ctx, A, check, L, jobs, label):
    """fp64_reshade over every viewport of a list of (scene, viewports)
    jobs, each viewport rendered alone: the largest and the pixel-weighted
    mean colour error against the FP64 re-shade of the same winners."""
    from swegl_amd.device import DeviceScene
    from swegl_amd.host import SceneDescHolder

    mx, tot, px, exact, same = 0, 0.0, 0, 0.0, True
    for s, vps in jobs:
        ds = DeviceScene(ctx, SceneDescHolder(s))
        for vp in vps:
            e = fp64_reshade(ctx, A, check, L, ds, s, vp)
            mx = max(mx, e["max"])
            tot += e["mean"] * e["pixels"]
            exact += e.get("frac_exact", 1.0) * e["pixels"]
            px += e["pixels"]
            same = same and e["depth_and_ids_bit_exact"]
        ds.close()
    return {"workload": label, "max": mx, "mean": tot / max(px, 1), "pixels": px,
            "frac_exact": exact / max(px, 1), "depth_and_ids_bit_exact": same}


def plain_vs_batched(ctx, A, check, L, pairs, iters, plain_mode="graph", tunings=None):
    """Render every (scene, viewports) pair resident in HBM two ways:
    plain = one graph replay per scene (all its viewports; plain_mode
    "eager": one swegl_render call per scene instead), batched = one graph
    replay for the whole set.  A frame is one viewport.  Each form carries
    its HBM roofline: the jobs' algorithmic bytes over the measured time.
    tunings: {"plain": entry, "batched": entry} of tuning.json applied
    before each form, each form timed with the settings tuned for it."""
    from swegl_amd.device import DeviceScene, Target
    from swegl_amd.host import SceneDescHolder

    count = len(pairs)
    ds = [DeviceScene(ctx, SceneDescHolder(s)) for s, _ in pairs]
    keep, descs = [], []
    for s, vps in pairs:
        arr = (A.ViewportDesc * len(vps))()
        for i, vp in enumerate(vps):
            col = Target(ctx, A.TARGET_COLOR, vp.m_screen.w, vp.m_screen.h)
            z = Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)
            keep += [col, z]
            arr[i] = vp.desc(col, z)
        descs.append(arr)
    jobs = (A.Job * count)()
    for k in range(count):
        jobs[k] = A.Job(ds[k].h, descs[k], len(descs[k]))
    n_frames = sum(len(d) for d in descs)

    plain_fn = L.swegl_render_graph if plain_mode == "graph" else L.swegl_render
    other_fn = L.swegl_render if plain_mode == "graph" else L.swegl_render_graph

    def plain():
        for k in range(count):
            check(plain_fn(ctx.h, ds[k].h, descs[k], len(descs[k]), None), "render_plain")

    def plain_other():
        for k in range(count):
            check(other_fn(ctx.h, ds[k].h, descs[k], len(descs[k]), None), "render_plain")

    def batched():
        check(L.swegl_render_batch_graph(ctx.h, jobs, count, None), "render_batch_graph")

    st = (A.FrameStats * 1)()
    check(L.swegl_render_batch(ctx.h, jobs, count, st), "render_batch")
    frags = st[0].fragments
    out = {"scenes": count, "frames": n_frames, "triangles": int(sum(s.n_triangles() for s, _ in pairs)),
           "pixels": int(sum(vp.m_w * vp.m_h for _, v in pairs for vp in v)), "fragments": int(frags)}
    tunings = tunings or {}
    for name, fn in (("plain", plain), ("plain_other", plain_other), ("batched", batched)):
        if tunings.get(name if name != "plain_other" else "plain"):
            ctx.apply_tuning(tunings[name if name != "plain_other" else "plain"])
        fn()
        fn()
        ctx.sync()
        a = time.perf_counter()
        for _ in range(iters):
            fn()
        ctx.sync()
        el = time.perf_counter() - a
        out[name] = {"frames_per_s": n_frames * iters / el, "gfragments_per_s": frags * iters / el / 1e9,
                     "ms_per_batch": el * 1e3 / iters}
    out["plain"]["path"] = "graph replay per scene" if plain_mode == "graph" else "swegl_render per scene"
    # the other single-render path beside it (eager renders pipeline one
    # scene's front end under the previous scene's raster)
    other = out.pop("plain_other")
    other["path"] = "swegl_render per scene" if plain_mode == "graph" else "graph replay per scene"
    out["plain_" + ("eager" if plain_mode == "graph" else "graph")] = other
    out["batched_over_plain_" + ("eager" if plain_mode == "graph" else "graph")] = (
        out["batched"]["frames_per_s"] / other["frames_per_s"])
    for name in ("plain", "batched"):
        if tunings.get(name):
            out[name]["tuning"] = {k: tunings[name][k] for k in ("tile", "bin_limit", "raster_groups")}
    out["batched_over_plain"] = out["batched"]["frames_per_s"] / out["plain"]["frames_per_s"]
    # algorithmic bytes of the whole set (shaded pixels read back from the depth planes)
    nbytes = 0
    zk = iter(keep[1::2])
    for s, vps in pairs:
        nbytes += job_bytes(s, vps, [shaded_pixels(next(zk)) for _ in vps])
    out["algorithmic_bytes_per_batch"] = nbytes
    for name in ("plain", "batched"):
        out[name]["roofline"] = roofline_of(nbytes, out[name]["ms_per_batch"] * 1e-3)
    for d in ds:
        d.close()
    return out


def bench_cube(ctx, A, check, L, seed, frames):
    """Workload 'cube': make_cube(1, 0) at seeded Euler angles, 256^2 texture,
    one sun, 640x480, plain path (one render call per frame, eager), and the
    same frames as one graph replay each."""
    import numpy as np
    from swegl_amd import scenes
    from swegl_amd.device import DeviceScene, Target
    from swegl_amd.host import SceneDescHolder

    scene, vps = scenes.cube_scene(seed)
    vp = vps[0]
    ds = DeviceScene(ctx, SceneDescHolder(scene))
    col = Target(ctx, A.TARGET_COLOR, vp.m_screen.w, vp.m_screen.h)
    z = Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)
    d = (A.ViewportDesc * 1)()
    d[0] = vp.desc(col, z)
    st = (A.FrameStats * 1)()
    check(L.swegl_render(ctx.h, ds.h, d, 1, st), "swegl_render")
    for _ in range(5):
        check(L.swegl_render(ctx.h, ds.h, d, 1, None), "swegl_render")
    ctx.sync()
    a = time.perf_counter()
    for _ in range(frames):
        check(L.swegl_render(ctx.h, ds.h, d, 1, None), "swegl_render")
    ctx.sync()
    el = time.perf_counter() - a
    # the same frames as one graph replay each (the captured launch sequence)
    for _ in range(3):
        check(L.swegl_render_graph(ctx.h, ds.h, d, 1, None), "swegl_render_graph")
    ctx.sync()
    a = time.perf_counter()
    for _ in range(frames):
        check(L.swegl_render_graph(ctx.h, ds.h, d, 1, None), "swegl_render_graph")
    ctx.sync()
    el_graph = time.perf_counter() - a
    lat = []
    for _ in range(100):
        t = time.perf_counter()
        check(L.swegl_render(ctx.h, ds.h, d, 1, None), "swegl_render")
        ctx.sync()
        lat.append((time.perf_counter() - t) * 1e3)
    ds.close()
    return {"workload": f"cube 12 triangles, 640x480, tex 256^2, one sun, seed {seed}",
            "frames_per_s": frames / el, "graph_frames_per_s": frames / el_graph,
            "gfragments_per_s": st[0].fragments * frames / el / 1e9,
            "fragments_per_frame": int(st[0].fragments),
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))}}, (scene, vps)


def bench_demo(ctx, A, check, L, seed, frames):
    """The reference demo's scene (src/test_1.cpp:50-148: torus and sphere of
    20,200 triangles, a stretched cube, translucent triangles) at 800x600
    with 3 transparency layers, as its frame loop draws it: frames per second
    on the device, and with each frame read back to host memory (colour and
    depth, the demo's SDL surface and z-buffer); the same without layers."""
    from swegl_amd import scenes
    from swegl_amd.device import DeviceScene, Target
    from swegl_amd.host import SceneDescHolder

    scene = scenes.demo_scene(seed)
    ds = DeviceScene(ctx, SceneDescHolder(scene))
    out = {"workload": f"demo scene {scene.n_triangles()} triangles, 800x600, seed {seed}"}
    for layers in (3, 0):
        vp = scenes.demo_viewport(layers=layers)
        col = Target(ctx, A.TARGET_COLOR, 800, 600)
        z = Target(ctx, A.TARGET_FLOAT, 800, 600)
        d = (A.ViewportDesc * 1)()
        d[0] = vp.desc(col, z)
        for _ in range(5):
            check(L.swegl_render(ctx.h, ds.h, d, 1, None), "swegl_render")
        ctx.sync()
        a = time.perf_counter()
        for _ in range(frames):
            check(L.swegl_render(ctx.h, ds.h, d, 1, None), "swegl_render")
        ctx.sync()
        dev = frames / (time.perf_counter() - a)
        a = time.perf_counter()
        for _ in range(frames // 4):
            check(L.swegl_render(ctx.h, ds.h, d, 1, None), "swegl_render")
            col.read()
            z.read()
        shown = (frames // 4) / (time.perf_counter() - a)
        out[f"layers_{layers}"] = {"frames_per_s": dev, "frames_per_s_read_back": shown}
    ds.close()
    return out, scene


def bench_batch(ctx, A, check, L, count, seed, iters, plain_mode="graph", tunings=None):
    """Workload 'batch': `count` height fields (T log-uniform in [1e3, 2e5],
    square framebuffers of 256..1024, DoF on half)."""
    from swegl_amd import scenes

    t0 = time.perf_counter()
    pairs = scenes.batch_scenes(seed, count=count)
    gen_s = time.perf_counter() - t0
    out = plain_vs_batched(ctx, A, check, L, pairs, iters, plain_mode, tunings)
    out["generate_s"] = round(gen_s, 1)
    return out, pairs


def bench_heavy(ctx, A, check, L, count, seed, iters, plain_mode="graph", tunings=None):
    """Workload 'heavy': `count` scenes of 500 stacked 1000-triangle
    clusters (500,000 triangles), 4096^2 texture, sun + 8 point lights,
    each seen by 4 viewports of 1920x1080 with the 9-tap DoF."""
    from swegl_amd import scenes

    t0 = time.perf_counter()
    pairs = scenes.heavy_scenes(seed, count=count)
    gen_s = time.perf_counter() - t0
    out = plain_vs_batched(ctx, A, check, L, pairs, iters, plain_mode, tunings)
    out["generate_s"] = round(gen_s, 1)
    out["workload"] = (f"{count} scenes x 500,000 triangles (500 clusters of 1000, stacks of 8-20), "
                       "4 x 1920x1080 viewports each, tex 4096^2, sun + 8 point lights, 9-tap DoF r_max 8")
    return out


def bench_served(ctx, A, check, L, pairs, big, n_req, seed, lanes, max_batch, fracs, plain_mode="graph",
                 tunings=None):
    """Workload 'served': n_req requests, 95% drawn from the batch scenes and
    5% of the 1M-triangle 1080p shape, each with a seeded camera pose,
    admitted against a 200 GB budget.  Reports the plain path (each request
    rendered alone, one graph replay each), the saturated served rate, and
    p50/p99 latency at Poisson arrival rates that are fractions of it.
    tunings: {"plain": entry, "batched": entry} as in plain_vs_batched, the
    batched entry for the server."""
    tunings = tunings or {}
    import numpy as np
    from swegl_amd.device import DeviceScene, Server, Target
    from swegl_amd.host import SceneDescHolder

    rng = np.random.default_rng(seed)
    allp = list(pairs) + [big]
    ds = [DeviceScene(ctx, SceneDescHolder(s)) for s, _ in allp]
    pools, keep = {}, []
    # Request i writes target set i mod the pool size of its viewport size.
    # The trace runner submits in arrival order and holds back a request
    # whose targets an earlier one still writes, so the pools bound how
    # many requests wait: four full batches per lane of small ones, 64
    # large ones (5 % of the stream), so that a saturated run keeps the
    # lanes' batches full.
    slots = 4 * lanes * max_batch + 8

    def target(vp, i):
        key = (vp.m_w, vp.m_h)
        pool = pools.setdefault(key, [])
        j = i % (slots if key != (big[1][0].m_w, big[1][0].m_h) else 64)
        while len(pool) <= j:
            pool.append((Target(ctx, A.TARGET_COLOR, vp.m_w, vp.m_h),
                         Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)))
        return pool[j]

    descs = (A.ViewportDesc * n_req)()
    jobs = (A.Job * n_req)()
    which = np.where(rng.random(n_req) < 0.05, len(pairs), rng.integers(0, len(pairs), n_req))
    base = {}
    for i in range(n_req):
        k = int(which[i])
        s, vps = allp[k]
        vp = vps[0]
        if k not in base:
            base[k] = vp.m_camera.m_viewmatrix.copy()
        vp.m_camera.m_viewmatrix[...] = base[k]
        vp.m_camera.rotate_y(float(rng.uniform(-0.3, 0.3)))
        col, z = target(vp, i)
        descs[i] = vp.desc(col, z)
        jobs[i] = A.Job(ds[k].h, C.pointer(descs[i]), 1)
    n_big = int((which == len(pairs)).sum())

    # plain: every request alone, in arrival order (graphs captured first);
    # the warm-up render of each scene's first request gives its shaded
    # pixels, which stand for every request of that scene in the texel term
    plain_fn = L.swegl_render_graph if plain_mode == "graph" else L.swegl_render
    fv = {}
    for k in range(len(allp)):
        i = int(np.argmax(which == k)) if (which == k).any() else None
        if i is not None:
            check(plain_fn(ctx.h, ds[k].h, C.pointer(descs[i]), 1, None), "warm")
            ctx.sync()
            fv[k] = shaded_pixels(target(allp[k][1][0], i)[1])
    ctx.sync()
    req_bytes = sum(job_bytes(allp[int(k)][0], allp[int(k)][1][:1], [fv[int(k)]]) for k in which)
    if tunings.get("plain"):
        ctx.apply_tuning(tunings["plain"])
        for k in fv:   # warm the graphs of the plain settings
            i = int(np.argmax(which == k))
            check(plain_fn(ctx.h, ds[k].h, C.pointer(descs[i]), 1, None), "warm")
        ctx.sync()
    a = time.perf_counter()
    for i in range(n_req):
        check(plain_fn(ctx.h, ds[int(which[i])].h, C.pointer(descs[i]), 1, None), "plain")
    ctx.sync()
    plain_fps = n_req / (time.perf_counter() - a)
    # the other single-render path beside it
    other_fn = L.swegl_render if plain_mode == "graph" else L.swegl_render_graph
    for k in fv:
        check(other_fn(ctx.h, ds[k].h, C.pointer(descs[int(np.argmax(which == k))]), 1, None), "warm")
    ctx.sync()
    a = time.perf_counter()
    for i in range(n_req):
        check(other_fn(ctx.h, ds[int(which[i])].h, C.pointer(descs[i]), 1, None), "plain")
    ctx.sync()
    other_fps = n_req / (time.perf_counter() - a)

    plain_s = n_req / plain_fps
    other_name = "eager" if plain_mode == "graph" else "graph"
    out = {"requests": n_req, "large_requests": n_big, "lanes": lanes, "max_batch": max_batch,
           "hbm_budget_bytes": 200 * 10**9, "algorithmic_bytes": req_bytes,
           "plain_alone": {"frames_per_s": plain_fps, "roofline": roofline_of(req_bytes, plain_s),
                           "path": "graph replay per request" if plain_mode == "graph" else "swegl_render per request"},
           "plain_alone_" + other_name: {
               "frames_per_s": other_fps,
               "path": "swegl_render per request" if plain_mode == "graph" else "graph replay per request"}}
    if tunings.get("batched"):
        ctx.apply_tuning(tunings["batched"])
    srv = Server(ctx, 200 * 10**9, max_batch=max_batch, lanes=lanes)
    # saturated: everything queued at t=0
    lat = srv.run_trace(jobs, np.zeros(n_req))
    st = srv.stats()
    cap_fps = n_req / lat.max()
    out["saturated"] = {"frames_per_s": cap_fps, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                        "p99_ms": float(np.percentile(lat, 99) * 1e3), "batches": st["batches"],
                        "roofline": roofline_of(req_bytes, float(lat.max()))}
    out["served_over_plain"] = cap_fps / plain_fps
    out["served_over_plain_" + other_name] = cap_fps / other_fps
    rates = []
    for f in fracs:
        rate = f * cap_fps
        arr = np.cumsum(rng.exponential(1.0 / rate, n_req))
        b0 = srv.stats()["batches"]
        lat = srv.run_trace(jobs, arr)
        el = float(arr[-1] + lat[-1])
        rates.append({"arrival_per_s": rate, "load": f, "frames_per_s": n_req / max(el, 1e-9),
                      "p50_ms": float(np.percentile(lat, 50) * 1e3),
                      "p99_ms": float(np.percentile(lat, 99) * 1e3),
                      "batches": srv.stats()["batches"] - b0})
    out["rates"] = rates
    out["peak_bytes_in_flight"] = srv.stats()["peak_bytes_in_flight"]
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tile", default=None, help="WxH raster tile, e.g. 32x32")
    ap.add_argument("--raster-groups", type=int, default=0, help="raster workgroups per CU (0 = tuned)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=256, help="scenes in the batched workload (0 = skip)")
    ap.add_argument("--batch-iters", type=int, default=10)
    ap.add_argument("--heavy", type=int, default=64, help="scenes in the heavy workload (0 = skip)")
    ap.add_argument("--demo", type=int, default=1, help="the demo scene leg (0 = skip)")
    ap.add_argument("--cube", type=int, default=2000, help="frames of the cube leg (0 = skip)")
    ap.add_argument("--heavy-iters", type=int, default=3)
    ap.add_argument("--requests", type=int, default=20000, help="served requests per rate (0 = skip)")
    ap.add_argument("--lanes", type=int, default=3)
    ap.add_argument("--max-batch", type=int, default=64)
    ap.add_argument("--plain", choices=("graph", "eager"), default="graph",
                    help="plain form of the batch and heavy legs: one graph replay or one swegl_render per scene")
    ap.add_argument("--reshade", type=int, default=1, help="FP64 re-shade error per workload (0 = skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also copy a target of this many MiB 3 times (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the render path is single-GPU in this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from swegl_amd import _abi as A
    from swegl_amd import scenes
    from swegl_amd.device import Context, DeviceScene, Target, check, lib
    from swegl_amd.host import SceneDescHolder

    scene, vps = scenes.heightfield_scene(args.seed)
    vp = vps[0]
    ctx = Context(0)
    from swegl_amd import autotune
    tuning = autotune.load()
    if "heightfield" in tuning:
        ctx.apply_tuning(tuning["heightfield"])
    if args.tile:
        tw, th = (int(x) for x in args.tile.lower().split("x"))
        ctx.set_tile(tw, th)
    if args.raster_groups:
        ctx.set_raster_groups(args.raster_groups)
    info = ctx.info()
    ds = DeviceScene(ctx, SceneDescHolder(scene))
    surf = Target(ctx, A.TARGET_COLOR, vp.m_screen.w, vp.m_screen.h)
    zt = Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)
    desc = (A.ViewportDesc * 1)()
    desc[0] = vp.desc(surf, zt)
    L = lib()

    # The plain path, one frame per render call (swegl_render: the frame's
    # kernels launched one by one); the same frame as one replayed graph
    # (swegl_render_graph) is timed beside it.
    def frame(stats=None):
        check(L.swegl_render(ctx.h, ds.h, desc, 1, stats), "swegl_render")

    def frame_graph(stats=None):
        check(L.swegl_render_graph(ctx.h, ds.h, desc, 1, stats), "swegl_render_graph")

    if args.calibrate_copy:
        side = int((args.calibrate_copy * 1024 * 1024 // 4) ** 0.5)
        ca, cb = Target(ctx, A.TARGET_UINT, side, side), Target(ctx, A.TARGET_UINT, side, side)
        for _ in range(3):
            check(L.swegl_target_copy(ctx.h, cb.h, ca.h), "swegl_target_copy")
        ctx.sync()
    st = (A.FrameStats * 1)()
    frame(st)
    stats = {k: getattr(st[0], k) for k, _ in A.FrameStats._fields_}
    for _ in range(args.warmup):
        frame()
    ctx.sync()

    # timed region: exactly K frames, synchronised on both sides
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    ctx.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    fps = 1e3 / ms
    for _ in range(args.warmup):
        frame_graph()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame_graph()
    ctx.sync()
    graph_fps = args.steps / (time.perf_counter() - t0)

    # per-frame latency (submit -> complete), frames one at a time
    lat = []
    for _ in range(max(20, min(args.steps, 200))):
        a = time.perf_counter()
        frame()
        ctx.sync()
        lat.append((time.perf_counter() - a) * 1e3)
    lat = np.array(lat)

    # per-kernel HIP-event timing on the library stream (launched one by one)
    ctx.set_profiling(True)
    for _ in range(10):
        frame()
    ctx.sync()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    dom = max(kt.items(), key=lambda kv: kv[1][0])
    f_vis = None
    # shaded pixels of this frame (for the texel term)
    z = zt.read()
    f_vis = int((z.view(np.uint32) != 0x7F7F7F7F).sum())
    frame_bytes, raster_bytes = algorithmic_bytes(scene, vp, f_vis)
    dom_bytes = raster_bytes if dom[0] == "raster" else None
    achieved = dom_bytes / (dom[1][0] * 1e-3) / 1e9 if dom_bytes else None
    # HBM traffic and VALU instruction count of the dominant kernel per launch,
    # from the committed rocprofv3 summary of this workload
    traffic, valu_instr = None, None
    tfile = os.path.join(ROOT, "profiles", "heightfield_1080p_kernels.json")
    if os.path.exists(tfile):
        try:
            with open(tfile) as f:
                tj = json.load(f)
            traffic = tj.get(dom[0])
            if dom[0] == "raster":
                valu_instr = tj.get("raster_valu_instr")
        except (OSError, ValueError):
            traffic = None
    valu = None
    if valu_instr and f_vis:
        t = dom[1][0] * 1e-3   # live average launch duration, seconds
        valu = {"kernel": dom[0], "valu_instr_per_launch": valu_instr,
                "achieved": valu_instr / t / 1e9, "peak": VALU_PEAK_G, "unit": "G wave64 VALU instr/s",
                "frac": valu_instr / t / 1e9 / VALU_PEAK_G,
                "lane_instr_per_shaded_pixel": valu_instr * 64 / f_vis,
                "shaded_pixels_per_s": f_vis / t / 1e9,
                "shaded_pixels_per_s_at_valu_peak": VALU_PEAK_G * f_vis / valu_instr}

    out = {
        "metric": METRIC,
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded height field, random texture)",
        "config": {"workload": "heightfield 999,698 triangles, 1920x1080, tex 2048^2, "
                               "sun + 4 point lights, 9-tap DoF r_max 8",
                   "path": "plain (one render call per frame)", "tile": "%dx%d" % ctx.get_tile(), "seed": args.seed,
                   "tuning": tuning.get("heightfield") and {k: tuning["heightfield"][k] for k in
                                                            ("tile", "bin_limit", "raster_groups")}},
        "graph_frames_per_s": graph_fps,
        "gfragments_per_s": stats["fragments"] * fps / 1e9,
        "fragments_per_frame": stats["fragments"],
        "shaded_pixels": f_vis,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "frame_algorithmic_bytes": frame_bytes,
        "frame_hbm_frac": frame_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "kernels_ms": {k: round(v[0], 5) for k, v in kt.items()},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": traffic, "algorithmic_bytes": dom_bytes},
        "valu_roofline": valu,
        "device": info,
    }
    cube = None
    if args.cube:
        out["cube"], cube = bench_cube(ctx, A, check, L, args.seed, args.cube)
    if args.demo:
        out["demo"], demo_scene = bench_demo(ctx, A, check, L, args.seed, 400)
    # plain legs (one render per scene or request) with the single-render
    # tuning, batched and served legs with the batch tuning
    tunings = {"plain": tuning.get("heightfield"), "batched": tuning.get("batch")}
    if args.batch:
        out["batch"], pairs = bench_batch(ctx, A, check, L, args.batch, args.seed + 1, args.batch_iters,
                                          args.plain, tunings)
        if args.requests:
            out["served"] = bench_served(ctx, A, check, L, pairs, (scene, vps), args.requests,
                                         args.seed + 2, args.lanes, args.max_batch, (0.3, 0.6, 0.9), args.plain,
                                         tunings)
    if args.heavy:
        out["heavy"] = bench_heavy(ctx, A, check, L, args.heavy, args.seed + 3, args.heavy_iters, args.plain,
                                   dict(tunings, batched=tuning.get("heavy", tuning.get("batch"))))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, vps, args.cpu_seconds, args.cpu_threads)
        if cube is not None:
            out["cpu_baseline_cube"] = cpu_baseline(cube[0], cube[1], 4.0, args.cpu_threads)
        if args.demo:
            from swegl_amd import scenes as _sc
            out["cpu_baseline_demo"] = cpu_baseline(demo_scene, [_sc.demo_viewport(layers=3)], 4.0,
                                                    args.cpu_threads)
    if args.reshade:
        e = fp64_reshade(ctx, A, check, L, ds, scene, vp)
        e["workload"] = "heightfield 999,698 triangles, 1920x1080 (post shader off)"
        per = {"heightfield": e}
        if cube is not None:
            per["cube"] = fp64_reshade_jobs(ctx, A, check, L, [cube], "cube 640x480")
        if args.batch:
            per["batch"] = fp64_reshade_jobs(ctx, A, check, L, pairs[:8], "first 8 scenes of the mixed batch")
            # served requests are mixed-batch scenes and the large shape at per-request poses
            rng = np.random.default_rng(args.seed + 4)
            reqs = []
            for s2, v2 in [pairs[8], pairs[9], pairs[10], (scene, vps)]:
                keep_view = v2[0].m_camera.m_viewmatrix.copy()
                v2[0].m_camera.rotate_y(float(rng.uniform(-0.3, 0.3)))
                reqs.append((s2, v2[:1], keep_view))
            per["served"] = fp64_reshade_jobs(ctx, A, check, L, [(a, b) for a, b, _ in reqs],
                                              "3 mixed-batch requests and 1 large request at per-request poses")
            for _, v2, kv in reqs:
                v2[0].m_camera.m_viewmatrix[...] = kv
        if args.heavy:
            hv = scenes.heavy_scenes(args.seed + 3, count=1)
            per["heavy"] = fp64_reshade_jobs(ctx, A, check, L, hv, "one heavy scene, its 4 viewports of 1920x1080")
        worst = max(v["max"] for v in per.values())
        out["fp64_reshade"] = dict(per["heightfield"], per_workload=per, max_over_workloads=worst)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
