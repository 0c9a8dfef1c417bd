"""Autotuner for the raster knobs: tile size, bin limit (tiles per binned
triangle) and persistent raster workgroups per CU.

For each workload class it renders the workload with every candidate
setting (warm-up, then timed repetitions, median), keeps the fastest and
writes swegl_amd/tuning.json:
    {"heightfield": {"tile": [32, 32], "bin_limit": 16, "raster_groups": 4,
                     "ms": ..., "candidates": [...]},
     "batch": {...}, "device": {...}}
Run on the GPU:  python -m swegl_amd.autotune [--quick]
"""
import argparse
import itertools
import json
import os
import time

import numpy as np

from . import _abi as A
from . import scenes
from .device import Context, DeviceScene, Target, check, lib
from .host import SceneDescHolder

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_PATH = os.path.join(HERE, "tuning.json")

TILES = [(16, 16), (32, 16), (32, 32), (64, 32), (64, 64)]
BIN_LIMITS = [4, 16, 64]
GROUPS = [2, 4, 8]


def load():
    """The tuning file as a dict, or {} when there is none."""
    try:
        with open(TUNING_PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _jobs(ctx, pairs):
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
    jobs = (A.Job * len(pairs))()
    for k, (s, vps) in enumerate(pairs):
        jobs[k] = A.Job(ds[k].h, descs[k], len(vps))
    return jobs, (ds, keep, descs)


def _time(ctx, fn, reps):
    fn()
    fn()
    ctx.sync()
    ts = []
    for _ in range(reps):
        a = time.perf_counter()
        fn()
        ctx.sync()
        ts.append((time.perf_counter() - a) * 1e3)
    return float(np.median(ts))


def tune(ctx, name, pairs, reps, space):
    jobs, keep = _jobs(ctx, pairs)
    L = lib()
    n = len(pairs)

    def run():
        check(L.swegl_render_batch_graph(ctx.h, jobs, n, None), "render_batch_graph")

    results = []
    ctx.set_tile_auto(False)   # measure each tile size as given
    for tile, limit, groups in space:
        ctx.set_tile(*tile)
        ctx.set_bin_limit(limit)
        ctx.set_raster_groups(groups)
        ms = _time(ctx, run, reps)
        results.append({"tile": list(tile), "bin_limit": limit, "raster_groups": groups, "ms": ms})
        print(f"{name}: tile {tile} limit {limit} groups {groups}: {ms:.3f} ms", flush=True)
    ctx.set_tile_auto(True)
    best = min(results, key=lambda r: r["ms"])
    out = dict(best)
    out["candidates"] = results
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="smaller search space")
    ap.add_argument("--reps", type=int, default=7)
    args = ap.parse_args()
    ctx = Context(0)
    space = list(itertools.product(TILES, BIN_LIMITS, GROUPS))
    if args.quick:
        space = list(itertools.product(TILES, [16], [2, 4, 8])) + [((32, 32), 4, 4), ((32, 32), 64, 4)]
    out = {"device": ctx.info()}
    out["heightfield"] = tune(ctx, "heightfield", [scenes.heightfield_scene(0)], args.reps, space)
    out["batch"] = tune(ctx, "batch", scenes.batch_scenes(1, count=64), max(3, args.reps // 2), space)
    with open(TUNING_PATH, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", TUNING_PATH)


if __name__ == "__main__":
    main()
