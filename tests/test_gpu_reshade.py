"""Colour error of the GPU's FP32 shading against the oracle's FP64
re-shade of the same winners (pixel_shaders.cpp:88-384 in double), per
workload.  Depth and triangle ids must match the oracle bit for bit first,
so both shades see the same triangle at every pixel.  The post shader is
off: the comparison is of the shaded colour itself."""
import json

import numpy as np
import pytest

import oracle
from swegl_amd import host, scenes
from swegl_amd.device import Context

from _helpers import compare, gpu_render

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _error(ctx, name, scene, vps):
    for vp in vps:
        vp.set_post_shader(host.post_shader_t())
    want = oracle.render(scene, vps, reshade64=True)
    got = gpu_render(ctx, scene, vps)
    agg = {"max": 0, "sum": 0.0, "pixels": 0}
    for g, w, vp in zip(got, want, vps):
        compare(g, w)
        rect = g.pixels.reshape(vp.m_screen.h, vp.m_screen.w)[vp.m_y:vp.m_y + vp.m_h, vp.m_x:vp.m_x + vp.m_w]
        e = oracle.reshade_error(rect, g.tid, w.color64)
        agg["max"] = max(agg["max"], e["max"])
        agg["sum"] += e["mean"] * e["pixels"]
        agg["pixels"] += e["pixels"]
    res = {"workload": name, "max_lsb": agg["max"], "mean_lsb": agg["sum"] / max(1, agg["pixels"]),
           "pixels": agg["pixels"]}
    print("fp64-reshade", json.dumps(res))
    return res


def test_reshade_error_cube(ctx):
    for seed in range(4):
        sc, vps = scenes.cube_scene(seed=seed)
        r = _error(ctx, f"cube 640x480 seed {seed}", sc, vps)
        assert r["max_lsb"] <= 1 and r["mean_lsb"] < 0.005


def test_reshade_error_heightfield_1080p(ctx):
    sc, vps = scenes.heightfield_scene(0)
    r = _error(ctx, "heightfield 999,698 triangles 1920x1080", sc, vps)
    assert r["pixels"] > 1_000_000
    assert r["max_lsb"] <= 2 and r["mean_lsb"] < 0.02


def test_reshade_error_batch_sample(ctx):
    pairs = scenes.batch_scenes(1, count=8)
    for k, (sc, vps) in enumerate(pairs):
        r = _error(ctx, f"batch scene {k}", sc, vps)
        assert r["max_lsb"] <= 2 and r["mean_lsb"] < 0.02
