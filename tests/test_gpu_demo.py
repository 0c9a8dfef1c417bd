"""GPU parity of the reference demo's scene and frame loop (src/test_1.cpp:
50-148, 324-387): procedural meshes (make_tore, make_sphere, make_tri,
make_cube), translucent triangles over 3 transparency layers, 800x600, and
the C++ program tests/cpp/demo that runs the loop through the C++ API with
the frame read back into a host surface each time."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from swegl_amd import host, scenes
from swegl_amd.device import Context

from _helpers import compare, gpu_render

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("layers", [3, 0])
def test_demo_scene(ctx, layers):
    s = scenes.demo_scene()
    for turn in (0.0, 0.4, -0.7):
        vp = scenes.demo_viewport(layers=layers)
        vp.m_camera.rotate_y(turn)
        want = oracle.render(s, [vp])[0]
        got = gpu_render(ctx, s, [vp])[0]
        compare(got, want, colour_tol=1)
        assert (want.tid != 0xFFFFFFFF).sum() > 30000


def test_cpp_demo_loop(tmp_path):
    exe = os.path.join(HERE, "cpp", "demo")
    assert os.path.exists(exe), "build() compiles tests/cpp/demo"
    s = scenes.demo_scene(seed=4)
    texs = []
    for k, im in enumerate(s.images):
        p = tmp_path / f"t{k}.raw"
        p.write_bytes(np.ascontiguousarray(im.bitmap).tobytes())
        texs.append(str(p))
    w, h, frames = 400, 300, 4
    out_p, out_z = tmp_path / "px.raw", tmp_path / "z.raw"
    r = subprocess.run([exe] + texs + [str(frames), str(w), str(h), str(out_p), str(out_z)],
                       check=True, timeout=120, capture_output=True, text=True)
    print(r.stdout.strip())
    vp = scenes.demo_viewport(w, h, 3)
    for _ in range(frames - 1):
        vp.m_camera.rotate_y(0.001)
    want = oracle.render(s, [vp])[0]

    class F:
        pass
    got = F()
    got.pixels = np.frombuffer(out_p.read_bytes(), np.uint32).reshape(h, w)
    got.zbuffer = np.frombuffer(out_z.read_bytes(), np.float32).reshape(h, w)
    got.tid = None
    compare(got, want, colour_tol=1, check_tid=False)
