"""The vertex stage against the oracle, field by field: the scene fields the
reference's render writes (v_world, normal_world, v_viewport and the yes
marks, swegl/data/model.hpp:20-29; vertex_shaders.hpp:16-88 and the marking
loops of renderer.cpp:83-185), read back from the device with
swegl_debug_vertex_state and compared bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A
from swegl_amd import host, scenes
from swegl_amd.device import Context, DeviceScene, Target, check, lib
from swegl_amd.host import SceneDescHolder

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def device_state(ctx, scene, vp):
    ds = DeviceScene(ctx, SceneDescHolder(scene))
    col = Target(ctx, A.TARGET_COLOR, vp.m_screen.w, vp.m_screen.h)
    z = Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)
    d = (A.ViewportDesc * 1)()
    d[0] = vp.desc(col, z)
    n = scene.n_vertices()
    out = {k: np.zeros((n, 3), np.float32) for k in ("v_world", "normal_world", "v_viewport")}
    yes = np.zeros(n, np.uint8)
    check(lib().swegl_debug_vertex_state(ctx.h, ds.h, d, A.fptr(out["v_world"]), A.fptr(out["normal_world"]),
                                         A.fptr(out["v_viewport"]), yes.ctypes.data_as(C.POINTER(C.c_uint8))),
          "swegl_debug_vertex_state")
    out["yes"] = yes
    ds.close()
    return out


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def cases():
    s, vps = scenes.cube_scene(3)
    yield "cube", s, vps[0]
    s, vps = scenes.heightfield_scene(5, n=60, width=320, height=200, tex=64, dof=False)
    yield "height field", s, vps[0]
    s = scenes.demo_scene(tore_precision=24, sphere_precision=24)
    yield "demo", s, scenes.demo_viewport(400, 300, 0)
    # a triangle crossing the near plane (cut by fill_triangle) and one behind the camera
    from test_oracle_scene import flat_scene, view
    tri = [[-0.4, -0.3, 0.0], [0.4, -0.3, 0.0], [0.0, -0.3, 2.9]]
    behind = [[-0.4, 0.3, 3.5], [0.4, 0.3, 3.5], [0.0, 0.5, 3.6]]
    yield "near plane", flat_scene([tri, behind], double_sided=True), view(80, 60)


@pytest.mark.parametrize("name", ["cube", "height field", "demo", "near plane"])
def test_vertex_state_matches_the_oracle(ctx, name):
    for n, scene, vp in cases():
        if n != name:
            continue
        want = oracle.vertex_state(scene, vp)
        got = device_state(ctx, scene, vp)
        assert want["yes"].sum() > 0
        assert np.array_equal(got["yes"], want["yes"]), np.argwhere(got["yes"] != want["yes"])[:5]
        for k in ("v_world", "normal_world", "v_viewport"):
            bad = np.argwhere(got[k].view(np.uint32) != want[k].view(np.uint32))
            assert bad.size == 0, (k, bad[:5].tolist(), got[k][bad[0][0]], want[k][bad[0][0]])
