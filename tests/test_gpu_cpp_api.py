"""The C++ API (include/swegl/*.hpp) is a drop-in for the reference's
render call: a program written like the reference demo renders through
libswegl_hip.so and matches the oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
from swegl_amd import scenes
from swegl_amd.host import (IDENTITY, Surface, material_t, pixel_colors,
                            pixel_shader_light_and_texture, post_shader_t, rotate_x,
                            rotate_y, rotate_z, scene_t, texture_t, viewport_t)

from _helpers import compare

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "render_cube")


@pytest.mark.parametrize("seed", [0, 5])
def test_cpp_program_matches_oracle(tmp_path, seed):
    assert os.path.exists(EXE), "build() compiles tests/cpp/render_cube"
    rng = np.random.default_rng(seed)
    tex = rng.integers(0, 1 << 32, size=(256, 256), dtype=np.uint64).astype(np.uint32)
    rot = [float(np.float32(x)) for x in rng.uniform(0, 2 * math.pi, 3)]
    sun = scenes.normal_t(0.3, -0.8, -0.4)
    w, h = 320, 240
    (tmp_path / "tex.raw").write_bytes(tex.tobytes())
    out_p, out_z = tmp_path / "px.raw", tmp_path / "z.raw"
    args = [EXE, str(tmp_path / "tex.raw")] + [repr(r) for r in rot] + \
           [repr(float(v)) for v in (0.3, -0.8, -0.4)] + [str(w), str(h), str(out_p), str(out_z)]
    subprocess.run(args, check=True, timeout=120)
    got_px = np.frombuffer(out_p.read_bytes(), np.uint32).reshape(h, w)
    got_z = np.frombuffer(out_z.read_bytes(), np.float32).reshape(h, w)

    s = scene_t()
    s.images = [texture_t(tex)]
    s.materials = [material_t(pixel_colors(255, 255, 255, 255), 1, 1, 0, False)]
    cube = scenes.make_cube(1.0, 0)
    R = rotate_z(rotate_y(rotate_x(IDENTITY.copy(), rot[0]), rot[1]), rot[2])
    cube.rotation = R
    s.nodes, s.root_nodes = [cube], [0]
    s.ambient_light_intensity = float(np.float32(0.1))
    s.sun_direction = sun
    s.sun_intensity = 1.0
    vp = viewport_t(0, 0, w, h, Surface(w, h), pixel_shader_light_and_texture(), 0)
    vp.m_camera.translate(0, 0, -3)
    vp.set_post_shader(post_shader_t())
    want = oracle.render(s, [vp])[0]

    class F:
        pass
    got = F()
    got.pixels, got.zbuffer, got.tid = got_px, got_z, None
    compare(got, want, colour_tol=1, check_tid=False)


def test_cpp_scenes_at_a_reused_address_render_their_own_meshes():
    """scene_t objects built one after another in one stack slot, each with
    another quad, then one scene's index array replaced and its vertices
    edited in place with a generation bump: every render's coverage is that
    of its own geometry, as the oracle renders it."""
    exe = os.path.join(HERE, "cpp", "scene_reuse")
    assert os.path.exists(exe), "build() compiles tests/cpp/scene_reuse"
    w, h = 160, 120
    out = subprocess.run([exe, str(w), str(h)], check=True, timeout=120, capture_output=True, text=True)
    got = [int(x) for x in out.stdout.split()]

    from swegl_amd.host import node_t, pixel_shader_t, primitive_t
    from swegl_amd import _abi as A

    def expect(half, indices, scale=1.0):
        xs, ys = [-half, half, -half, half], [-half, -half, half, half]
        pos = np.array([[x * scale, y * scale, 0] for x, y in zip(xs, ys)], np.float32)
        prim = primitive_t(pos, np.tile([0, 0, 1], (4, 1)), np.full((4, 2), 0.5), indices, A.TRIANGLES, 0)
        s = scene_t()
        s.materials = [material_t(pixel_colors(200, 100, 50, 255), 1, 1, -1, False)]
        s.nodes, s.root_nodes = [node_t([prim])], [0]
        s.ambient_light_intensity = 1.0
        vp = viewport_t(0, 0, w, h, Surface(w, h), pixel_shader_t(), 0)
        vp.m_camera.translate(0, 0, -3)
        vp.set_post_shader(post_shader_t())
        fr = oracle.render(s, [vp])[0]
        return int((fr.zbuffer.view(np.uint32) != 0x7F7F7F7F).sum())

    quad = [0, 1, 2, 2, 1, 3]
    want = [expect(0.25, quad), expect(0.5, quad), expect(1.0, quad), expect(1.0, [0, 1, 2]),
            expect(1.0, [0, 1, 2], scale=0.5)]
    assert len(set(want)) == 5
    assert got == want
