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
