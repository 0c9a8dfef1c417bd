"""GPU parity of scenes read from glTF/GLB files (src/data/gltf.cpp) with node
hierarchies (vertex_shaders.hpp:16-28) and animation (model.hpp:146-177): the
scene is uploaded once, then posed with scene_t::animate and
swegl_scene_update_pose, and every frame is compared with the oracle."""
import numpy as np
import pytest

import oracle
from gltf_writer import UINT, USHORT, Gltf, png
from swegl_amd import gltf, host, scenes
from swegl_amd.device import Context, DeviceScene
from swegl_amd.host import SceneDescHolder

from _helpers import compare, gpu_render

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def cube_prims(g, material):
    node = scenes.make_cube(1.0, material)
    prims = []
    for p in node.primitives:
        prims.append(g.primitive(p.positions, p.normals, p.tex_coords[:, ::-1], idx=p.indices,
                                 mode=p.mode, material=material))
    return prims


def grid_prim(g, n, rng, material):
    xs = np.linspace(-1, 1, n, dtype=np.float32)
    X, Z = np.meshgrid(xs, xs)
    Y = (0.15 * np.sin(3 * X) * np.cos(2 * Z)).astype(np.float32)
    pos = np.stack([X, Y, Z], -1).reshape(-1, 3)
    nor = np.tile(np.array([0, 1, 0], np.float32), (n * n, 1))
    uv = np.stack([(X + 1) / 2, (Z + 1) / 2], -1).reshape(-1, 2).astype(np.float32)
    a = np.arange(n - 1)[:, None] * n + np.arange(n - 1)[None, :]
    idx = np.stack([a, a + n, a + 1, a + 1, a + n, a + n + 1], -1).reshape(-1)
    return g.primitive(pos, nor, uv, idx=idx, idx_type=USHORT if n * n < 65536 else UINT,
                       material=material)


def animated_glb(path, seed=0, grid=60):
    rng = np.random.default_rng(seed)
    g = Gltf()
    for k in range(2):
        px = rng.integers(0, 256, (64, 48, 4))
        px[..., 3] = 255
        g.png_image(png(px, 6, 8, interlace=bool(k)))
        g.add("textures", {"source": k})
        g.add("materials", {"pbrMetallicRoughness": {"baseColorTexture": {"index": k}},
                            "doubleSided": bool(k)})
    g.add("meshes", {"primitives": cube_prims(g, 0)})
    g.add("meshes", {"primitives": [grid_prim(g, grid, rng, 1)]})
    g.add("meshes", {"primitives": cube_prims(g, 1)})   # a mesh goes to one node only
    g.add("nodes", {"children": [1, 3], "rotation": [0.9, 0.1, 0.3, 0.2]})           # 0: root
    g.add("nodes", {"mesh": 0, "translation": [0.8, 0, 0], "scale": [0.6, 0.6, 0.6],
                    "children": [2]})                                                   # 1
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = [0, 1.2, 0.3]
    m[:3, :3] *= 0.5
    g.add("nodes", {"matrix": [float(v) for v in m.T.reshape(-1)], "mesh": 2})       # 2
    g.add("nodes", {"mesh": 1, "translation": [0, -0.7, 0]})                           # 3
    g.animation([
        (0, "rotation", [0, 1, 2, 3], [[1, 0, 0, 0], [0.7, 0.7, 0, 0], [0.5, 0.5, 0.5, 0.5],
                                       [1, 0, 0, 0]], None),
        (1, "translation", [0, 1.5, 3], [[0.8, 0, 0], [0, 0.5, 0.4], [0.8, 0, 0]], None),
        (2, "scale", [0, 3], [[1, 1, 1], [0.3, 2, 1]], "LINEAR"),
        (3, "translation", [0.5, 2.5], [[0, -0.7, 0], [0.2, -0.9, 0.1]], None),
    ])
    g.write_glb(str(path))


def lit(scene, rng):
    scene.ambient_light_intensity = 0.15
    scene.sun_direction = scenes.seeded_sun(rng)
    scene.sun_intensity = 0.9
    scene.point_source_lights = [host.point_source_light(rng.uniform(-2, 2, 3).astype(np.float32),
                                                         1.5)]


def view(shader, w=320, h=240):
    surf = host.Surface(w, h)
    vp = host.viewport_t(0, 0, w, h, surf, shader, 0)
    vp.m_camera.translate(0, 0.3, -4)
    scenes.set_fov(vp.m_camera, np.float32(w / h))
    vp.set_post_shader(host.post_shader_t())
    return vp


def test_animated_hierarchy_resident_on_the_device(ctx, tmp_path):
    path = tmp_path / "anim.glb"
    animated_glb(path)
    scene = gltf.load_scene(str(path))
    lit(scene, np.random.default_rng(1))
    assert scene.n_triangles() == 2 * 12 + 2 * 59 * 59
    ds = DeviceScene(ctx, SceneDescHolder(scene))
    try:
        drawn = 0
        for t in (0.0, 0.37, 1.25, 2.9, 4.6):
            scene.animate(t)
            ds.update_pose(SceneDescHolder(scene))
            for sh in (host.pixel_shader_light_and_texture(), host.pixel_shader_texture()):
                vp = view(sh)
                want = oracle.render(scene, [vp])[0]
                got = gpu_render(ctx, scene, [vp], dscene=ds)[0]
                compare(got, want, colour_tol=1)
                drawn += int((want.tid != 0xFFFFFFFF).sum())
        assert drawn > 100000
    finally:
        ds.close()


def test_loaded_scene_in_a_batch_with_generated_ones(ctx, tmp_path):
    """A loaded scene is an ordinary scene: batched with generated scenes it
    renders as it does alone."""
    from _helpers import Batch
    path = tmp_path / "anim.glb"
    animated_glb(path, seed=3, grid=40)
    scene = gltf.load_scene(str(path))
    lit(scene, np.random.default_rng(2))
    scene.animate(1.7)
    pairs = [(scene, [view(host.pixel_shader_light_and_texture())])]
    pairs += [scenes.cube_scene(s, 160, 120) for s in range(3)]
    b = Batch(ctx, pairs)
    b.render()
    for (sc, vps), fr in zip(pairs, b.frames()):
        want = oracle.render(sc, vps)
        for g, w in zip(fr, want):
            compare(g, w, colour_tol=1)


def test_cpp_program_loads_animates_and_renders(tmp_path):
    """tests/cpp/render_asset: swegl::load_scene + scene_t::animate +
    swegl::render through the C++ API, against the oracle on the same file."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "render_asset")
    assert os.path.exists(exe), "build() compiles tests/cpp/render_asset"
    path = tmp_path / "anim.glb"
    animated_glb(path, seed=5, grid=50)
    w, h = 256, 192
    for t in (0.6, 2.2):
        out_p, out_z = tmp_path / "px.raw", tmp_path / "z.raw"
        subprocess.run([exe, str(path), repr(t), str(w), str(h), str(out_p), str(out_z)], check=True,
                       timeout=120)
        scene = gltf.load_scene(str(path))
        scene.ambient_light_intensity = float(np.float32(0.15))
        scene.sun_direction = scenes.normal_t(0.3, -0.8, -0.4)
        scene.sun_intensity = float(np.float32(0.9))
        scene.point_source_lights = [host.point_source_light(np.array([0.5, 1.0, -1.5], np.float32), 1.5)]
        scene.animate(t)
        vp = host.viewport_t(0, 0, w, h, host.Surface(w, h), host.pixel_shader_light_and_texture(), 0)
        vp.m_camera.translate(0, 0.3, -4)
        vp.set_post_shader(host.post_shader_t())
        want = oracle.render(scene, [vp])[0]
        assert (want.tid != 0xFFFFFFFF).sum() > 2000

        class F:
            pass
        got = F()
        got.pixels = np.frombuffer(out_p.read_bytes(), np.uint32).reshape(h, w)
        got.zbuffer = np.frombuffer(out_z.read_bytes(), np.float32).reshape(h, w)
        got.tid = None
        compare(got, want, colour_tol=1, check_tid=False)
