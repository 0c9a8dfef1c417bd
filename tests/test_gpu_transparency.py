"""GPU parity of transparency layers (renderer.cpp:486-550, viewport.cpp:43-86,
colors.cpp:39-57, post_shaders.hpp:22-47) against the oracle: layered quads
of known alpha, offset viewports (whose flatten reads the surface origin),
and textured scenes whose random texel alpha makes most fragments
transparent, with 1 to 4 layers."""
import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A
from swegl_amd import host, scenes
from swegl_amd.device import Context
from swegl_amd.host import pixel_colors

from _helpers import compare, gpu_render
from test_oracle_transparency import BG, quads_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def check(ctx, scene, vps):
    for vp in vps:
        vp.m_screen.pixels[:] = BG
    want = oracle.render(scene, vps)
    got = gpu_render(ctx, scene, vps)
    for g, w in zip(got, want):
        compare(g, w, colour_tol=1)
    return got, want


@pytest.mark.parametrize("layers", [1, 2, 3, 4])
def test_layered_quads(ctx, layers):
    quads = [(0.5, 0.5, pixel_colors(0, 0, 200, 100)), (-0.5, 0.7, pixel_colors(200, 0, 0, 90)),
             (0.0, 0.3, pixel_colors(0, 200, 0, 255)), (0.2, 0.9, pixel_colors(90, 90, 90, 60)),
             (-0.2, 0.4, pixel_colors(10, 200, 30, 140))]
    scene = quads_scene(quads)
    surf = host.Surface(64, 48)
    vp = host.viewport_t(0, 0, 64, 48, surf, host.pixel_shader_t(), layers)
    vp.m_camera.translate(0, 0, -3)
    vp.set_post_shader(host.post_shader_t())
    got, want = check(ctx, scene, [vp])
    assert (want[0].pixels != BG).sum() > 500


def test_offset_viewport(ctx):
    scene = quads_scene([(0.0, 2.0, pixel_colors(200, 100, 50, 128)), (0.3, 0.5, pixel_colors(9, 9, 200, 255))])
    surf = host.Surface(48, 30)
    vp = host.viewport_t(20, 10, 16, 12, surf, host.pixel_shader_t(), 2)
    vp.m_camera.translate(0, 0, -3)
    vp.set_post_shader(host.post_shader_t())
    check(ctx, scene, [vp])


@pytest.mark.parametrize("seed,layers", [(0, 3), (4, 1), (6, 4)])
def test_textured_cube_with_layers(ctx, seed, layers):
    """Random texel alpha: almost every fragment is transparent."""
    scene, vps = scenes.cube_scene(seed)
    vps[0].transparency_layer_count = layers
    vps[0].m_got_transparency = True
    check(ctx, scene, vps)


def test_height_field_with_layers_and_other_shaders(ctx):
    scene, vps = scenes.heightfield_scene(3, n=40, width=160, height=120, tex=64, dof=False)
    vps[0].transparency_layer_count = 2
    # the light-only shaders take their alpha from the light word, so they
    # shade every fragment in order; the others shade the survivors at the end
    for sh in (host.pixel_shader_light_and_texture(), host.pixel_shader_texture(),
               host.pixel_shader_light_and_texture(host.pixel_shader_lights_flat, host.pixel_shader_texture),
               host.pixel_shader_lights_phong(), host.pixel_shader_lights_flat(), host.pixel_shader_t()):
        vps[0].m_pixel_shader = sh
        check(ctx, scene, vps)


def test_layers_next_to_plain_viewports(ctx):
    """A viewport with layers between two plain ones on the same surface keeps
    the reference's order of effects."""
    scene, _ = scenes.cube_scene(2)
    surf = host.Surface(200, 120)
    vps = []
    for k, (x, layers) in enumerate([(0, 0), (60, 2), (120, 0)]):
        vp = host.viewport_t(x, 20, 80, 80, surf, host.pixel_shader_light_and_texture(), layers)
        vp.m_camera.translate(0, 0, -3)
        scenes.set_fov(vp.m_camera, np.float32(1.0))
        vp.set_post_shader(host.post_shader_t())
        vps.append(vp)
    check(ctx, scene, vps)


def crowded_scene(n, seed):
    """n small triangles of random translucent materials over a few pixels
    near the centre, one primitive each (draw order = index order)."""
    rng = np.random.default_rng(seed)
    s = host.scene_t()
    s.materials = [host.material_t(pixel_colors(*rng.integers(0, 256, 3), int(a)))
                   for a in rng.integers(20, 256, 16)]
    c = rng.uniform(-0.1, 0.1, (n, 1, 3))
    tri = (c + rng.uniform(-0.3, 0.3, (n, 3, 3))).astype(np.float32)
    tri[..., 2] = rng.uniform(-0.5, 0.5, (n, 1))
    mids = rng.integers(0, 16, n)
    prims = [host.primitive_t(tri[k], np.tile([0, 0, 1], (3, 1)), np.zeros((3, 2)), [0, 1, 2],
                              A.TRIANGLES, int(mids[k])) for k in range(n)]
    s.nodes = [host.node_t(prims)]
    s.root_nodes = [0]
    return s


@pytest.mark.parametrize("n", [300, 6000])
def test_crowded_tiles(ctx, n):
    """6000 triangles over the centre tiles exceed the sorted list in LDS
    (2048 entries) and take the walk over all triangles; 300 take the sorted
    bin list."""
    s = crowded_scene(n, n)
    surf = host.Surface(64, 48)
    vp = host.viewport_t(0, 0, 64, 48, surf, host.pixel_shader_t(), 3)
    vp.m_camera.translate(0, 0, -3)
    vp.set_post_shader(host.post_shader_t())
    got, want = check(ctx, s, [vp])
    assert (want[0].pixels != 0).sum() > 40


def test_big_triangles_merge_into_the_tile_lists(ctx):
    """With a bin limit of 1 tile, every triangle over two tiles or more goes
    to the job's big list, which the layer raster merges into each tile's
    sorted list."""
    scene, vps = scenes.cube_scene(9)
    vps[0].transparency_layer_count = 2
    ctx.set_bin_limit(1)
    try:
        check(ctx, scene, vps)
    finally:
        ctx.set_bin_limit(16)


def test_layered_jobs_in_batches_and_graph_replays(ctx):
    """Layered viewports batched with plain ones (one graph, replayed three
    times with eager renders between) equal each job rendered alone."""
    from _helpers import Batch, same_frame
    pairs = []
    for k in range(4):
        s, vps = scenes.cube_scene(20 + k, 160, 120)
        vps[0].transparency_layer_count = k % 3
        pairs.append((s, vps))
    d = scenes.demo_scene(1, 16, 16)
    pairs.append((d, [scenes.demo_viewport(200, 150, 3)]))
    alone = [gpu_render(ctx, s, vps)[0] for s, vps in pairs]
    b = Batch(ctx, pairs)
    for rep in range(3):
        b.render(graph=True)
        got = b.frames()
        for k in range(len(pairs)):
            assert same_frame(got[k][0], alone[k]), (rep, k)
        gpu_render(ctx, *pairs[0])     # eager work between the replays
    want = oracle.render(*pairs[-1])[0]
    compare(b.frames()[-1][0], want, colour_tol=1)


def test_layered_graphs_keyed_by_pixel_count(ctx):
    """Layered viewports of 17x17 and 24x24 have the same 3x3 grid of 8x8
    layer tiles but need 2 and 3 blocks of the flatten and copy passes: each
    takes a graph of its own, and every replay equals the eager render."""
    from _helpers import Batch, same_frame
    s, _ = scenes.cube_scene(5, 64, 48)
    rng = np.random.default_rng(5)
    pairs = []
    for w in (17, 24):
        surf = host.Surface(64, 48)
        surf.pixels[:] = rng.integers(0, 1 << 32, (48, 64), dtype=np.uint64).astype(np.uint32)
        vp = host.viewport_t(10, 8, w, w, surf, host.pixel_shader_light_and_texture(), 2)
        vp.m_camera.translate(0, 0, -3)
        vp.set_post_shader(host.post_shader_t())
        pairs.append((s, [vp]))
    alone = [gpu_render(ctx, sc, v)[0] for sc, v in pairs]
    batches = [Batch(ctx, [p]) for p in pairs]
    for rep in range(2):
        for k, b in enumerate(batches):
            b.targets[0][0][0].write(pairs[k][1][0].m_screen.pixels)
            b.render(graph=True)
            assert same_frame(b.frames()[0][0], alone[k]), (rep, k)
    compare(alone[1], oracle.render(*pairs[1])[0], colour_tol=1)
