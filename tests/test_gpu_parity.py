"""GPU parity: the HIP render path through the C ABI against the CPU oracle on
the same seeded inputs.  Coverage, triangle id and depth are bit-exact;
colour within 1 LSB per channel (depth of field within 2 LSB)."""
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


@pytest.mark.parametrize("seed", range(8))
def test_cube_640x480(ctx, seed):
    scene, vps = scenes.cube_scene(seed)
    want = oracle.render(scene, vps)[0]
    got = gpu_render(ctx, scene, vps)[0]
    mx, nd = compare(got, want)
    assert got.stats["fragments"] == want.stats["fragments"]
    assert got.stats["drawn_subtriangles"] == want.stats["drawn_subtriangles"]
    assert got.stats["nan_fragments"] == 0
    print(f"cube seed {seed}: max colour diff {mx} at {nd} px")


@pytest.mark.parametrize("seed,n,w,h,dof", [(1, 64, 320, 240, False), (2, 128, 256, 256, True),
                                            (3, 200, 640, 360, True)])
def test_heightfield_small(ctx, seed, n, w, h, dof):
    scene, vps = scenes.heightfield_scene(seed, n=n, width=w, height=h, tex=256, dof=dof)
    want = oracle.render(scene, vps)[0]
    got = gpu_render(ctx, scene, vps)[0]
    compare(got, want, colour_tol=2 if dof else 1)
    assert got.stats["fragments"] == want.stats["fragments"]


@pytest.mark.parametrize("taps,rmax,fdepth", [("box", 5, 5.0), ("box", 16, 3.0), ("tap9", 16, 1.5),
                                              ("tap9", 1, 5.0)])
def test_dof_variants(ctx, taps, rmax, fdepth):
    """Box and 9-tap DoF, radii 1..16 (16 = the largest halo), focal depths
    with fdepth - 1 below and above 1 (both division paths)."""
    scene, vps = scenes.heightfield_scene(9, n=120, width=300, height=200, tex=128, dof=False)
    vp = vps[0]
    vp.set_post_shader(host.post_shader_depth_box(2.0, fdepth, vp, max_radius=rmax, taps=taps))
    want = oracle.render(scene, vps)[0]
    got = gpu_render(ctx, scene, vps)[0]
    compare(got, want, colour_tol=2)
    # blurred pixels exist, so the test exercises the taps
    assert (want.zbuffer.view(np.uint32) != 0x7F7F7F7F).sum() > 1000


def test_heightfield_1m_1080p(ctx):
    """The full 999,698-triangle 1920x1080 height field with 9-tap DoF."""
    scene, vps = scenes.heightfield_scene(0)
    want = oracle.render(scene, vps)[0]
    got = gpu_render(ctx, scene, vps)[0]
    mx, nd = compare(got, want, colour_tol=2)
    assert got.stats["fragments"] == want.stats["fragments"]
    assert got.stats["drawn_subtriangles"] == want.stats["drawn_subtriangles"]
    print(f"heightfield 1M: fragments {got.stats['fragments']}, colour max diff {mx} at {nd} px")


@pytest.mark.parametrize("seed,clusters,w,h,tex", [(1, 40, 480, 270, 512), (2, 120, 640, 360, 1024)])
def test_heavy_small(ctx, seed, clusters, w, h, tex):
    """Stacked 1000-triangle clusters, 4 viewports from different cameras,
    sun + 8 point lights, 9-tap DoF."""
    scene, vps = scenes.heavy_scene(seed, n_clusters=clusters, width=w, height=h, tex=tex)
    want = oracle.render(scene, vps)
    got = gpu_render(ctx, scene, vps)
    for g, wf in zip(got, want):
        compare(g, wf, colour_tol=2)
    assert got[0].stats["fragments"] == sum(f.stats["fragments"] for f in want)


def test_heavy_full_scene(ctx):
    """One full heavy scene: 500,000 triangles, 4 x 1920x1080, 4096^2 texture."""
    scene, vps = scenes.heavy_scene(5)
    want = oracle.render(scene, vps)
    got = gpu_render(ctx, scene, vps)
    for g, wf in zip(got, want):
        compare(g, wf, colour_tol=2)
    assert got[0].stats["fragments"] == sum(f.stats["fragments"] for f in want)
    assert got[0].stats["drawn_subtriangles"] == sum(f.stats["drawn_subtriangles"] for f in want)


@pytest.mark.parametrize("limit", [0, 1, 4])
def test_bin_limit_paths_agree(ctx, limit):
    """Triangles over the bin limit take the per-viewport list that every
    tile tests; forcing it (limit 0) gives the same bits."""
    scene, vps = scenes.heightfield_scene(4, n=160, width=400, height=300, tex=128, dof=False)
    want = oracle.render(scene, vps)[0]
    ctx.set_bin_limit(limit)
    try:
        got = gpu_render(ctx, scene, vps)[0]
    finally:
        ctx.set_bin_limit(16)
    compare(got, want)
    if limit == 0:
        assert got.stats["big_triangles"] == got.stats["drawn_subtriangles"] or got.stats["big_triangles"] > 0


@pytest.mark.parametrize("tile", [(8, 8), (16, 16), (32, 16), (64, 32), (64, 64)])
def test_tile_sizes_agree(ctx, tile):
    """Each tile size as set (automatic tile shrinking off)."""
    scene, vps = scenes.heightfield_scene(6, n=150, width=333, height=211, tex=128, dof=True)
    want = oracle.render(scene, vps)[0]
    ctx.set_tile(*tile)
    ctx.set_tile_auto(False)
    try:
        got = gpu_render(ctx, scene, vps)[0]
    finally:
        ctx.set_tile(32, 32)
        ctx.set_tile_auto(True)
    compare(got, want, colour_tol=2)


def test_large_triangles_split_over_lanes(ctx):
    """Few large triangles per tile (a 24x24-vertex grid over 640x480): the
    depth replay deals row chunks of each triangle to different lanes."""
    scene, vps = scenes.heightfield_scene(13, n=24, width=640, height=480, tex=256, dof=False)
    want = oracle.render(scene, vps)[0]
    got = gpu_render(ctx, scene, vps)[0]
    compare(got, want)
    assert got.stats["fragments"] == want.stats["fragments"]
