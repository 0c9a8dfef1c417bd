"""The asset loader (src/data/gltf.cpp, src/misc/image.cpp) and node animation
(scene_t::animate, swegl/data/model.hpp:146-177), on files written by
tests/gltf_writer.py: every expected value follows from what was written and
from the file formats' specifications or the reference's reading of them."""
import ctypes as C

import numpy as np
import pytest

from gltf_writer import FLOAT, UINT, USHORT, Gltf, bmp_as_read, png
from swegl_amd import _abi as A
from swegl_amd import gltf, host
from swegl_amd.device import SweglError

F = np.float32


# ------------------------------------------------------------------ PNG ---
def scale8(v, depth):
    """A sample at `depth` bits as 8 bits: 16-bit samples keep their high
    byte, gray below 8 bits is scaled to the full range."""
    return v >> 8 if depth == 16 else v * 255 // ((1 << depth) - 1)


def expected_png(s, ctype, depth, plte=None, trns=None):
    s = np.asarray(s, np.int64)
    h, w = s.shape[:2]
    s = s.reshape(h, w, -1)
    out = np.zeros((h, w), np.uint64)
    for y in range(h):
        for x in range(w):
            v = s[y, x]
            if ctype == 0:
                g = scale8(v[0], depth)
                a = 0 if trns is not None and v[0] == (trns[0] << 8 | trns[1]) else 255
                r = g, g, g, a
            elif ctype == 2:
                a = 255
                if trns is not None and all(v[k] == (trns[2 * k] << 8 | trns[2 * k + 1]) for k in range(3)):
                    a = 0
                r = scale8(v[0], depth), scale8(v[1], depth), scale8(v[2], depth), a
            elif ctype == 3:
                i = int(v[0])
                rgb = tuple(plte[i]) if i < len(plte) else (0, 0, 0)
                a = trns[i] if trns is not None and i < len(trns) else 255
                r = (*rgb, a)
            elif ctype == 4:
                g = scale8(v[0], depth)
                r = g, g, g, scale8(v[1], depth)
            else:
                r = tuple(scale8(v[k], depth) for k in range(4))
            out[y, x] = host.pixel_colors(r[2], r[1], r[0], r[3])
    return out.astype(np.uint32)


PNG_CASES = [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4),
             (3, 8), (4, 8), (4, 16), (6, 8), (6, 16)]


@pytest.mark.parametrize("ctype,depth", PNG_CASES)
@pytest.mark.parametrize("interlace", [False, True])
def test_png_decoding(tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth + 7 * interlace)
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    for k, (h, w) in enumerate([(13, 11), (1, 1), (3, 2), (1, 9), (17, 5)]):
        top = (1 << depth) - 1
        plte = trns = None
        if ctype == 3:
            n = min(top + 1, 200)
            plte = rng.integers(0, 256, (n, 3))
            top = min(top, n + 2)              # a few indices past the palette
            trns = list(rng.integers(0, 256, max(1, n // 2)))
        s = rng.integers(0, top + 1, (h, w, ch))
        if ctype in (0, 2) and k % 2 == 0:
            key = s[0, 0]
            trns = sum(([int(v) >> 8, int(v) & 255] for v in key), [])
        f = tmp_path / f"img{k}.png"
        f.write_bytes(png(s, ctype, depth, interlace, plte, bytes(trns) if trns is not None else None,
                          first_filter=k))
        got = gltf.read_image_file(str(f)).bitmap
        want = expected_png(s, ctype, depth, plte, trns)
        assert got.shape == (h, w)
        assert np.array_equal(got, want), (k, np.argwhere(got != want)[:4])


def test_png_rejects_damaged_files(tmp_path):
    good = png(np.zeros((2, 2, 3), int), 2, 8)
    for bad in (good[:-20], good[:8] + b"\x00" * 30, good.replace(b"IHDR", b"IHDQ")):
        f = tmp_path / "bad.png"
        f.write_bytes(bad)
        with pytest.raises(SweglError, match="EINVAL"):
            gltf.read_image_file(str(f))


# ------------------------------------------------------------------ BMP ---
@pytest.mark.parametrize("w,h", [(1, 1), (2, 3), (3, 2), (4, 4), (5, 3), (7, 1)])
def test_bmp_reading(tmp_path, w, h):
    rng = np.random.default_rng(w * 10 + h)
    rgb = rng.integers(0, 256, (h, w, 3))
    f = tmp_path / "t.bmp"
    f.write_bytes(bmp_as_read(rgb))
    got = gltf.read_image_file(str(f)).bitmap
    want = (255 << 24) | (rgb[..., 0] << 16) | (rgb[..., 1] << 8) | rgb[..., 2]
    assert np.array_equal(got, want.astype(np.uint32))


def test_image_formats_not_read(tmp_path):
    for name, code in (("a.jpg", "ENOTSUP"), ("a.tga", "ENOTSUP"), ("missing.png", "EINVAL")):
        if name != "missing.png":
            (tmp_path / name).write_bytes(b"\xff\xd8\xff")
        with pytest.raises(SweglError, match=code):
            gltf.read_image_file(str(tmp_path / name))


# ----------------------------------------------------------------- glTF ---
POS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], F)
NOR = np.array([[0, 0, 1], [0, 0, 1], [0, 1, 0]], F)
UV = np.array([[0.25, 0.5], [0.75, 0.125], [1.5, -2.0]], F)


def rz90_scaled_matrix():
    """T(5,6,7) . Rz(90) . S(2,3,4) in glTF's column-major order."""
    m = np.array([[0, -3, 0, 5], [2, 0, 0, 6], [0, 0, 4, 7], [0, 0, 0, 1]], F)
    return [float(v) for v in m.T.reshape(-1)]


def build_quirks(g, image_refs):
    mat_tex = g.add("textures", {"source": 1})
    g.add("materials", {"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.4, 0.6, 0.5],
                                                 "metallicFactor": 0.25, "roughnessFactor": 0.75,
                                                 "baseColorTexture": {"index": mat_tex}},
                        "doubleSided": True})
    g.add("materials", {"doubleSided": True})            # no pbrMetallicRoughness
    g.add("materials", {"pbrMetallicRoughness": {"baseColorFactor": [1.0, 0.5, 0.0, 1.0]}})
    image_refs(g)
    g.add("meshes", {"name": "no primitives"})
    g.add("meshes", {"primitives": [g.primitive(POS, NOR, UV, idx=[65536, 1, 65538], idx_type=UINT,
                                                material=0, uv_stride=12)]})
    g.add("meshes", {"primitives": [g.primitive(POS, idx=[0, 1, 2, 0], idx_type=USHORT, mode=5,
                                                material=1),
                                    g.primitive(POS, idx=[0, 1], mode=1)]})
    g.add("nodes", {"rotation": [0, 0, 0, 1], "children": [1, 3]})          # 0
    g.add("nodes", {"matrix": rz90_scaled_matrix(), "mesh": 0, "children": [2]})   # 1
    g.add("nodes", {"translation": [1, 2, 3], "scale": [0.5, 0.5, 0.5], "rotation": [1, 0, 0, 0],
                    "mesh": 1})                                               # 2
    g.add("nodes", {"mesh": 0})                                               # 3: mesh already taken
    g.add("nodes", {"translation": [0, 0, -1]})                               # 4: a second root
    g.animation([(2, "translation", [2, 0, 1], [[4, 4, 4], [0, 0, 0], [2, 0, -2]], None),
                 (4, "scale", [0, 3], [[1, 1, 1], [4, 1, 0.5]], "LINEAR"),
                 (4, "rotation", [0, 1], [[0, 0, 0, 1], [1, 0, 0, 0]], None),
                 (0, "weights", [0, 1], [[0], [1]], None)])


def tex_pixels(seed, h=4, w=3):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 4))


def quirks_files(tmp_path, kind):
    g = Gltf()
    if kind == "glb":
        def images(g):
            for k in range(2):
                g.png_image(png(tex_pixels(k), 6, 8))
        build_quirks(g, images)
        path = tmp_path / "quirks.glb"
        g.write_glb(str(path))
    else:
        def images(g):
            for k in range(2):
                (tmp_path / f"t{k}.png").write_bytes(png(tex_pixels(k), 6, 8))
                g.add("images", {"uri": f"t{k}.png"})
        build_quirks(g, images)
        path = tmp_path / "quirks.gltf"
        g.write_gltf(str(path), "quirks.bin")
    return str(path)


@pytest.fixture(params=["glb", "gltf"])
def quirks(tmp_path, request):
    return gltf.load_scene(quirks_files(tmp_path, request.param))


def test_textures_and_materials(quirks):
    s = quirks
    assert len(s.images) == 2
    for k in range(2):
        assert np.array_equal(s.images[k].bitmap, expected_png(tex_pixels(k), 6, 8))
    m0, m1, m2 = s.materials
    # (unsigned char)(255 * f): truncation, in b, g, r, a order
    assert m0.color == host.pixel_colors(int(F(255) * F(0.6)), int(F(255) * F(0.4)),
                                         int(F(255) * F(0.2)), 127)
    assert m0.color == host.pixel_colors(153, 102, 51, 127)
    assert (m0.metallic, m0.roughness, m0.texture_idx, m0.double_sided) == (0.25, 0.75, 1, True)
    # without pbrMetallicRoughness: white, one-sided even with doubleSided
    assert (m1.color, m1.metallic, m1.roughness, m1.texture_idx, m1.double_sided) == \
        (0xFFFFFFFF, 1.0, 1.0, -1, False)
    assert m2.color == host.pixel_colors(0, 127, 255, 255)


def test_meshes(quirks):
    s = quirks
    # mesh 0 has no primitives and is skipped: node "mesh": 0 is the second mesh
    p = s.nodes[1].primitives
    assert len(p) == 1
    p = p[0]
    assert len(p.positions) == 5                       # two zero vertices past the accessor
    assert np.array_equal(p.positions[:3], POS) and not p.positions[3:].any()
    assert np.array_equal(p.normals[:3], NOR)
    assert np.array_equal(p.tex_coords[:3], UV[:, ::-1])   # read transposed
    assert list(p.indices) == [0, 1, 2]               # 32-bit indices: the low 16 bits
    assert (p.mode, p.material_id) == (A.TRIANGLES, 0)
    q = s.nodes[2].primitives
    assert [(x.mode, x.material_id) for x in q] == [(A.TRIANGLE_STRIP, 1), (A.LINES, -1)]
    assert not q[0].normals.any() and not q[0].tex_coords.any()
    assert q[1].n_triangles() == 0
    assert s.nodes[3].primitives == []                  # the mesh went to node 1


def test_nodes_and_hierarchy(quirks):
    s = quirks
    assert s.root_nodes == [0, 4]
    assert [n.children_idx for n in s.nodes] == [[1, 3], [2], [], [], []]
    # quaternion (x, y, z, w) = (0, 0, 0, 1) with x taken as the scalar part
    assert np.array_equal(s.nodes[0].rotation, np.diag([-1, -1, 1, 1]).astype(F))
    assert np.array_equal(s.nodes[2].rotation, np.eye(4, dtype=F))        # (1, 0, 0, 0)
    assert np.array_equal(s.nodes[2].translation, [1, 2, 3])
    assert np.array_equal(s.nodes[2].scale, [0.5, 0.5, 0.5])
    # matrix T . Rz(90) . S(2, 3, 4): translation, row lengths as scale,
    # columns divided by them
    n1 = s.nodes[1]
    assert np.array_equal(n1.translation, [5, 6, 7])
    assert np.array_equal(n1.scale, [3, 2, 4])
    want = np.array([[0, -1.5, 0, 0], [F(2) / F(3), 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], F)
    assert np.array_equal(n1.rotation, want)


def test_from_quaternion_closed_forms():
    assert np.array_equal(host.from_quaternion(1, 0, 0, 0), np.eye(4, dtype=F))
    # 90 degrees about z: q = (cos 45, 0, 0, sin 45) with the scalar first
    c = F(np.sqrt(0.5))
    m = host.from_quaternion(c, 0, 0, c)
    assert np.allclose(m[:3, :3], [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-6)
    m = host.from_quaternion(c, c, 0, 0)     # about x
    assert np.allclose(m[:3, :3], [[1, 0, 0], [0, 0, -1], [0, 1, 0]], atol=1e-6)


def test_animation_read(quirks):
    s = quirks
    assert len(s.animations) == 1
    an = s.animations[0]
    assert an.end_time == 3.0
    ch = an.channels
    assert [(c.node_idx, c.path) for c in ch] == [(2, A.PATH_TRANSLATION), (4, A.PATH_SCALE),
                                                  (4, A.PATH_ROTATION), (0, A.PATH_NONE)]
    assert list(ch[0].times) == [0, 1, 2]             # sorted by time
    assert np.array_equal(ch[0].values[:, :3], [[0, 0, 0], [2, 0, -2], [4, 4, 4]])
    assert np.array_equal(ch[0].values[:, 3], [1, 1, 1])


@pytest.mark.parametrize("t,trans,scale", [
    (0.5, [1, 0, -1], None),
    (1.0, [2, 0, -2], None),                  # on a keyframe
    (1.5, [3, 2, 1], [2.5, 1, 0.75]),
    (3.75, [1.5, 0, -1.5], [1.75, 1, 0.875]),  # fmod over end_time 3
    (3.0, [0, 0, 0], [1, 1, 1]),
    (-0.5, [0, 0, 0], [1, 1, 1]),              # before the first keyframe: the first
])
def test_animate_closed_forms(quirks, t, trans, scale):
    s = quirks
    r0 = s.nodes[0].rotation.copy()
    s.animate(t)
    assert np.array_equal(s.nodes[2].translation, np.array(trans, F))
    if scale is not None:
        assert np.array_equal(s.nodes[4].scale, np.array(scale, F))
    assert np.array_equal(s.nodes[0].rotation, r0)      # the weights channel does nothing


def test_animate_rotation_is_normalised():
    """Halfway weights 0.25 / 0.75 between (0, 0, 0, 1) and (1, 0, 0, 0): the
    frame (0.75, 0, 0, 0.25) normalised is (3, 0, 0, 1) / sqrt(10); with the
    scalar first that is a rotation about z with cos = 0.8, sin = 0.6."""
    s = host.scene_t()
    s.nodes = [host.node_t()]
    s.root_nodes = [0]
    s.animations = [host.animation_t([host.animation_channel_t(
        0, A.PATH_ROTATION, [0, 1], [[0, 0, 0, 1], [1, 0, 0, 0]])], 3.0)]
    s.animate(3.75)
    assert np.allclose(s.nodes[0].rotation[:3, :3], [[0.8, -0.6, 0], [0.6, 0.8, 0], [0, 0, 1]],
                       atol=2e-7)


def test_loaded_scene_uploads_and_matches_a_hand_built_one(quirks):
    """The loaded scene and the same scene built by hand give identical
    oracle frames (the loader's scene is an ordinary scene_t)."""
    import oracle
    s = quirks
    s.ambient_light_intensity = 0.2
    s.sun_intensity = 0.8
    s.sun_direction = np.array([0, 0, -1], F)
    hand = host.scene_t()
    hand.__dict__.update({k: v for k, v in s.__dict__.items()})
    hand.nodes = []
    for n in s.nodes:
        m = host.node_t([host.primitive_t(p.positions.copy(), p.normals.copy(), p.tex_coords.copy(),
                                          p.indices.copy(), p.mode, p.material_id) for p in n.primitives])
        m.scale, m.rotation, m.translation = n.scale.copy(), n.rotation.copy(), n.translation.copy()
        m.children_idx = list(n.children_idx)
        hand.nodes.append(m)
    outs = []
    for sc in (s, hand):
        surf = host.Surface(64, 48)
        vp = host.viewport_t(0, 0, 64, 48, surf, host.pixel_shader_t(), 0)
        vp.m_camera.translate(0, 0, -6)
        vp.set_post_shader(host.post_shader_t())
        outs.append(oracle.render(sc, [vp])[0])
    assert np.array_equal(outs[0].pixels, outs[1].pixels)
    assert np.array_equal(outs[0].tid, outs[1].tid)


def write_min(tmp_path, mutate):
    g = Gltf()
    g.add("meshes", {"primitives": [g.primitive(POS, idx=[0, 1, 2])]})
    g.add("nodes", {"mesh": 0})
    mutate(g)
    p = tmp_path / "m.glb"
    g.write_glb(str(p))
    return str(p)


@pytest.mark.parametrize("mutate,code", [
    (lambda g: g.j["meshes"][0]["primitives"].append(g.primitive(POS, idx=[0, 1, 9])), "EINVAL"),
    (lambda g: g.j["nodes"].append({"mesh": 5}), "EINVAL"),
    (lambda g: g.j["accessors"][0].update(count=1000), "EINVAL"),
    (lambda g: g.j["nodes"].append({"children": [7]}), "EINVAL"),
    (lambda g: g.add("images", {"bufferView": 0, "mimeType": "image/jpeg"}), "ENOTSUP"),
    (lambda g: g.j["accessors"][0].pop("bufferView"), "ENOTSUP"),
])
def test_malformed_files_are_refused(tmp_path, mutate, code):
    with pytest.raises(SweglError, match=code):
        gltf.load_scene(write_min(tmp_path, mutate))


def test_file_kinds(tmp_path):
    with pytest.raises(SweglError, match="ENOTSUP"):
        gltf.load_scene(str(tmp_path / "scene.obj"))
    with pytest.raises(SweglError, match="EINVAL"):
        gltf.load_scene(str(tmp_path / "missing.glb"))
    (tmp_path / "bad.gltf").write_text('{"asset": {"version": "2.0"}, "nodes": [}')
    with pytest.raises(SweglError, match="EINVAL"):
        gltf.load_scene(str(tmp_path / "bad.gltf"))
