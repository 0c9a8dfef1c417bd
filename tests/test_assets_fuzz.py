"""Damaged files never crash the asset loader: seeded byte mutations and
truncations of valid GLB, glTF, PNG and BMP files either load or raise
SweglError (the reference's loader has undefined behaviour on such input;
this one checks every read against its buffer)."""
import numpy as np
import pytest

from gltf_writer import UINT, Gltf, bmp_as_read, png
from swegl_amd import gltf
from swegl_amd.device import SweglError

from test_assets import build_quirks, tex_pixels


def glb_bytes(tmp_path):
    g = Gltf()

    def images(g):
        for k in range(2):
            g.png_image(png(tex_pixels(k), 6, 8))
    build_quirks(g, images)
    p = tmp_path / "q.glb"
    g.write_glb(str(p))
    return p.read_bytes()


def mutate(rng, data):
    b = bytearray(data)
    kind = rng.integers(0, 4)
    if kind == 0:                                  # flip a few bytes
        for _ in range(rng.integers(1, 8)):
            b[rng.integers(0, len(b))] = rng.integers(0, 256)
    elif kind == 1:                                # truncate
        b = b[:rng.integers(0, len(b))]
    elif kind == 2:                                # overwrite a 4-byte word with an extreme value
        i = rng.integers(0, max(1, len(b) - 4))
        b[i:i + 4] = int(rng.choice([0, 0x7FFFFFFF, 0xFFFFFFFF, 0x80000000])).to_bytes(4, "little")
    else:                                          # duplicate a slice
        i, j = sorted(rng.integers(0, len(b), 2))
        b = b[:j] + b[i:j] + b[j:]
    return bytes(b)


def try_load(fn, path):
    try:
        fn(str(path))
        return "ok"
    except SweglError:
        return "refused"


@pytest.mark.parametrize("seed", range(4))
def test_mutated_glb(tmp_path, seed):
    rng = np.random.default_rng(seed)
    base = glb_bytes(tmp_path)
    outcomes = set()
    for k in range(150):
        p = tmp_path / f"m{k}.glb"
        p.write_bytes(mutate(rng, base))
        outcomes.add(try_load(gltf.load_scene, p))
    assert "refused" in outcomes


@pytest.mark.parametrize("seed", range(2))
def test_mutated_gltf_json(tmp_path, seed):
    rng = np.random.default_rng(100 + seed)
    g = Gltf()

    def images(g):
        for k in range(2):
            (tmp_path / f"t{k}.png").write_bytes(png(tex_pixels(k), 6, 8))
            g.add("images", {"uri": f"t{k}.png"})
    build_quirks(g, images)
    g.write_gltf(str(tmp_path / "q.gltf"), "q.bin")
    base = (tmp_path / "q.gltf").read_bytes()
    for k in range(150):
        (tmp_path / "m.gltf").write_bytes(mutate(rng, base))
        try_load(gltf.load_scene, tmp_path / "m.gltf")


@pytest.mark.parametrize("ctype,depth,interlace", [(6, 8, False), (3, 4, True), (0, 16, True), (2, 8, False)])
def test_mutated_png(tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 10 + depth)
    ch = {0: 1, 2: 3, 3: 1, 6: 4}[ctype]
    s = rng.integers(0, 1 << depth, (9, 13, ch))
    plte = rng.integers(0, 256, (16, 3)) if ctype == 3 else None
    base = png(s, ctype, depth, interlace, plte)
    for k in range(200):
        p = tmp_path / "m.png"
        p.write_bytes(mutate(rng, base))
        try_load(gltf.read_image_file, p)


def test_mutated_bmp(tmp_path):
    rng = np.random.default_rng(7)
    base = bmp_as_read(rng.integers(0, 256, (5, 7, 3)))
    for k in range(200):
        p = tmp_path / "m.bmp"
        p.write_bytes(mutate(rng, base))
        try_load(gltf.read_image_file, p)
