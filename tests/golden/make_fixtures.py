"""Write tests/golden/oracle_frames.json: SHA-256 digests of oracle frames
(surface pixels, z-buffer bits, winning draw sequence) for seeded scenes.

The reference renderer cannot be built in this image (freon/Matrix.hpp and
SDL.h are absent), so these digests are produced by the oracle itself.  They
pin the oracle's behaviour across machines and changes; the tests that tie
the oracle to the reference's algorithm are tests/test_oracle_*.py.
Run:  python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

CASES = [
    ("cube", dict(seed=s)) for s in range(4)
] + [
    ("heightfield", dict(seed=11, n=64, width=160, height=120, tex=64, dof=False)),
    ("heightfield", dict(seed=12, n=128, width=256, height=192, tex=128, dof=True)),
    ("demo", dict(seed=0, precision=20, width=200, height=150, layers=3)),
    ("cube_layers", dict(seed=1, layers=2)),
    ("cube_flat_nearest", dict(seed=2)),
]


def digest(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


def frames():
    import oracle
    from swegl_amd import scenes
    out = []
    for kind, kw in CASES:
        if kind == "cube":
            sc, vps = scenes.cube_scene(**kw)
        elif kind == "heightfield":
            sc, vps = scenes.heightfield_scene(**kw)
        elif kind == "demo":
            sc = scenes.demo_scene(kw["seed"], kw["precision"], kw["precision"])
            vps = [scenes.demo_viewport(kw["width"], kw["height"], kw["layers"])]
        elif kind == "cube_layers":
            sc, vps = scenes.cube_scene(kw["seed"])
            vps[0].transparency_layer_count = kw["layers"]
        else:
            from swegl_amd import host
            sc, vps = scenes.cube_scene(kw["seed"])
            vps[0].m_pixel_shader = host.pixel_shader_light_and_texture(host.pixel_shader_lights_flat,
                                                                         host.pixel_shader_texture)
        fr = oracle.render(sc, vps)[0]
        out.append({"kind": kind, "args": kw, "pixels": digest(fr.pixels),
                    "zbuffer": digest(fr.zbuffer), "tid": digest(fr.tid),
                    "fragments": int(fr.stats["fragments"])})
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_frames.json")
    with open(path, "w") as f:
        json.dump(frames(), f, indent=1)
    print("wrote", path)
