"""Pins of the oracle's FP64 re-shade (the reference's shading model of
pixel_shaders.cpp:88-384 and pixel_shaders.hpp:159-178 carried in double,
for the same winner of each pixel), which measures the colour error of
FP32 shading."""
import numpy as np
import pytest

import oracle
from swegl_amd import host, scenes

from test_oracle_scene import A_, B_, C_, flat_scene, view


def _frame(scene, vp):
    fr = oracle.render(scene, [vp], reshade64=True)[0]
    return fr, fr.pixels.reshape(vp.m_screen.h, vp.m_screen.w)


def test_reshade_matches_when_light_is_exactly_one():
    """Ambient 0.5 + head-on sun 0.5 gives light 65536/65536 = 1, so the
    combine keeps the texel; a constant texture makes both shades the texel
    colour exactly."""
    fr, px = _frame(flat_scene([[A_, B_, C_]]), view())
    drawn = fr.tid != 0xFFFFFFFF
    assert drawn.sum() > 50
    want = host.pixel_colors(100, 150, 200, 255)
    assert (px[drawn] == want).all()
    assert (fr.color64[drawn] == want).all()
    assert (fr.color64[~drawn] == 0).all()


def test_reshade_closed_form_ambient_only():
    """light < 1 branch: channel = trunc(c * L / 65536) with
    L = trunc(65536 * ambient); c * L is exact in both precisions."""
    s = flat_scene([[A_, B_, C_]])
    s.sun_intensity = 0.0
    s.ambient_light_intensity = 0.3
    fr, px = _frame(s, view())
    drawn = fr.tid != 0xFFFFFFFF
    L = int(np.float32(65536) * np.float32(0.3))
    want = [(100 * L) >> 16, (150 * L) >> 16, (200 * L) >> 16]
    for img in (px, fr.color64):
        ch = img[drawn].view(np.uint8).reshape(-1, 4)
        for k in range(3):   # pixel_colors(100, 150, 200, a) keeps 100 in byte 0
            assert (ch[:, k] == want[k]).all()
        assert (ch[:, 3] == 255).all()


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_reshade_error_small_on_cube(seed):
    sc, vps = scenes.cube_scene(seed=seed)
    fr, px = _frame(sc, vps[0])
    e = oracle.reshade_error(px, fr.tid, fr.color64)
    assert e["pixels"] > 10_000
    assert e["max"] <= 1 and e["mean"] < 0.005


def test_reshade_error_small_with_point_lights():
    """Sun + 4 point lights with specular: FP32 stays within 2 LSB of the
    FP64 shade, and a wrong sign or a dropped term would not."""
    sc, vps = scenes.heightfield_scene(0, n=160, width=320, height=180)
    for vp in vps:
        vp.set_post_shader(host.post_shader_t())
    fr, px = _frame(sc, vps[0])
    e = oracle.reshade_error(px, fr.tid, fr.color64)
    assert e["pixels"] > 20_000
    assert e["max"] <= 2 and e["mean"] < 0.02
