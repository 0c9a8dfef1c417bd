"""Pins of the oracle's shading and depth-of-field against closed forms and
an independent brute force (pixel_shaders.cpp:159-205, 348-384;
pixel_shaders.hpp:159-178; post_shaders.hpp:51-131 with the build's DoF spec).
"""
import ctypes as C
import math
import random

import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A

L = oracle.lib()
f32 = np.float32


def word(b, g, r, a):
    return b | (g << 8) | (r << 16) | (a << 24)


def cround(x):
    """C round(): halves away from zero."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def chans(w):
    return [(w >> (8 * k)) & 255 for k in range(4)]


def bil(tex, tx, ty):
    tex = np.ascontiguousarray(tex, np.uint32)
    return L.oracle_bilinear(A.u32ptr(tex), tex.shape[1], tex.shape[0], tx, ty)


# ------------------------------------------------------------ bilinear
def test_bilinear_constant_texture():
    tex = np.full((16, 16), word(10, 200, 33, 255), np.uint32)
    rng = random.Random(1)
    for _ in range(200):
        assert bil(tex, rng.uniform(0.6, 15.4), rng.uniform(0.6, 15.4)) == word(10, 200, 33, 255)


def test_bilinear_texel_centres_and_transposed_axes():
    """At (row+0.5, col+0.5) the sample is exactly texel[row][col]; the row
    index comes from t.x (pixel_shaders.cpp:352-353)."""
    rng = np.random.default_rng(2)
    tex = rng.integers(0, 1 << 32, (8, 12), dtype=np.uint64).astype(np.uint32)
    for r in range(8):
        for c in range(12):
            assert bil(tex, r + 0.5, c + 0.5) == tex[r, c]


def test_bilinear_weights_on_gradient():
    """A texture whose blue channel is 4*col: sampling between texel centres
    interpolates linearly (weights (u-u1) and (u2-u))."""
    tex = np.zeros((4, 64), np.uint32)
    for c in range(64):
        tex[:, c] = word(4 * c, 0, 0, 255)
    for c in range(1, 62):
        for f in (0.125, 0.25, 0.375, 0.5, 0.625, 0.875):
            got = chans(bil(tex, 1.5, c + 0.5 + f))[0]
            want = 4 * (c + f)
            assert got == cround(want), (c, f, got, want)


def test_bilinear_wraps_and_truncates():
    """Coordinates wrap modulo the size; (int) truncation (not floor) picks
    texel 0 and 1 for t in (-0.5, 0.5) (pixel_shaders.cpp:364-373)."""
    tex = np.zeros((4, 8), np.uint32)
    for c in range(8):
        tex[:, c] = word(c * 30, 0, 0, 255)
    # t.y = 7.75: between column 7 and column 0 (wrap)
    got = chans(bil(tex, 1.5, 7.75))[0]
    assert got == cround(7 * 30 * 0.75 + 0 * 0.25)
    # t.y = 0.25: u1 = -0.25 -> (int) 0, u = floor(0.75) = 0: weight of col 0
    # is u-u1 = 0.25 and of col 1 is u2-u = 0.75
    got = chans(bil(tex, 1.5, 0.25))[0]
    assert got == cround(0 * 0.25 + 30 * 0.75)


# --------------------------------------------------------------- phong
def phong(center, normal, cam, ambient, sun, sun_I, lights):
    arr = (A.PointLight * max(1, len(lights)))()
    for i, (p, I) in enumerate(lights):
        arr[i].position[:] = p
        arr[i].intensity = I
    f = lambda v: np.asarray(v, np.float32)
    c, n, k, s = f(center), f(normal), f(cam), f(sun)
    return L.oracle_phong(A.fptr(c), A.fptr(n), A.fptr(k), ambient, A.fptr(s), sun_I, arr, len(lights))


def test_phong_ambient_and_sun():
    """65536*(ambient + max(0, -N.sun)*sun_I), truncated (pixel_shaders.cpp:166-171,204)."""
    assert phong([0, 0, 0], [0, 1, 0], [0, 5, 0], 0.25, [0, 0, 1], 0.0, []) == 16384
    assert phong([0, 0, 0], [0, 1, 0], [0, 5, 0], 0.25, [0, -1, 0], 0.5, []) == 49152
    assert phong([0, 0, 0], [0, 1, 0], [0, 5, 0], 0.25, [0, 1, 0], 0.5, []) == 16384   # sun behind
    s = np.array([0.6, -0.8, 0.0], np.float32)
    want = int(65536 * (f32(0.1) + f32(f32(0.8) * f32(0.7))))
    assert phong([0, 0, 0], [0, 1, 0], [0, 5, 0], 0.1, s, 0.7, []) in (want, want - 1, want + 1)


def test_phong_point_light_closed_form():
    """Light straight above at distance 2 and camera above: a = 1, d2 = 4,
    specular 1^32*32/2/d2 = 4 (pixel_shaders.cpp:176-202)."""
    got = phong([0, 0, 0], [0, 1, 0], [0, 5, 0], 0.125, [0, 0, 1], 0.0, [([0, 2, 0], 1.0)])
    assert got == int(65536 * (0.125 + 0.25 + 4.0))
    # camera off the mirror direction: R = (0,1,0), C = normalize(3,4,0): s = 0.8
    got = phong([0, 0, 0], [0, 1, 0], [3, 4, 0], 0.125, [0, 0, 1], 0.0, [([0, 2, 0], 1.0)])
    spec = float(f32(0.8)) ** 32 * 16 / 4
    assert abs(got - 65536 * (0.125 + 0.25 + spec)) <= 2


def test_phong_cutoffs():
    # diffuse = I/d2 below 0.05 contributes nothing; just above it does
    below = phong([0, 0, 0], [0, 1, 0], [0, 5, 0], 0.0, [0, 0, 1], 0.0, [([0, 2, 0], 0.199)])
    above = phong([0, 0, 0], [0, 1, 0], [0, 5, 0], 0.0, [0, 0, 1], 0.0, [([0, 2, 0], 0.201)])
    assert below == 0
    assert above > 65536 * 0.05
    # light behind the surface (alignment < 0): nothing
    assert phong([0, 0, 0], [0, 1, 0], [0, 5, 0], 0.0, [0, 0, 1], 0.0, [([0, -2, 0], 4.0)]) == 0
    # specular facing away (s <= 0): diffuse only
    got = phong([0, 0, 0], [0, 1, 0], [0, -5, 0.1], 0.0, [0, 0, 1], 0.0, [([0, 2, 0], 1.0)])
    assert got == int(65536 * 0.25)


def test_phong_lights_accumulate_in_order():
    lights = [([0.5, 2, 0], 1.0), ([-1, 3, 1], 2.0), ([2, 1, -1], 0.5)]
    tot = phong([0, 0, 0], [0, 1, 0], [0, 5, 0], 0.1, [0, 0, 1], 0.0, lights)
    parts = [phong([0, 0, 0], [0, 1, 0], [0, 5, 0], 0.0, [0, 0, 1], 0.0, [l]) for l in lights]
    assert abs(tot - (sum(parts) + 6553.6)) <= 4


# ------------------------------------------------------------- combine
def test_combine_closed_forms():
    c = word(200, 100, 50, 77)
    assert chans(L.oracle_combine(c, 32768)) == [100, 50, 25, 77]      # light 0.5
    assert L.oracle_combine(c, 65536) == c                              # light 1
    got = chans(L.oracle_combine(c, 65536 * 16))                       # sqrt(sqrt(16)) = 2
    assert got == [255 - (55 // 2), 255 - (155 // 2), 255 - (205 // 2), 77]
    assert chans(L.oracle_combine(c, 0)) == [0, 0, 0, 77]
    assert chans(L.oracle_combine(c, -2**31)) == [0, 0, 0, 77]          # (float)INT_MIN scaling


# ----------------------------------------------------------------- DoF
MAXZ = np.frombuffer(np.uint32(0x7F7F7F7F).tobytes(), np.float32)[0]


def dof(src, z, fd, fdepth, rmax, taps9):
    h, w = src.shape
    out = np.zeros_like(src)
    L.oracle_dof(A.u32ptr(np.ascontiguousarray(src)), A.fptr(np.ascontiguousarray(z)), w, h, fd,
                 fdepth, rmax, int(taps9), A.u32ptr(out))
    return out


def dof_brute(src, z, fd, fdepth, rmax, taps9):
    """The DoF specification restated independently in Python."""
    h, w = src.shape
    bf = np.zeros((h, w), np.float64)
    for y in range(h):
        for x in range(w):
            t = f32(abs(f32(fd) - f32(z[y, x])))
            if fdepth == 1:
                u = f32(0.5)
            elif t <= 1:
                u = f32(0)
            elif t >= fdepth:
                u = f32(1)
            else:
                u = f32(f32(t - f32(1)) / f32(f32(fdepth) - f32(1)))
            bf[y, x] = 0.0 if u <= 0 else (rmax if u >= rmax else float(f32(f32(rmax) * u)))
    out = src.copy()
    for y in range(h):
        for x in range(w):
            r = int(bf[y, x])
            if r == 0:
                continue
            if taps9:
                taps = [(x + a, y + b) for b in (-r, 0, r) for a in (-r, 0, r)]
            else:
                taps = [(i, j) for j in range(y - r, y + r) for i in range(x - r, x + r)]
            acc = [0, 0, 0]
            n = 0
            for i, j in taps:
                if 0 <= i < w and 0 <= j < h and bf[j, i] != 0:
                    n += 1
                    for k in range(3):
                        acc[k] += (int(src[j, i]) >> (8 * k)) & 255
            if n:
                out[y, x] = word(acc[0] // n, acc[1] // n, acc[2] // n, 255)
    return out


def test_dof_in_focus_is_identity():
    rng = np.random.default_rng(3)
    src = rng.integers(0, 1 << 32, (12, 16), dtype=np.uint64).astype(np.uint32)
    z = np.full((12, 16), 5.5, np.float32)   # |5 - 5.5| <= 1: radius 0
    assert np.array_equal(dof(src, z, 5.0, 5.0, 8, True), src)
    assert np.array_equal(dof(src, z, 5.0, 5.0, 5, False), src)


def test_dof_constant_image_and_background():
    src = np.full((10, 10), word(9, 99, 199, 17), np.uint32)
    z = np.full((10, 10), MAXZ, np.float32)   # background: max radius
    out = dof(src, z, 5.0, 5.0, 5, False)
    assert (out == word(9, 99, 199, 255)).all()


@pytest.mark.parametrize("taps9,rmax", [(False, 5), (True, 8), (True, 3)])
def test_dof_brute_force(taps9, rmax):
    rng = np.random.default_rng(4 + rmax)
    src = rng.integers(0, 1 << 32, (14, 19), dtype=np.uint64).astype(np.uint32)
    z = rng.uniform(0.5, 12, (14, 19)).astype(np.float32)
    z[rng.random((14, 19)) < 0.2] = MAXZ
    z[rng.random((14, 19)) < 0.2] = 5.25
    assert np.array_equal(dof(src, z, 5.0, 5.0, rmax, taps9), dof_brute(src, z, 5.0, 5.0, rmax, taps9))
