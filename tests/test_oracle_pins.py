"""Pins of the CPU oracle (oracle/refcpu.cpp) against what the mathematics
fixes: closed forms, exact special cases, invariants and brute force on tiny
inputs.  The reference cannot be built here (it needs freon/Matrix.hpp and
SDL.h, absent from the image), so these checks are what ties the oracle to
the reference's algorithm; each names the reference lines it pins.
"""
import ctypes as C
import math
import random
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle
from swegl_amd import _abi as A

L = oracle.lib()
f32 = np.float32


def fl(x):
    return float(f32(x))


# ------------------------------------------------------------ literals
def test_double_literal_compares_equal_float_compares():
    """renderer.cpp:65,281,286,318,489 and pixel_shaders.cpp:181 compare a
    float with a double literal; the device uses float literals.  Equal iff
    no float lies in [lit, float(lit)), i.e. float(lit) is the least float
    above lit."""
    for lit in (0.001, 0.05):
        up = f32(lit)
        assert float(up) > lit
        assert float(np.nextafter(up, f32(0))) < lit


def test_cvtt_x86_semantics():
    """(int)f on x86-64 (cvttss2si): truncation, INT_MIN when out of range."""
    assert L.oracle_cvtt(2.9) == 2
    assert L.oracle_cvtt(-2.9) == -2
    assert L.oracle_cvtt(float("nan")) == -2**31
    assert L.oracle_cvtt(3e9) == -2**31
    assert L.oracle_cvtt(-3e9) == -2**31
    assert L.oracle_cvtt(2147483520.0) == 2147483520
    assert L.oracle_cvtt(-2147483648.0) == -2**31


# ---------------------------------------------------------- transforms
def _xf(v, m):
    v = np.asarray(v, np.float32)
    m = np.ascontiguousarray(m, np.float32)
    out = np.zeros(3, np.float32)
    L.oracle_transform(A.fptr(v), A.fptr(m), A.fptr(out))
    return out


def test_transform_exact_cases():
    """transform (points.cpp:8-13): identity, translation, axis permutation
    and power-of-two scales are exact in float."""
    v = np.array([1.25, -3.5, 7.0], np.float32)
    assert np.array_equal(_xf(v, np.eye(4)), v)
    T = np.eye(4, dtype=np.float32)
    T[:3, 3] = [0.5, 2.0, -4.0]
    assert np.array_equal(_xf(v, T), v + T[:3, 3])
    P = np.zeros((4, 4), np.float32)
    P[0, 1] = 1
    P[1, 2] = -1
    P[2, 0] = 2
    P[3, 3] = 1
    assert np.array_equal(_xf(v, P), np.array([v[1], -v[2], 2 * v[0]], np.float32))


def test_transform_random_close_to_exact():
    rng = np.random.default_rng(1)
    for _ in range(200):
        m = rng.uniform(-2, 2, (4, 4)).astype(np.float32)
        v = rng.uniform(-5, 5, 3).astype(np.float32)
        want = m[:3, :3].astype(np.float64) @ v.astype(np.float64) + m[:3, 3]
        got = _xf(v, m)
        assert np.allclose(got, want, rtol=0, atol=4e-6 * (1 + np.abs(want).max()))


def _vertex(v, n, M, RS, view, proj):
    args = [np.asarray(a, np.float32).copy() for a in (v, n, M, RS, view, proj)]
    outs = [np.zeros(3, np.float32) for _ in range(3)]
    L.oracle_vertex(*[A.fptr(a) for a in args + outs])
    return outs


def test_vertex_stage_projection_closed_form():
    """vertex_shaders.hpp:35-71: frustum x = P00*x_c/|z'|, y = P11*y_c/|z'|,
    z' = P22*z_c + P23; no divide when z' == 0."""
    view = np.eye(4, dtype=np.float32)
    view[2, 2] = -1
    proj = np.eye(4, dtype=np.float32)
    proj[0, 0], proj[1, 1] = 1.0, f32(4.0 / 3.0)
    proj[2, 2], proj[2, 3], proj[3, 2] = f32(10 / 9.5), f32(-5 / 9.5), 1
    rng = np.random.default_rng(2)
    for _ in range(100):
        w = rng.uniform(-2, 2, 3).astype(np.float32)
        w[2] = rng.uniform(-9, -0.6)   # in front of the camera
        vw, nw, fr = _vertex(w, [0, 1, 0], np.eye(4), np.eye(4), view, proj)
        zc = -float(w[2])
        zp = float(proj[2, 2]) * zc + float(proj[2, 3])
        assert abs(fr[2] - zp) <= 2e-6 * abs(zp) + 1e-7
        assert abs(fr[0] - float(w[0]) / abs(zp)) <= 1e-5 * (1 + abs(fr[0]))
        assert abs(fr[1] - float(proj[1, 1]) * float(w[1]) / abs(zp)) <= 1e-5 * (1 + abs(fr[1]))
    # z' exactly 0 at the near distance 0.5: x and y are not divided
    w = np.array([0.75, -0.25, -0.5], np.float32)
    _, _, fr = _vertex(w, [0, 1, 0], np.eye(4), np.eye(4), view, proj)
    assert fr[2] == 0.0
    assert fr[0] == f32(0.75) and fr[1] == f32(f32(4 / 3) * f32(-0.25))


def test_vertex_stage_normals():
    """normal_world = normalise(rotate(n, scale(R, S))): unit length, parallel
    to R*diag(S)*n, and ignores translation."""
    rng = np.random.default_rng(3)
    for _ in range(100):
        a = rng.uniform(0, 2 * math.pi)
        R = np.eye(4, dtype=np.float32)
        R[0, 0], R[0, 1], R[1, 0], R[1, 1] = math.cos(a), -math.sin(a), math.sin(a), math.cos(a)
        S = rng.uniform(0.5, 2, 3)
        RS = R.copy()
        RS[:, :3] *= S.astype(np.float32)
        RS[:3, 3] = [5, 6, 7]
        n = rng.standard_normal(3).astype(np.float32)
        _, nw, _ = _vertex([0, 0, 0], n, np.eye(4), RS, np.eye(4), np.eye(4))
        want = RS[:3, :3].astype(np.float64) @ n
        want /= np.linalg.norm(want)
        assert abs(np.linalg.norm(nw.astype(np.float64)) - 1) < 1e-6
        assert np.allclose(nw, want, atol=2e-6)


# -------------------------------------------------------- interpolator
def _interp(d, z1, z2, m, steps):
    out = np.zeros(2, np.float32)
    L.oracle_interp(d, z1, z2, m, steps, A.fptr(out))
    return float(out[0]), float(out[1])


def test_interpolator_constant_depth_is_exact():
    """interpolator.hpp:77-103: with z1 == z2 the value is z1 exactly."""
    rng = random.Random(4)
    for _ in range(300):
        z = fl(rng.uniform(0.01, 20))
        d = fl(rng.uniform(0.5, 300))
        m = fl(rng.uniform(0, 1))
        k = rng.randrange(0, int(d) + 1)
        assert _interp(d, z, z, m, k)[1] == z


def test_interpolator_start_and_perspective_closed_form():
    """u after displacement m and k steps is the perspective-correct
    parameter t/z2 / ((1-t)/z1 + t/z2) with t = (m+k)/d; value = z1+(z2-z1)u
    is the harmonic interpolation of depth."""
    assert _interp(1.0, 2.0, 3.0, 0.0, 0) == (0.0, 2.0)
    rng = random.Random(5)
    for _ in range(500):
        z1, z2 = fl(rng.uniform(0.05, 10)), fl(rng.uniform(0.05, 10))
        d = fl(rng.uniform(1, 400))
        m = fl(rng.uniform(0, 1))
        k = rng.randrange(0, max(1, int(d - m)))
        u, val = _interp(d, z1, z2, m, k)
        t = (m + k) / d
        u_ex = (t / z2) / ((1 - t) / z1 + t / z2)
        z_ex = 1.0 / ((1 - t) / z1 + t / z2)
        assert abs(u - u_ex) <= 2e-5 * (1 + k), (d, z1, z2, m, k)
        assert abs(val - z_ex) <= 2e-5 * (1 + k) * max(z1, z2)


# -------------------------------------------------------------- raster
def _raster(v, w, h):
    v = np.asarray(v, np.float32).reshape(9)
    cover = np.zeros((h, w), np.uint32)
    zl = np.zeros((h, w), np.float32)
    L.oracle_raster_triangle(A.fptr(v), w, h, A.u32ptr(cover), A.fptr(zl))
    return cover, zl


def _exact_coverage(v, w, h):
    """Pixel (x,y) is filled iff ceil(ytop) <= y < ceil(ybot) and
    ceil(xl(y)) <= x < ceil(xr(y)) with exact edge abscissae, clipped to the
    viewport: fill_triangle_2/fill_half_triangle (renderer.cpp:361-558)."""
    P = sorted([(Fr(float(v[i][1])), Fr(float(v[i][0]))) for i in range(3)])
    (y0, x0), (y1, x1), (y2, x2) = P
    cov = np.zeros((h, w), bool)
    near = np.zeros((h, w), bool)
    if math.ceil(y0) == math.ceil(y2):
        return cov, near

    def at(ya, xa, yb, xb, y):
        return xa + (xb - xa) * (y - ya) / (yb - ya)

    for y in range(max(math.ceil(y0), 0), min(math.ceil(y2), h)):
        xl = at(y0, x0, y2, x2, y)
        xs = at(y0, x0, y1, x1, y) if y < y1 else at(y1, x1, y2, x2, y)
        a, b = min(xl, xs), max(xl, xs)
        lo, hi = max(math.ceil(a), 0), min(math.ceil(b), w)
        cov[y, lo:max(lo, hi)] = True
        for e in (a, b):   # pixels whose boundary sits within 1e-3 of an edge
            c = math.ceil(e)
            if abs(e - c) < Fr(1, 1000) or abs(e - (c - 1)) < Fr(1, 1000):
                for xx in (c - 1, c):
                    if 0 <= xx < w:
                        near[y, xx] = True
    return cov, near


@pytest.mark.parametrize("seed", range(40))
def test_raster_coverage_matches_exact_rule(seed):
    rng = random.Random(seed)
    w = h = 48
    v = [[fl(rng.uniform(-4, 52)), fl(rng.uniform(-4, 52)), fl(rng.uniform(1, 5))] for _ in range(3)]
    cover, _ = _raster(v, w, h)
    assert cover.max() <= 1
    cov, near = _exact_coverage(v, w, h)
    diff = (cover.astype(bool) != cov) & ~near
    assert not diff.any(), np.argwhere(diff)[:5]


def test_raster_dyadic_triangles_exact():
    """Vertices on a 1/4 grid with power-of-two edge heights: every float
    step is exact, so coverage equals the exact rule with no tolerance."""
    rng = random.Random(7)
    for _ in range(60):
        y0 = rng.randrange(0, 16)
        v = [[rng.randrange(0, 160) / 4, y0, 2.0], [rng.randrange(0, 160) / 4, y0 + 8, 2.0],
             [rng.randrange(0, 160) / 4, y0 + 16, 2.0]]
        cover, zl = _raster(v, 48, 48)
        cov, _ = _exact_coverage(v, 48, 48)
        assert np.array_equal(cover.astype(bool), cov)
        assert (zl[cov] == 2.0).all()


def test_raster_depth_is_harmonic_interpolation():
    rng = random.Random(8)
    for _ in range(30):
        v = [[fl(rng.uniform(0, 40)), fl(rng.uniform(0, 40)), fl(rng.uniform(0.5, 8))] for _ in range(3)]
        cover, zl = _raster(v, 40, 40)
        P = np.array(v, np.float64)
        for (y, x) in np.argwhere(cover > 0)[::7]:
            # barycentric of the pixel; depth is 1/(interpolated 1/z) along the
            # DDA's edge and span parameters, which for a triangle equals the
            # screen-space barycentric interpolation of 1/z.
            A_ = np.array([[P[0, 0], P[1, 0], P[2, 0]], [P[0, 1], P[1, 1], P[2, 1]], [1, 1, 1]])
            lam = np.linalg.solve(A_, [x, y, 1])
            z = 1.0 / (lam @ (1 / P[:, 2]))
            assert abs(zl[y, x] - z) <= 1e-3 * z, (v, x, y)


def test_shared_edges_cover_once():
    """Fill convention: a triangulated quad fan around a shared centre covers
    each interior pixel once (half-open spans and rows)."""
    rng = random.Random(9)
    for _ in range(20):
        c = [fl(rng.uniform(20, 28)), fl(rng.uniform(20, 28)), 1.0]
        ring = [[4.3, 3.7, 1.0], [44.6, 5.1, 1.0], [43.2, 44.4, 1.0], [3.9, 42.8, 1.0]]
        total = np.zeros((48, 48), np.uint32)
        for i in range(4):
            cv, _ = _raster([c, ring[i], ring[(i + 1) % 4]], 48, 48)
            total += cv
        assert total.max() <= 1 or (total > 1).sum() <= 4
        inner = total[12:36, 12:36]
        assert (inner >= 1).all() or (inner == 0).sum() <= 4
