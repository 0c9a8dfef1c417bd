"""Pins of the oracle's float sequences against the reference's literal
statements, on inputs where a plausible restatement gives other bits.

Each expected value below is computed in this file by IEEE float32 stepping
of the reference's own statements (numpy float32 scalars and
`np.add.accumulate`, which adds strictly in sequence), independently of the
oracle's code.  Every test also checks that the inputs discriminate: the
closed forms a restatement could be tempted by (x = x0 + ratio*k instead of
`x += ratio` per row, ta = ta0 + k*ts instead of `ta += ts` per step, one
normalisation instead of three, the hierarchy product in the other order)
give different results on them, so an oracle taking any of those forms fails.

  renderer.cpp:361-460   fill_triangle_2: y-sort, long side displaced to the
                         first row, halves, long side carried across halves
  renderer.cpp:462-558   fill_half_triangle: x += ratio per row (:553-554),
                         per-span interpolator (:474-486), Step() per row (:555-556)
  interpolator.hpp:52-104 InitSelf / DisplaceStartingPoint / Step
  vertex_shaders.hpp:16-33 original_to_world: parent * local, recursively
  vertex_shaders.hpp:61-71 normal_world = rotate(n, scale(rotation, scale)).normalize()
                         assigned to a normal_t: three normalisations
                         (points.hpp:80-90, 139-150; points.cpp:44-49)
"""
import ctypes as C
import math

import numpy as np

import oracle
from swegl_amd import _abi as A
from swegl_amd import host

f32 = np.float32


# ------------------------------------------------ reference restatements ---
class Interp:
    """interpolator_g<1> (interpolator.hpp:52-104) in float32 scalars.
    closed=True is the alternative the pins must reject: the k-th step
    computed as ta0 + k*ts, ba0 + k*bs."""

    def __init__(self, dist, z1, z2, closed=False):
        self.closed = closed
        self.v0 = f32(z1)
        self.v1 = f32(f32(z2) - f32(z1))
        alphastep = f32(f32(1.0) / f32(dist))
        self.ba = f32(f32(1.0) / f32(z1))
        invz2 = f32(f32(1.0) / f32(z2))
        self.ta = f32(f32(0.0) * self.ba)
        self.ts = f32(f32(f32(1.0) * invz2 - self.ta) * alphastep)
        self.bs = f32(f32(invz2 - self.ba) * alphastep)
        self.u = f32(0.0)
        self.k = 0

    def displace(self, move):
        self.ta = f32(self.ta + f32(self.ts * f32(move)))
        self.ba = f32(self.ba + f32(self.bs * f32(move)))
        self.ta0, self.ba0 = self.ta, self.ba
        self.u = f32(self.ta / self.ba)

    def step(self):
        self.k += 1
        if self.closed:
            self.ta = f32(self.ta0 + f32(self.ts * f32(self.k)))
            self.ba = f32(self.ba0 + f32(self.bs * f32(self.k)))
        else:
            self.ta = f32(self.ta + self.ts)
            self.ba = f32(self.ba + self.bs)
        self.u = f32(self.ta / self.ba)

    def value(self):
        return f32(self.v0 + f32(self.v1 * self.u))

    def run(self, n):
        """value() at this point and after each of the next n-1 steps, then
        n steps taken (the per-pixel loop of fill_half_triangle)."""
        if n <= 0:
            return np.zeros(0, np.float32)
        if self.closed:
            k = np.arange(self.k, self.k + n, dtype=np.float32)
            ta = np.where(k == 0, self.ta, (self.ta0 + self.ts * k).astype(np.float32)).astype(np.float32)
            ba = np.where(k == 0, self.ba, (self.ba0 + self.bs * k).astype(np.float32)).astype(np.float32)
        else:
            ta = np.add.accumulate(np.concatenate([[self.ta], np.full(n - 1, self.ts, np.float32)]).astype(np.float32))
            ba = np.add.accumulate(np.concatenate([[self.ba], np.full(n - 1, self.bs, np.float32)]).astype(np.float32))
        u = (ta / ba).astype(np.float32)
        return (self.v0 + (self.v1 * u).astype(np.float32)).astype(np.float32)


class Side:
    """line_side: x stepped by ratio per row, with its interpolator."""

    def __init__(self, ratio, x, it, closed_x):
        self.ratio, self.x, self.it, self.closed_x = ratio, x, it, closed_x
        self.x0, self.k = x, 0

    def step(self):
        self.k += 1
        if self.closed_x:
            self.x = f32(self.x0 + f32(self.ratio * f32(self.k)))
        else:
            self.x = f32(self.x + self.ratio)
        self.it.step()


def _cvtt_ceil(x):
    c = math.ceil(float(x))
    return c if -2 ** 31 <= c < 2 ** 31 else -2 ** 31


def fill_triangle_2(v, w, h, closed_x=False, closed_interp=False, restart_long=False):
    """fill_triangle_2 + fill_half_triangle (renderer.cpp:361-558) over a
    w x h viewport at (0, 0): per-pixel fragment count and the last z, as
    oracle_raster_triangle reports them."""
    cover = np.zeros((h, w), np.uint32)
    zlast = np.zeros((h, w), np.float32)
    p = [tuple(f32(c) for c in v[3 * i:3 * i + 3]) for i in range(3)]
    v0, v1, v2 = p
    if v1[1] < v0[1]:
        v0, v1 = v1, v0
    if v2[1] < v1[1]:
        v1, v2 = v2, v1
    if v1[1] < v0[1]:
        v0, v1 = v1, v0
    y0, y1, y2 = (_cvtt_ceil(q[1]) for q in (v0, v1, v2))
    if y0 == y2:
        return cover, zlast
    vy, vx = 0, 0
    ratio = f32(f32(v2[0] - v0[0]) / f32(v2[1] - v0[1]))
    it = Interp(f32(v2[1] - v0[1]), v0[2], v2[2], closed_interp)
    ys = vy if y0 < vy else y0
    it.displace(f32(f32(ys) - v0[1]))
    long_side = Side(ratio, f32(v0[0] + f32(ratio * f32(f32(ys) - v0[1]))), it, closed_x)

    def half(y, y_end, left, right):
        for yy in range(y, y_end):
            x1 = max(_cvtt_ceil(left.x), vx)
            x2 = min(_cvtt_ceil(right.x), vx + w)
            if x1 < x2:
                q = Interp(f32(right.x - left.x), left.it.value(), right.it.value(), closed_interp)
                q.displace(f32(f32(x1) - left.x))
                cover[yy, x1:x2] += 1
                zlast[yy, x1:x2] = q.run(x2 - x1)
            left.step()
            right.step()

    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        if y1 >= vy:
            r = f32(f32(v1[0] - v0[0]) / f32(v1[1] - v0[1]))
            s_it = Interp(f32(v1[1] - v0[1]), v0[2], v1[2], closed_interp)
            y = max(y0, vy)
            y_end = min(y1, vy + h)
            s_it.displace(f32(f32(y) - v0[1]))
            short = Side(r, f32(v0[0] + f32(r * f32(f32(y) - v0[1]))), s_it, closed_x)
            right_long = long_side.ratio > short.ratio
            half(y, y_end, short if right_long else long_side, long_side if right_long else short)
        if y1 < vy + h:
            if restart_long:   # the mistake the carry pin rejects
                yl = max(y1, vy)
                it = Interp(f32(v2[1] - v0[1]), v0[2], v2[2], closed_interp)
                it.displace(f32(f32(yl) - v0[1]))
                long_side = Side(ratio, f32(v0[0] + f32(ratio * f32(f32(yl) - v0[1]))), it, closed_x)
            r = f32(f32(v2[0] - v1[0]) / f32(v2[1] - v1[1]))
            s_it = Interp(f32(v2[1] - v1[1]), v1[2], v2[2], closed_interp)
            y = max(y1, vy)
            y_end = min(y2, vy + h)
            s_it.displace(f32(f32(y) - v1[1]))
            short = Side(r, f32(v1[0] + f32(r * f32(f32(y) - v1[1]))), s_it, closed_x)
            right_long = long_side.ratio < short.ratio
            half(y, y_end, short if right_long else long_side, long_side if right_long else short)
    return cover, zlast


def oracle_raster(v, w, h):
    cover = np.zeros((h, w), np.uint32)
    zlast = np.zeros((h, w), np.float32)
    tri = (C.c_float * 9)(*[float(x) for x in v])
    oracle.lib().oracle_raster_triangle(tri, w, h, A.u32ptr(cover), A.fptr(zlast))
    return cover, zlast


def tall_triangles(seed, n, w, h):
    """Triangles spanning most of a w x h viewport's rows with non-dyadic
    edge slopes: hundreds of sequential adds per edge."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ys = np.sort(rng.uniform(-0.4 * h, 1.2 * h, 3)).astype(np.float32)
        ys[0] = f32(rng.uniform(-0.3 * h, 0.1 * h))
        ys[2] = f32(rng.uniform(0.85 * h, 1.3 * h))
        xs = rng.uniform(-0.2 * w, 1.2 * w, 3).astype(np.float32)
        zs = rng.uniform(0.5, 9.0, 3).astype(np.float32)
        order = rng.permutation(3)
        v = []
        for i in order:
            v += [xs[i], ys[i], zs[i]]
        out.append(np.array(v, np.float32))
    return out


# ------------------------------------------------------------------ tests ---
def test_interpolator_steps_are_sequential_adds():
    """oracle_interp == InitSelf, DisplaceStartingPoint, Step^k restated in
    float32; the closed form differs on a good share of these inputs."""
    rng = np.random.default_rng(7)
    L = oracle.lib()
    out = (C.c_float * 2)()
    differ = 0
    for _ in range(300):
        d = f32(rng.uniform(50, 2000))
        z1, z2 = (f32(x) for x in rng.uniform(0.3, 9.5, 2))
        m = f32(rng.uniform(0, 3))
        steps = int(rng.integers(200, 2500))
        L.oracle_interp(float(d), float(z1), float(z2), float(m), steps, out)
        want = Interp(d, z1, z2)
        want.displace(m)
        alt = Interp(d, z1, z2, closed=True)
        alt.displace(m)
        for _ in range(steps):
            want.step()
            alt.step()
        assert f32(out[0]).view(np.uint32) == want.u.view(np.uint32)
        assert f32(out[1]).view(np.uint32) == want.value().view(np.uint32)
        differ += alt.value().view(np.uint32) != want.value().view(np.uint32)
    assert differ >= 30, f"closed-form interpolator differs on only {differ} of 300 inputs"


def test_dda_coverage_and_depth_follow_the_sequential_float_steps():
    """Coverage and last depth of tall triangles: the oracle's raster equals
    the float32 restatement of renderer.cpp:361-558 bit for bit, and the
    closed forms of x and of the interpolators change coverage or depth on
    these triangles."""
    w, h = 700, 900
    x_differs = z_differs = 0
    for v in tall_triangles(11, 24, w, h):
        got_c, got_z = oracle_raster(v, w, h)
        want_c, want_z = fill_triangle_2(v, w, h)
        assert np.array_equal(got_c, want_c), "coverage differs from the sequential DDA"
        drawn = want_c > 0
        assert np.array_equal(got_z[drawn].view(np.uint32), want_z[drawn].view(np.uint32)), \
            "depth differs from the sequential interpolators"
        alt_c, _ = fill_triangle_2(v, w, h, closed_x=True)
        x_differs += not np.array_equal(alt_c, want_c)
        _, alt_z = fill_triangle_2(v, w, h, closed_interp=True)
        z_differs += not np.array_equal(alt_z[drawn].view(np.uint32), want_z[drawn].view(np.uint32))
    assert x_differs >= 4, f"closed-form x changes coverage of only {x_differs} triangles"
    assert z_differs >= 12, f"closed-form interpolators change depth of only {z_differs} triangles"


def test_dda_long_side_carries_across_halves():
    """The long side's x and interpolator continue from the upper half into
    the lower half (renderer.cpp:399-409, 439-459); a restatement that
    re-derives the long side at the lower half's first row gives another
    coverage or depth on these triangles."""
    w, h = 640, 800
    differs = 0
    for v in tall_triangles(23, 16, w, h):
        got_c, got_z = oracle_raster(v, w, h)
        want_c, want_z = fill_triangle_2(v, w, h)
        assert np.array_equal(got_c, want_c)
        alt_c, alt_z = fill_triangle_2(v, w, h, restart_long=True)
        drawn = want_c > 0
        differs += not (np.array_equal(alt_c, want_c) and
                        np.array_equal(alt_z[drawn].view(np.uint32), want_z[drawn].view(np.uint32)))
    assert differs >= 8, f"restarting the long side changes only {differs} of 16 triangles"


def _normalize(v):
    x, y, z = (f32(a) for a in v)
    l = f32(np.sqrt(f32(f32(f32(x * x) + f32(y * y)) + f32(z * z))))
    if l != 0:
        x, y, z = f32(x / l), f32(y / l), f32(z / l)
    return np.array([x, y, z], np.float32)


def _rotate(n, m):
    """rotate(v, m) before the normal_t constructor (points.cpp:44-49)."""
    return [f32(f32(f32(m[r, 0] * n[0]) + f32(m[r, 1] * n[1])) + f32(m[r, 2] * n[2])) for r in range(3)]


def test_normal_world_is_normalised_three_times():
    """normal_world = rotate(...) [normal_t ctor] .normalize() then
    normal_t::operator= : three normalisations.  On the inputs where one,
    two and three normalisations give three different vectors, the oracle
    must give the third."""
    rng = np.random.default_rng(5)
    L = oracle.lib()
    ident = host.IDENTITY.astype(np.float32).reshape(16)
    proj = ident.copy()
    distinct = 0
    for _ in range(3000):
        rs = np.eye(4, dtype=np.float32)
        rs[:3, :3] = rng.uniform(-2.5, 2.5, (3, 3)).astype(np.float32)
        n = rng.uniform(-1, 1, 3).astype(np.float32)
        n1 = _normalize(_rotate(n, rs))
        n2 = _normalize(n1)
        n3 = _normalize(n2)
        vw, nw, vf = (np.zeros(3, np.float32) for _ in range(3))
        v = np.zeros(3, np.float32)
        L.oracle_vertex(A.fptr(v), A.fptr(n), A.fptr(ident), A.fptr(rs.reshape(16).copy()), A.fptr(ident),
                        A.fptr(proj), A.fptr(vw), A.fptr(nw), A.fptr(vf))
        assert np.array_equal(nw.view(np.uint32), n3.view(np.uint32))
        if not np.array_equal(n1.view(np.uint32), n2.view(np.uint32)) and \
                not np.array_equal(n2.view(np.uint32), n3.view(np.uint32)):
            distinct += 1
    assert distinct >= 10, f"only {distinct} inputs separate one, two and three normalisations"


def _hierarchy_scene():
    """A 3-level hierarchy with exact matrices: parent rotated 90 degrees
    about z (x -> y) and lifted by 5 in z; child scaled by 2 and moved by
    (1,0,0); grandchild moved by (0,1,0).  One vertex per node at (1,0,0)
    with normal (1,0,0)."""
    def node_with_point():
        p = host.primitive_t([[1, 0, 0], [1, 0, 0], [1, 0, 0]], [[1, 0, 0]] * 3, [[0, 0]] * 3, [0, 1, 2])
        return host.node_t([p])

    sc = host.scene_t()
    parent, child, grand = node_with_point(), node_with_point(), node_with_point()
    parent.rotation = np.array([[0, -1, 0, 0], [1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32)
    parent.translation = np.array([0, 0, 5], np.float32)
    child.scale = np.array([2, 2, 2], np.float32)
    child.translation = np.array([1, 0, 0], np.float32)
    grand.translation = np.array([0, 1, 0], np.float32)
    parent.children_idx = [1]
    child.children_idx = [2]
    sc.nodes = [parent, child, grand]
    sc.root_nodes = [0]
    sc.materials = [host.material_t()]
    return sc


def test_hierarchy_product_is_parent_times_local():
    """original_to_world_matrix = parent_matrix * local (vertex_shaders.hpp:18),
    applied down the tree.  Every product here is exact, so the expected
    points follow by hand: parent (1,0,0) -> (0,1,5); child: local gives
    (3,0,0), parent maps it to (0,3,5); grandchild: (1,1,0) -> (3,2,0) ->
    (-2,3,5).  local * parent would give (1,2,10) for the child."""
    sc = _hierarchy_scene()
    surf = host.Surface(16, 16)
    vp = host.viewport_t(0, 0, 16, 16, surf, host.pixel_shader_t())
    vp.set_post_shader(host.post_shader_t())
    st = oracle.vertex_state(sc, vp)
    vw = st["v_world"]
    assert vw[0].tolist() == [0.0, 1.0, 5.0]
    assert vw[3].tolist() == [0.0, 3.0, 5.0]
    assert vw[6].tolist() == [-2.0, 3.0, 5.0]


def test_normal_world_uses_the_node_own_rotation_and_scale():
    """camera_to_frustum rotates normals by scale(node.rotation, node.scale)
    (vertex_shaders.hpp:63): the node's own rotation, not the product with
    its parents, so the child's normal stays (1,0,0) under a parent rotated
    90 degrees, while the parent's own normal turns to (0,1,0)."""
    sc = _hierarchy_scene()
    surf = host.Surface(16, 16)
    vp = host.viewport_t(0, 0, 16, 16, surf, host.pixel_shader_t())
    vp.set_post_shader(host.post_shader_t())
    nw = oracle.vertex_state(sc, vp)["normal_world"]
    assert nw[0].tolist() == [0.0, 1.0, 0.0]
    assert nw[3].tolist() == [1.0, 0.0, 0.0]
    assert nw[6].tolist() == [1.0, 0.0, 0.0]


def test_clip_vertex_lerps_then_renormalises_the_normal():
    """make_clip_vertex (renderer.cpp:290-297): v_world and tex are a + (b - a)*cut
    in float32; the normal is normalize(a + (b - a)*cut) unless the two normals
    are equal (then a's, as it is), and world_to_viewport then normalises it
    three more times (vertex_shaders.hpp:54-71).  The test restates those
    statements in float32; leaving out the lerp's normalisation, or
    normalising equal normals too, gives other normal_world bits on these
    inputs, and the lerp as a*(1-cut) + b*cut other positions."""
    rng = np.random.default_rng(19)
    L = oracle.lib()
    ident = np.eye(4, dtype=np.float32).reshape(16).copy()
    vpm = np.array([320, 320, -240, 240], np.float32)
    out = np.zeros(11, np.float32)
    raw_differs = eq_differs = lerp_differs = 0
    for trial in range(3000):
        rs = np.eye(4, dtype=np.float32)
        rs[:3, :3] = rng.uniform(-2, 2, (3, 3)).astype(np.float32)
        a = rng.uniform(-3, 3, 8).astype(np.float32)
        b = rng.uniform(-3, 3, 8).astype(np.float32)
        equal = trial % 4 == 0
        if equal:
            b[3:6] = a[3:6]
        cut = f32(rng.uniform(0.01, 0.99))
        L.oracle_clip_vertex(A.fptr(a), A.fptr(b), float(cut), A.fptr(rs.reshape(16).copy()), A.fptr(ident),
                             A.fptr(ident), A.fptr(vpm), A.fptr(out))
        vw = (a[:3] + ((b[:3] - a[:3]) * cut).astype(np.float32)).astype(np.float32)
        tex = (a[6:] + ((b[6:] - a[6:]) * cut).astype(np.float32)).astype(np.float32)
        assert np.array_equal(out[:3].view(np.uint32), vw.view(np.uint32))
        assert np.array_equal(out[3:5].view(np.uint32), tex.view(np.uint32))
        alt_vw = (a[:3] * f32(f32(1) - cut) + b[:3] * cut).astype(np.float32)
        lerp_differs += not np.array_equal(alt_vw.view(np.uint32), vw.view(np.uint32))
        lerp = (a[3:6] + ((b[3:6] - a[3:6]) * cut).astype(np.float32)).astype(np.float32)
        n_clip = a[3:6].copy() if equal else _normalize(lerp)

        def world(n):
            return _normalize(_normalize(_normalize(_rotate(n, rs))))

        want = world(n_clip)
        assert np.array_equal(out[5:8].view(np.uint32), want.view(np.uint32))
        if equal:
            eq_differs += not np.array_equal(world(_normalize(a[3:6])).view(np.uint32), want.view(np.uint32))
        else:
            raw_differs += not np.array_equal(world(lerp).view(np.uint32), want.view(np.uint32))
    assert raw_differs >= 10, f"an unnormalised clip normal differs on only {raw_differs} inputs"
    assert eq_differs >= 3, f"normalising equal normals differs on only {eq_differs} inputs"
    assert lerp_differs >= 100, f"a*(1-cut) + b*cut differs on only {lerp_differs} inputs"
