"""Seeded synthetic scenes for the benchmark workloads.

Every generator is deterministic in its seed: numpy's PCG64 bit stream and
IEEE +,-,*,/,sqrt,floor only (no transcendental functions except through
the library's own host helpers, which use the C library's cosf/sinf).

Workloads
  cube      : make_cube(1, 0) rotated by seeded Euler angles, 256^2 texture,
              one sun, 640x480, no depth of field.
  heightfield: n x n grid on [-1,1]^2 with 4-octave value noise (amplitude
              0.2), finite-difference normals, 2048^2 texture, sun + 4 point
              lights, 1920x1080, 9-tap depth of field (r_max 8).
  batch     : 256 height fields, T log-uniform in [1e3, 2e5], square
              framebuffers, 512^2/1024^2 textures, 0-4 point lights, DoF on
              for half of them.
"""
import math

import numpy as np

from . import _abi as A
from .host import (IDENTITY, Surface, material_t, node_t, pixel_colors,
                   pixel_shader_light_and_texture, point_source_light,
                   post_shader_depth_box, post_shader_t, primitive_t, rotate_x,
                   rotate_y, rotate_z, scene_t, texture_t, viewport_t)


def f32(x):
    return np.float32(x)


def normal_t(x, y, z):
    """normal_t(x,y,z): normalised with the reference's float steps."""
    x, y, z = f32(x), f32(y), f32(z)
    l = np.sqrt(f32(f32(x * x) + f32(y * y)) + f32(z * z), dtype=np.float32)
    if l != 0:
        x, y, z = x / l, y / l, z / l
    return np.array([x, y, z], np.float32)


# --------------------------------------------------------------- make_cube
_CUBE = [  # (positions sign, tex, normal) per face: model.hpp:254-325
    ([(-1, 1, 1), (1, 1, 1), (1, 1, -1), (-1, 1, -1)],
     [(0.0, 0.0), (0.5, 0.0), (0.5, 0.333), (0.0, 0.333)], (0, 1, 0)),
    ([(-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)],
     [(0.5, 0.0), (1.0, 0.0), (1.0, 0.333), (0.5, 0.333)], (0, 0, 1)),
    ([(1, -1, 1), (1, -1, -1), (1, 1, -1), (1, 1, 1)],
     [(0.0, 0.333), (0.5, 0.333), (0.5, 0.667), (0.0, 0.667)], (1, 0, 0)),
    ([(1, -1, -1), (-1, -1, -1), (-1, 1, -1), (1, 1, -1)],
     [(0.5, 0.667), (1.0, 0.667), (1.0, 1.0), (0.5, 1.0)], (0, 0, -1)),
    ([(-1, -1, -1), (-1, -1, 1), (-1, 1, 1), (-1, 1, -1)],
     [(0.5, 0.333), (1.0, 0.333), (1.0, 0.667), (0.5, 0.667)], (-1, 0, 0)),
    ([(-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1)],
     [(0.0, 0.667), (0.5, 0.667), (0.5, 1.0), (0.0, 1.0)], (0, -1, 0)),
]


def make_cube(size, material_idx):
    """make_cube (swegl/data/model.hpp:254-325): six 4-vertex fans."""
    h = f32(size) / f32(2.0)
    node = node_t()
    for signs, tex, n in _CUBE:
        pos = np.array([[s * h for s in p] for p in signs], np.float32)
        nor = np.tile(normal_t(*n), (4, 1))
        uv = np.array(tex, np.float32)   # double literals stored to float
        node.primitives.append(primitive_t(pos, nor, uv, [0, 1, 2, 3], A.TRIANGLE_FAN,
                                           material_idx))
    return node


def make_tri(size, material_idx):
    """make_tri (swegl/data/model.hpp:233-252): one triangle in the z = 0 plane."""
    s = f32(size)
    pos = np.array([[0, 0, 0], [s, 0, 0], [0, s, 0]], np.float32)
    nor = np.tile(normal_t(0, 0, 1), (3, 1))
    uv = np.array([[0, 0], [0, 1], [1, 0]], np.float32)
    return node_t([primitive_t(pos, nor, uv, [0, 1, 2], A.TRIANGLES, material_idx)])


def _strips(fill, precision, material_idx):
    n = precision + 1
    V = n * 2 * n
    pos = np.zeros((V, 3), np.float32)
    nor = np.zeros((V, 3), np.float32)
    uv = np.zeros((V, 2), np.float32)
    idx = np.zeros(V, np.uint32)
    fill(pos, nor, uv, idx)
    node = node_t()
    for k in range(n):
        sl = slice(2 * n * k, 2 * n * (k + 1))
        node.primitives.append(primitive_t(pos[sl], nor[sl], uv[sl], idx[sl], A.TRIANGLE_STRIP,
                                           material_idx))
    return node


def make_tore(precision, material_idx):
    """make_tore (swegl/data/model.hpp:327-391): precision + 1 strips around a
    torus of radii 2 and 0.8 about the y axis (library's swegl_make_tore)."""
    from .device import check, lib

    def fill(pos, nor, uv, idx):
        check(lib().swegl_make_tore(int(precision), A.fptr(pos), A.fptr(nor), A.fptr(uv),
                                    A.u32ptr(idx)), "make_tore")
    return _strips(fill, int(precision), material_idx)


def make_sphere(precision, radius, material_idx):
    """make_sphere (swegl/data/model.hpp:393-464) (library's swegl_make_sphere)."""
    from .device import check, lib

    def fill(pos, nor, uv, idx):
        check(lib().swegl_make_sphere(int(precision), float(f32(radius)), A.fptr(pos), A.fptr(nor),
                                      A.fptr(uv), A.u32ptr(idx)), "make_sphere")
    return _strips(fill, int(precision), material_idx)


def calculate_face_normals(primitive):
    """calculate_face_normals (swegl/data/model.hpp:181-231), in place."""
    from .device import check, lib
    check(lib().swegl_calculate_face_normals(primitive.mode, A.fptr(primitive.positions),
                                             A.fptr(primitive.normals), len(primitive.positions),
                                             A.u32ptr(primitive.indices), len(primitive.indices)),
          "calculate_face_normals")


def demo_scene(seed=0, tore_precision=100, sphere_precision=100):
    """The reference demo's scene (src/test_1.cpp:50-148): a torus, a cube
    stretched along x, a sphere, three overlapping triangles of which two are
    translucent, two small cubes at the point lights; the three textures
    (dice.bmp, tex.bmp, mercator.bmp there) are seeded random images of
    256 x 256 here."""
    rng = np.random.default_rng(seed)
    s = scene_t()
    s.images = [random_texture(rng, 256, 256) for _ in range(3)]
    for im in s.images:
        im.bitmap |= np.uint32(0xFF000000)
    cols = [(128, 128, 128, 255, 0), (128, 128, 128, 255, 1), (128, 128, 128, 255, 2),
            (128, 128, 255, 255, -1), (255, 128, 255, 255, -1), (128, 128, 255, 100, -1),
            (128, 255, 128, 100, -1), (255, 128, 128, 100, -1)]
    s.materials = [material_t(pixel_colors(b, g, r, a), 1, 1, t) for (b, g, r, a, t) in cols]
    s.ambient_light_intensity = float(f32(0.2))
    s.sun_direction = normal_t(*normal_t(1.0, -1.0, -1.0))
    s.sun_intensity = float(f32(0.3))
    s.point_source_lights = [point_source_light(np.array([0.0, 3.0, 0.0], np.float32), 0.6),
                             point_source_light(np.array([0.5, 2.0, 0.0], np.float32), 100.0)]
    tore = make_tore(tore_precision, 1)
    tore.rotation = rotate_z(IDENTITY.copy(), 0.5)
    tore.translation = np.array([0, 0, -2.5], np.float32)
    cube = make_cube(1.0, 0)
    cube.scale = np.array([2, 1, 1], np.float32)
    sphere = make_sphere(sphere_precision, 2.0, 2)
    sphere.translation = np.array([3, 0, -1], np.float32)
    tri = make_tri(1, 5)
    tri.translation = np.array([1, 0.5, 2.1], np.float32)
    cube2 = make_cube(0.1, 4)
    cube2.translation = s.point_source_lights[0].position.copy()
    cube3 = make_cube(0.1, 4)
    cube3.translation = s.point_source_lights[1].position.copy()
    tri2 = make_tri(1, 6)
    tri2.translation = np.array([1, 0.5, 2.0], np.float32)
    tri3 = make_tri(1, 7)
    tri3.translation = np.array([1, 0.5, 2.2], np.float32)
    s.nodes = [tore, cube, sphere, tri, cube2, cube3, tri2, tri3]
    s.root_nodes = list(range(len(s.nodes)))
    return s


def demo_viewport(width=800, height=600, layers=3):
    """The demo's viewport (src/test_1.cpp:355-363): full window, Phong +
    bilinear, `layers` transparency layers, null post shader, camera moved
    back and turned."""
    vp = viewport_t(0, 0, width, height, Surface(width, height), pixel_shader_light_and_texture(),
                    layers)
    vp.set_post_shader(post_shader_t())
    vp.m_camera.translate(1, 2, -5)
    vp.m_camera.rotate_y(-0.2)
    vp.m_camera.rotate_x(-0.3)
    return vp


def random_texture(rng, w, h):
    """w x h pixel_colors with every byte i.i.d. uniform in [0, 255]."""
    return texture_t(rng.integers(0, 1 << 32, size=(h, w), dtype=np.uint64).astype(np.uint32))


def set_fov(camera, aspect, fov_deg=60.0):
    """Horizontal FOV through the public projection matrix, keeping
    P11 = aspect * P00 (camera.cpp:21-24 convention)."""
    p00 = f32(1.0 / math.tan(math.radians(fov_deg) / 2.0))
    camera.m_projectionmatrix[0, 0] = p00
    camera.m_projectionmatrix[1, 1] = f32(f32(aspect) * p00)


def seeded_sun(rng):
    v = rng.standard_normal(3).astype(np.float32)
    v[1] = -abs(v[1]) - f32(0.5)   # mostly from above
    return normal_t(*v)


# ------------------------------------------------------------------- cube
def cube_scene(seed=0, width=640, height=480):
    """Workload 'cube': returns (scene, [viewport])."""
    rng = np.random.default_rng(seed)
    a, b, c = (float(f32(x)) for x in rng.uniform(0.0, 2.0 * math.pi, 3))
    s = scene_t()
    s.images = [random_texture(rng, 256, 256)]
    s.materials = [material_t(pixel_colors(255, 255, 255, 255), 1, 1, 0, False)]
    cube = make_cube(1.0, 0)
    R = IDENTITY.copy()
    R = rotate_x(R, a)
    R = rotate_y(R, b)
    R = rotate_z(R, c)
    cube.rotation = R
    s.nodes = [cube]
    s.root_nodes = [0]
    s.ambient_light_intensity = float(f32(0.1))
    s.sun_direction = seeded_sun(rng)
    s.sun_intensity = 1.0
    surface = Surface(width, height)
    vp = viewport_t(0, 0, width, height, surface, pixel_shader_light_and_texture(), 0)
    vp.m_camera.translate(0, 0, -3)
    set_fov(vp.m_camera, f32(1.0 * width / height))
    vp.set_post_shader(post_shader_t())
    return s, [vp]


# ------------------------------------------------------------ height field
def value_noise(n, rng, octaves=4, base=4):
    """Sum of `octaves` bilinear-smoothstep value-noise layers on an n x n
    grid over [0,1]^2, frequencies base*2^k, weights 2^-k, normalised."""
    g = np.arange(n, dtype=np.float64) / (n - 1)
    total = np.zeros((n, n), np.float64)
    wsum = 0.0
    for k in range(octaves):
        f = base << k
        lat = rng.uniform(-1.0, 1.0, (f + 1, f + 1))
        x = g * f
        i = np.minimum(np.floor(x).astype(np.int64), f - 1)
        t = x - i
        s = t * t * (3.0 - 2.0 * t)
        a = lat[i][:, i]
        b = lat[i + 1][:, i]
        c = lat[i][:, i + 1]
        d = lat[i + 1][:, i + 1]
        sx = s[:, None]
        sz = s[None, :]
        layer = (a * (1 - sx) + b * sx) * (1 - sz) + (c * (1 - sx) + d * sx) * sz
        w = 0.5 ** k
        total += w * layer
        wsum += w
    return total / wsum


def heightfield_mesh(n, rng, amplitude=0.2, material=0):
    """n x n vertex grid, x and z in [-1,1], y = amplitude * noise.
    Triangles (a, c, b) and (a, d, c) per cell, counter-clockwise from +y."""
    h = amplitude * value_noise(n, rng)
    coord = np.linspace(-1.0, 1.0, n)
    X, Z = np.meshgrid(coord, coord, indexing="ij")   # [i, j]: x = coord[i], z = coord[j]
    pos = np.stack([X, h, Z], -1).reshape(-1, 3).astype(np.float32)
    dx = np.gradient(h, coord, axis=0)
    dz = np.gradient(h, coord, axis=1)
    nor = np.stack([-dx, np.ones_like(h), -dz], -1)
    nor /= np.sqrt((nor * nor).sum(-1, keepdims=True))
    nor = nor.reshape(-1, 3).astype(np.float32)
    gi = np.arange(n, dtype=np.float64) / (n - 1)
    U, Vv = np.meshgrid(gi, gi, indexing="ij")
    uv = np.stack([U, Vv], -1).reshape(-1, 2).astype(np.float32)
    i = np.arange(n - 1)
    I, J = np.meshgrid(i, i, indexing="ij")
    a = (I * n + J).reshape(-1)
    b = ((I + 1) * n + J).reshape(-1)
    c = ((I + 1) * n + J + 1).reshape(-1)
    d = (I * n + J + 1).reshape(-1)
    tris = np.stack([np.stack([a, c, b], -1), np.stack([a, d, c], -1)], 1).reshape(-1)
    node = node_t([primitive_t(pos, nor, uv, tris.astype(np.uint32), A.TRIANGLES, material)])
    return node


def heightfield_scene(seed=0, n=708, width=1920, height=1080, tex=2048, n_lights=4, dof=True,
                      dof_radius=8, light_intensity=(0.5, 2.0)):
    """Workload 'heightfield' (n=708 gives 999,698 triangles)."""
    rng = np.random.default_rng(seed)
    s = scene_t()
    s.images = [random_texture(rng, tex, tex)]
    s.materials = [material_t(pixel_colors(255, 255, 255, 255), 1, 1, 0, False)]
    s.nodes = [heightfield_mesh(n, rng)]
    s.root_nodes = [0]
    s.ambient_light_intensity = float(f32(0.1))
    s.sun_direction = seeded_sun(rng)
    s.sun_intensity = float(f32(0.7))
    for _ in range(n_lights):
        p = rng.uniform(-2.0, 2.0, 3).astype(np.float32)
        I = f32(rng.uniform(*light_intensity))
        s.point_source_lights.append(point_source_light(p, I))
    surface = Surface(width, height)
    vp = viewport_t(0, 0, width, height, surface, pixel_shader_light_and_texture(), 0)
    cam = vp.m_camera
    cam.translate(0.0, 1.1, -1.9)
    cam.rotate_x(-0.55)
    set_fov(cam, f32(1.0 * width / height))
    if dof:
        vp.set_post_shader(post_shader_depth_box(5.0, 5.0, vp, max_radius=dof_radius, taps="tap9"))
    else:
        vp.set_post_shader(post_shader_t())
    return s, [vp]


def batch_scenes(seed=0, count=256, tmin=1_000, tmax=200_000,
                 sides=(256, 384, 512, 768, 1024)):
    """Workload 'batch': list of (scene, [viewport]) pairs."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        t = math.exp(rng.uniform(math.log(tmin), math.log(tmax)))
        n = int(round(math.sqrt(t / 2.0))) + 1
        side = int(sides[rng.integers(0, len(sides))])
        tex = 512 if rng.integers(0, 2) == 0 else 1024
        nl = int(rng.integers(0, 5))
        dof = (k % 2) == 0
        sub_seed = int(rng.integers(0, 1 << 31))
        s, vps = heightfield_scene(sub_seed, n=n, width=side, height=side, tex=tex, n_lights=nl,
                                   dof=dof)
        out.append((s, vps))
    return out


# ------------------------------------------------------------------ heavy
def _project(cam, p):
    """(x', y', z') of world points p (..., 3): view, then projection rows
    0-2 (vertex_shaders.hpp:35-71 before the divide by |z'|)."""
    V = cam.m_viewmatrix.astype(np.float64)
    P = cam.m_projectionmatrix.astype(np.float64)
    h = np.concatenate([p, np.ones(p.shape[:-1] + (1,))], -1)
    c = h @ V.T
    c[..., 3] = 1.0
    return (c @ P.T)[..., :3]


def cluster_mesh(rng, cam, width, n_clusters=500, quads=(25, 20), depth=(-2.0, 1.0),
                 edge_px=(0.5, 4.0), stack=(8, 20), tex_side=4096):
    """`n_clusters` patches of quads[0] x quads[1] cells (2 triangles each),
    facing the camera, piled in stacks of stack[0]..stack[1] patches at
    random depths over one screen point each, so covered pixels see that
    many layers.  Each patch's cell edge projects to edge_px[0]..edge_px[1]
    pixels; its texture coordinates cover about 1.5 texels per pixel of a
    tex_side^2 texture.  Returns (positions, normals, uvs, indices)."""
    qx, qy = quads
    nvx, nvy = qx + 1, qy + 1
    sizes = []
    while sum(sizes) < n_clusters:
        sizes.append(int(rng.integers(stack[0], stack[1] + 1)))
    sizes[-1] -= sum(sizes) - n_clusters
    centres = rng.uniform(-0.8, 0.8, (len(sizes), 2))
    pos, nor, uv, idx = [], [], [], []
    gx, gy = np.meshgrid(np.arange(nvx, dtype=np.float64) - qx / 2, np.arange(nvy, dtype=np.float64) - qy / 2,
                         indexing="ij")
    i = np.arange(qx)
    I, J = np.meshgrid(i, np.arange(qy), indexing="ij")
    a = (I * nvy + J).reshape(-1)
    b = ((I + 1) * nvy + J).reshape(-1)
    c = ((I + 1) * nvy + J + 1).reshape(-1)
    d = (I * nvy + J + 1).reshape(-1)
    cell_tris = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], 1).reshape(-1)
    base = 0
    for s_i, n_s in enumerate(sizes):
        for _ in range(n_s):
            zw = rng.uniform(*depth)
            sx, sy = centres[s_i] + rng.uniform(-0.01, 0.01, 2)
            # world x, y landing on (sx, sy): x', y' are affine in x, y at fixed depth
            p0 = _project(cam, np.array([0.0, 0.0, zw]))
            px = _project(cam, np.array([1.0, 0.0, zw]))
            py = _project(cam, np.array([0.0, 1.0, zw]))
            zabs = abs(p0[2])
            xw = (sx * zabs - p0[0]) / (px[0] - p0[0])
            yw = (sy * zabs - p0[1]) / (py[1] - p0[1])
            e = rng.uniform(*edge_px)
            cell = e * 2.0 * zabs / (width * (px[0] - p0[0]))   # world size of an e-pixel edge
            bump = rng.uniform(-0.3, 0.3, gx.shape) * cell
            P = np.stack([xw + gx * cell, yw + gy * cell, zw + bump], -1).reshape(-1, 3)
            tilt = rng.uniform(-0.4, 0.4, 2)
            N = np.stack([np.full(gx.shape, tilt[0]), np.full(gx.shape, tilt[1]), np.ones(gx.shape)], -1)
            N = N + rng.uniform(-0.15, 0.15, N.shape)
            N /= np.sqrt((N * N).sum(-1, keepdims=True))
            span = 1.5 * e * np.array([qx, qy]) / tex_side
            off = rng.uniform(0.0, 1.0, 2)
            T = np.stack([off[0] + (gx + qx / 2) / qx * span[0], off[1] + (gy + qy / 2) / qy * span[1]], -1)
            pos.append(P.reshape(-1, 3))
            nor.append(N.reshape(-1, 3))
            uv.append(T.reshape(-1, 2))
            idx.append(cell_tris + base)
            base += nvx * nvy
    return (np.concatenate(pos).astype(np.float32), np.concatenate(nor).astype(np.float32),
            np.concatenate(uv).astype(np.float32), np.concatenate(idx).astype(np.uint32))


def heavy_scene(seed=0, n_clusters=500, n_viewports=4, width=1920, height=1080, tex=4096, n_lights=8,
                dof=True, dof_radius=8):
    """Workload 'heavy' (one scene): n_clusters x 1000 triangles in stacked
    patches (500 clusters = 500,000 triangles), a tex^2 texture, sun +
    n_lights point lights, n_viewports viewports of width x height onto
    their own surfaces, from cameras a little apart, 9-tap depth of field."""
    rng = np.random.default_rng(seed)
    s = scene_t()
    s.images = [random_texture(rng, tex, tex)]
    s.materials = [material_t(pixel_colors(255, 255, 255, 255), 1, 1, 0, False)]
    vps = []
    for k in range(n_viewports):
        surface = Surface(width, height)
        vp = viewport_t(0, 0, width, height, surface, pixel_shader_light_and_texture(), 0)
        cam = vp.m_camera
        cam.translate(float(f32(0.12 * (k - (n_viewports - 1) / 2.0))), 0.0, -3.0)
        set_fov(cam, f32(1.0 * width / height))
        vps.append(vp)
    pos, nor, uv, idx = cluster_mesh(rng, vps[0].m_camera, width, n_clusters=n_clusters, tex_side=tex)
    s.nodes = [node_t([primitive_t(pos, nor, uv, idx, A.TRIANGLES, 0)])]
    s.root_nodes = [0]
    s.ambient_light_intensity = float(f32(0.1))
    s.sun_direction = seeded_sun(rng)
    s.sun_intensity = float(f32(0.5))
    for _ in range(n_lights):
        p = np.array([rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.5, 2.5)], np.float32)
        s.point_source_lights.append(point_source_light(p, f32(rng.uniform(0.5, 2.0))))
    # focus on the middle of the depth range (z' of a point at world z = -0.5)
    zmid = float(abs(_project(vps[0].m_camera, np.array([0.0, 0.0, -0.5]))[2]))
    for vp in vps:
        if dof:
            vp.set_post_shader(post_shader_depth_box(zmid, 2.0, vp, max_radius=dof_radius, taps="tap9"))
        else:
            vp.set_post_shader(post_shader_t())
    return s, vps


def heavy_scenes(seed=0, count=64, **kw):
    """Workload 'heavy': `count` heavy scenes (defaults: 500k triangles,
    4 x 1920x1080 viewports, 4096^2 texture, sun + 8 point lights, DoF)."""
    rng = np.random.default_rng(seed)
    return [heavy_scene(int(rng.integers(0, 1 << 31)), **kw) for _ in range(count)]
