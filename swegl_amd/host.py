"""Python mirror of swegl's scene / camera / viewport / render API.

Names and argument meanings follow the reference:
  scene_t, node_t, primitive_t, material_t, point_source_light  (swegl/data/model.hpp)
  texture_t                                                    (swegl/data/texture.hpp)
  camera_t                                                     (swegl/projection/camera.hpp)
  viewport_t                                                   (swegl/render/viewport.hpp)
  pixel_shader_light_and_texture<lights_phong, texture_bilinear> (swegl/render/pixel_shaders.hpp)
  post_shader_t, post_shader_depth_box                         (swegl/render/post_shaders.hpp)
  render(scene, viewports...)                                  (swegl/render/renderer.hpp:27-34)

The device work goes through the C ABI (libswegl_hip.so); there is no CPU
fallback.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from .device import Context, DeviceScene, Target, lib, check

IDENTITY = np.eye(4, dtype=np.float32)


def pixel_colors(b, g, r, a):
    """swegl::pixel_colors{b,g,r,a} as its 32-bit word (colors.hpp:9-33)."""
    return (int(b) & 255) | ((int(g) & 255) << 8) | ((int(r) & 255) << 16) | ((int(a) & 255) << 24)


# ---------------------------------------------------------------- data ---
class primitive_t:
    def __init__(self, positions, normals, tex_coords, indices, mode=A.TRIANGLES,
                 material_id=0):
        self.positions = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        self.normals = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        self.tex_coords = np.ascontiguousarray(tex_coords, np.float32).reshape(-1, 2)
        self.indices = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        self.mode = int(mode)
        self.material_id = int(material_id)
        assert len(self.positions) == len(self.normals) == len(self.tex_coords)

    def n_triangles(self):
        n = len(self.indices)
        if self.mode == A.TRIANGLES:
            return n // 3
        if self.mode in (A.TRIANGLE_STRIP, A.TRIANGLE_FAN):
            return max(0, n - 2)
        return 0   # points and lines are not drawn


class node_t:
    def __init__(self, primitives=None):
        self.scale = np.ones(3, np.float32)
        self.rotation = IDENTITY.copy()
        self.translation = np.zeros(3, np.float32)
        self.primitives = list(primitives or [])
        self.children_idx = []


class material_t:
    def __init__(self, color=pixel_colors(255, 255, 255, 255), metallic=1.0, roughness=1.0,
                 texture_idx=-1, double_sided=False):
        self.color, self.metallic, self.roughness = int(color), metallic, roughness
        self.texture_idx, self.double_sided = int(texture_idx), bool(double_sided)


class texture_t:
    """Mip 0 only (texture.hpp:29-38); bitmap is (height, width) uint32."""

    def __init__(self, bitmap):
        self.bitmap = np.ascontiguousarray(bitmap, np.uint32)
        assert self.bitmap.ndim == 2

    @property
    def width(self):
        return self.bitmap.shape[1]

    @property
    def height(self):
        return self.bitmap.shape[0]


class point_source_light:
    def __init__(self, position, intensity):
        self.position = np.asarray(position, np.float32)
        self.intensity = float(np.float32(intensity))


class animation_channel_t:
    """animation_channel_t (model.hpp:95-122): node, path (A.PATH_*) and the
    keyframes sorted by time: times (n,) and values (n, 4) float32."""

    def __init__(self, node_idx, path, times, values):
        self.node_idx, self.path = int(node_idx), int(path)
        self.times = np.ascontiguousarray(times, np.float32).reshape(-1)
        self.values = np.ascontiguousarray(values, np.float32).reshape(-1, 4)
        assert len(self.times) == len(self.values) and len(self.times) > 0


class animation_t:
    """animation_t (model.hpp:123-127)."""

    def __init__(self, channels=None, end_time=None):
        self.channels = list(channels or [])
        self.end_time = float(np.float32(end_time if end_time is not None else
                                         max([c.times[-1] for c in self.channels], default=0.0)))


def from_quaternion(q0, q1, q2, q3):
    """matrix44_t::from_quaternion (swegl/projection/matrix44.hpp:34-43); q0
    is the scalar part."""
    m = np.zeros((4, 4), np.float32)
    lib().swegl_matrix_from_quaternion(A.fptr(m), q0, q1, q2, q3)
    return m


class scene_t:
    def __init__(self):
        self.nodes = []
        self.root_nodes = []
        self.ambient_light_intensity = 0.0
        self.sun_direction = np.array([0, -1, 0], np.float32)
        self.sun_intensity = 0.0
        self.default_material = material_t()
        self.point_source_lights = []
        self.materials = []
        self.images = []
        self.animations = []
        self.generation = 0   # bump after editing meshes/textures in place

    def animate(self, elapsed_seconds):
        """scene_t::animate (model.hpp:146-177): pose the animated nodes for
        elapsed_seconds (each animation loops over its end_time)."""
        if not self.animations:
            return
        nodes = (A.NodeDesc * max(1, len(self.nodes)))()
        for i, nd in enumerate(self.nodes):
            nodes[i].scale[:] = [float(x) for x in np.asarray(nd.scale, np.float32)]
            nodes[i].rotation[:] = [float(x) for x in np.asarray(nd.rotation, np.float32).reshape(16)]
            nodes[i].translation[:] = [float(x) for x in np.asarray(nd.translation, np.float32)]
        keep = []
        anims = (A.Animation * len(self.animations))()
        for a, an in enumerate(self.animations):
            chans = (A.AnimChannel * max(1, len(an.channels)))()
            for k, ch in enumerate(an.channels):
                steps = (A.AnimStep * len(ch.times))()
                for i in range(len(ch.times)):
                    steps[i].time = float(ch.times[i])
                    steps[i].value[:] = [float(v) for v in ch.values[i]]
                chans[k] = A.AnimChannel(ch.node_idx, ch.path, len(ch.times), steps)
                keep.append(steps)
            keep.append(chans)
            anims[a] = A.Animation(an.end_time, len(an.channels), chans)
        check(lib().swegl_animate(anims, len(self.animations), float(elapsed_seconds), nodes,
                                  len(self.nodes)), "swegl_animate")
        for i, nd in enumerate(self.nodes):
            nd.scale = np.array(nodes[i].scale[:], np.float32)
            nd.rotation = np.array(nodes[i].rotation[:], np.float32).reshape(4, 4)
            nd.translation = np.array(nodes[i].translation[:], np.float32)

    def n_triangles(self):
        return sum(p.n_triangles() for n in self.nodes for p in n.primitives)

    def n_vertices(self):
        return sum(len(p.positions) for n in self.nodes for p in n.primitives)


def _mat(m):
    return (C.c_float * 16)(*np.asarray(m, np.float32).reshape(16).tolist())


def _material_desc(m):
    return A.MaterialDesc(m.color, m.metallic, m.roughness, m.texture_idx, int(m.double_sided))


class SceneDescHolder:
    """Builds a swegl_scene_desc and keeps every array it points to alive."""

    def __init__(self, scene):
        keep = []
        prims = []
        nodes = (A.NodeDesc * max(1, len(scene.nodes)))()
        children = []
        for i, nd in enumerate(scene.nodes):
            d = nodes[i]
            d.scale[:] = [float(x) for x in np.asarray(nd.scale, np.float32)]
            d.rotation[:] = [float(x) for x in np.asarray(nd.rotation, np.float32).reshape(16)]
            d.translation[:] = [float(x) for x in np.asarray(nd.translation, np.float32)]
            d.first_primitive, d.n_primitives = len(prims), len(nd.primitives)
            d.first_child, d.n_children = len(children), len(nd.children_idx)
            children.extend(nd.children_idx)
            for p in nd.primitives:
                keep += [p.positions, p.normals, p.tex_coords, p.indices]
                prims.append(A.PrimitiveDesc(len(p.positions), A.fptr(p.positions),
                                             A.fptr(p.normals), A.fptr(p.tex_coords),
                                             len(p.indices), A.u32ptr(p.indices), p.mode,
                                             p.material_id))
        prim_arr = (A.PrimitiveDesc * max(1, len(prims)))(*prims)
        roots = np.asarray(scene.root_nodes, np.int32)
        child_arr = np.asarray(children if children else [0], np.int32)
        mats = (A.MaterialDesc * max(1, len(scene.materials)))(
            *[_material_desc(m) for m in scene.materials])
        texs = (A.TextureDesc * max(1, len(scene.images)))()
        for i, t in enumerate(scene.images):
            keep.append(t.bitmap)
            texs[i] = A.TextureDesc(A.u32ptr(t.bitmap), t.width, t.height)
        lights = (A.PointLight * max(1, len(scene.point_source_lights)))()
        for i, l in enumerate(scene.point_source_lights):
            lights[i].position[:] = [float(x) for x in l.position]
            lights[i].intensity = l.intensity
        keep += [nodes, prim_arr, roots, child_arr, mats, texs, lights]
        d = A.SceneDesc()
        d.n_nodes, d.nodes = len(scene.nodes), nodes
        d.n_root_nodes, d.root_nodes = len(roots), A.i32ptr(roots)
        d.n_child_indices, d.child_indices = len(children), A.i32ptr(child_arr)
        d.n_primitives, d.primitives = len(prims), prim_arr
        d.n_materials, d.materials = len(scene.materials), mats
        d.default_material = _material_desc(scene.default_material)
        d.n_textures, d.textures = len(scene.images), texs
        d.n_point_lights, d.point_lights = len(scene.point_source_lights), lights
        d.ambient_light_intensity = scene.ambient_light_intensity
        d.sun_direction[:] = [float(x) for x in np.asarray(scene.sun_direction, np.float32)]
        d.sun_intensity = scene.sun_intensity
        self.desc = d
        self._keep = keep


# ---------------------------------------------------------- projection ---
def rotate_x(m, a):
    """matrix44_t::rotate_x in place (src/projection/matrix44.cpp:12-29)."""
    m = np.ascontiguousarray(m, np.float32)
    lib().swegl_matrix_rotate_x(A.fptr(m), a)
    return m


def rotate_y(m, a):
    m = np.ascontiguousarray(m, np.float32)
    lib().swegl_matrix_rotate_y(A.fptr(m), a)
    return m


def rotate_z(m, a):
    m = np.ascontiguousarray(m, np.float32)
    lib().swegl_matrix_rotate_z(A.fptr(m), a)
    return m


class camera_t:
    """camera_t (src/projection/camera.cpp)."""

    def __init__(self, aspect_ratio):
        self.m_viewmatrix = np.zeros((4, 4), np.float32)
        self.m_projectionmatrix = np.zeros((4, 4), np.float32)
        self.m_center = np.zeros(3, np.float32)
        lib().swegl_camera_init(float(aspect_ratio), A.fptr(self.m_viewmatrix),
                                A.fptr(self.m_projectionmatrix), A.fptr(self.m_center))

    def rotate_x(self, a):
        lib().swegl_camera_rotate(A.fptr(self.m_viewmatrix), 0, a)

    def rotate_y(self, a):
        lib().swegl_camera_rotate(A.fptr(self.m_viewmatrix), 1, a)

    def rotate_z(self, a):
        lib().swegl_camera_rotate(A.fptr(self.m_viewmatrix), 2, a)

    def translate(self, x, y, z):
        lib().swegl_camera_translate(A.fptr(self.m_viewmatrix), A.fptr(self.m_center), x, y, z)

    def position(self):
        return self.m_center.copy()


# ------------------------------------------------------------- shaders ---
# The pixel shader plugins of swegl/render/pixel_shaders.hpp.  The GPU cannot
# run virtual calls, so each type names the kernel variant that computes it.
class pixel_shader_t:
    """pixel_shader_t: shades with the material colour (pixel_shaders.hpp:15-29)."""

    def kind(self):
        return A.SHADER_DEFAULT


class pixel_shader_lights_flat(pixel_shader_t):
    """One light value per triangle at its centroid (pixel_shaders.cpp:33-84)."""

    def kind(self):
        return A.SHADER_LIGHTS_FLAT


class pixel_shader_lights_phong(pixel_shader_t):
    """Per-pixel Phong (pixel_shaders.cpp:88-205)."""

    def kind(self):
        return A.SHADER_LIGHTS_PHONG


class pixel_shader_texture(pixel_shader_t):
    """Nearest texel (pixel_shaders.cpp:209-281)."""

    def kind(self):
        return A.SHADER_TEXTURE_NEAREST


class pixel_shader_texture_bilinear(pixel_shader_t):
    """Bilinear texel (pixel_shaders.cpp:284-384)."""

    def kind(self):
        return A.SHADER_TEXTURE_BILINEAR


_COMBINED = {("flat", "nearest"): A.SHADER_FLAT_NEAREST, ("flat", "bilinear"): A.SHADER_FLAT_BILINEAR,
             ("phong", "nearest"): A.SHADER_PHONG_NEAREST, ("phong", "bilinear"): A.SHADER_PHONG_BILINEAR}
_LIGHT = {pixel_shader_lights_flat: "flat", pixel_shader_lights_phong: "phong"}
_TEX = {pixel_shader_texture: "nearest", pixel_shader_texture_bilinear: "bilinear"}


class pixel_shader_light_and_texture(pixel_shader_t):
    """pixel_shader_light_and_texture<L, T> (pixel_shaders.hpp:121-179)."""

    def __init__(self, L=pixel_shader_lights_phong, T=pixel_shader_texture_bilinear):
        if L not in _LIGHT or T not in _TEX:
            raise TypeError("pixel_shader_light_and_texture takes a light shader type and a texture shader type")
        self.L, self.T = L, T

    def kind(self):
        return _COMBINED[(_LIGHT[self.L], _TEX[self.T])]


class post_shader_t:
    """No-op without transparency (post_shaders.hpp:15-48)."""
    kind = A.POST_NULL


class post_shader_depth_box(post_shader_t):
    """Depth of field (post_shaders.hpp:51-131); max_radius 5 in the reference.
    taps='tap9' selects the 3x3 {-r,0,r} tap set."""

    def __init__(self, dist, depth, vp=None, max_radius=5, taps="box"):
        self.focal_distance = float(dist)
        self.focal_depth = float(depth)
        self.max_radius = int(max_radius)
        self.kind = A.POST_DEPTH_TAP9 if taps == "tap9" else A.POST_DEPTH_BOX


class Surface:
    """SDL_Surface-compatible host surface: w, h, pitch, pixels (BGRA words)."""

    BytesPerPixel = 4

    def __init__(self, w, h):
        self.w, self.h, self.pitch = w, h, 4 * w
        self.pixels = np.zeros((h, w), np.uint32)


class viewport_t:
    """viewport_t(x, y, w, h, screen, pixel_shader, transparency_layer_count)
    (viewport.hpp:40-44, viewport.cpp:14-40)."""

    def __init__(self, x, y, w, h, screen, pixel_shader, transparency_layer_count=0):
        self.m_x, self.m_y, self.m_w, self.m_h = x, y, w, h
        self.m_screen = screen
        self.m_zbuffer = np.zeros(w * h, np.float32)
        self.m_camera = camera_t(np.float32(1.0 * w / h))
        self.m_pixel_shader = pixel_shader
        self.m_post_shader = None
        self.m_got_transparency = transparency_layer_count > 0
        self.transparency_layer_count = transparency_layer_count

    def set_post_shader(self, ps):
        self.m_post_shader = ps

    def camera(self):
        return self.m_camera

    def desc(self, color, depth, triangle_id=None):
        d = A.ViewportDesc()
        d.x, d.y, d.w, d.h = self.m_x, self.m_y, self.m_w, self.m_h
        d.color = color.h if color is not None else None
        d.depth = depth.h if depth is not None else None
        d.triangle_id = triangle_id.h if triangle_id is not None else None
        d.view[:] = [float(v) for v in self.m_camera.m_viewmatrix.reshape(16)]
        d.projection[:] = [float(v) for v in self.m_camera.m_projectionmatrix.reshape(16)]
        d.camera_position[:] = [float(v) for v in self.m_camera.m_center]
        d.pixel_shader = self.m_pixel_shader.kind()
        ps = self.m_post_shader
        if ps is None:
            d.post_shader = -1   # the reference dereferences a null post shader: rejected
        else:
            d.post_shader = ps.kind
            if ps.kind != A.POST_NULL:
                d.focal_distance, d.focal_depth = ps.focal_distance, ps.focal_depth
                d.dof_max_radius = ps.max_radius
        d.transparency_layer_count = self.transparency_layer_count
        return d


# -------------------------------------------------------------- render ---
_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class _SurfaceTargets:
    pass


def render(scene, *viewports, ctx=None, stats=None, read_back=True):
    """swegl::render(scene, vp...) (renderer.hpp:27-34) on the GPU.

    Meshes and textures are uploaded on first use (and again whenever
    scene.generation changes); poses and lights are refreshed every call.
    With read_back, every viewport's surface pixels and m_zbuffer are copied
    back to the host as the reference leaves them there.
    """
    ctx = ctx or default_context()
    holder = SceneDescHolder(scene)
    ds = getattr(scene, "_device", None)
    if ds is None or getattr(scene, "_device_gen", None) != scene.generation or ds.ctx is not ctx:
        ds = DeviceScene(ctx, holder)
        scene._device, scene._device_gen = ds, scene.generation
    else:
        ds.update_pose(holder)
    descs = (A.ViewportDesc * len(viewports))()
    surfaces = []
    for i, vp in enumerate(viewports):
        s = vp.m_screen
        t = getattr(s, "_device", None)
        if t is None or t.ctx is not ctx:
            t = Target(ctx, A.TARGET_COLOR, s.w, s.h)
            t.write(s.pixels)
            s._device = t
        z = getattr(vp, "_device_z", None)
        if z is None or z.ctx is not ctx:
            z = Target(ctx, A.TARGET_FLOAT, vp.m_w, vp.m_h)
            vp._device_z = z
        descs[i] = vp.desc(t, z)
        surfaces.append((vp, s, t, z))
    st = (A.FrameStats * 1)()
    check(lib().swegl_render(ctx.h, ds.h, descs, len(viewports), st), "swegl_render")
    if read_back:
        for vp, s, t, z in surfaces:
            s.pixels[...] = t.read()
            vp.m_zbuffer[...] = z.read().reshape(-1)
    else:
        ctx.sync()
    if stats is not None:
        stats.update({k: getattr(st[0], k) for k, _ in A.FrameStats._fields_})
    return ctx
