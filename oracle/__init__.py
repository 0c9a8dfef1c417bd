"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from swegl_amd import _abi as A
from swegl_amd.host import SceneDescHolder

_HERE = os.path.dirname(os.path.abspath(__file__))
# SWEGL_ORACLE_LIB points the binding at another build of refcpu.cpp: the
# planted-mistake test (tests/test_oracle_mutations.py) loads mutated copies
# through it to show that the pin tests reject each mistake.
LIB_PATH = os.environ.get("SWEGL_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None


class OracleViewport(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("w", C.c_int32), ("h", C.c_int32),
                ("surface", A.U32P), ("surface_w", C.c_int32), ("surface_h", C.c_int32),
                ("pitch", C.c_int32), ("zbuffer", A.FP), ("triangle_id", A.U32P),
                ("view", C.c_float * 16), ("projection", C.c_float * 16),
                ("camera_position", C.c_float * 3), ("post_shader", C.c_int32),
                ("focal_distance", C.c_float), ("focal_depth", C.c_float),
                ("dof_max_radius", C.c_int32), ("color64", A.U32P), ("pixel_shader", C.c_int32),
                ("transparency_layer_count", C.c_int32)]


class OracleStats(C.Structure):
    _fields_ = [("fragments", C.c_uint64), ("nan_fragments", C.c_uint64),
                ("drawn_subtriangles", C.c_uint64), ("texel_oob", C.c_uint64),
                ("clipped_triangles", C.c_uint64)]


def build():
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_render.restype = C.c_int
        L.oracle_render.argtypes = [C.POINTER(A.SceneDesc), C.POINTER(OracleViewport), C.c_uint32,
                                    C.POINTER(OracleStats)]
        L.oracle_transform.argtypes = [A.FP, A.FP, A.FP]
        L.oracle_vertex.argtypes = [A.FP] * 9
        L.oracle_clip_vertex.argtypes = [A.FP, A.FP, C.c_float, A.FP, A.FP, A.FP, A.FP, A.FP]
        L.oracle_interp.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, A.FP]
        L.oracle_raster_triangle.restype = C.c_int
        L.oracle_raster_triangle.argtypes = [A.FP, C.c_int, C.c_int, A.U32P, A.FP]
        L.oracle_bilinear.restype = C.c_uint32
        L.oracle_bilinear.argtypes = [A.U32P, C.c_int, C.c_int, C.c_float, C.c_float]
        L.oracle_phong.restype = C.c_int32
        L.oracle_phong.argtypes = [A.FP, A.FP, A.FP, C.c_float, A.FP, C.c_float,
                                   C.POINTER(A.PointLight), C.c_int]
        L.oracle_combine.restype = C.c_uint32
        L.oracle_combine.argtypes = [C.c_uint32, C.c_int32]
        L.oracle_dof.argtypes = [A.U32P, A.FP, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                 C.c_int, A.U32P]
        L.oracle_cvtt.restype = C.c_int32
        L.oracle_cvtt.argtypes = [C.c_float]
        _lib = L
    return _lib


class Frame:
    """Oracle output of one viewport: surface, zbuffer (h, w), triangle ids."""

    def __init__(self, pixels, zbuffer, tid, stats, color64=None):
        self.pixels, self.zbuffer, self.tid, self.stats = pixels, zbuffer, tid, stats
        self.color64 = color64


def _viewports(viewports, surfaces, reshade64):
    n = len(viewports)
    ovs = (OracleViewport * n)()
    keep = []
    shared = {}
    for i, vp in enumerate(viewports):
        s = vp.m_screen
        if surfaces is not None:
            pix = surfaces[i]
        else:
            key = id(s)
            if key not in shared:
                shared[key] = s.pixels.copy()
            pix = shared[key]
        z = np.zeros((vp.m_h, vp.m_w), np.float32)
        tid = np.zeros((vp.m_h, vp.m_w), np.uint32)
        c64 = np.zeros((vp.m_h, vp.m_w), np.uint32) if reshade64 else None
        keep += [pix, z, tid, c64]
        o = ovs[i]
        o.x, o.y, o.w, o.h = vp.m_x, vp.m_y, vp.m_w, vp.m_h
        o.surface = A.u32ptr(pix)
        o.surface_w, o.surface_h, o.pitch = s.w, s.h, s.pitch
        o.zbuffer = A.fptr(z)
        o.triangle_id = A.u32ptr(tid)
        o.color64 = A.u32ptr(c64) if reshade64 else None
        o.view[:] = [float(v) for v in vp.m_camera.m_viewmatrix.reshape(16)]
        o.projection[:] = [float(v) for v in vp.m_camera.m_projectionmatrix.reshape(16)]
        o.camera_position[:] = [float(v) for v in vp.m_camera.m_center]
        o.pixel_shader = vp.m_pixel_shader.kind()
        o.transparency_layer_count = vp.transparency_layer_count
        ps = vp.m_post_shader
        o.post_shader = ps.kind if ps is not None else -1
        if ps is not None and ps.kind != A.POST_NULL:
            o.focal_distance, o.focal_depth = ps.focal_distance, ps.focal_depth
            o.dof_max_radius = ps.max_radius
    return ovs, keep


def render(scene, viewports, surfaces=None, reshade64=False):
    """swegl::render(scene, vp...) on the CPU oracle.  Each viewport renders
    into a copy of its surface's current pixels (or `surfaces[i]`).  With
    `reshade64` each frame also carries `color64` (h, w): the FP64 re-shade
    of every pixel's final winner, before the post shader."""
    holder = SceneDescHolder(scene)
    n = len(viewports)
    ovs, keep = _viewports(viewports, surfaces, reshade64)
    st = (OracleStats * n)()
    rc = lib().oracle_render(C.byref(holder.desc), ovs, n, st)
    if rc != 0:
        raise RuntimeError(f"oracle_render failed: {A.ERRNAMES.get(rc, rc)}")
    frames = []
    for i in range(n):
        d = {k: getattr(st[i], k) for k, _ in OracleStats._fields_}
        frames.append(Frame(keep[4 * i], keep[4 * i + 1], keep[4 * i + 2], d, keep[4 * i + 3]))
    return frames


def vertex_state(scene, viewport):
    """The scene fields render() writes (model.hpp:20-29) after rendering
    `viewport`: v_world, normal_world, v_viewport (n, 3) and yes (n,), the
    vertices in node, primitive, vertex order."""
    holder = SceneDescHolder(scene)
    ovs, keep = _viewports([viewport], None, False)
    n = scene.n_vertices()
    out = {k: np.zeros((max(n, 1), 3), np.float32) for k in ("v_world", "normal_world", "v_viewport")}
    yes = np.zeros(max(n, 1), np.uint8)
    L = lib()
    L.oracle_vertex_state.argtypes = [C.c_void_p, C.c_void_p, A.FP, A.FP, A.FP, C.POINTER(C.c_uint8)]
    rc = L.oracle_vertex_state(C.byref(holder.desc), ovs, A.fptr(out["v_world"]), A.fptr(out["normal_world"]),
                               A.fptr(out["v_viewport"]), yes.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc != 0:
        raise RuntimeError(f"oracle_vertex_state failed: {A.ERRNAMES.get(rc, rc)}")
    out = {k: v[:n] for k, v in out.items()}
    out["yes"] = yes[:n]
    return out


def reshade_error(color, tid, color64):
    """Max and mean absolute difference (8-bit LSB, over the B, G and R
    channels of every pixel some triangle drew) between FP32-shaded colours
    `color` (h, w) and the FP64 re-shade `color64` of the same winners."""
    drawn = tid != 0xFFFFFFFF
    n = int(drawn.sum())
    if n == 0:
        return {"max": 0, "mean": 0.0, "pixels": 0}
    a = color[drawn].view(np.uint8).reshape(-1, 4)[:, :3].astype(np.int16)
    b = color64[drawn].view(np.uint8).reshape(-1, 4)[:, :3].astype(np.int16)
    d = np.abs(a - b)
    return {"max": int(d.max()), "mean": float(d.mean()), "pixels": n,
            "frac_exact": float((d.max(axis=1) == 0).mean())}
