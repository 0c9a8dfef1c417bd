"""Loading of libswegl_hip.so and thin handle wrappers over its C ABI.

The product path has no fallback: if the HIP library is missing or was not
built, :func:`lib` raises.
"""
import ctypes as C
import os

from . import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libswegl_hip.so")
_lib = None


class SweglError(RuntimeError):
    def __init__(self, rc, what):
        self.rc = rc
        msg = _last_error()
        super().__init__(f"{what} failed: {A.ERRNAMES.get(rc, rc)}" + (f" ({msg})" if msg else ""))


def lib():
    """Return the loaded C library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "swegl_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
            "swegl_ctx_destroy": (C.c_int, [vp]),
            "swegl_ctx_synchronize": (C.c_int, [vp]),
            "swegl_last_error": (C.c_char_p, []),
            "swegl_device_info": (C.c_int, [vp, C.POINTER(C.c_int), C.c_char_p, C.c_int,
                                             C.POINTER(C.c_uint64)]),
            "swegl_ctx_set_tile": (C.c_int, [vp, C.c_int, C.c_int]),
            "swegl_ctx_get_tile": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "swegl_ctx_set_profiling": (C.c_int, [vp, C.c_int]),
            "swegl_ctx_set_bin_limit": (C.c_int, [vp, C.c_uint32]),
            "swegl_ctx_set_raster_groups": (C.c_int, [vp, C.c_uint32]),
            "swegl_ctx_set_tile_auto": (C.c_int, [vp, C.c_int]),
            "swegl_ctx_kernel_times": (C.c_int, [vp, C.c_char_p, C.c_int, C.POINTER(C.c_double),
                                                 C.POINTER(C.c_uint64), C.c_int,
                                                 C.POINTER(C.c_int)]),
            "swegl_target_create": (C.c_int, [vp, C.c_int, C.c_int32, C.c_int32, C.POINTER(vp)]),
            "swegl_target_destroy": (C.c_int, [vp, vp]),
            "swegl_target_read": (C.c_int, [vp, vp, vp, C.c_int32]),
            "swegl_target_write": (C.c_int, [vp, vp, vp, C.c_int32]),
            "swegl_target_fill": (C.c_int, [vp, vp, C.c_uint32]),
            "swegl_target_copy": (C.c_int, [vp, vp, vp]),
            "swegl_scene_upload": (C.c_int, [vp, C.POINTER(A.SceneDesc), C.POINTER(vp)]),
            "swegl_scene_update_pose": (C.c_int, [vp, vp, C.POINTER(A.SceneDesc)]),
            "swegl_scene_destroy": (C.c_int, [vp, vp]),
            "swegl_scene_bytes": (C.c_int, [vp, vp, C.POINTER(C.c_uint64)]),
            "swegl_render": (C.c_int, [vp, vp, C.POINTER(A.ViewportDesc), C.c_uint32,
                                       C.POINTER(A.FrameStats)]),
            "swegl_render_graph": (C.c_int, [vp, vp, C.POINTER(A.ViewportDesc), C.c_uint32,
                                             C.POINTER(A.FrameStats)]),
            "swegl_render_noalloc": (C.c_int, [vp, vp, C.POINTER(A.ViewportDesc), C.c_uint32,
                                               C.POINTER(A.FrameStats)]),
            "swegl_render_batch": (C.c_int, [vp, C.POINTER(A.Job), C.c_uint32,
                                             C.POINTER(A.FrameStats)]),
            "swegl_render_batch_graph": (C.c_int, [vp, C.POINTER(A.Job), C.c_uint32,
                                                   C.POINTER(A.FrameStats)]),
            "swegl_server_create": (C.c_int, [vp, C.POINTER(A.ServerConfig), C.POINTER(vp)]),
            "swegl_server_destroy": (C.c_int, [vp]),
            "swegl_server_submit": (C.c_int, [vp, C.POINTER(A.Job), C.POINTER(C.c_uint64)]),
            "swegl_server_poll": (C.c_int, [vp, C.POINTER(C.c_uint64), C.c_uint32,
                                            C.POINTER(C.c_uint32)]),
            "swegl_server_drain": (C.c_int, [vp]),
            "swegl_server_get_stats": (C.c_int, [vp, C.POINTER(A.ServerStats)]),
            "swegl_server_run_trace": (C.c_int, [vp, C.POINTER(A.Job), C.POINTER(C.c_double),
                                                 C.c_uint32, C.POINTER(C.c_double)]),
            "swegl_matrix_rotate_x": (None, [A.FP, C.c_float]),
            "swegl_matrix_rotate_y": (None, [A.FP, C.c_float]),
            "swegl_matrix_rotate_z": (None, [A.FP, C.c_float]),
            "swegl_matrix_translate": (None, [A.FP, C.c_float, C.c_float, C.c_float]),
            "swegl_camera_init": (None, [C.c_float, A.FP, A.FP, A.FP]),
            "swegl_camera_translate": (None, [A.FP, A.FP, C.c_float, C.c_float, C.c_float]),
            "swegl_camera_rotate": (None, [A.FP, C.c_int, C.c_float]),
            "swegl_matrix_from_quaternion": (None, [A.FP, C.c_float, C.c_float, C.c_float,
                                                    C.c_float]),
            "swegl_asset_load": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
            "swegl_asset_scene": (C.c_int, [vp, C.POINTER(A.SceneDesc)]),
            "swegl_asset_animations": (C.c_int, [vp, C.POINTER(C.POINTER(A.Animation)),
                                                 C.POINTER(C.c_uint32)]),
            "swegl_asset_destroy": (None, [vp]),
            "swegl_animate": (C.c_int, [C.POINTER(A.Animation), C.c_uint32, C.c_float,
                                        C.POINTER(A.NodeDesc), C.c_uint32]),
            "swegl_image_read": (C.c_int, [C.c_char_p, C.POINTER(A.U32P), C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint32)]),
            "swegl_image_free": (None, [A.U32P]),
            "swegl_debug_vertex_state": (C.c_int, [vp, vp, C.POINTER(A.ViewportDesc), A.FP, A.FP, A.FP,
                                                   C.POINTER(C.c_uint8)]),
            "swegl_make_tore": (C.c_int, [C.c_uint32, A.FP, A.FP, A.FP, A.U32P]),
            "swegl_make_sphere": (C.c_int, [C.c_uint32, C.c_float, A.FP, A.FP, A.FP, A.U32P]),
            "swegl_calculate_face_normals": (C.c_int, [C.c_int32, A.FP, A.FP, C.c_uint32, A.U32P,
                                                       C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _last_error():
    try:
        s = lib().swegl_last_error()
        return s.decode() if s else ""
    except Exception:
        return ""


def check(rc, what):
    if rc != A.OK:
        raise SweglError(rc, what)
    return rc


class Context:
    """swegl_ctx: one HIP stream on one device."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().swegl_ctx_create(device, C.byref(h)), "swegl_ctx_create")
        self.h = h

    def close(self):
        if self.h:
            lib().swegl_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().swegl_ctx_synchronize(self.h), "swegl_ctx_synchronize")

    def info(self):
        n = C.c_int()
        arch = C.create_string_buffer(128)
        mem = C.c_uint64()
        check(lib().swegl_device_info(self.h, C.byref(n), arch, 128, C.byref(mem)),
              "swegl_device_info")
        return {"cus": n.value, "arch": arch.value.decode(), "hbm_bytes": mem.value}

    def set_tile(self, w, h):
        check(lib().swegl_ctx_set_tile(self.h, w, h), "swegl_ctx_set_tile")

    def get_tile(self):
        w, h = C.c_int(), C.c_int()
        check(lib().swegl_ctx_get_tile(self.h, C.byref(w), C.byref(h)), "swegl_ctx_get_tile")
        return w.value, h.value

    def set_bin_limit(self, tiles):
        check(lib().swegl_ctx_set_bin_limit(self.h, int(tiles)), "swegl_ctx_set_bin_limit")

    def set_raster_groups(self, per_cu):
        check(lib().swegl_ctx_set_raster_groups(self.h, int(per_cu)), "swegl_ctx_set_raster_groups")

    def set_tile_auto(self, on):
        check(lib().swegl_ctx_set_tile_auto(self.h, int(on)), "swegl_ctx_set_tile_auto")

    def apply_tuning(self, t):
        """Apply one entry of swegl_amd/tuning.json: {"tile": [w, h],
        "bin_limit": n, "raster_groups": g}."""
        self.set_tile(*t["tile"])
        self.set_bin_limit(t["bin_limit"])
        self.set_raster_groups(t["raster_groups"])

    def set_profiling(self, on):
        check(lib().swegl_ctx_set_profiling(self.h, int(on)), "swegl_ctx_set_profiling")

    def kernel_times(self):
        cap = 64
        names = C.create_string_buffer(8192)
        ms = (C.c_double * cap)()
        counts = (C.c_uint64 * cap)()
        n = C.c_int()
        check(lib().swegl_ctx_kernel_times(self.h, names, 8192, ms, counts, cap, C.byref(n)),
              "swegl_ctx_kernel_times")
        nm = names.raw.split(b"\0")
        return {nm[i].decode(): (ms[i], counts[i]) for i in range(n.value)}


class Target:
    """A device surface: colour (BGRA words), float plane or uint plane."""

    def __init__(self, ctx, kind, width, height):
        h = C.c_void_p()
        check(lib().swegl_target_create(ctx.h, kind, width, height, C.byref(h)),
              "swegl_target_create")
        self.ctx, self.h, self.kind, self.width, self.height = ctx, h, kind, width, height

    def close(self):
        if self.h and self.ctx.h:
            lib().swegl_target_destroy(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def read(self):
        import numpy as np
        dt = {A.TARGET_COLOR: np.uint32, A.TARGET_FLOAT: np.float32, A.TARGET_UINT: np.uint32}
        out = np.empty((self.height, self.width), dt[self.kind])
        check(lib().swegl_target_read(self.ctx.h, self.h, out.ctypes.data_as(C.c_void_p),
                                      self.width * 4), "swegl_target_read")
        return out

    def write(self, arr):
        import numpy as np
        a = np.ascontiguousarray(arr)
        assert a.shape == (self.height, self.width) and a.itemsize == 4
        check(lib().swegl_target_write(self.ctx.h, self.h, a.ctypes.data_as(C.c_void_p),
                                       self.width * 4), "swegl_target_write")

    def fill(self, word):
        check(lib().swegl_target_fill(self.ctx.h, self.h, word), "swegl_target_fill")


class DeviceScene:
    """A scene resident in HBM (swegl_scene_upload)."""

    def __init__(self, ctx, desc_holder):
        h = C.c_void_p()
        check(lib().swegl_scene_upload(ctx.h, C.byref(desc_holder.desc), C.byref(h)),
              "swegl_scene_upload")
        self.ctx, self.h = ctx, h

    def update_pose(self, desc_holder):
        check(lib().swegl_scene_update_pose(self.ctx.h, self.h, C.byref(desc_holder.desc)),
              "swegl_scene_update_pose")

    def nbytes(self):
        b = C.c_uint64()
        check(lib().swegl_scene_bytes(self.ctx.h, self.h, C.byref(b)), "swegl_scene_bytes")
        return b.value

    def close(self):
        if self.h and self.ctx.h:
            lib().swegl_scene_destroy(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """swegl_server: continuous batching of render requests over lanes."""

    def __init__(self, ctx, budget_bytes, max_batch=256, lanes=2):
        h = C.c_void_p()
        cfg = A.ServerConfig(int(budget_bytes), int(max_batch), int(lanes))
        check(lib().swegl_server_create(ctx.h, C.byref(cfg), C.byref(h)), "swegl_server_create")
        self.ctx, self.h = ctx, h

    def submit(self, job):
        t = C.c_uint64()
        check(lib().swegl_server_submit(self.h, C.byref(job), C.byref(t)), "swegl_server_submit")
        return t.value

    def poll(self, cap=1024):
        buf = (C.c_uint64 * cap)()
        n = C.c_uint32()
        check(lib().swegl_server_poll(self.h, buf, cap, C.byref(n)), "swegl_server_poll")
        return list(buf[:n.value])

    def drain(self):
        check(lib().swegl_server_drain(self.h), "swegl_server_drain")

    def stats(self):
        st = A.ServerStats()
        check(lib().swegl_server_get_stats(self.h, C.byref(st)), "swegl_server_get_stats")
        return {k: getattr(st, k) for k, _ in A.ServerStats._fields_}

    def run_trace(self, jobs, arrival_s):
        import numpy as np
        n = len(arrival_s)
        arr = np.ascontiguousarray(arrival_s, np.float64)
        lat = np.zeros(n, np.float64)
        check(lib().swegl_server_run_trace(self.h, jobs, arr.ctypes.data_as(C.POINTER(C.c_double)), n,
                                           lat.ctypes.data_as(C.POINTER(C.c_double))),
              "swegl_server_run_trace")
        return lat

    def close(self):
        if self.h:
            lib().swegl_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
