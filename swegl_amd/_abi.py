"""ctypes mirror of include/swegl_hip.h (the C ABI of libswegl_hip.so).

Only data layouts and constants live here; loading the library is done by
:mod:`swegl_amd.device`, which raises when the HIP library is missing.
"""
import ctypes as C

OK = 0
EIO, EAGAIN, ENOMEM, ENODEV, EINVAL, ERANGE, ENOTSUP = -5, -11, -12, -19, -22, -34, -95
ERRNAMES = {0: "OK", -5: "EIO", -11: "EAGAIN", -12: "ENOMEM", -19: "ENODEV",
            -22: "EINVAL", -34: "ERANGE", -95: "ENOTSUP"}

# primitive_t::index_mode_t (swegl/data/model.hpp:33-42); points and lines draw nothing
POINTS, LINES, LINE_LOOP, LINE_STRIP = 0, 1, 2, 3
TRIANGLES, TRIANGLE_STRIP, TRIANGLE_FAN = 4, 5, 6

# animation_channel_t::path_t (swegl/data/model.hpp:97-105)
PATH_SCALE, PATH_ROTATION, PATH_TRANSLATION, PATH_NONE = 0, 1, 2, 3

SHADER_PHONG_BILINEAR = 1
SHADER_PHONG_NEAREST = 2
SHADER_FLAT_BILINEAR = 3
SHADER_FLAT_NEAREST = 4
SHADER_DEFAULT = 5
SHADER_TEXTURE_NEAREST = 6
SHADER_TEXTURE_BILINEAR = 7
SHADER_LIGHTS_FLAT = 8
SHADER_LIGHTS_PHONG = 9
MAX_TRANSPARENCY_LAYERS = 4
POST_NULL, POST_DEPTH_BOX, POST_DEPTH_TAP9 = 0, 1, 2

TARGET_COLOR, TARGET_FLOAT, TARGET_UINT = 0, 1, 2

FP = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)


class PrimitiveDesc(C.Structure):
    _fields_ = [("n_vertices", C.c_uint32), ("positions", FP), ("normals", FP),
                ("tex_coords", FP), ("n_indices", C.c_uint32), ("indices", U32P),
                ("mode", C.c_int32), ("material_id", C.c_int32)]


class NodeDesc(C.Structure):
    _fields_ = [("scale", C.c_float * 3), ("rotation", C.c_float * 16),
                ("translation", C.c_float * 3), ("first_primitive", C.c_uint32),
                ("n_primitives", C.c_uint32), ("first_child", C.c_uint32),
                ("n_children", C.c_uint32)]


class MaterialDesc(C.Structure):
    _fields_ = [("color", C.c_uint32), ("metallic", C.c_float), ("roughness", C.c_float),
                ("texture_idx", C.c_int32), ("double_sided", C.c_int32)]


class TextureDesc(C.Structure):
    _fields_ = [("bitmap", U32P), ("width", C.c_uint32), ("height", C.c_uint32)]


class PointLight(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("intensity", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("nodes", C.POINTER(NodeDesc)),
                ("n_root_nodes", C.c_uint32), ("root_nodes", I32P),
                ("n_child_indices", C.c_uint32), ("child_indices", I32P),
                ("n_primitives", C.c_uint32), ("primitives", C.POINTER(PrimitiveDesc)),
                ("n_materials", C.c_uint32), ("materials", C.POINTER(MaterialDesc)),
                ("default_material", MaterialDesc),
                ("n_textures", C.c_uint32), ("textures", C.POINTER(TextureDesc)),
                ("n_point_lights", C.c_uint32), ("point_lights", C.POINTER(PointLight)),
                ("ambient_light_intensity", C.c_float), ("sun_direction", C.c_float * 3),
                ("sun_intensity", C.c_float)]


class ViewportDesc(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("w", C.c_int32), ("h", C.c_int32),
                ("color", C.c_void_p), ("depth", C.c_void_p), ("triangle_id", C.c_void_p),
                ("view", C.c_float * 16), ("projection", C.c_float * 16),
                ("camera_position", C.c_float * 3), ("pixel_shader", C.c_int32),
                ("post_shader", C.c_int32), ("focal_distance", C.c_float),
                ("focal_depth", C.c_float), ("dof_max_radius", C.c_int32),
                ("transparency_layer_count", C.c_int32)]


class AnimStep(C.Structure):
    _fields_ = [("time", C.c_float), ("value", C.c_float * 4)]


class AnimChannel(C.Structure):
    _fields_ = [("node", C.c_int32), ("path", C.c_int32), ("n_steps", C.c_uint32),
                ("steps", C.POINTER(AnimStep))]


class Animation(C.Structure):
    _fields_ = [("end_time", C.c_float), ("n_channels", C.c_uint32),
                ("channels", C.POINTER(AnimChannel))]


class FrameStats(C.Structure):
    _fields_ = [("fragments", C.c_uint64), ("nan_fragments", C.c_uint64),
                ("drawn_subtriangles", C.c_uint64), ("bin_entries", C.c_uint64),
                ("texel_oob", C.c_uint64), ("big_triangles", C.c_uint64)]


class Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("viewports", C.POINTER(ViewportDesc)),
                ("n_viewports", C.c_uint32)]


def fptr(a):
    return a.ctypes.data_as(FP)


def u32ptr(a):
    return a.ctypes.data_as(U32P)


def i32ptr(a):
    return a.ctypes.data_as(I32P)


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch", C.c_uint32), ("lanes", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("bytes_in_flight", C.c_uint64), ("peak_bytes_in_flight", C.c_uint64),
                ("largest_batch", C.c_uint32)]
