"""glTF 2.0 / GLB scenes and texture files (swegl::load_scene,
src/data/gltf.cpp:420-464; read_image_file, src/misc/image.cpp:20-28).

The file is read by the library's loader (swegl_asset_load), which follows
the reference's reading of a file, and copied into the host mirror
(:class:`swegl_amd.host.scene_t` with its animations) so that it can be
edited, posed with ``scene.animate(t)`` and rendered like any other scene.
"""
import ctypes as C
import os

import numpy as np

from . import _abi as A
from .device import check, lib
from .host import (animation_channel_t, animation_t, material_t, node_t, primitive_t, scene_t,
                   texture_t)


def _array(ptr, n, dtype, shape):
    if n == 0:
        return np.zeros(shape, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True).reshape(shape)


def load_scene(filename):
    """swegl::load_scene(filename): a .glb or .gltf file as a scene_t
    (lights, ambient and sun left for the caller to set)."""
    h = C.c_void_p()
    check(lib().swegl_asset_load(os.fsencode(filename), C.byref(h)), f"load_scene({filename})")
    try:
        d = A.SceneDesc()
        check(lib().swegl_asset_scene(h, C.byref(d)), "swegl_asset_scene")
        s = scene_t()
        s.images = []
        for i in range(d.n_textures):
            t = d.textures[i]
            s.images.append(texture_t(_array(t.bitmap, t.width * t.height, np.uint32,
                                             (t.height, t.width))))
        s.materials = [material_t(m.color, m.metallic, m.roughness, m.texture_idx,
                                  bool(m.double_sided))
                       for m in (d.materials[i] for i in range(d.n_materials))]
        prims = []
        for i in range(d.n_primitives):
            p = d.primitives[i]
            nv = p.n_vertices
            prims.append(primitive_t(_array(p.positions, 3 * nv, np.float32, (nv, 3)),
                                     _array(p.normals, 3 * nv, np.float32, (nv, 3)),
                                     _array(p.tex_coords, 2 * nv, np.float32, (nv, 2)),
                                     _array(p.indices, p.n_indices, np.uint32, (p.n_indices,)),
                                     p.mode, p.material_id))
        children = _array(d.child_indices, d.n_child_indices, np.int32, (d.n_child_indices,))
        s.nodes = []
        for i in range(d.n_nodes):
            nd = d.nodes[i]
            n = node_t(prims[nd.first_primitive:nd.first_primitive + nd.n_primitives])
            n.scale = np.array(nd.scale[:], np.float32)
            n.rotation = np.array(nd.rotation[:], np.float32).reshape(4, 4)
            n.translation = np.array(nd.translation[:], np.float32)
            n.children_idx = [int(c) for c in children[nd.first_child:nd.first_child + nd.n_children]]
            s.nodes.append(n)
        s.root_nodes = [int(r) for r in _array(d.root_nodes, d.n_root_nodes, np.int32,
                                                (d.n_root_nodes,))]
        pa = C.POINTER(A.Animation)()
        na = C.c_uint32()
        check(lib().swegl_asset_animations(h, C.byref(pa), C.byref(na)), "swegl_asset_animations")
        for a in range(na.value):
            an = pa[a]
            chans = []
            for k in range(an.n_channels):
                ch = an.channels[k]
                steps = [ch.steps[i] for i in range(ch.n_steps)]
                chans.append(animation_channel_t(ch.node, ch.path, [st.time for st in steps],
                                                 [st.value[:] for st in steps]))
            s.animations.append(animation_t(chans, an.end_time))
        return s
    finally:
        lib().swegl_asset_destroy(h)


def read_image_file(filename):
    """read_image_file(filename) for .png and .bmp files: a texture_t."""
    px = A.U32P()
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().swegl_image_read(os.fsencode(filename), C.byref(px), C.byref(w), C.byref(h)),
          f"read_image_file({filename})")
    try:
        return texture_t(_array(px, w.value * h.value, np.uint32, (h.value, w.value)))
    finally:
        lib().swegl_image_free(px)
