"""CPU tests of the product library that need no GPU: every symbol declared
in include/swegl_hip.h is exported, host matrix/camera helpers follow the
reference closed forms, input validation errors come back as codes, and
graceful failure without a device."""
import ctypes as C
import math
import os
import re

import numpy as np

from swegl_amd import _abi as A
from swegl_amd.device import LIB_PATH, lib
from swegl_amd.host import camera_t, rotate_x, rotate_y, rotate_z

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if hdr.endswith(".h"):
            text = open(os.path.join(ROOT, "include", hdr)).read()
            names |= set(re.findall(r"\b(swegl_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(LIB_PATH)
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_rotations_closed_form():
    """matrix44_t::rotate_x/y/z (matrix44.cpp:12-67) on the identity."""
    for fn, (i, j) in ((rotate_x, (1, 2)), (rotate_y, (0, 2)), (rotate_z, (0, 1))):
        for a in (0.0, 0.3, -1.2, math.pi / 2):
            m = fn(np.eye(4, dtype=np.float32), a)
            c, s = np.float32(math.cos(a)), np.float32(math.sin(a))
            assert abs(m[i, i] - c) <= 1e-7 and abs(m[j, j] - c) <= 1e-7
            assert abs(m[i, j] - s) <= 1e-7 and abs(m[j, i] + s) <= 1e-7
            assert np.linalg.det(m.astype(np.float64)) == pytest_approx(1.0)
    # rotation composes: x(a) then x(b) == x(a+b) to rounding
    m = rotate_x(rotate_x(np.eye(4, dtype=np.float32), 0.4), 0.5)
    assert np.allclose(m, rotate_x(np.eye(4, dtype=np.float32), 0.9), atol=1e-6)


def pytest_approx(v, tol=1e-5):
    class _A:
        def __eq__(self, o):
            return abs(o - v) <= tol
    return _A()


def test_camera_constructor_and_translate():
    """camera_t(aspect) (camera.cpp:8-35) and translate (camera.cpp:50-58)."""
    cam = camera_t(np.float32(640 / 480))
    P = cam.m_projectionmatrix
    assert P[0, 0] == 1.0 and P[1, 1] == np.float32(640 / 480)
    assert P[2, 2] == np.float32(10 / 9.5) and P[2, 3] == np.float32(-5 / 9.5) and P[3, 2] == 1
    assert cam.m_viewmatrix[2, 2] == -1
    cam2 = camera_t(0.5)
    assert cam2.m_projectionmatrix[0, 0] == 2.0 and cam2.m_projectionmatrix[1, 1] == 1.0
    cam.translate(1, 2, -5)
    assert np.array_equal(cam.m_viewmatrix[:3, 3], np.array([-1, -2, 5], np.float32))
    assert np.array_equal(cam.position(), np.array([1, 2, 5], np.float32))   # local z flips
    cam.rotate_y(0.3)   # camera rotate(a) applies rotate(-a) to the view
    want = rotate_y(np.eye(4, dtype=np.float32), -0.3)
    assert abs(cam.m_viewmatrix[0, 0] - want[0, 0]) < 1e-7


def test_context_without_device_fails_cleanly():
    """On a machine without a gfx950 device the context call returns an
    error code (no crash); on the GPU box it succeeds."""
    h = C.c_void_p()
    rc = lib().swegl_ctx_create(0, C.byref(h))
    if rc == 0:
        lib().swegl_ctx_destroy(h)
    else:
        assert rc in (A.ENODEV, A.EIO)
        assert lib().swegl_last_error()


def test_null_arguments_are_rejected():
    L = lib()
    assert L.swegl_render(None, None, None, 0, None) == A.EINVAL
    assert L.swegl_ctx_synchronize(None) == A.EINVAL
    assert L.swegl_scene_upload(None, None, None) == A.EINVAL
