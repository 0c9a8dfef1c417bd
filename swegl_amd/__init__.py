"""swegl_amd: swegl's render path on MI355X (gfx950).

The compute path is libswegl_hip.so (HIP kernels behind a C ABI,
include/swegl_hip.h); this package is its Python host mirror:
  swegl_amd.host    scene_t / camera_t / viewport_t / render(...)
  swegl_amd.device  library loading, contexts, device surfaces and scenes
  swegl_amd.scenes  seeded synthetic workloads
"""
from . import _abi  # noqa: F401
