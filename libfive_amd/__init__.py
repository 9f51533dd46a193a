"""MI355X tape evaluation and dual-contouring meshing, drop-in for libfive's
C API on that path (see include/libfive.h)."""
from .shape import Shape, GpuError, X, Y, Z, autotune_report, render_mesh_batch  # noqa: F401
from . import ffi  # noqa: F401


def gpu_available():
    return bool(ffi.lib().libfive_gpu_available())
