"""futspace_amd: MI355X (gfx950) Voxel Space renderer behind the futspace C-ABI.

The renderer lives in libfutspace.so (C host + HIP kernels); this package
is a thin ctypes harness over it plus the seeded scene generator.
"""
from .renderer import Context, FutspaceError, Map, Session, View, sky_color  # noqa: F401
from ._native import LIB_PATH, NativeLibraryMissing  # noqa: F401

__all__ = ["Context", "Map", "View", "Session", "sky_color", "FutspaceError",
           "NativeLibraryMissing", "LIB_PATH"]
