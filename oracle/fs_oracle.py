"""ctypes wrapper of libfsoracle.so, the CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FSO_LIB selects another build of the same source (`make -C oracle asan`)
LIB_PATH = os.environ.get("FSO_LIB") or os.path.join(_HERE, "libfsoracle.so")


class Camera(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("height", ctypes.c_float),
                ("angle", ctypes.c_float), ("horizon", ctypes.c_float),
                ("distance", ctypes.c_float), ("fov", ctypes.c_float),
                ("sky_color", ctypes.c_uint32)]


class Depth(ctypes.Structure):
    _fields_ = [("d", ctypes.c_float), ("z0", ctypes.c_float), ("scale", ctypes.c_float),
                ("n_slices", ctypes.c_int32)]


P = ctypes.c_void_p
F = ctypes.c_float
I = ctypes.c_int32
U = ctypes.c_uint32
CP = ctypes.POINTER(Camera)
DP = ctypes.POINTER(Depth)

_SIG = {
    "fso_i32_f32": (I, [F]),
    "fso_u32_f32": (U, [F]),
    "fso_mod_i32": (I, [I, I]),
    "fso_to_rgba": (None, [U, P]),
    "fso_from_rgba": (U, [F, F, F, F]),
    "fso_mix": (U, [F, U, F, U]),
    "fso_scale": (U, [U, F]),
    "fso_black": (U, []),
    "fso_get_zs_count": (I, [F, F, F]),
    "fso_get_zs": (I, [F, F, F, P, I]),
    "fso_slices": (I, [CP, DP, P, I]),
    "fso_get_h_line": (None, [F, CP, I, P]),
    "fso_png_height": (F, [P, I, I, F, F]),
    "fso_png_height_filtered": (F, [P, I, I, F, F]),
    "fso_png_color": (U, [P, I, I, F, F]),
    "fso_png_color_filtered": (U, [P, I, I, F, F]),
    "fso_render_off": (ctypes.c_int, [CP, DP, P, I, I, P, I, I, I, I, P]),
    "fso_render_painter": (ctypes.c_int, [CP, DP, P, I, I, P, I, I, I, I, P, P]),
    "fso_count_visible": (ctypes.c_int64, [CP, DP, P, I, I, I, I]),
    "fso_render_painter_mt": (ctypes.c_int, [CP, DP, P, I, I, P, I, I, I, I, P, ctypes.c_int]),
    "fso_max_threads": (ctypes.c_int, []),
    "fso_vec3_rotate": (None, [ctypes.c_int, F, P, P]),
    "fso_sun": (None, [F, F, P]),
    "fso_shadowmap": (ctypes.c_int, [P, I, I, P, I, I, P, I, P]),
    "fso_shadow_counts": (ctypes.c_int, [P, I, I, P, I, P]),
    "fso_expo2": (F, [F, F]),
    "fso_function_coloring": (U, [F, F]),
    "fso_math_shadowmap": (ctypes.c_int, [P, P]),
    "fso_render_state": (ctypes.c_int, [CP, I, I, P, I, I, P, I, I, P, I, I, P]),
    "fso_sky": (U, [F]),
    "fso_render_nearest": (ctypes.c_int, [CP, P, P, I, I, I, I, I, P]),
    "fso_render_nearest_painter": (ctypes.c_int, [CP, P, P, I, I, I, I, I, P]),
    "fso_render_f32": (ctypes.c_int, [CP, DP, P, P, I, I, P, I, I, I, I, F, P, P]),
    "fso_render_f32_vjp": (ctypes.c_int, [CP, DP, P, P, I, I, I, I, P, P, P, P]),
    "fso_render2": (ctypes.c_int, [CP, P, P, I, I, I, I, P]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def camera(view):
    return Camera(view.x, view.y, view.height, view.angle, view.horizon, view.distance,
                  view.fov, view.sky_color & 0xFFFFFFFF)


def depth(view):
    return Depth(view.d, view.z0, view.scale, view.n_slices)


def render_off(view, shaded, alt):
    """The reference's scan/scatter/fill formulation (literal restatement)."""
    c, a = _u32(shaded), _i32(alt)
    out = np.empty((view.h, view.w), dtype=np.uint32)
    rc = lib().fso_render_off(ctypes.byref(camera(view)), ctypes.byref(depth(view)),
                              c.ctypes.data, c.shape[0], c.shape[1], a.ctypes.data,
                              a.shape[0], a.shape[1], view.h, view.w, out.ctypes.data)
    assert rc == 0
    return out


def render_painter(view, shaded, alt, with_src=False):
    c, a = _u32(shaded), _i32(alt)
    out = np.empty((view.h, view.w), dtype=np.uint32)
    src = np.empty((view.h, view.w), dtype=np.int32) if with_src else None
    rc = lib().fso_render_painter(ctypes.byref(camera(view)), ctypes.byref(depth(view)),
                                  c.ctypes.data, c.shape[0], c.shape[1], a.ctypes.data,
                                  a.shape[0], a.shape[1], view.h, view.w, out.ctypes.data,
                                  src.ctypes.data if with_src else None)
    assert rc == 0
    return (out, src) if with_src else out


def render_painter_mt(view, shaded, alt, threads=0):
    c, a = _u32(shaded), _i32(alt)
    out = np.empty((view.h, view.w), dtype=np.uint32)
    rc = lib().fso_render_painter_mt(ctypes.byref(camera(view)), ctypes.byref(depth(view)),
                                     c.ctypes.data, c.shape[0], c.shape[1], a.ctypes.data,
                                     a.shape[0], a.shape[1], view.h, view.w, out.ctypes.data,
                                     threads)
    assert rc == 0
    return out


def count_visible(view, alt):
    a = _i32(alt)
    return int(lib().fso_count_visible(ctypes.byref(camera(view)), ctypes.byref(depth(view)),
                                       a.ctypes.data, a.shape[0], a.shape[1], view.h, view.w))


def sun(sun_height, sun_ang):
    out = np.empty(3, dtype=np.float32)
    lib().fso_sun(sun_height, sun_ang, out.ctypes.data)
    return out


def vec3_rotate(axis, angle, v):
    v = _f32(v)
    out = np.empty(3, dtype=np.float32)
    lib().fso_vec3_rotate(axis, angle, v.ctypes.data, out.ctypes.data)
    return out


def shadowmap(color, alt, sun_vec, size=1024):
    c, a, s = _u32(color), _i32(alt), _f32(sun_vec)
    out = np.empty((size, size), dtype=np.uint32)
    rc = lib().fso_shadowmap(c.ctypes.data, c.shape[0], c.shape[1], a.ctypes.data, a.shape[0],
                             a.shape[1], s.ctypes.data, size, out.ctypes.data)
    assert rc == 0
    return out


def shadow_counts(alt, sun_vec, size=1024):
    """Per-texel shadowing step count of the shading pass (fut/effects.fut:116-122)."""
    a, s = _i32(alt), _f32(sun_vec)
    out = np.empty((size, size), dtype=np.int32)
    rc = lib().fso_shadow_counts(a.ctypes.data, a.shape[0], a.shape[1], s.ctypes.data, size,
                                 out.ctypes.data)
    assert rc == 0
    return out


def shade_factor(counts):
    """The relaxation's per-texel shade factor sqrt(1 / (4*count + 1)), in f32."""
    f = np.float32
    return np.sqrt(f(1.0) / (f(4.0) * np.asarray(counts).astype(np.float32) + f(1.0))).astype(
        np.float32)


def math_shadowmap(sun_vec):
    """The #math mode's shaded map (function_coloring over expo2), 1024^2."""
    s = _f32(sun_vec)
    out = np.empty((1024, 1024), dtype=np.uint32)
    assert lib().fso_math_shadowmap(s.ctypes.data, out.ctypes.data) == 0
    return out


def render2(view, color, alt):
    """render2 (fut/voxel_renderer.fut:257-305): the ray march over the
    altitude and unshaded colour maps; view supplies the camera and frame
    size."""
    c, a = _u32(color), _i32(alt)
    out = np.empty((view.h, view.w), dtype=np.uint32)
    rc = lib().fso_render2(ctypes.byref(camera(view)), c.ctypes.data, a.ctypes.data,
                           a.shape[0], a.shape[1], view.h, view.w, out.ctypes.data)
    assert rc == 0
    return out


def render_state(view, mode, smoothing_on, shaded, alt, sun_vec):
    """The state's render entry (fut/interactive.fut:173-183): mode "math"
    or "png", smoothing on/off, z_0 = 0 and d = 0.001; view supplies the
    camera and frame size (its depth fields are not used)."""
    c, a, s = _u32(shaded), _i32(alt), _f32(sun_vec)
    out = np.empty((view.h, view.w), dtype=np.uint32)
    rc = lib().fso_render_state(ctypes.byref(camera(view)), 1 if mode == "png" else 0,
                                1 if smoothing_on else 0, c.ctypes.data, c.shape[0], c.shape[1],
                                a.ctypes.data, a.shape[0], a.shape[1], s.ctypes.data,
                                view.h, view.w, out.ctypes.data)
    assert rc == 0
    return out


def update_map(color, height, sun_height=0.1, sun_ang=0.1, size=1024, bits=8):
    """update_map (fut/interactive.fut:188-198): mask heights, shade colours."""
    alt = (_i32(height) & ((1 << bits) - 1)).astype(np.int32)
    return shadowmap(color, alt, sun(sun_height, sun_ang), size), alt


def get_zs(c, d, z0):
    n = lib().fso_get_zs_count(c, d, z0)
    out = np.empty(max(n, 1), dtype=np.float32)
    lib().fso_get_zs(c, d, z0, out.ctypes.data, n)
    return out[:n]


def get_h_line(z, view, w):
    out = np.empty(4, dtype=np.float32)
    lib().fso_get_h_line(z, ctypes.byref(camera(view)), w, out.ctypes.data)
    return out


def png_height_filtered(alt, x, y):
    a = _i32(alt)
    return lib().fso_png_height_filtered(a.ctypes.data, a.shape[0], a.shape[1], x, y)


def png_color_filtered(col, x, y):
    c = _u32(col)
    return lib().fso_png_color_filtered(c.ctypes.data, c.shape[0], c.shape[1], x, y)


def render_nearest(cam_fields, color, alt, l, m, sky, painter=False):
    cam = Camera(*cam_fields, 0)
    c, a = _i32(color), _i32(alt)
    out = np.empty((l, m), dtype=np.int32)
    fn = lib().fso_render_nearest_painter if painter else lib().fso_render_nearest
    rc = fn(ctypes.byref(cam), c.ctypes.data, a.ctypes.data, c.shape[0], c.shape[1], l, m,
            sky, out.ctypes.data)
    assert rc == 0
    return out


def render_f32(view, planes, shade_factor, alt, sky_scale):
    cp, sf, a = _f32(planes), _f32(shade_factor), _i32(alt)
    img = np.empty((3, view.h, view.w), dtype=np.float32)
    src = np.empty((view.h, view.w), dtype=np.int32)
    rc = lib().fso_render_f32(ctypes.byref(camera(view)), ctypes.byref(depth(view)),
                              cp.ctypes.data, sf.ctypes.data, cp.shape[1], cp.shape[2],
                              a.ctypes.data, a.shape[0], a.shape[1], view.h, view.w,
                              sky_scale, img.ctypes.data, src.ctypes.data)
    assert rc == 0
    return img, src


def render_f32_vjp(view, planes, shade_factor, src, gimg):
    cp, sf = _f32(planes), _f32(shade_factor)
    s, g = _i32(src), _f32(gimg)
    gcp = np.empty_like(cp)
    gsky = ctypes.c_float()
    rc = lib().fso_render_f32_vjp(ctypes.byref(camera(view)), ctypes.byref(depth(view)),
                                  cp.ctypes.data, sf.ctypes.data, cp.shape[1], cp.shape[2],
                                  view.h, view.w, s.ctypes.data, g.ctypes.data, gcp.ctypes.data,
                                  ctypes.byref(gsky))
    assert rc == 0
    return gcp, gsky.value
