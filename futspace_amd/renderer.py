"""Python face of libfutspace: a context, resident maps, views and renders.

Mirrors the reference's interface for the render path: the state-machine
entries of fut/interactive_entrypoints.fut through `Session`, and the
explicit-camera extension entries through `Context.render`.  Device
buffers are torch tensors (memory and streams only); every frame is
produced by the HIP kernels behind the C-ABI.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native
from ._native import FsxView


class FutspaceError(RuntimeError):
    pass


@dataclass
class View:
    """One camera (fut/voxel_renderer.fut:5-12) plus frame and depth params.

    d, z0 and scale are the reference's invz_param1, invz_param2 and
    magic_number (fut/voxel_renderer.fut:103-104, :217); scale <= 0 means
    f32(w/2) and n_slices <= 0 derives the slice count with get_zs.
    """

    x: float
    y: float
    height: float
    angle: float
    horizon: float
    distance: float = 800.0
    fov: float = 1.2
    sky_color: int = 0
    h: int = 1024
    w: int = 1024
    d: float = 0.001
    z0: float = 0.0
    scale: float = 0.0
    n_slices: int = 0
    map: int = 0

    def to_c(self) -> FsxView:
        return FsxView(self.x, self.y, self.height, self.angle, self.horizon, self.distance,
                       self.fov, self.sky_color & 0xFFFFFFFF, self.h, self.w, self.d, self.z0,
                       self.scale, self.n_slices, self.map, 0)


# `main` of tests/futspace.fut:127-145 (its depth fields are fixed by the
# program: z_0 = 1, delta = 0.005, scale 240)
STATELESS_MAIN = View(x=512.0, y=800.0, height=78.0, angle=0.0, horizon=100.0, distance=800.0,
                      fov=1.0, sky_color=0xFF9090E0, h=400, w=800)


def sky_color(sun_height: float) -> int:
    """argb.scale 0xFF9090e0 sun_height, as `step` sets it (fut/interactive.fut:163)."""
    return int(_native.futspace().fsx_sky(ctypes.c_float(sun_height)))


class Context:
    """A futhark_context on one HIP device."""

    def __init__(self, device=None):
        self.lib = _native.futspace()
        cfg = self.lib.futhark_context_config_new()
        if device is not None:
            self.lib.futhark_context_config_set_device(cfg, f"#{int(device)}".encode())
        self.ptr = self.lib.futhark_context_new(cfg)
        self.lib.futhark_context_config_free(cfg)
        if not self.ptr:
            raise FutspaceError("futhark_context_new failed (no usable HIP device?)")

    def error(self) -> str:
        p = self.lib.futhark_context_get_error(self.ptr)
        if not p:
            return ""
        msg = ctypes.cast(p, ctypes.c_char_p).value.decode(errors="replace")
        _native.libc().free(p)
        return msg

    def check(self, rc, what):
        if rc != 0:
            raise FutspaceError(f"{what} failed: {self.error()}")

    def sync(self):
        self.check(self.lib.futhark_context_sync(self.ptr), "futhark_context_sync")

    def set_stream(self, stream):
        """Run later device work on `stream` (a torch.cuda.Stream or raw handle)."""
        handle = getattr(stream, "cuda_stream", stream)
        self.check(self.lib.fsx_context_set_stream(self.ptr, ctypes.c_void_p(handle or 0)),
                   "fsx_context_set_stream")

    def close(self):
        if self.ptr:
            self.lib.futhark_context_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- extension entries -------------------------------------------------

    def map(self, color, height, height_bits=8, shade_size=0, sun_height=0.1, sun_ang=0.1):
        return Map(self, color, height, height_bits, shade_size, sun_height, sun_ang)

    def render(self, m, view: View, out=None):
        """Render `view` of map `m`.

        out=None returns a host numpy uint32 [h][w] frame; a torch CUDA
        int32/uint32 tensor of shape [h][w] (or [h][>=w]) is written in
        place, asynchronously on the context stream.
        """
        cv = view.to_c()
        if out is None:
            frame = np.empty((view.h, view.w), dtype=np.uint32)
            self.check(self.lib.fsx_render_host(self.ptr, m.ptr, ctypes.byref(cv),
                                                frame.ctypes.data), "fsx_render_host")
            return frame
        if out.dim() != 2 or out.shape[0] != view.h or out.shape[1] < view.w:
            raise ValueError(f"output {tuple(out.shape)} does not hold a {view.h}x{view.w} frame")
        pitch = out.stride(0)
        if out.stride(1) != 1:
            raise ValueError("output rows must be contiguous")
        self.check(self.lib.fsx_render(self.ptr, m.ptr, ctypes.byref(cv),
                                       ctypes.c_void_p(out.data_ptr()), pitch), "fsx_render")
        return out

    def render2(self, m, view: View, out=None):
        """render2 (fsx_render2): the reference's ray-march renderer over
        the map's heights and unshaded colours.  out=None returns a host
        numpy uint32 [h][w] frame; a torch CUDA int32/uint32 [h][>=w] tensor
        is written in place, asynchronously on the context stream."""
        import torch
        own = out is None
        if own:
            out = torch.empty((view.h, view.w), dtype=torch.int32, device="cuda")
        if out.dim() != 2 or out.shape[0] != view.h or out.shape[1] < view.w or out.stride(1) != 1:
            raise ValueError(f"output {tuple(out.shape)} does not hold a {view.h}x{view.w} frame")
        cv = view.to_c()
        self.check(self.lib.fsx_render2(self.ptr, m.ptr, ctypes.byref(cv),
                                        ctypes.c_void_p(out.data_ptr()), out.stride(0)),
                   "fsx_render2")
        if own:
            self.sync()
            return out.cpu().numpy().view(np.uint32)
        return out

    def render_batch(self, maps, views, out, offsets):
        """One batched launch: views[i] into out.flat[offsets[i]:...]."""
        n = len(views)
        cviews = (FsxView * n)(*[v.to_c() for v in views])
        cmaps = (ctypes.c_void_p * len(maps))(*[m.ptr for m in maps])
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        self.check(self.lib.fsx_render_batch(self.ptr, cmaps, len(maps), cviews, n,
                                             ctypes.c_void_p(out.data_ptr()), offs.ctypes.data),
                   "fsx_render_batch")
        return out

    def batch(self, maps, views, out, offsets):
        """A prepared batch that can be run, captured into a HIP graph and
        replayed (fsx_batch_*)."""
        return Batch(self, maps, views, out, offsets)

    def render_stateless(self, color, height, view: View = None):
        """The stateless program of tests/futspace.fut (fsx_render_stateless):
        i32 colour and height maps [q][r] (numpy or CUDA tensors); view=None
        is the program's `main` (camera (512, 800, 78, 0, 100, 800, fov 1),
        sky 0xFF9090e0, 400 x 800).  Returns the [h][w] int32 frame."""
        import torch
        if view is None:
            view = STATELESS_MAIN
        c = torch.as_tensor(np.ascontiguousarray(color, dtype=np.int32)).cuda() \
            if not isinstance(color, torch.Tensor) else color.contiguous()
        h = torch.as_tensor(np.ascontiguousarray(height, dtype=np.int32)).cuda() \
            if not isinstance(height, torch.Tensor) else height.contiguous()
        if c.shape != h.shape or c.dim() != 2:
            raise ValueError("colour and height maps must be [q][r] arrays of one shape")
        out = torch.empty((view.h, view.w), dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        cv = view.to_c()
        self.check(self.lib.fsx_render_stateless(
            self.ptr, ctypes.c_void_p(c.data_ptr()), ctypes.c_void_p(h.data_ptr()),
            c.shape[0], c.shape[1], ctypes.byref(cv), ctypes.c_void_p(out.data_ptr())),
            "fsx_render_stateless")
        self.sync()
        return out.cpu().numpy()

    def set_tuning(self, tile_columns=0, wg_per_cu=0, chunk=0):
        """Render kernel shape of this context (0 = automatic); see
        fsx_set_tuning."""
        self.check(self.lib.fsx_set_tuning(self.ptr, tile_columns, wg_per_cu, chunk),
                   "fsx_set_tuning")

    def tuning(self):
        """(tile_columns, wg_per_cu, chunk) of this context."""
        out = (ctypes.c_int32 * 3)()
        self.check(self.lib.fsx_get_tuning(self.ptr, out), "fsx_get_tuning")
        return tuple(out)

    def profile(self, enable=True):
        self.check(self.lib.fsx_profile_enable(self.ptr, 1 if enable else 0),
                   "fsx_profile_enable")

    def profile_read(self):
        """(summed kernel milliseconds, launches) since the last read."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        self.check(self.lib.fsx_profile_read(self.ptr, ctypes.byref(ms), ctypes.byref(n)),
                   "fsx_profile_read")
        return ms.value, n.value


class Batch:
    """fsx_batch: many views over resident maps in one persistent launch."""

    def __init__(self, ctx, maps, views, out, offsets):
        self.ctx = ctx
        self.maps = list(maps)           # keep the maps alive with the batch
        self.out = out
        n = len(views)
        cviews = (FsxView * n)(*[v.to_c() for v in views])
        cmaps = (ctypes.c_void_p * len(maps))(*[m.ptr for m in maps])
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        p = ctypes.c_void_p()
        ctx.check(ctx.lib.fsx_batch_new(ctx.ptr, cmaps, len(maps), cviews, n,
                                        ctypes.c_void_p(out.data_ptr()), offs.ctypes.data,
                                        ctypes.byref(p)), "fsx_batch_new")
        self.ptr = p.value

    def run(self):
        self.ctx.check(self.ctx.lib.fsx_batch_run(self.ctx.ptr, self.ptr), "fsx_batch_run")

    def capture(self):
        self.ctx.check(self.ctx.lib.fsx_batch_capture(self.ctx.ptr, self.ptr),
                       "fsx_batch_capture")

    def replay(self):
        self.ctx.check(self.ctx.lib.fsx_batch_replay(self.ctx.ptr, self.ptr), "fsx_batch_replay")

    def close(self):
        if self.ptr and self.ctx.ptr:
            self.ctx.lib.fsx_batch_free(self.ctx.ptr, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Map:
    """A resident map pair with its shaded colour map (extension entry)."""

    def __init__(self, ctx, color, height, height_bits=8, shade_size=0, sun_height=0.1,
                 sun_ang=0.1):
        color = np.ascontiguousarray(color, dtype=np.uint32)
        height = np.ascontiguousarray(height, dtype=np.int32)
        if color.shape != height.shape or color.ndim != 2:
            raise ValueError("colour and height maps must have the same 2-D shape")
        self.ctx = ctx
        self.shape = color.shape
        self.height_bits = height_bits
        ptr = ctypes.c_void_p()
        ctx.check(ctx.lib.fsx_map_new(ctx.ptr, ctypes.byref(ptr), color.ctypes.data,
                                      height.ctypes.data, color.shape[0], color.shape[1],
                                      height_bits, shade_size, sun_height, sun_ang),
                  "fsx_map_new")
        self.ptr = ptr.value
        self.shade_size = ctx.lib.fsx_map_shade_size(self.ptr)

    def shaded(self):
        out = np.empty((self.shade_size, self.shade_size), dtype=np.uint32)
        self.ctx.check(self.ctx.lib.fsx_map_shaded(self.ctx.ptr, self.ptr, out.ctypes.data),
                       "fsx_map_shaded")
        return out

    def close(self):
        if self.ptr and self.ctx.ptr:
            self.ctx.lib.fsx_map_free(self.ctx.ptr, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SDLK = dict({chr(c): c for c in range(ord("a"), ord("z") + 1)},
            **{str(d): ord(str(d)) for d in range(10)},
            UP=1073741906, LEFT=1073741904, DOWN=1073741905, RIGHT=1073741903)


@dataclass
class Session:
    """The reference's interactive state through the generated entries.

    Follows the call pattern of c/interactive.c and the lys loop: init,
    update_map, key, step, render, text_content.
    """

    ctx: Context
    seed: int = 0
    state: int = field(default=None, init=False)

    def __post_init__(self):
        s = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.futhark_entry_init(self.ctx.ptr, ctypes.byref(s), self.seed),
                       "futhark_entry_init")
        self.state = s.value

    def _replace(self, fn, *args, what):
        s = ctypes.c_void_p()
        self.ctx.check(fn(self.ctx.ptr, ctypes.byref(s), *args, self.state), what)
        self.ctx.lib.futhark_free_opaque_state(self.ctx.ptr, self.state)
        self.state = s.value

    def update_map(self, color, height):
        lib, ctx = self.ctx.lib, self.ctx.ptr
        color = np.ascontiguousarray(color, dtype=np.uint32)
        height = np.ascontiguousarray(height, dtype=np.int32)
        c = lib.futhark_new_u32_2d(ctx, color.ctypes.data, color.shape[0], color.shape[1])
        h = lib.futhark_new_i32_2d(ctx, height.ctypes.data, height.shape[0], height.shape[1])
        if not c or not h:
            raise FutspaceError(f"futhark_new_*_2d failed: {self.ctx.error()}")
        try:
            self._replace(lib.futhark_entry_update_map, ctypes.c_void_p(c), ctypes.c_void_p(h),
                          what="futhark_entry_update_map")
        finally:
            lib.futhark_free_u32_2d(ctx, c)
            lib.futhark_free_i32_2d(ctx, h)

    def key(self, down: bool, key):
        code = SDLK[key] if isinstance(key, str) else int(key)
        self._replace(self.ctx.lib.futhark_entry_key, 0 if down else 1, code,
                      what="futhark_entry_key")

    def step(self, dt=0.016):
        self._replace(self.ctx.lib.futhark_entry_step, ctypes.c_float(dt), what="futhark_entry_step")

    def resize(self, h, w):
        self._replace(self.ctx.lib.futhark_entry_resize, h, w, what="futhark_entry_resize")

    def mouse(self, buttons, x, y):
        self._replace(self.ctx.lib.futhark_entry_mouse, buttons, x, y, what="futhark_entry_mouse")

    def wheel(self, dx, dy):
        self._replace(self.ctx.lib.futhark_entry_wheel, dx, dy, what="futhark_entry_wheel")

    def render(self):
        lib, ctx = self.ctx.lib, self.ctx.ptr
        arr = ctypes.c_void_p()
        self.ctx.check(lib.futhark_entry_render(ctx, ctypes.byref(arr), self.state),
                       "futhark_entry_render")
        try:
            shape = lib.futhark_shape_u32_2d(ctx, arr)
            h, w = int(shape[0]), int(shape[1])
            out = np.empty((h, w), dtype=np.uint32)
            self.ctx.check(lib.futhark_values_u32_2d(ctx, arr, out.ctypes.data),
                           "futhark_values_u32_2d")
        finally:
            lib.futhark_free_u32_2d(ctx, arr)
        return out

    def text_content(self):
        vals = [ctypes.c_float() for _ in range(9)]
        self.ctx.check(self.ctx.lib.futhark_entry_text_content(
            self.ctx.ptr, *[ctypes.byref(v) for v in vals], self.state), "text_content")
        return tuple(v.value for v in vals)

    def close(self):
        if self.state and self.ctx.ptr:
            self.ctx.lib.futhark_free_opaque_state(self.ctx.ptr, self.state)
        self.state = None
