"""ctypes declarations of libfutspace.so and libfsscene.so.

The HIP library is the product: nothing here computes a frame.  If the
library is not built the import of the renderer fails loudly instead of
falling back to any CPU path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfutspace.so")
SCENE_PATH = os.path.join(_HERE, "libfsscene.so")


class NativeLibraryMissing(RuntimeError):
    pass


class FsxView(ctypes.Structure):
    """struct fsx_view of include/libfutspace.h."""

    _fields_ = [
        ("x", ctypes.c_float),
        ("y", ctypes.c_float),
        ("height", ctypes.c_float),
        ("angle", ctypes.c_float),
        ("horizon", ctypes.c_float),
        ("distance", ctypes.c_float),
        ("fov", ctypes.c_float),
        ("sky_color", ctypes.c_uint32),
        ("h", ctypes.c_int32),
        ("w", ctypes.c_int32),
        ("d", ctypes.c_float),
        ("z0", ctypes.c_float),
        ("scale", ctypes.c_float),
        ("n_slices", ctypes.c_int32),
        ("map", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class FssPose(ctypes.Structure):
    _fields_ = [
        ("x", ctypes.c_float),
        ("y", ctypes.c_float),
        ("height", ctypes.c_float),
        ("angle", ctypes.c_float),
        ("horizon", ctypes.c_float),
    ]


class FssRng(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint64), ("inc", ctypes.c_uint64)]


P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
U32 = ctypes.c_uint32
F32 = ctypes.c_float
F64 = ctypes.c_double

# name -> (restype, argtypes); the opaque structs are passed as void*.
FUTSPACE_SIGNATURES = {
    "futhark_context_config_new": (P, []),
    "futhark_context_config_free": (None, [P]),
    "futhark_context_config_set_debugging": (None, [P, ctypes.c_int]),
    "futhark_context_config_set_profiling": (None, [P, ctypes.c_int]),
    "futhark_context_config_set_logging": (None, [P, ctypes.c_int]),
    "futhark_context_config_set_device": (None, [P, ctypes.c_char_p]),
    "futhark_context_config_select_device_interactively": (None, [P]),
    "futhark_context_new": (P, [P]),
    "futhark_context_free": (None, [P]),
    "futhark_context_sync": (ctypes.c_int, [P]),
    "futhark_context_get_error": (P, [P]),
    "futhark_context_clear_caches": (ctypes.c_int, [P]),
    "futhark_context_pause_profiling": (None, [P]),
    "futhark_context_unpause_profiling": (None, [P]),
    "futhark_debugging_report": (None, [P]),
    "futhark_new_i32_2d": (P, [P, P, I64, I64]),
    "futhark_free_i32_2d": (ctypes.c_int, [P, P]),
    "futhark_values_i32_2d": (ctypes.c_int, [P, P, P]),
    "futhark_shape_i32_2d": (ctypes.POINTER(I64), [P, P]),
    "futhark_new_u32_2d": (P, [P, P, I64, I64]),
    "futhark_free_u32_2d": (ctypes.c_int, [P, P]),
    "futhark_values_u32_2d": (ctypes.c_int, [P, P, P]),
    "futhark_shape_u32_2d": (ctypes.POINTER(I64), [P, P]),
    "futhark_free_opaque_state": (ctypes.c_int, [P, P]),
    "futhark_entry_resize": (ctypes.c_int, [P, ctypes.POINTER(P), I32, I32, P]),
    "futhark_entry_key": (ctypes.c_int, [P, ctypes.POINTER(P), I32, I32, P]),
    "futhark_entry_mouse": (ctypes.c_int, [P, ctypes.POINTER(P), I32, I32, I32, P]),
    "futhark_entry_wheel": (ctypes.c_int, [P, ctypes.POINTER(P), I32, I32, P]),
    "futhark_entry_render": (ctypes.c_int, [P, ctypes.POINTER(P), P]),
    "futhark_entry_init": (ctypes.c_int, [P, ctypes.POINTER(P), U32]),
    "futhark_entry_text_content": (ctypes.c_int, [P] + [ctypes.POINTER(F32)] * 9 + [P]),
    "futhark_entry_step": (ctypes.c_int, [P, ctypes.POINTER(P), F32, P]),
    "futhark_entry_update_map": (ctypes.c_int, [P, ctypes.POINTER(P), P, P, P]),
    "fsx_map_new": (ctypes.c_int, [P, ctypes.POINTER(P), P, P, I64, I64, I32, I32, F32, F32]),
    "fsx_map_free": (ctypes.c_int, [P, P]),
    "fsx_map_shaded": (ctypes.c_int, [P, P, P]),
    "fsx_map_shade_size": (I32, [P]),
    "fsx_render": (ctypes.c_int, [P, P, ctypes.POINTER(FsxView), P, I64]),
    "fsx_render_host": (ctypes.c_int, [P, P, ctypes.POINTER(FsxView), P]),
    "fsx_render_batch": (ctypes.c_int, [P, P, I32, ctypes.POINTER(FsxView), I32, P, P]),
    "fsx_batch_new": (ctypes.c_int, [P, P, I32, ctypes.POINTER(FsxView), I32, P, P,
                                     ctypes.POINTER(P)]),
    "fsx_batch_run": (ctypes.c_int, [P, P]),
    "fsx_batch_capture": (ctypes.c_int, [P, P]),
    "fsx_batch_replay": (ctypes.c_int, [P, P]),
    "fsx_batch_free": (ctypes.c_int, [P, P]),
    "fsx_set_tuning": (ctypes.c_int, [P, I32, I32, I32]),
    "fsx_get_tuning": (ctypes.c_int, [P, P]),
    "fsx_context_set_stream": (ctypes.c_int, [P, P]),
    "fsx_context_stream": (P, [P]),
    "fsx_map_shade_factor": (ctypes.c_int, [P, P, P]),
    "fsx_render_f32": (ctypes.c_int, [P, P, ctypes.POINTER(FsxView), P, P, I32, I32, F32, P, P]),
    "fsx_render_f32_vjp": (ctypes.c_int, [P, ctypes.POINTER(FsxView), P, P, I32, I32, P, P, I32,
                                          P, P]),
    "fsx_l2_loss": (ctypes.c_int, [P, P, P, I64, I32, P, P]),
    "fsx_adam_step": (ctypes.c_int, [P, P, P, P, P, I64, F64, F64, F64, F64, I32, F32, F32]),
    "fsx_render_f32_l2_step": (ctypes.c_int, [P, P, ctypes.POINTER(FsxView), P, P, I32, I32, F32,
                                              P, I32, P, P, P, P, P, P]),
    "fsx_render_f32_l2_step_noalloc": (ctypes.c_int, [P, P, ctypes.POINTER(FsxView), P, P, I32, I32, F32,
                                              P, I32, P, P, P, P, P, P]),
    "fsx_render_stateless": (ctypes.c_int, [P, P, P, I32, I32, ctypes.POINTER(FsxView), P]),
    "fsx_render2": (ctypes.c_int, [P, P, ctypes.POINTER(FsxView), P, I64]),
    "fsx_context_device": (ctypes.c_int, [P]),
    "fsx_version": (ctypes.c_char_p, []),
    "fsx_sky": (U32, [F32]),
    "fsx_selftest": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
    "fsx_profile_enable": (ctypes.c_int, [P, ctypes.c_int]),
    "fsx_profile_read": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(I64)]),
}

SCENE_SIGNATURES = {
    "fss_heightmap": (ctypes.c_int, [ctypes.c_uint64, I32, ctypes.c_double, I32, P]),
    "fss_colormap": (ctypes.c_int, [ctypes.c_uint64, P, I32, I32, I32, I32, P]),
    "fss_flythrough": (ctypes.c_int, [ctypes.c_uint64, I32, I32, I32, F32, F32, F32, F32, P]),
    "fss_rng_seed": (None, [ctypes.POINTER(FssRng), ctypes.c_uint64, ctypes.c_uint64]),
    "fss_rng_u32": (U32, [ctypes.POINTER(FssRng)]),
    "fss_rng_int": (I32, [ctypes.POINTER(FssRng), I32, I32]),
    "fss_rng_unit": (ctypes.c_double, [ctypes.POINTER(FssRng)]),
}


def _declare(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_futspace = None
_scene = None
_libc = None


def _share_torch_hip_runtime():
    """Load torch's HIP runtime before libfutspace when torch is present.

    torch ships its own libamdhip64.so.7 / libhsa-runtime64.so.1.  Once
    they are loaded, libfutspace's NEEDED entries bind to them by soname,
    so the process has one HIP runtime and torch tensors' device pointers
    and streams are valid in the renderer.  Loading libfutspace first would
    pull in /opt/rocm's copies as well and leave torch without a device.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def futspace():
    """The HIP renderer library; raises if it has not been built."""
    global _futspace
    if _futspace is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C futspace_amd/csrc)")
        _share_torch_hip_runtime()
        _futspace = _declare(ctypes.CDLL(LIB_PATH), FUTSPACE_SIGNATURES)
    return _futspace


def scene():
    global _scene
    if _scene is None:
        if not os.path.exists(SCENE_PATH):
            raise NativeLibraryMissing(f"{SCENE_PATH} is missing: build it first")
        _scene = _declare(ctypes.CDLL(SCENE_PATH), SCENE_SIGNATURES)
    return _scene


def libc():
    global _libc
    if _libc is None:
        _libc = ctypes.CDLL(None)
        _libc.free.argtypes = [ctypes.c_void_p]
        _libc.free.restype = None
    return _libc
