"""The C-ABI library loads, exports what its header declares, links the
reference host's call shapes, and refuses to run without a GPU (no GPU
needed: no device compute is called)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "libfutspace.h")
LIB = os.path.join(ROOT, "futspace_amd", "libfutspace.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b((?:futhark|fsx)_\w+)\s*\(", text)
    return sorted(set(names))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], check=True, capture_output=True,
                         text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_header_declares_the_generated_interface():
    names = set(declared_functions())
    for required in ["futhark_entry_render", "futhark_entry_init", "futhark_entry_update_map",
                     "futhark_entry_step", "futhark_entry_key", "futhark_entry_resize",
                     "futhark_entry_mouse", "futhark_entry_wheel", "futhark_entry_text_content",
                     "futhark_new_u32_2d", "futhark_new_i32_2d", "futhark_free_u32_2d",
                     "futhark_free_i32_2d", "futhark_values_u32_2d", "futhark_shape_u32_2d",
                     "futhark_free_opaque_state", "futhark_context_config_new",
                     "futhark_context_config_free", "futhark_context_new",
                     "futhark_context_free", "futhark_context_sync",
                     "futhark_context_get_error"]:
        assert required in names, required


def test_library_exports_every_declared_function():
    missing = [n for n in declared_functions() if n not in exported_symbols()]
    assert not missing, missing


def test_ctypes_table_covers_the_header():
    from futspace_amd import _native
    missing = [n for n in declared_functions() if n not in _native.FUTSPACE_SIGNATURES]
    assert not missing, missing


def test_library_loads_and_binds_every_symbol():
    from futspace_amd import _native
    lib = _native.futspace()
    for name in declared_functions():
        assert getattr(lib, name) is not None
    assert lib.fsx_version().decode().startswith("futspace-mi355x")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_reference_host_call_shapes_compile_and_link(tmp_path):
    exe = tmp_path / "conformance"
    cmd = ["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
           f"-I{os.path.join(ROOT, 'include')}", os.path.join(ROOT, "tests", "abi_conformance.c"),
           f"-L{os.path.dirname(LIB)}", "-lfutspace", "-L/opt/rocm/lib",
           "-Wl,--allow-shlib-undefined", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_sky_colour_is_host_side_and_matches_the_oracle():
    from futspace_amd import _native
    from oracle import fs_oracle as O
    lib = _native.futspace()
    for h in [0.0, 0.1, 0.5, 0.99, 1.0, 1.7, -0.2]:
        assert lib.fsx_sky(ctypes.c_float(h)) == O.lib().fso_sky(h), h


def test_no_device_means_no_context():
    """Without a GPU the library refuses to create a context: there is no
    CPU fallback for the renderer."""
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from futspace_amd import Context, FutspaceError
    with pytest.raises(FutspaceError):
        Context(device=0)


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "futspace_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".c", ".h", ".hip")):
                text = open(os.path.join(dirpath, f)).read()
                assert "fs_oracle" not in text and "import oracle" not in text, f
                assert "from oracle" not in text, f


def test_every_entry_refuses_a_null_context():
    """Every int-returning entry of the header returns nonzero for a NULL
    context (before touching HIP, so no GPU is needed); the array
    constructors, the error getter and the stream getter return NULL."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    decls = re.findall(r"\n(int|[^\n;(]*\*)\s*((?:futhark|fsx)_\w+)\s*\(([^;]*?)\);", text)
    lib = ctypes.CDLL(LIB)
    checked = 0
    for ret, name, args in decls:
        nargs = 0 if args.strip() in ("", "void") else args.count(",") + 1
        if nargs == 0 or not args.lstrip().startswith("struct futhark_context *"):
            continue
        fn = getattr(lib, name)
        argv = [ctypes.c_void_p(None)] * nargs
        if ret.strip() == "int":
            fn.restype = ctypes.c_int
            assert fn(*argv) != 0, name
        else:
            fn.restype = ctypes.c_void_p
            assert fn(*argv) is None, name
        checked += 1
    assert checked >= 46
    for name in ("futhark_context_free", "futhark_context_pause_profiling",
                 "futhark_context_unpause_profiling", "futhark_debugging_report"):
        getattr(lib, name).restype = None
        getattr(lib, name)(ctypes.c_void_p(None))
    lib.fsx_map_shade_size.restype = ctypes.c_int32
    assert lib.fsx_map_shade_size(ctypes.c_void_p(None)) == 0
