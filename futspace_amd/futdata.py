"""Futhark's binary value format, and the reference's map conversions.

A binary value is b'b', the format version 2, the rank (one byte), the
element type as four characters (b' i32', b' u32', b' f32', b'  u8' ...),
rank little-endian u64 dimensions and the elements little-endian
(the format futhark's generated programs read and write with -b, as
tools/png2data.py writes it and tools/data2png.py reads it).

png_to_i32 packs an RGB image the way tools/png2data.py:36-57 does
(R << 16 | G << 8 | B, no alpha); freeimage_pack the way the interactive
host loads maps (c/freeimage_futspace.h:47-53: alpha 0xFF).

`python -m futspace_amd.futdata render < colour.in height.in > frame.data`
runs the stateless program of tests/futspace.fut on the GPU, like
`./futspace -b` in tests/Makefile:3; `python -m futspace_amd.futdata bench`
times it over the input sweeps of its `futhark bench` header
(tests/futspace.fut:1-22).
"""
import struct
import sys

import numpy as np

_TYPES = {
    b"  i8": np.int8, b"  u8": np.uint8, b" i16": np.int16, b" u16": np.uint16,
    b" i32": np.int32, b" u32": np.uint32, b" i64": np.int64, b" u64": np.uint64,
    b" f32": np.float32, b" f64": np.float64, b"bool": np.bool_,
}
_NAMES = {np.dtype(v): k for k, v in _TYPES.items()}


def write_value(f, array):
    """Write one array in the binary format."""
    a = np.asarray(array)
    if a.ndim:
        a = np.ascontiguousarray(a)
    name = _NAMES.get(a.dtype)
    if name is None:
        raise TypeError(f"no Futhark binary type for {a.dtype}")
    f.write(b"b")
    f.write(struct.pack("<BB", 2, a.ndim))
    f.write(name)
    for d in a.shape:
        f.write(struct.pack("<Q", d))
    f.write(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())


def read_value(f):
    """Read one array in the binary format (None at end of input)."""
    head = f.read(1)
    if not head:
        return None
    if head != b"b":
        raise ValueError("not a Futhark binary value")
    version, rank = struct.unpack("<BB", f.read(2))
    if version != 2:
        raise ValueError(f"binary format version {version} is not 2")
    name = f.read(4)
    if name not in _TYPES:
        raise ValueError(f"unknown element type {name!r}")
    dims = struct.unpack(f"<{rank}Q", f.read(8 * rank)) if rank else ()
    dt = np.dtype(_TYPES[name]).newbyteorder("<")
    count = int(np.prod(dims)) if rank else 1
    data = np.frombuffer(f.read(count * dt.itemsize), dtype=dt, count=count)
    return data.astype(dt.newbyteorder("="), copy=False).reshape(dims)


def png_to_i32(rgb):
    """[h][w][3] u8 -> [h][w] i32 packed R << 16 | G << 8 | B."""
    rgb = np.asarray(rgb, dtype=np.int32)
    return (rgb[..., 0] << 16) | (rgb[..., 1] << 8) | rgb[..., 2]


def freeimage_pack(rgb):
    """[h][w][3] u8 -> [h][w] u32 0xFF000000 | R << 16 | G << 8 | B."""
    return (np.uint32(0xFF000000) | png_to_i32(rgb).astype(np.uint32))


def i32_to_rgb(frame):
    """[h][w] i32/u32 -> [h][w][3] u8, as tools/data2png.py unpacks frames."""
    a = np.asarray(frame).astype(np.uint32)
    return np.stack([(a >> 16) & 255, (a >> 8) & 255, a & 255], axis=-1).astype(np.uint8)


# The sweeps of tests/futspace.fut:1-22: the program as a function of the
# depth (camera distance, d_*), the frame's columns (m_*) and its rows
# (l_*), each from 10 to 10^6.  The datasets themselves are not in the
# reference tree; seeded 1024^2 maps stand in for them.
SWEEP = (10, 100, 1000, 10000, 100000, 500000, 1000000)


def bench_stateless(reps=10, seed=3, size=1024, out=None):
    """Time fsx_render_stateless on `main`'s camera (tests/futspace.fut:
    127-145) with one of distance, columns or rows swept over SWEEP (the
    others at main's 800, 800 and 400); frames stay in device memory.
    Prints one JSON line per case; frames taller than the library's limit
    are reported as refused."""
    import ctypes
    import dataclasses
    import json

    import torch

    from . import scene
    from .renderer import STATELESS_MAIN, Context

    out = out or sys.stdout
    col, hts = scene.scene_maps(seed, size)
    colour = torch.from_numpy((np.asarray(col, np.uint32) & 0xFFFFFF).astype(np.int32)).cuda()
    height = torch.from_numpy((np.asarray(hts, np.uint32) & 0xFF).astype(np.int32)).cuda()
    ctx = Context()
    stream = torch.cuda.Stream()
    ctx.set_stream(stream)
    cases = ([("d", n, dataclasses.replace(STATELESS_MAIN, distance=float(n))) for n in SWEEP] +
             [("m", n, dataclasses.replace(STATELESS_MAIN, w=n)) for n in SWEEP] +
             [("l", n, dataclasses.replace(STATELESS_MAIN, h=n)) for n in SWEEP])
    for name, n, v in cases:
        frame = torch.empty((v.h, v.w), dtype=torch.int32, device="cuda")
        cv = v.to_c()

        def call():
            return ctx.lib.fsx_render_stateless(
                ctx.ptr, ctypes.c_void_p(colour.data_ptr()), ctypes.c_void_p(height.data_ptr()),
                size, size, ctypes.byref(cv), ctypes.c_void_p(frame.data_ptr()))

        rec = {"input": f"{name}_{n}", "distance": v.distance, "rows": v.h, "cols": v.w}
        if call() != 0:
            rec["refused"] = ctx.error()
        else:
            ctx.sync()
            with torch.cuda.stream(stream):
                call()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(reps):
                    call()
                e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            rec.update({"ms_per_frame": ms, "mpixel_per_s": v.h * v.w / ms / 1e3})
        del frame
        out.write(json.dumps(rec) + "\n")
        out.flush()
    ctx.close()


def _main(argv):
    if len(argv) >= 2 and argv[1] == "bench":
        bench_stateless()
        return 0
    if len(argv) < 2 or argv[1] != "render":
        sys.stderr.write("usage: python -m futspace_amd.futdata render < colour.in height.in "
                         "> frame.data\n       python -m futspace_amd.futdata bench\n")
        return 2
    colour = read_value(sys.stdin.buffer)
    height = read_value(sys.stdin.buffer)
    if colour is None or height is None:
        sys.stderr.write("futdata: expected two [q][r]i32 values on stdin\n")
        return 1
    from .renderer import Context
    ctx = Context()
    frame = ctx.render_stateless(colour.astype(np.int32), height.astype(np.int32))
    write_value(sys.stdout.buffer, frame.astype(np.int32))
    sys.stdout.buffer.flush()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(_main(sys.argv))
