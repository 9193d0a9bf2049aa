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
`./futspace -b` in tests/Makefile:3.
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


def _main(argv):
    if len(argv) < 2 or argv[1] != "render":
        sys.stderr.write("usage: python -m futspace_amd.futdata render < colour.in height.in "
                         "> frame.data\n")
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
