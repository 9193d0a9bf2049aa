"""Futhark binary values and the reference's map packings (no GPU)."""
import io
import os

import numpy as np
import pytest

from futspace_amd import futdata


def test_layout_matches_png2data():
    """tools/png2data.py:50-57 writes b'b', 2, 2, b' i32', u64 height,
    u64 width, then the i32 data."""
    a = (np.arange(6 * 5, dtype=np.int32).reshape(6, 5) * 7919) - 100000
    f = io.BytesIO()
    futdata.write_value(f, a)
    ref = io.BytesIO()
    ref.write(b"b")
    ref.write(np.int8(2).tobytes())
    ref.write(np.int8(2).tobytes())
    ref.write(b" i32")
    ref.write(np.uint64(6).tobytes())
    ref.write(np.uint64(5).tobytes())
    ref.write(a.astype("<i4").tobytes())
    assert f.getvalue() == ref.getvalue()


@pytest.mark.parametrize("dtype", [np.int32, np.uint32, np.float32, np.uint8, np.int8,
                                   np.int64, np.float64, np.uint16])
@pytest.mark.parametrize("shape", [(3, 4), (7,), (2, 3, 3), ()])
def test_round_trip(dtype, shape):
    rng = np.random.default_rng(1)
    a = (rng.random(shape) * 100).astype(dtype)
    f = io.BytesIO()
    futdata.write_value(f, a)
    futdata.write_value(f, a)
    f.seek(0)
    for _ in range(2):
        b = futdata.read_value(f)
        assert b.dtype == a.dtype and b.shape == a.shape and np.array_equal(a, b)
    assert futdata.read_value(f) is None


def test_rejects_other_formats():
    with pytest.raises(ValueError):
        futdata.read_value(io.BytesIO(b"x\x02\x02 i32"))
    with pytest.raises(ValueError):
        futdata.read_value(io.BytesIO(b"b\x01\x02 i32"))
    with pytest.raises(ValueError):
        futdata.read_value(io.BytesIO(b"b\x02\x02 q32"))


def test_map_packings_of_the_reference_data():
    """The golden crops are packed like the interactive host loads maps
    (alpha 0xFF); png2data's packing is the same without alpha."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "maps.npz"))
    colour = g["c1w_colour"]
    rgb = futdata.i32_to_rgb(colour)
    assert np.array_equal(futdata.freeimage_pack(rgb), colour)
    assert np.array_equal(futdata.png_to_i32(rgb), (colour & 0xFFFFFF).astype(np.int32))
