"""Bit-exact parity of the HIP renderer with the CPU oracle (needs a GPU).

Every frame, shaded colour map and state-machine output must equal the
oracle's bit for bit: the renderer is integer-exact by construction (f32
in the reference's operation order, no contraction, x86 cast semantics).
"""
import math

import numpy as np
import pytest

from futspace_amd import View, scene, sky_color
from oracle import fs_oracle as O

pytestmark = pytest.mark.gpu


def _diff_report(got, want):
    bad = np.argwhere(got != want)
    if len(bad) == 0:
        return ""
    r, c = bad[0]
    return (f"{len(bad)} of {got.size} pixels differ; first at row {r} col {c}: "
            f"got {int(got[r, c]):#010x} want {int(want[r, c]):#010x}")


def _check(got, want):
    assert got.shape == want.shape
    msg = _diff_report(got, want)
    assert not msg, msg


@pytest.fixture(scope="module")
def small_scene():
    col, hts = scene.scene_maps(11, 128)
    return col, hts


def test_shade_matches_oracle(ctx, small_scene):
    col, hts = small_scene
    m = ctx.map(col, hts)
    shaded, _ = O.update_map(col, hts)
    _check(m.shaded(), shaded)


@pytest.mark.parametrize("sun_h,sun_a", [(0.1, 0.1), (0.6, 2.0), (-0.3, 0.7), (1.2, -2.5)])
def test_shade_sun_angles(ctx, small_scene, sun_h, sun_a):
    col, hts = small_scene
    m = ctx.map(col, hts, sun_height=sun_h, sun_ang=sun_a)
    shaded, _ = O.update_map(col, hts, sun_h, sun_a)
    _check(m.shaded(), shaded)


SMALL_VIEWS = [
    # x, y, height, angle, horizon, distance, fov, h, w, d, z0, scale, n
    (64.3, 100.7, 260.0, 0.4, 60.0, 300.0, 1.2, 96, 128, 0.01, 0.5, 0.0, 0),
    (64.0, 64.0, 200.0, 0.0, 50.0, 200.0, 1.2, 64, 80, 0.01, 0.5, 0.0, 0),      # yaw 0: integer x
    (10.5, 3.25, 300.0, 1.5707964, 90.0, 400.0, 0.8, 100, 77, 0.005, 1.0, 120.0, 0),
    (-40.2, 900.1, 180.0, 3.0, 20.0, 250.0, 1.5, 33, 31, 0.02, 0.25, 0.0, 0),
    (5.0, 5.0, 140.0, -2.2, 150.0, 500.0, 1.2, 120, 64, 0.001, 0.0, 0.0, 0),    # z0 = 0
    (70.0, 30.0, 400.0, 0.9, -50.0, 600.0, 1.1, 64, 64, 0.0, 2.0, 0.0, 300),     # d = 0, explicit n
    (33.3, 77.7, 250.0, 5.0, 300.0, 700.0, 1.2, 257, 129, 0.003, 0.4, 64.0, 0),
]


@pytest.mark.parametrize("vi", range(len(SMALL_VIEWS)))
def test_small_views_bitexact(ctx, small_scene, vi):
    col, hts = small_scene
    x, y, hh, a, hor, dist, fov, h, w, d, z0, sc, n = SMALL_VIEWS[vi]
    v = View(x=x, y=y, height=hh, angle=a, horizon=hor, distance=dist, fov=fov,
             sky_color=sky_color(0.1), h=h, w=w, d=d, z0=z0, scale=sc, n_slices=n)
    m = ctx.map(col, hts)
    shaded, alt = O.update_map(col, hts)
    _check(ctx.render(m, v), O.render_off(v, shaded, alt))


@pytest.mark.parametrize("rows,cols", [(2, 1), (3, 5), (96, 80), (128, 64)])
def test_non_power_of_two_maps(ctx, rows, cols):
    hts = (np.arange(rows * cols, dtype=np.int32).reshape(rows, cols) * 37) % 256
    col = (0xFF000000 | (np.arange(rows * cols, dtype=np.uint32).reshape(rows, cols) * 2654435761
                         % 0xFFFFFF)).astype(np.uint32)
    m = ctx.map(col, hts)
    shaded, alt = O.update_map(col, hts)
    _check(m.shaded(), shaded)
    v = View(x=0.37, y=0.61, height=300.0, angle=2.2, horizon=200.0, distance=400.0, fov=1.2,
             sky_color=sky_color(0.1), h=70, w=90, d=0.004, z0=0.3, scale=0.0)
    _check(ctx.render(m, v), O.render_off(v, shaded, alt))


def test_zero_colour_texels_are_transparent(ctx):
    """Colour-0 texels are `empty` for fill_vline and sky
    (fut/voxel_renderer.fut:78-81, :248)."""
    col, hts = scene.scene_maps(12, 64)
    col = col.copy()
    col[::3, :] = 0
    col[:, ::5] = 0
    m = ctx.map(col, hts)
    shaded, alt = O.update_map(col, hts)
    v = View(x=20.5, y=40.2, height=220.0, angle=0.3, horizon=80.0, distance=200.0, fov=1.2,
             sky_color=sky_color(0.1), h=128, w=96, d=0.01, z0=0.5)
    _check(ctx.render(m, v), O.render_off(v, shaded, alt))


def test_sixteen_bit_heights(ctx):
    hts = scene.heightmap(13, 128, 0.6, 600)
    col = scene.colormap(13, hts, 600)
    m = ctx.map(col, hts, height_bits=16, shade_size=128)
    shaded, alt = O.update_map(col, hts, size=128, bits=16)
    _check(m.shaded(), shaded)
    v = View(x=50.0, y=90.0, height=700.0, angle=1.0, horizon=120.0, distance=1000.0, fov=1.2,
             sky_color=sky_color(0.1), h=144, w=160, d=0.002, z0=0.5)
    _check(ctx.render(m, v), O.render_off(v, shaded, alt))


def test_plain_1024_bitexact(ctx):
    col, hts = scene.plain_1024_maps()
    m = ctx.map(col, hts)
    shaded, alt = O.update_map(col, hts)
    _check(m.shaded(), shaded)
    v = scene.plain_1024_view()
    _check(ctx.render(m, v), O.render_off(v, shaded, alt))


@pytest.fixture(scope="module")
def plain_2048(ctx):
    col, hts = scene.plain_2048_maps()
    m = ctx.map(col, hts)
    shaded, alt = O.update_map(col, hts)
    return m, shaded, alt, scene.plain_2048_views(hts)


@pytest.mark.parametrize("frame", [0, 8, 31, 64, 97, 119])
def test_plain_2048_flythrough_frames(ctx, plain_2048, frame):
    m, shaded, alt, views = plain_2048
    v = views[frame]
    _check(ctx.render(m, v), O.render_painter_mt(v, shaded, alt))


def test_render_into_torch_tensor(ctx, small_scene):
    torch = pytest.importorskip("torch")
    col, hts = small_scene
    m = ctx.map(col, hts)
    v = View(x=64.3, y=100.7, height=260.0, angle=0.4, horizon=60.0, distance=300.0, fov=1.2,
             sky_color=sky_color(0.1), h=96, w=128, d=0.01, z0=0.5)
    out = torch.zeros((96, 136), dtype=torch.int32, device="cuda")
    ctx.set_stream(torch.cuda.current_stream())
    ctx.render(m, v, out)
    torch.cuda.synchronize()
    ctx.set_stream(None)
    got = out.cpu().numpy().view(np.uint32)
    shaded, alt = O.update_map(col, hts)
    _check(got[:, :128].copy(), O.render_off(v, shaded, alt))
    assert not got[:, 128:].any(), "render wrote past the frame width"


@pytest.mark.parametrize("which", [0, 1, 2, 3, 4, 5, 6, 7])
def test_device_arithmetic_selftest(ctx, which):
    """The kernels' shortcuts (byte/255 by reciprocal + FMA correction, sqrt
    on [0,1] without scaling, median-based quantisation, row clamp, the
    bucket-table quantisation of mixed channels, and the slice table's
    Newton reciprocal and Markstein divisions by the frame width) agree with
    the correctly rounded formulas on every input they can see."""
    import ctypes
    bad, first = ctypes.c_int64(-1), ctypes.c_int64(-1)
    ctx.check(ctx.lib.fsx_selftest(ctx.ptr, which, ctypes.byref(bad), ctypes.byref(first)),
              "fsx_selftest")
    assert bad.value == 0, f"{bad.value} mismatches, first input bits {first.value:#x}"


def test_reference_map_data_bitexact(ctx):
    """The reference's C1W/D1 maps (a 256^2 crop packed as the reference
    host loads them, tests/golden/maps.npz)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "maps.npz"))
    col, hts = g["c1w_colour"], g["d1_height"]
    m = ctx.map(col, hts, shade_size=256)
    shaded, alt = O.update_map(col, hts, size=256)
    _check(m.shaded(), shaded)
    for v in [View(x=128.5, y=200.25, height=130.0, angle=0.3, horizon=10.0, distance=300.0,
                   fov=1.2, sky_color=sky_color(0.1), h=120, w=160, d=0.004, z0=0.3),
              View(x=40.7, y=31.1, height=140.0, angle=2.9, horizon=-20.0, distance=500.0,
                   fov=1.2, sky_color=sky_color(0.1), h=90, w=200, d=0.002, z0=0.5, scale=240.0),
              View(x=250.0, y=128.0, height=110.0, angle=-1.2, horizon=30.0, distance=250.0,
                   fov=0.9, sky_color=sky_color(0.1), h=150, w=100, d=0.01, z0=1.0)]:
        want = O.render_off(v, shaded, alt)
        assert (want == v.sky_color).any() and (want != v.sky_color).any()
        _check(ctx.render(m, v), want)


@pytest.mark.parametrize("shape", [(8, 2, 512), (8, 2, 256), (8, 1, 512), (4, 2, 1024),
                                   (4, 1, 2048)])
def test_kernel_shapes_bitexact(ctx, small_scene, shape):
    """Every tile shape the autotuner may pick renders the same frames."""
    col, hts = small_scene
    m = ctx.map(col, hts)
    shaded, alt = O.update_map(col, hts)
    ctx.set_tuning(*shape)
    try:
        for vi in (0, 1, 3, 6):
            x, y, hh, a, hor, dist, fov, h, w, d, z0, sc, n = SMALL_VIEWS[vi]
            v = View(x=x, y=y, height=hh, angle=a, horizon=hor, distance=dist, fov=fov,
                     sky_color=sky_color(0.1), h=h, w=w, d=d, z0=z0, scale=sc, n_slices=n)
            _check(ctx.render(m, v), O.render_off(v, shaded, alt))
        v = View(x=70.0, y=30.0, height=400.0, angle=0.9, horizon=-50.0, distance=600.0,
                 fov=1.1, sky_color=sky_color(0.1), h=300, w=61, d=0.0005, z0=0.2,
                 n_slices=2600)
        _check(ctx.render(m, v), O.render_off(v, shaded, alt))
    finally:
        ctx.set_tuning(0, 0, 0)


@pytest.mark.parametrize("h,w", [(2160, 72), (4096, 24), (4320, 20), (7680, 9)])
def test_tall_frames(ctx, small_scene, h, w):
    """Frames whose 8-column stripe does not fit in LDS take 4-column
    tiles (with shorter chunks as the frame grows; the tallest, up to the
    7680-row limit, 512-thread tiles): still bit-exact."""
    col, hts = small_scene
    m = ctx.map(col, hts)
    shaded, alt = O.update_map(col, hts)
    v = View(x=40.3, y=90.6, height=230.0, angle=1.3, horizon=h * 0.4, distance=400.0, fov=1.2,
             sky_color=sky_color(0.1), h=h, w=w, d=0.004, z0=0.5, scale=h / 3.0)
    _check(ctx.render(m, v), O.render_off(v, shaded, alt))


@pytest.mark.parametrize("h,w,seed", [(7681, 13, 0), (12000, 9, 1)])
def test_frames_taller_than_a_tile(ctx, small_scene, h, w, seed):
    """Frames of more than 7680 rows take the one-thread-per-column kernel:
    bit-exact, with colour-0 (transparent) texels in the map as well."""
    col, hts = small_scene
    col = np.array(col, copy=True)
    if seed:
        col[np.random.default_rng(seed).random(col.shape) < 0.02] = 0
    m = ctx.map(col, hts)
    shaded, alt = O.update_map(col, hts)
    v = View(x=40.3, y=90.6, height=230.0, angle=1.3, horizon=h * 0.4, distance=400.0, fov=1.2,
             sky_color=sky_color(0.1), h=h, w=w, d=0.004, z0=0.5, scale=h / 3.0)
    _check(ctx.render(m, v), O.render_off(v, shaded, alt))


def test_rows_beyond_the_limit_are_rejected(ctx, small_scene):
    col, hts = small_scene
    m = ctx.map(col, hts)
    v = View(x=40.3, y=90.6, height=230.0, angle=1.3, horizon=100.0, distance=400.0, fov=1.2,
             h=(1 << 22) + 1, w=1, d=0.004, z0=0.5)
    with pytest.raises(RuntimeError):
        ctx.render(m, v)


def test_tuning_belongs_to_its_context(ctx):
    """fsx_set_tuning changes the kernel shape of its own context only; each
    context renders the same frame bit for bit in its shape."""
    from futspace_amd import Context
    col, hts = scene.scene_maps(5, 256)
    v = View(x=120.3, y=17.7, height=240.0, angle=2.2, horizon=90.0, distance=500.0, fov=1.2,
             sky_color=sky_color(0.1), h=400, w=300, d=0.002, z0=0.5)
    shaded, alt = O.update_map(col, hts)
    want = O.render_off(v, shaded, alt)
    base = ctx.tuning()
    c2 = Context(device=0)
    try:
        assert c2.tuning() == base
        c2.set_tuning(4, 1, 1024)
        assert c2.tuning() == (4, 1, 1024)
        assert ctx.tuning() == base
        for c in (ctx, c2):
            _check(c.render(c.map(col, hts), v), want)
    finally:
        c2.close()
    assert ctx.tuning() == base
