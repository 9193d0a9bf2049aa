"""Pins of the CPU oracle (no GPU).

The oracle restates the reference's f32 algorithm; these tests check it
against what fixes the answer independently of its own code: the
reference's z-series specification (tests/javascript_zs.py), closed forms
(yaw-0 and yaw-pi/2 frustum lines, constant and ramp height maps, the
flat-map horizon row, the z0 = 0 uniform frame, the closed-form sun vector,
a single spike's shadow counted by integer arithmetic), and two
independent formulations of the same render that must agree bit for bit
(the reference's scan/scatter/fill restatement and a front-to-back
painter, for both the png renderer and the stateless program of
tests/futspace.fut).  The f32 relaxation of the differentiable path is
tied to the u32 renderer's visibility and checked against finite
differences of itself.

matte's argb formulas (mix, to_rgba, from_rgba, scale) are not vendored in
the reference and cannot be fetched: only their algebra is checked here,
so colour channels are "parity unpinned" against matte itself.
"""
import math

import numpy as np
import pytest

from futspace_amd import View, scene
from oracle import fs_oracle as O

INT_MIN = -(2 ** 31)
L = O.lib()


# ---- scalar semantics --------------------------------------------------

def test_i32_f32_follows_x86_truncation():
    cases = {float("nan"): INT_MIN, float("inf"): INT_MIN, float("-inf"): INT_MIN,
             3e9: INT_MIN, -3e9: INT_MIN, 2147483520.0: 2147483520, -2147483648.0: INT_MIN,
             -1.5: -1, 1.9: 1, -0.0: 0, 0.75: 0}
    for v, want in cases.items():
        assert L.fso_i32_f32(v) == want, v


def test_u32_f32_goes_through_64_bit_conversion():
    cases = {float("nan"): 0, 3e9: 3000000000, -1.0: 0xFFFFFFFF, 5e9: 5000000000 & 0xFFFFFFFF,
             255.9: 255, float("inf"): 0}
    for v, want in cases.items():
        assert L.fso_u32_f32(v) == want, v


def test_mod_is_floor_mod():
    for a in range(-20, 21):
        for m in (-7, -3, -1, 1, 2, 5, 64):
            assert L.fso_mod_i32(a, m) == a % m, (a, m)


# ---- depth slices ------------------------------------------------------

def _javascript_zs(limit=800.0, z=1.0, deltaz=1.0, step=0.005):
    """tests/javascript_zs.py:2-8 in double precision."""
    out = []
    while z <= limit:
        out.append(z)
        deltaz += step
        z += deltaz
    return np.array(out)


def test_get_zs_matches_the_z_series_specification():
    want = _javascript_zs()
    got = O.get_zs(800.0, 0.005, 1.0)
    assert len(got) == len(want) == 400
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=0)
    assert got[:3].tolist() == pytest.approx([1.0, 2.005, 3.015], rel=1e-6)
    assert got[-1] == pytest.approx(799.0, rel=1e-6)


@pytest.mark.parametrize("c,d,z0,n", [(800.0, 0.001, 0.0, 1265), (1000.0, 0.001, 0.0, 1414),
                                      (800.0, 0.005, 1.0, 400), (1000.0, 0.0019092, 0.0, 1024)])
def test_get_zs_closed_form(c, d, z0, n):
    z = O.get_zs(c, d, z0)
    assert len(z) == n
    i = np.arange(1, n + 1, dtype=np.float64)
    closed = i * np.float64(np.float32(z0)) + np.float64(np.float32(d)) * i * (i - 1) / 2
    np.testing.assert_allclose(z, closed, rtol=3e-6, atol=1e-30)
    # the count is the last term within the far distance
    assert closed[-1] <= c * (1 + 1e-6)
    nxt = (n + 1) * np.float32(z0) + np.float32(d) * (n + 1) * n / 2
    assert nxt > c * (1 - 1e-6)


def test_get_h_line_yaw_zero_and_quarter_turn():
    w = 640
    v = View(x=10.5, y=20.25, height=100.0, angle=0.0, horizon=0.0, fov=1.2)
    sx, sy, dx, dy = O.get_h_line(7.0, v, w)
    f = np.float32
    assert (sx, sy) == (f(-7.0) + f(10.5), f(f(-1.2) * f(7.0)) + f(20.25))
    assert dy == 0.0 and dx == f(f(14.0) / f(w))
    # a quarter turn exchanges the axes: the line runs along y
    v.angle = float(np.float32(math.pi / 2))
    sx, sy, dx, dy = O.get_h_line(7.0, v, w)
    assert abs(dx) < 1e-5 and dy == pytest.approx(-14.0 / w, rel=1e-5)
    assert sx == pytest.approx(10.5 - 1.2 * 7.0, abs=1e-4)
    assert sy == pytest.approx(20.25 + 7.0, abs=1e-4)


# ---- samplers ----------------------------------------------------------

def test_bilinear_height_constant_map_and_integer_points():
    alt = np.full((8, 8), 7, dtype=np.int32)
    rng = np.random.default_rng(1)
    for x, y in rng.uniform(-20, 20, (50, 2)):
        if x == math.floor(x) or y == math.floor(y):
            continue
        assert O.png_height_filtered(alt, x, y) == pytest.approx(7.0, rel=1e-6)
    for x, y in [(3.0, 2.5), (1.25, 4.0), (5.0, 5.0), (-2.0, 0.5)]:
        assert O.png_height_filtered(alt, x, y) == 0.0


def test_bilinear_height_linear_ramp():
    r, c = np.mgrid[0:16, 0:16]
    alt = (3 * c + 5 * r).astype(np.int32)
    rng = np.random.default_rng(2)
    for x, y in rng.uniform(0.01, 14.99, (100, 2)):
        assert O.png_height_filtered(alt, x, y) == pytest.approx(3 * x + 5 * y, rel=1e-5)


def test_bilinear_height_wraps_with_floor_mod():
    rng = np.random.default_rng(3)
    alt = rng.integers(0, 256, (5, 7)).astype(np.int32)
    for x, y in rng.uniform(0.01, 6.99, (40, 2)):
        for kx, ky in [(-7, 0), (14, -5), (-21, 10)]:
            assert O.png_height_filtered(alt, x + kx, y + ky) == pytest.approx(
                O.png_height_filtered(alt, x, y), rel=1e-4, abs=1e-3)


# ---- renderer: two formulations and closed forms -----------------------

def _random_case(seed):
    rng = np.random.default_rng(seed)
    rows, cols = [(64, 64), (96, 80), (3, 5), (128, 128), (40, 72), (64, 32)][seed % 6]
    alt = rng.integers(0, 256, (rows, cols)).astype(np.int32)
    shaded = (np.uint32(0xFF000000) | rng.integers(0, 1 << 24, (rows, cols)).astype(np.uint32))
    if seed % 2:
        shaded[rng.random((rows, cols)) < 0.2] = 0     # transparent texels
    v = View(x=float(rng.uniform(-50, 150)), y=float(rng.uniform(-50, 150)),
             height=float(rng.uniform(100, 400)), angle=float(rng.uniform(-7, 7)),
             horizon=float(rng.uniform(-80, 250)), distance=float(rng.uniform(50, 600)),
             fov=float(rng.uniform(0.5, 1.8)), sky_color=0xFF9090E0,
             h=int(rng.integers(1, 120)), w=int(rng.integers(1, 120)),
             d=float(rng.choice([0.0, 0.001, 0.01])), z0=float(rng.choice([0.0, 0.3, 1.0])),
             scale=float(rng.choice([0.0, 64.0, 240.0])),
             n_slices=int(rng.choice([0, 0, 50, 700])))
    return v, shaded, alt


@pytest.mark.parametrize("seed", range(24))
def test_scan_scatter_fill_equals_painter(seed):
    v, shaded, alt = _random_case(seed)
    a = O.render_off(v, shaded, alt)
    b = O.render_painter(v, shaded, alt)
    assert np.array_equal(a, b)
    assert np.array_equal(O.render_painter_mt(v, shaded, alt, 2), b)
    # transparent samples never reach the frame: a pixel is a colour or the sky
    assert not np.any(a == 0)


def test_flat_map_horizon_row_is_closed_form():
    """Flat terrain at h0: rows fall monotonically with depth, so every
    column shows sky exactly above the farthest slice's projected row."""
    h0, n, H = 100.0, 500, 300
    alt = np.full((64, 64), int(h0), dtype=np.int32)
    shaded = np.full((64, 64), 0xFF336699, dtype=np.uint32)
    d, z0 = scene.depth_for(n)
    v = View(x=20.3, y=41.7, height=180.0, angle=0.83, horizon=40.0, distance=1000.0,
             fov=1.2, sky_color=0xFF000001, h=H, w=257, d=d, z0=z0, scale=200.0, n_slices=n)
    frame = O.render_off(v, shaded, alt)
    f = np.float32
    i = f(n)
    z_last = (i / f(2.0)) * (f(2.0) * f(z0) + (i - f(1.0)) * f(d))
    rel = (f(180.0) - f(h0)) * ((f(1.0) / z_last) * f(200.0)) + f(40.0)
    row = max(0, min(H, int(rel)))
    sky_rows = (frame == v.sky_color).sum(axis=0)
    assert np.all(sky_rows == row), (row, np.unique(sky_rows))
    assert np.all(frame[:row] == v.sky_color) and not np.any(frame[row:] == v.sky_color)


def test_z0_zero_gives_a_uniform_frame():
    """Slice 1 at z = 0 projects through inv_z = inf to row 0 (x86 cast of
    +-inf or NaN), covering every column with the texel under the camera."""
    col, hts = scene.scene_maps(5, 64)
    shaded, alt = O.update_map(col, hts, size=64)
    for x, y in [(10.5, 20.25), (33.7, 1.1)]:
        v = View(x=x, y=y, height=300.0, angle=1.1, horizon=100.0, distance=400.0, fov=1.2,
                 sky_color=0xFF0000FF, h=50, w=60, d=0.01, z0=0.0)
        frame = O.render_off(v, shaded, alt)
        assert np.all(frame == frame[0, 0])
        c = O.png_color_filtered(shaded, x, y)
        assert frame[0, 0] == (c if c else v.sky_color)


# ---- shading -----------------------------------------------------------

def test_sun_closed_form_and_rotations():
    for h, a in [(0.1, 0.1), (0.6, 2.0), (-0.3, 0.7), (1.2, -2.5)]:
        s = O.sun(h, a)
        want = [math.cos(a) * -math.sin(h), math.cos(h), -math.sin(a) * -math.sin(h)]
        np.testing.assert_allclose(s, want, atol=3e-7)
        r = O.vec3_rotate(1, a, O.vec3_rotate(2, h, [0.0, 1.0, 0.0]))
        np.testing.assert_allclose(r, s, atol=3e-7)


def test_flat_map_has_no_shadow():
    rng = np.random.default_rng(4)
    alt = np.full((16, 16), 90, dtype=np.int32)
    col = (np.uint32(0xFF000000) | rng.integers(0, 1 << 24, (16, 16)).astype(np.uint32))
    sv = O.sun(0.4, 1.3)
    assert not O.shadow_counts(alt, sv, 16).any()
    shaded = O.shadowmap(col, alt, sv, 16)
    black = L.fso_black()
    assert black == 0xFF000000
    for y in range(16):
        for x in range(16):
            assert shaded[y, x] == L.fso_mix(0.0, black, 1.0, int(col[y, x]))


def test_single_spike_shadow_counted_by_hand():
    """Sun straight along -x (sun_height = pi/2, sun_ang = 0): ray steps of
    4 texels; only texels on the spike's row whose distance to it is a
    multiple of 4 (mod the 64-texel wrap) see it, once per lap."""
    M, sx, sy, HH = 64, 21, 37, 200
    alt = np.zeros((M, M), dtype=np.int32)
    alt[sy, sx] = HH
    sv = O.sun(float(np.float32(math.pi / 2)), 0.0)
    counts = O.shadow_counts(alt, sv, M)
    want = np.zeros((M, M), dtype=np.int32)
    for x in range(M):
        if x == sx or (x - sx) % 4:
            continue
        r = ((x - sx) // 4) % 16
        want[sy, x] = sum(1 for dist in range(1, 256) if dist % 16 == r)
    assert np.array_equal(counts, want)
    col = np.full((M, M), 0xFF80C040, dtype=np.uint32)
    shaded = O.shadowmap(col, alt, sv, M)
    for x in range(M):
        assert shaded[sy, x] == L.fso_mix(4.0 * want[sy, x], L.fso_black(), 1.0, 0xFF80C040)


# ---- matte algebra (parity unpinned against matte itself) --------------

def test_argb_algebra():
    rng = np.random.default_rng(5)
    for c in rng.integers(0, 1 << 32, 64, dtype=np.uint64).astype(np.uint32):
        c = int(c)
        out = (ctypes_floats := np.empty(4, dtype=np.float32))
        L.fso_to_rgba(c, out.ctypes.data)
        back = L.fso_from_rgba(*[float(x) for x in ctypes_floats])
        for sh in (0, 8, 16, 24):
            assert abs(((back >> sh) & 255) - ((c >> sh) & 255)) <= 1
        m = L.fso_mix(0.3, c, 0.7, c)          # mixing a colour with itself
        for sh in (0, 8, 16, 24):
            assert abs(((m >> sh) & 255) - ((c >> sh) & 255)) <= 1
        assert L.fso_scale(c, 1.0) >> 24 == c >> 24          # scale keeps alpha
    assert L.fso_sky(0.1) == L.fso_scale(0xFF9090E0, 0.1)


# ---- stateless program, tests/futspace.fut -----------------------------

@pytest.mark.parametrize("seed", range(6))
def test_stateless_program_scan_equals_painter(seed):
    rng = np.random.default_rng(50 + seed)
    q, r = [(64, 64), (40, 72), (128, 128)][seed % 3]
    col = rng.integers(-(2 ** 31), 2 ** 31 - 1, (q, r)).astype(np.int32)
    if seed % 2:
        col[rng.random((q, r)) < 0.3] = 0
    alt = rng.integers(0, 256, (q, r)).astype(np.int32)
    cam = (float(rng.uniform(0, r)), float(rng.uniform(0, q)), float(rng.uniform(50, 300)),
           float(rng.uniform(-4, 4)), float(rng.uniform(0, 200)), float(rng.uniform(50, 500)),
           float(rng.uniform(0.5, 1.5)))
    l, m = int(rng.integers(1, 90)), int(rng.integers(1, 90))
    a = O.render_nearest(cam, col, alt, l, m, 0x7F9090E0)
    b = O.render_nearest(cam, col, alt, l, m, 0x7F9090E0, painter=True)
    assert np.array_equal(a, b)


# ---- f32 relaxation ----------------------------------------------------

def _relax_case():
    col, hts = scene.scene_maps(17, 64)
    shaded, alt = O.update_map(col, hts, size=64)
    v = View(x=20.3, y=41.7, height=230.0, angle=0.9, horizon=70.0, distance=300.0, fov=1.2,
             sky_color=0xFF102030, h=90, w=110, d=0.004, z0=0.3)
    return shaded, alt, v


def _planes(c):
    c = np.asarray(c, dtype=np.uint32)
    return np.stack([(c >> 16) & 255, (c >> 8) & 255, c & 255]).astype(np.float32) / np.float32(255)


def test_relaxation_follows_the_u32_renderer():
    """With unit shade factors and planes = the shaded map's channels, the
    relaxation shows sky exactly where the u32 frame does and colours
    within the u32 path's three byte truncations."""
    shaded, alt, v = _relax_case()
    assert not np.any(shaded == 0)
    frame = O.render_off(v, shaded, alt)
    img, src = O.render_f32(v, _planes(shaded), np.ones(shaded.shape, np.float32), alt, 1.0)
    assert np.array_equal(src < 0, frame == v.sky_color)
    ch = _planes(frame)
    terrain = src >= 0
    assert terrain.any()
    assert np.abs(img[:, terrain] - ch[:, terrain]).max() * 255 <= 3.0
    sky = np.array([144, 144, 224], np.float32) / np.float32(255)
    assert np.allclose(img[:, ~terrain], sky[:, None])


def test_relaxation_vjp_matches_finite_differences():
    shaded, alt, v = _relax_case()
    rng = np.random.default_rng(6)
    planes = _planes(shaded)
    sf = O.shade_factor(O.shadow_counts(alt, O.sun(0.1, 0.1), 64))
    target = np.clip(planes + rng.normal(0, 0.03, planes.shape), 0, 1).astype(np.float32)
    timg, _ = O.render_f32(v, target, sf, alt, 0.25)

    def loss(p, s=0.2):
        img, _ = O.render_f32(v, p, sf, alt, s)
        return float(((img.astype(np.float64) - timg) ** 2).sum())

    img, src = O.render_f32(v, planes, sf, alt, 0.2)
    gcp, gsky = O.render_f32_vjp(v, planes, sf, src, (2.0 * (img - timg)).astype(np.float32))
    h = 1e-2
    cand = np.argwhere((np.abs(gcp) > 1e-3 * np.abs(gcp).max()) & (planes > 5 * h))
    for ch, y, x in cand[rng.choice(len(cand), 20, replace=False)]:
        p = planes.copy()
        p[ch, y, x] += h
        lp = loss(p)
        p[ch, y, x] -= 2 * h
        lm = loss(p)
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(float(gcp[ch, y, x]), rel=2e-2, abs=1e-4), (ch, y, x)
    # the sky term is linear: its gradient is the sky pixels' own sum
    g = 2.0 * (img.astype(np.float64) - timg)
    sky = np.array([144, 144, 224]) / 255.0
    assert gsky == pytest.approx(float((g[:, src < 0] * sky[:, None]).sum()), rel=1e-5)
    fd = (loss(planes, 0.2 + h) - loss(planes, 0.2 - h)) / (2 * h)
    assert gsky == pytest.approx(fd, rel=1e-3)


# ---- the reference's own map data (tests/golden/maps.npz) ---------------

def _golden():
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "maps.npz"))
    return g["c1w_colour"], g["d1_height"], g["flat_height"]


def test_golden_maps_are_packed_like_the_reference_host():
    colour, height, flat = _golden()
    assert colour.shape == height.shape == (256, 256)
    assert np.all(colour >> 24 == 0xFF)
    h = height.view(np.uint32)
    assert np.all(h >> 24 == 0xFF)              # FreeImage packing of the height PNG
    assert not np.any(flat.view(np.uint32) & 0xFF)   # flat_height.png is flat at 0


# D1's heights in this crop span 0..91: low cameras and horizons
GOLDEN_VIEWS = [
    View(x=128.5, y=200.25, height=130.0, angle=0.3, horizon=10.0, distance=300.0, fov=1.2,
         sky_color=0xFF9090E0, h=120, w=160, d=0.004, z0=0.3),
    View(x=40.7, y=31.1, height=140.0, angle=2.9, horizon=-20.0, distance=500.0, fov=1.2,
         sky_color=0xFF9090E0, h=90, w=200, d=0.002, z0=0.5, scale=240.0),
    View(x=250.0, y=128.0, height=110.0, angle=-1.2, horizon=30.0, distance=250.0, fov=0.9,
         sky_color=0xFF9090E0, h=150, w=100, d=0.01, z0=1.0),
]


@pytest.mark.parametrize("vi", range(3))
def test_golden_map_renders_agree(vi):
    colour, height, _ = _golden()
    shaded, alt = O.update_map(colour, height, size=256)
    assert alt.max() <= 255 and alt.min() >= 0
    v = GOLDEN_VIEWS[vi]
    a = O.render_off(v, shaded, alt)
    assert np.array_equal(a, O.render_painter(v, shaded, alt))
    assert (a != v.sky_color).any() and (a == v.sky_color).any()


def test_golden_flat_map_has_no_shadow():
    colour, _, flat = _golden()
    shaded, alt = O.update_map(colour, flat, size=64)
    assert not O.shadow_counts(alt, O.sun(0.1, 0.1), 64).any()


# ---- the state's render: #math mode and #on smoothing -------------------

def test_state_render_modes_reduce_to_the_sample_under_the_camera():
    """render fixes z_0 = 0 (fut/voxel_renderer.fut:103): the first slice is
    at z = 0, so for both modes and both smoothing branches each frame is the
    colour under the camera (the sky when it is 0)."""
    col, hts = scene.scene_maps(5, 64)
    shaded, alt = O.update_map(col, hts, size=64)
    rng = np.random.default_rng(12)
    for _ in range(4):
        h, a = float(rng.uniform(-0.5, 1.2)), float(rng.uniform(-3, 3))
        sv = O.sun(h, a)
        v = View(x=float(rng.uniform(-900, 900)), y=float(rng.uniform(-900, 900)),
                 height=float(rng.uniform(50, 400)), angle=float(rng.uniform(-4, 4)),
                 horizon=float(rng.uniform(-100, 300)), distance=float(rng.uniform(50, 900)),
                 fov=1.2, sky_color=0xFF102030, h=23, w=31)
        mm = O.math_shadowmap(sv)
        for mode in ("png", "math"):
            off = O.render_state(v, mode, False, shaded, alt, sv)
            on = O.render_state(v, mode, True, shaded, alt, sv)
            assert np.array_equal(off, on)
            assert np.all(off == off[0, 0])
            if mode == "math":
                c = mm[int(np.trunc(v.y)) % 1024, int(np.trunc(v.x)) % 1024]
            else:
                c = O.png_color_filtered(shaded, v.x, v.y)
            assert off[0, 0] == (c if c else v.sky_color)
        assert np.array_equal(O.render_state(v, "png", False, shaded, alt, sv),
                              O.render_off(View(**{**v.__dict__, "d": 0.001, "z0": 0.0,
                                                   "scale": 0.0, "n_slices": 0}), shaded, alt))


def test_math_samplers():
    f = np.float32
    for x, y in [(3.5, -7.25), (1023.0, 2.0), (-1023.5, 100.0), (1024.0, 0.5), (0.0, -1024.0)]:
        want = (f(f(f(x) * f(x)) - f(f(y) * f(y))) / f(1024)) if (-1024 < x < 1024 and
                                                                  -1024 < y < 1024) else f(0)
        assert L.fso_expo2(x, y) == pytest.approx(float(want), rel=1e-6, abs=1e-6)
    assert L.fso_function_coloring(64.0, 3.5) == 0xFF000000
    assert L.fso_function_coloring(-128.9, 3.5) == 0xFF000000      # truncation to -128
    assert L.fso_function_coloring(65.0, 3.5) == 0xFFFFFFFF
    # flat regions of expo2 never shadow: a sun straight up counts nothing
    mm = O.math_shadowmap(O.sun(0.0, 0.0))
    assert set(np.unique(mm).tolist()) <= {0xFF000000, 0xFFFFFFFF}


# ---- render2 (fut/voxel_renderer.fut:257-305) ----

def test_render2_inside_terrain_is_uniform():
    """A camera below the terrain hits at dist 0 in every pixel: the frame
    is the bilinear colour under the camera everywhere."""
    col, hts = scene.scene_maps(21, 64)
    alt = np.asarray(hts, dtype=np.int32) & 255
    v = View(x=20.37, y=41.61, height=-5.0, angle=1.1, horizon=30.0, distance=120.0, fov=1.2,
             h=24, w=32)
    img = O.render2(v, col, alt)
    want = O.png_color_filtered(col, np.float32(v.x), np.float32(v.y))
    assert np.all(img == want) and want != 0


def test_render2_hits_on_a_flat_map():
    """Heights 0 everywhere: a pixel is a hit iff some dist in 0 ... n has
    dist * vert_ang > height (f32, vert_ang = horizon/50 + (y/h) * 0.5),
    n = i32(distance / 4) included; hit pixels show the (non-zero) colour
    unless the hit point is integral."""
    rows = cols = 32
    col = np.full((rows, cols), 0xFF3366CC, dtype=np.uint32)
    alt = np.zeros((rows, cols), dtype=np.int32)
    f32 = np.float32
    for hor, height, dist in ((-20.0, 3.0, 101.0), (5.0, 7.5, 60.0), (-40.0, 1.25, 400.0)):
        v = View(x=3.3, y=7.7, height=height, angle=0.35, horizon=hor, distance=dist, fov=1.2,
                 h=48, w=40)
        img = O.render2(v, col, alt)
        n = int(f32(dist) / f32(4.0))
        d = np.arange(n + 1, dtype=np.float32)
        for y in range(v.h):
            va = (f32(hor) / f32(50.0)) + ((f32(y) / f32(v.h)) * f32(0.5))
            hit = bool(np.any((f32(height) - d * va) < f32(0.0)))
            if hit:
                assert np.all(img[y] != 0), (hor, y)
            else:
                assert np.all(img[y] == 0), (hor, y)
