"""Host-side logic without a GPU: the autotuner's choice and cache, the
seeded scene generator, the workloads' closed-form sizes, and bench.py's
single-GPU guard."""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from futspace_amd import View, autotune, scene
from futspace_amd._native import FsxView

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class StubCtx:
    """Stands in for a Context: kernel time depends on the tuning."""

    def __init__(self, cost):
        self.cost = cost
        self.shape = (0, 0, 0)
        self.runs = 0
        self.enabled = False
        self.ms = 0.0

    def set_tuning(self, *shape):
        self.shape = shape

    def sync(self):
        pass

    def profile(self, on=True):
        self.enabled = on

    def profile_read(self):
        ms, n = self.ms, self.runs
        self.ms, self.runs = 0.0, 0
        return ms, n

    def run(self):
        if self.shape == (4, 2, 1024):
            raise RuntimeError("unsupported shape")
        if self.enabled:
            self.ms += self.cost[self.shape]
            self.runs += 1


def test_autotuner_picks_the_fastest_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("FSX_TUNING_CACHE", str(tmp_path / "t.json"))
    cost = {c: 10.0 + i for i, c in enumerate(autotune.CANDIDATES)}
    cost[(4, 1, 2048)] = 3.0
    ctx = StubCtx(cost)
    best, timings = autotune.tune(ctx, lambda: ctx.run, "w1", reps=3)
    assert best == (4, 1, 2048) and ctx.shape == best
    assert (4, 2, 1024) not in timings          # failing shapes are skipped
    assert timings[autotune.CANDIDATES[0]] == pytest.approx(10.0)
    cached = json.load(open(tmp_path / "t.json"))
    assert any(k.endswith("|w1") for k in cached)
    ctx2 = StubCtx({c: 1.0 for c in autotune.CANDIDATES})
    best2, _ = autotune.tune(ctx2, lambda: ctx2.run, "w1", reps=3)
    assert best2 == best and ctx2.runs == 0     # served from the cache


def test_heightmap_is_seeded_and_in_range():
    a = scene.heightmap(3, 256)
    b = scene.heightmap(3, 256)
    c = scene.heightmap(4, 256)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() == 0 and a.max() == 255
    h16 = scene.heightmap(10, 128, 0.6, 1024)
    assert h16.min() == 0 and h16.max() == 1024


def test_colormap_is_opaque_and_height_banded():
    col, hts = scene.scene_maps(5, 256)
    assert np.all(col >> 24 == 0xFF)
    assert np.array_equal(col, scene.colormap(5, hts))
    # water (blue-dominant) at the lowest heights, snow (bright) at the top
    low, high = col[hts < 40], col[hts > 230]
    assert np.mean(low & 255) > np.mean((low >> 16) & 255)
    assert np.mean((high >> 16) & 255) > 180


def test_workload_cameras_stay_above_ground():
    col, hts = scene.plain_2048_maps()
    for v in scene.plain_2048_views(hts):
        g = hts[int(np.floor(v.y)) % 2048, int(np.floor(v.x)) % 2048]
        assert v.height >= g + scene.CLEARANCE - 1e-3
        assert (v.h, v.w, v.n_slices) == (1080, 1920, 2048)
    data = scene.batch_maps()
    views = scene.batched_views([h for _, h in data])
    assert len(views) == 256 and {v.map for v in views} == {0, 1, 2, 3}
    assert {(v.h, v.w) for v in views} == set(scene.BATCH_RESOLUTIONS)
    assert all(1024 <= v.n_slices <= 4096 for v in views)
    again = scene.batched_views([h for _, h in data])
    assert [v.to_c().x for v in again] == [v.to_c().x for v in views]


def test_depth_for_reaches_the_far_distance():
    for n in (1024, 2048, 4096):
        d, z0 = scene.depth_for(n)
        zn = n * z0 + d * n * (n - 1) / 2
        assert zn == pytest.approx(1000.0, rel=1e-5)


def test_algorithmic_bytes_closed_form():
    """B = 4*b_h*W*S + 20*W*H: 57,200,640 bytes for a 1080p frame with 2048
    slices, 12,902,400 at 800x600 / 1024 slices."""
    v = View(x=0, y=0, height=0, angle=0, horizon=0, h=1080, w=1920, n_slices=2048)
    assert scene.algorithmic_bytes(v, 1) == 57_200_640
    v = View(x=0, y=0, height=0, angle=0, horizon=0, h=600, w=800, n_slices=1024)
    assert scene.algorithmic_bytes(v, 1) == 3_276_800 + 9_600_000
    v = View(x=0, y=0, height=0, angle=0, horizon=0, h=2160, w=3840, n_slices=4096)
    assert scene.algorithmic_bytes(v, 2) == 4 * 2 * 3840 * 4096 + 20 * 3840 * 2160


def test_plain_1024_view_derives_1024_slices():
    v = scene.plain_1024_view()
    assert v.n_slices == 0 and scene._zs_count(v.distance, v.d, v.z0) == 1024


def test_view_struct_layout():
    assert ctypes.sizeof(FsxView) == 64
    v = View(x=1.5, y=2.5, height=3.0, angle=0.25, horizon=4.0, distance=5.0, fov=1.1,
             sky_color=0xFF9090E0, h=7, w=9, d=0.5, z0=0.75, scale=2.0, n_slices=11, map=3)
    c = v.to_c()
    assert (c.x, c.y, c.h, c.w, c.n_slices, c.map, c.sky_color) == (
        1.5, 2.5, 7, 9, 11, 3, 0xFF9090E0)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


class FskFrame(ctypes.Structure):
    """fsk_frame of futspace_amd/csrc/fs_internal.h (host-side layout)."""
    _fields_ = [(n, ctypes.c_float) for n in
                ("cam_x", "cam_y", "cam_h", "horizon", "Lx", "Ly", "Rx", "Ry", "d", "z0",
                 "scale", "fw")] + [("n", ctypes.c_int32), ("rows", ctypes.c_int32),
                                    ("cols", ctypes.c_int32), ("sky", ctypes.c_uint32),
                                    ("tame", ctypes.c_int32)]


def _tame_fn():
    lib = ctypes.CDLL(os.path.join(ROOT, "futspace_amd", "libfutspace.so"))
    fn = lib.fsk_frame_tame
    fn.restype = ctypes.c_int32
    fn.argtypes = [ctypes.POINTER(FskFrame)]
    return fn


def _frame(v):
    """The host's frame_of_view in f32 (glibc sinf/cosf differ from numpy's
    by at most an ulp, which the predicate's margins absorb)."""
    f32 = np.float32
    s, c, fov = f32(np.sin(f32(v.angle))), f32(np.cos(f32(v.angle))), f32(v.fov)
    f = FskFrame()
    f.cam_x, f.cam_y, f.cam_h, f.horizon = v.x, v.y, v.height, v.horizon
    f.Lx, f.Ly = (-c) - s * fov, s - c * fov
    f.Rx, f.Ry = c - s * fov, (-s) - c * fov
    f.d, f.z0 = v.d, v.z0
    f.scale = v.scale if v.scale > 0 else float(v.w // 2)
    f.fw = float(v.w)
    f.n = v.n_slices if v.n_slices > 0 else scene._zs_count(v.distance, v.d, v.z0)
    f.rows, f.cols = v.h, v.w
    return f


def _extremes(f):
    """max |x|, |y| over every sample the kernel evaluates (columns up to
    w + 7) and max |rel| over heights 0..65535, in f32 as the kernel forms
    them."""
    f32 = np.float32
    i = np.arange(1, f.n + 1, dtype=np.float32)
    z = (i / f32(2)) * ((f32(2) * f32(f.z0)) + ((i - f32(1)) * f32(f.d)))
    lx, ly, rx, ry = f32(f.Lx) * z, f32(f.Ly) * z, f32(f.Rx) * z, f32(f.Ry) * z
    dx, dy = (rx - lx) / f32(f.fw), (ry - ly) / f32(f.fw)
    sx, sy = lx + f32(f.cam_x), ly + f32(f.cam_y)
    col = np.arange(0, int(f.fw) + 8, dtype=np.float32)[None, :]
    x = sx[:, None] + col * dx[:, None]
    y = sy[:, None] + col * dy[:, None]
    iz = (f32(1) / z) * f32(f.scale)
    rel = [((f32(f.cam_h) - f32(h)) * iz) + f32(f.horizon) for h in (0.0, 65535.0)]
    return (float(np.abs(x).max()), float(np.abs(y).max()),
            float(max(np.abs(r).max() for r in rel)))


def test_tame_frames_of_the_workloads():
    tame = _tame_fn()
    _, hts = scene.plain_2048_maps()
    for v in scene.plain_2048_views(hts)[::10]:
        assert tame(ctypes.byref(_frame(v))) == 1
    assert tame(ctypes.byref(_frame(scene.plain_1024_view()))) == 1


def test_untame_frames():
    tame = _tame_fn()
    base = View(x=512.0, y=512.0, height=180.0, angle=0.3, horizon=150.0, distance=1000.0,
                fov=1.2, h=600, w=800, d=0.001, z0=0.465, scale=240.0)
    assert tame(ctypes.byref(_frame(base))) == 1
    for change in (dict(z0=0.0), dict(d=-1e-4), dict(x=float("nan")), dict(height=float("inf")),
                   dict(x=5e6), dict(y=-5e6), dict(horizon=2e9), dict(scale=1e9),
                   dict(z0=1e-7)):
        f = _frame(View(**{**base.__dict__, **change}))
        assert tame(ctypes.byref(f)) == 0, change


def test_tame_bound_holds_by_brute_force():
    """Random frames near the bounds: whenever the predicate says tame,
    every sample coordinate is below 2^22 and every row below 2^30 in
    magnitude, evaluated in f32 over all slices and columns."""
    tame = _tame_fn()
    rng = np.random.default_rng(31)
    accepted = 0
    for _ in range(120):
        v = View(x=float(rng.choice([1.0, -1.0]) * rng.uniform(3.0e6, 4.3e6)),
                 y=float(rng.uniform(-4e6, 4e6)), height=float(rng.uniform(-1e4, 1e4)),
                 angle=float(rng.uniform(0, 6.3)), horizon=float(rng.uniform(-1e8, 1e8)),
                 distance=1000.0, fov=float(rng.uniform(0.5, 1.5)), h=64, w=int(rng.integers(1, 64)),
                 d=float(rng.uniform(0.0, 0.05)), z0=float(10 ** rng.uniform(-3, 1)),
                 scale=float(10 ** rng.uniform(0, 3.5)), n_slices=int(rng.integers(1, 300)))
        f = _frame(v)
        if tame(ctypes.byref(f)):
            accepted += 1
            mx, my, mr = _extremes(f)
            assert mx < 2.0 ** 22 and my < 2.0 ** 22 and mr < 2.0 ** 30, (v, mx, my, mr)
    assert accepted >= 10
