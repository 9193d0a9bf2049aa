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
