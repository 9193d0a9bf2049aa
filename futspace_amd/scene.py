"""Seeded synthetic scenes and the benchmark workloads built from them.

Maps come from libfsscene (toroidal diamond-square heights, height-banded
colours with per-channel noise); cameras from its Catmull-Rom fly-through
or from seeded draws.  Workload definitions:

  plain_1024  800x600 frame over a 1024^2 map (seed 4), camera at
              (512, 512, 180), 1024 slices to distance 1000 at the
              reference's default d = 0.001 (z0 = 0.465 makes get_zs yield
              exactly 1024 slices), scale 240.  Seeds 1-3 are skipped: the
              first, largest diamond-square displacement lands on the map
              centre and puts the camera inside (seeds 1, 2) or right
              behind (seed 3) terrain higher than 180.
  plain_2048  120-frame fly-through over a 2048^2 map, 1920x1080, 2048
              slices, scale f32(w/2).
  batched     256 views over 1024^2, 2048^2, 2048^2 and 4096^2 maps,
              resolutions from {320x240, 640x480, 1280x720, 1920x1080},
              1024-4096 slices.
  diff_2048   1280x720 view of a 2048^2 map, 2048 slices.
  diff_8192   64 views of 3840x2160 over an 8192^2 map with 16-bit
              heights, 4096 slices.

Explicit slice counts n reaching distance D use z0 = D/(2n) and
d = D/(n(n-1)), so that z_n = n*z0 + d*n(n-1)/2 = D.  Generated cameras
keep at least CLEARANCE above the texel under them, as the reference's
terrain_collision keeps the camera out of the ground
(fut/interactive.fut:67-87).
"""
import ctypes
import math

import numpy as np

from . import _native
from ._native import FssPose, FssRng
from .renderer import View, sky_color

SUN_HEIGHT = 0.1
SUN_ANG = 0.1
PLAIN_1024_SEED = 4
PLAIN_2048_SEED = 2
CLEARANCE = 40.0


def above_ground(alt, heights, x, y):
    """max(alt, terrain under (x, y) + CLEARANCE), nearest texel with wrap."""
    rows, cols = heights.shape
    ground = float(heights[int(math.floor(y)) % rows, int(math.floor(x)) % cols])
    return max(alt, ground + CLEARANCE)


def heightmap(seed, size, roughness=0.6, max_value=255):
    out = np.empty((size, size), dtype=np.int32)
    rc = _native.scene().fss_heightmap(seed, size, roughness, max_value, out.ctypes.data)
    if rc:
        raise ValueError(f"heightmap({seed}, {size}) failed")
    return out


def colormap(seed, heights, max_value=255, noise=12):
    h, w = heights.shape
    out = np.empty((h, w), dtype=np.uint32)
    hh = np.ascontiguousarray(heights, dtype=np.int32)
    rc = _native.scene().fss_colormap(seed, hh.ctypes.data, h, w, max_value, noise, out.ctypes.data)
    if rc:
        raise ValueError("colormap failed")
    return out


def scene_maps(seed, size, max_value=255):
    hts = heightmap(seed, size, 0.6, max_value)
    return colormap(seed, hts, max_value, 12), hts


def flythrough(seed, n, map_size, alt_lo=120.0, alt_hi=400.0, horizon_mid=150.0,
               horizon_amp=30.0, n_ctrl=8):
    poses = (FssPose * n)()
    rc = _native.scene().fss_flythrough(seed, n, n_ctrl, map_size, alt_lo, alt_hi, horizon_mid,
                                        horizon_amp, poses)
    if rc:
        raise ValueError("flythrough failed")
    return [(p.x, p.y, p.height, p.angle, p.horizon) for p in poses]


class Rng:
    def __init__(self, seed, stream=0):
        self.r = FssRng()
        self.lib = _native.scene()
        self.lib.fss_rng_seed(ctypes.byref(self.r), seed, stream)

    def unit(self):
        return self.lib.fss_rng_unit(ctypes.byref(self.r))

    def int(self, lo, hi):
        return self.lib.fss_rng_int(ctypes.byref(self.r), lo, hi)

    def uniform(self, lo, hi):
        return lo + (hi - lo) * self.unit()


def depth_for(n, distance=1000.0):
    """(d, z0) putting n slices with get_zs' term formula out to `distance`."""
    z0 = float(np.float32(distance / (2.0 * n)))
    d = float(np.float32(distance / (n * (n - 1.0))))
    return d, z0


def plain_1024_maps():
    return scene_maps(PLAIN_1024_SEED, 1024)


def plain_2048_maps():
    return scene_maps(PLAIN_2048_SEED, 2048)


def plain_1024_view():
    return View(x=512.0, y=512.0, height=180.0, angle=0.0, horizon=150.0, distance=1000.0,
                fov=1.2, sky_color=sky_color(SUN_HEIGHT), h=600, w=800, d=0.001, z0=0.465,
                scale=240.0, n_slices=0)


def plain_2048_views(heights, n_frames=120, h=1080, w=1920, slices=2048, seed=PLAIN_2048_SEED):
    d, z0 = depth_for(slices)
    sky = sky_color(SUN_HEIGHT)
    views = []
    for x, y, alt, yaw, hor in flythrough(seed, n_frames, heights.shape[0]):
        alt = float(np.float32(above_ground(alt, heights, x, y)))
        views.append(View(x=x, y=y, height=alt, angle=yaw, horizon=hor, distance=1000.0,
                          fov=1.2, sky_color=sky, h=h, w=w, d=d, z0=z0, scale=0.0,
                          n_slices=slices))
    return views


BATCH_MAP_SIZES = (1024, 2048, 2048, 4096)
BATCH_MAP_SEEDS = (3, 4, 5, 6)
BATCH_RESOLUTIONS = ((240, 320), (480, 640), (720, 1280), (1080, 1920))


def batch_maps():
    """The batched workload's four (colour, height) maps: 1024^2, 2048^2,
    2048^2 and 4096^2, seeds 3-6."""
    return [scene_maps(s, n) for s, n in zip(BATCH_MAP_SEEDS, BATCH_MAP_SIZES)]


def batched_views(heights_list, n_views=256, seed=7):
    """heights_list: the (masked) heightmaps of the batch's maps."""
    map_sizes = [h.shape[0] for h in heights_list]
    r = Rng(seed, 7)
    sky = sky_color(SUN_HEIGHT)
    views = []
    for _ in range(n_views):
        m = r.int(0, len(map_sizes) - 1)
        h, w = BATCH_RESOLUTIONS[r.int(0, len(BATCH_RESOLUTIONS) - 1)]
        size = map_sizes[m]
        x = r.uniform(0.0, size)
        y = r.uniform(0.0, size)
        alt = r.uniform(100.0, 500.0)
        yaw = r.uniform(0.0, 2.0 * math.pi)
        hor = r.uniform(100.0, 300.0)
        slices = r.int(1024, 4096)
        alt = above_ground(alt, heights_list[m], x, y)
        d, z0 = depth_for(slices)
        views.append(View(x=float(np.float32(x)), y=float(np.float32(y)),
                          height=float(np.float32(alt)), angle=float(np.float32(yaw)),
                          horizon=float(np.float32(hor)), distance=1000.0, fov=1.2,
                          sky_color=sky, h=h, w=w, d=d, z0=z0, scale=0.0, n_slices=slices,
                          map=m))
    return views


def batch_offsets(views):
    offs, total = [], 0
    for v in views:
        offs.append(total)
        total += v.h * v.w
    return np.asarray(offs, dtype=np.int64), total


DIFF_2048_TARGET_SEED = 8
DIFF_2048_TEXEL_SEED = 9
DIFF_8192_SEED = 10
DIFF_8192_MAX_HEIGHT = 1024


def diff_2048_maps():
    """The differentiation workload's 2048^2 map: the plain 2048^2 pair."""
    return plain_2048_maps()


def diff_2048_view(heights):
    d, z0 = depth_for(2048)
    alt = above_ground(250.0, heights, 1024.0, 1024.0)
    return View(x=1024.0, y=1024.0, height=float(np.float32(alt)), angle=0.7, horizon=150.0,
                distance=1000.0, fov=1.2, sky_color=sky_color(SUN_HEIGHT), h=720, w=1280, d=d,
                z0=z0, scale=0.0, n_slices=2048)


def diff_8192_maps():
    """8192^2 16-bit heights (0..1024) and their colour map, seed 10."""
    hts = heightmap(DIFF_8192_SEED, 8192, 0.6, DIFF_8192_MAX_HEIGHT)
    return colormap(DIFF_8192_SEED, hts, DIFF_8192_MAX_HEIGHT, 12), hts


def perturbed_planes(planes, seed, sigma=8.0):
    """Colour planes plus seeded N(0, sigma^2) noise per channel (in byte
    units), clipped to [0, 1]."""
    rng = np.random.default_rng(seed)
    noisy = planes + rng.normal(0.0, sigma / 255.0, planes.shape).astype(np.float32)
    return np.clip(noisy, 0.0, 1.0).astype(np.float32)


def diff_8192_views(heights, n_views=64, seed=10):
    map_size = heights.shape[0]
    r = Rng(seed, 10)
    d, z0 = depth_for(4096)
    sky = sky_color(SUN_HEIGHT)
    views = []
    for i in range(n_views):
        t = 2.0 * math.pi * i / n_views
        rad = 0.25 * map_size
        x = 0.5 * map_size + rad * math.cos(t)
        y = 0.5 * map_size + rad * math.sin(t)
        alt = above_ground(r.uniform(300.0, 800.0), heights, x, y)
        # look toward the ring's centre: forward is (-sin a, -cos a) in map
        # coordinates (fut/voxel_renderer.fut:47-50)
        yaw = math.atan2(x - 0.5 * map_size, y - 0.5 * map_size)
        views.append(View(x=float(np.float32(x)), y=float(np.float32(y)),
                          height=float(np.float32(alt)), angle=float(np.float32(yaw)),
                          horizon=150.0, distance=1000.0, fov=1.2, sky_color=sky, h=2160,
                          w=3840, d=d, z0=z0, scale=0.0, n_slices=4096))
    return views


def algorithmic_bytes(view, height_bytes=1):
    """B = 4*b_h*W*S + (4*4 + 4)*W*H bytes of one view: four height texels
    per (slice, column) sample, at most one visible sample (four colour
    texels) per pixel, one u32 write per pixel."""
    if view.n_slices > 0:
        n = view.n_slices
    else:
        n = _zs_count(view.distance, view.d, view.z0)
    return 4 * height_bytes * view.w * n + 20 * view.w * view.h


def _zs_count(c, d, z0):
    f = np.float32
    sq = np.sqrt(f(f(d) - f(2.0) * f(z0)) ** 2 + f(f(8.0) * f(d)) * f(c), dtype=np.float32)
    num = f(f(sq - f(2.0) * f(z0)) + f(d))
    return int(np.floor(f(num / f(f(2.0) * f(d)))))
