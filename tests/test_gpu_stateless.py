"""The stateless program of tests/futspace.fut on the GPU (needs a GPU):
bit-exact against the oracle's restatement."""
import os
import subprocess
import sys

import numpy as np
import pytest

from futspace_amd import View, futdata
from futspace_amd.renderer import STATELESS_MAIN
from oracle import fs_oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _oracle(view, colour, height):
    cam = (view.x, view.y, view.height, view.angle, view.horizon, view.distance, view.fov)
    return O.render_nearest(cam, colour, height, view.h, view.w, view.sky_color & 0xFFFFFFFF)


def _maps(seed, q, r):
    rng = np.random.default_rng(seed)
    colour = rng.integers(-(2 ** 31), 2 ** 31 - 1, (q, r)).astype(np.int32)
    colour[rng.random((q, r)) < 0.15] = np.uint32(0xFF9090E0).view(np.int32)   # sky: empty
    colour[rng.random((q, r)) < 0.05] = 0                                       # solid
    height = rng.integers(-50, 300, (q, r)).astype(np.int32)
    return colour, height


@pytest.mark.parametrize("seed,q,r", [(1, 64, 64), (2, 96, 80), (3, 128, 256), (4, 3, 5)])
def test_random_maps_bitexact(ctx, seed, q, r):
    colour, height = _maps(seed, q, r)
    rng = np.random.default_rng(seed + 100)
    for _ in range(3):
        v = View(x=float(rng.uniform(-500, 500)), y=float(rng.uniform(-500, 500)),
                 height=float(rng.uniform(60, 400)), angle=float(rng.uniform(-7, 7)),
                 horizon=float(rng.uniform(-50, 250)), distance=float(rng.uniform(30, 900)),
                 fov=float(rng.uniform(0.5, 1.6)), sky_color=0xFF9090E0,
                 h=int(rng.integers(1, 300)), w=int(rng.integers(1, 300)))
        got = ctx.render_stateless(colour, height, v)
        want = _oracle(v, colour, height)
        assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} pixels differ"


def test_main_on_the_reference_maps(ctx):
    """`main` (400 x 800) on the C1W / D1 crop packed as tools/png2data.py
    packs maps (heights are not masked by this program)."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "maps.npz"))
    colour = (g["c1w_colour"] & 0xFFFFFF).astype(np.int32)
    height = (g["d1_height"].view(np.uint32) & 0xFFFFFF).astype(np.int32)
    got = ctx.render_stateless(colour, height)
    assert got.shape == (400, 800)
    assert np.array_equal(got, _oracle(STATELESS_MAIN, colour, height))
    # with the low byte as height the camera at 78 sees terrain and sky
    low = (height & 0xFF).astype(np.int32)
    got = ctx.render_stateless(colour, low)
    want = _oracle(STATELESS_MAIN, colour, low)
    assert np.array_equal(got, want)
    sky = np.uint32(0xFF9090E0).view(np.int32)
    assert (want == sky).any() and (want != sky).any()


def test_command_line_like_futspace_b(tmp_path):
    """cat colour.in height.in | python -m futspace_amd.futdata render >
    frame.data, as tests/Makefile:3 runs ./futspace -b."""
    colour, height = _maps(7, 64, 64)
    inp = tmp_path / "maps.in"
    with open(inp, "wb") as f:
        futdata.write_value(f, colour)
        futdata.write_value(f, height)
    with open(inp, "rb") as f:
        r = subprocess.run([sys.executable, "-m", "futspace_amd.futdata", "render"], stdin=f,
                           capture_output=True, cwd=ROOT, timeout=300)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    import io
    frame = futdata.read_value(io.BytesIO(r.stdout))
    assert frame.dtype == np.int32 and frame.shape == (400, 800)
    assert np.array_equal(frame, _oracle(STATELESS_MAIN, colour, height))


def test_bench_sweep_runs(ctx):
    """`python -m futspace_amd.futdata bench`: every case of the sweep
    (depth, columns and rows from 10 to 10^6) times a frame."""
    import io
    import json
    buf = io.StringIO()
    futdata.bench_stateless(reps=1, out=buf)
    recs = [json.loads(line) for line in buf.getvalue().splitlines()]
    assert len(recs) == 3 * len(futdata.SWEEP)
    for r in recs:
        assert r.get("ms_per_frame", 0) > 0, r


def test_frames_taller_than_a_tile(ctx):
    """The stateless program at 9000 rows (the column kernel), with
    sky-coloured (transparent) texels: bit-exact."""
    colour, height = _maps(5, 128, 128)
    v = View(x=40.0, y=-30.0, height=260.0, angle=0.7, horizon=3000.0, distance=300.0, fov=1.1,
             sky_color=0xFF9090E0, h=9000, w=11)
    assert np.array_equal(ctx.render_stateless(colour, height, v), _oracle(v, colour, height))
