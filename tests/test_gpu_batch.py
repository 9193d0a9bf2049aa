"""Batched persistent render and its HIP-graph replay (needs a GPU).

Every view of a batch must equal the plain single-view render, which the
render tests hold bit-exact against the oracle; graph replays must equal
the eager launch on every replay.
"""
import numpy as np
import pytest

from futspace_amd import View, autotune, scene, sky_color
from oracle import fs_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _views(n, maps_shapes, seed=5, max_h=300):
    rng = np.random.default_rng(seed)
    sky = sky_color(0.1)
    vs = []
    for i in range(n):
        m = int(rng.integers(len(maps_shapes)))
        rows, cols = maps_shapes[m]
        h = int(rng.integers(1, max_h))
        w = int(rng.integers(1, 301))
        nsl = int(rng.choice([0, 1, 7, 300, 2047, 2048, 2049, 4500]))
        vs.append(View(x=float(rng.uniform(0, cols)), y=float(rng.uniform(0, rows)),
                       height=float(rng.uniform(150, 450)), angle=float(rng.uniform(-7, 7)),
                       horizon=float(rng.uniform(-50, 300)), distance=float(rng.uniform(50, 900)),
                       fov=float(rng.uniform(0.6, 1.6)), sky_color=sky, h=h, w=w,
                       d=float(rng.uniform(0.0005, 0.01)), z0=float(rng.uniform(0.05, 2.0)),
                       scale=0.0, n_slices=nsl, map=m))
    return vs


def _run_batch(ctx, maps, views):
    offs, total = scene.batch_offsets(views)
    out = torch.zeros(max(total, 1), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    ctx.render_batch(maps, views, out, offs)
    ctx.sync()
    return out.cpu().numpy().view(np.uint32), offs


def _check_views(ctx, maps, views, flat, offs):
    for i, (v, o) in enumerate(zip(views, offs)):
        got = flat[o:o + v.h * v.w].reshape(v.h, v.w)
        want = ctx.render(maps[v.map], v)
        bad = int(np.count_nonzero(got != want))
        assert bad == 0, f"view {i} ({v.h}x{v.w}, {v.n_slices} slices): {bad} pixels differ"


@pytest.fixture(scope="module")
def pow2_maps(ctx):
    data = [scene.scene_maps(s, n) for s, n in ((3, 256), (4, 128), (5, 512))]
    return [ctx.map(c, h) for c, h in data], [h.shape for _, h in data]


@pytest.mark.parametrize("shape", [(0, 0, 0), (8, 1, 2048), (8, 1, 512), (8, 2, 512),
                                   (8, 2, 256), (4, 2, 1024), (4, 1, 1024)])
def test_batch_matches_single_views(ctx, pow2_maps, shape):
    maps, shapes = pow2_maps
    ctx.set_tuning(*shape)
    try:
        views = _views(40, shapes)
        flat, offs = _run_batch(ctx, maps, views)
        _check_views(ctx, maps, views, flat, offs)
    finally:
        ctx.set_tuning(0, 0, 0)


def test_batch_generic_maps_and_oracle(ctx):
    """Non-power-of-two maps take the generic kernel; a few views are also
    checked straight against the oracle."""
    data = [(np.uint32(0xFF000000) | (np.arange(96 * 80, dtype=np.uint32).reshape(96, 80)
                                       * 2654435761 % 0xFFFFFF)).astype(np.uint32),
            ((np.arange(96 * 80, dtype=np.int32).reshape(96, 80) * 37) % 256)]
    col2, hts2 = scene.scene_maps(6, 64)
    maps = [ctx.map(data[0], data[1]), ctx.map(col2, hts2)]
    views = _views(24, [(96, 80), (64, 64)], seed=9, max_h=150)
    flat, offs = _run_batch(ctx, maps, views)
    _check_views(ctx, maps, views, flat, offs)
    refs = [O.update_map(data[0], data[1]), O.update_map(col2, hts2)]
    for v, o in list(zip(views, offs))[:6]:
        shaded, alt = refs[v.map]
        got = flat[o:o + v.h * v.w].reshape(v.h, v.w)
        assert np.array_equal(got, O.render_off(v, shaded, alt))


def test_batch_graph_replays_match_eager(ctx, pow2_maps):
    maps, shapes = pow2_maps
    views = _views(30, shapes, seed=11)
    offs, total = scene.batch_offsets(views)
    out = torch.zeros(total, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    ctx.set_stream(stream)
    try:
        b = ctx.batch(maps, views, out, offs)
        with torch.cuda.stream(stream):
            b.run()
        torch.cuda.synchronize()
        eager = out.clone()
        b.capture()
        for rep in range(4):
            with torch.cuda.stream(stream):
                out.zero_()
            torch.cuda.synchronize()
            b.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, eager), f"replay {rep} differs from the eager launch"
        b.close()
    finally:
        ctx.set_stream(None)


def test_autotuner_picks_a_candidate(ctx, pow2_maps, tmp_path, monkeypatch):
    maps, shapes = pow2_maps
    monkeypatch.setenv("FSX_TUNING_CACHE", str(tmp_path / "tune.json"))
    views = _views(16, shapes, seed=13)
    offs, total = scene.batch_offsets(views)
    out = torch.zeros(total, dtype=torch.int32, device="cuda")

    def make_run():
        b = ctx.batch(maps, views, out, offs)
        return b.run

    try:
        best, timings = autotune.tune(ctx, make_run, "test-batch", reps=2)
        assert best in autotune.CANDIDATES and timings
        again, _ = autotune.tune(ctx, make_run, "test-batch", reps=2)
        assert again == best
    finally:
        ctx.set_tuning(0, 0, 0)


def test_batch_with_a_tall_view(ctx, pow2_maps):
    """One 2400-row view among small ones: the batch's shape follows its
    tallest view (4-column tiles); every view still equals its single
    render and the tall one equals the oracle."""
    maps, shapes = pow2_maps
    views = _views(12, shapes, seed=13, max_h=120)
    views.append(View(x=100.25, y=60.5, height=300.0, angle=2.4, horizon=900.0, distance=500.0,
                      fov=1.2, sky_color=0xFF9090E0, h=2400, w=40, d=0.003, z0=0.5,
                      scale=800.0, map=0))
    flat, offs = _run_batch(ctx, maps, views)
    _check_views(ctx, maps, views, flat, offs)
    col, hts = scene.scene_maps(3, 256)
    shaded, alt = O.update_map(col, hts)
    v, o = views[-1], offs[-1]
    assert np.array_equal(flat[o:o + v.h * v.w].reshape(v.h, v.w), O.render_off(v, shaded, alt))


def test_batch_sixteen_bit_heights(ctx):
    """Maps with 16-bit heights (power-of-two and not) in one batch: every
    view equals its single-view render."""
    maps, shapes = [], []
    for seed, size in ((21, 256), (22, 96)):
        hts = scene.heightmap(seed, 256, 0.6, 900)[:size, :size].copy()
        col = scene.colormap(seed, hts, 900)
        maps.append(ctx.map(col, hts, height_bits=16))
        shapes.append(hts.shape)
    views = _views(20, shapes, seed=23, max_h=200)
    for v in views:
        v.height += 600.0
    flat, offs = _run_batch(ctx, maps, views)
    _check_views(ctx, maps, views, flat, offs)


def test_batch_refuses_views_taller_than_a_tile(ctx, pow2_maps):
    """Batched views are tile renders: a view of more than 7680 rows is
    refused before anything runs (single renders of it take the column
    kernel instead)."""
    maps, _ = pow2_maps
    v = View(x=100.25, y=60.5, height=300.0, angle=2.4, horizon=900.0, distance=500.0, fov=1.2,
             sky_color=0xFF9090E0, h=7681, w=8, d=0.003, z0=0.5, map=0)
    out = torch.empty(v.h * v.w, dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError):
        ctx.batch(maps, [v], out, np.array([0], dtype=np.int64))
