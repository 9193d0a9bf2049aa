"""The benchmark's own workloads at full size, bit for bit (needs a GPU):
every frame of the 1920x1080 fly-through over the 2048^2 map and every
view of the 256-view batch over the 1024^2 / 2048^2 / 2048^2 / 4096^2
maps, against the oracle's multi-threaded painter, which the oracle tests
hold bit-equal to its scan / scatter / fill restatement of the reference."""
import numpy as np
import pytest

from futspace_amd import scene
from oracle import fs_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _differences(pairs):
    bad = []
    for i, (got, want) in enumerate(pairs):
        if got.shape != want.shape or not np.array_equal(got, want):
            bad.append((i, int(np.count_nonzero(got != want)) if got.shape == want.shape else -1))
    return bad


def test_every_frame_of_the_1080p_flythrough(ctx):
    col, hts = scene.plain_2048_maps()
    m = ctx.map(col, hts)
    shaded, alt = O.update_map(col, hts)
    assert np.array_equal(m.shaded(), shaded)
    views = scene.plain_2048_views(hts)
    assert len(views) == 120
    out = torch.empty((views[0].h, views[0].w), dtype=torch.int32, device="cuda")

    def pairs():
        for v in views:
            ctx.render(m, v, out)
            ctx.sync()      # the render runs on the context's stream, not torch's
            yield out.cpu().numpy().view(np.uint32), O.render_painter_mt(v, shaded, alt)

    bad = _differences(pairs())
    assert not bad, f"frames that differ (index, pixels): {bad[:8]}"


def test_every_view_of_the_256_view_batch(ctx):
    data = scene.batch_maps()
    maps = [ctx.map(c, h) for c, h in data]
    views = scene.batched_views([h for _, h in data])
    assert len(views) == 256
    offs, total = scene.batch_offsets(views)
    out = torch.zeros(total, dtype=torch.int32, device="cuda")
    b = ctx.batch(maps, views, out, offs)
    try:
        b.run()
        ctx.sync()
    finally:
        b.close()
    flat = out.cpu().numpy().view(np.uint32)
    del out
    refs = [O.update_map(c, h) for c, h in data]

    def pairs():
        for v, o in zip(views, offs):
            shaded, alt = refs[v.map]
            yield (flat[o:o + v.h * v.w].reshape(v.h, v.w),
                   O.render_painter_mt(v, shaded, alt))

    bad = _differences(pairs())
    assert not bad, f"views that differ (index, pixels): {bad[:8]}"
