"""Empty and degenerate inputs through the C-ABI (needs a GPU): frames
with no rows or no columns render nothing and touch no memory, the
state machine returns [0][w] / [h][0] frames, the differentiable entries
overwrite (or keep, when accumulating) their outputs, and batches refuse
empty views and empty view lists with a message."""
import numpy as np
import pytest

from futspace_amd import FutspaceError, Session, View, scene

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from futspace_amd.diff import DiffRenderer, planes_of  # noqa: E402

SHAPES = [(0, 16), (16, 0), (0, 0)]


@pytest.fixture(scope="module")
def small(ctx):
    col, hts = scene.scene_maps(5, 128)
    return ctx.map(col, hts, shade_size=128), col, hts


def _view(h, w):
    return View(x=64.3, y=100.7, height=260.0, angle=0.4, horizon=60.0, distance=300.0,
                fov=1.2, h=h, w=w, d=0.01, z0=0.5)


@pytest.mark.parametrize("h,w", SHAPES)
def test_empty_frames_write_nothing(ctx, small, h, w):
    m = small[0]
    v = _view(h, w)
    assert ctx.render(m, v).shape == (h, w)
    base = torch.full((4, 32), 7, dtype=torch.int32, device="cuda")
    out = base[:h, :w] if h == 0 else base[:, :0]
    v = _view(out.shape[0], out.shape[1])
    ctx.render(m, v, out)
    ctx.render2(m, v, out)
    ctx.sync()
    assert bool((base == 7).all())


@pytest.mark.parametrize("h,w", SHAPES)
def test_state_machine_empty_frames(ctx, small, h, w):
    _, col, hts = small
    s = Session(ctx)
    try:
        s.update_map(col, hts)
        s.step()
        s.resize(h, w)
        assert s.render().shape == (h, w)
    finally:
        s.close()


@pytest.mark.parametrize("accumulate", [False, True])
@pytest.mark.parametrize("h,w", SHAPES)
def test_differentiable_entries_on_empty_views(ctx, small, h, w, accumulate):
    m, col, _ = small
    r = DiffRenderer(ctx, m)
    tp = torch.from_numpy(planes_of(col)).cuda()
    v = _view(h, w)
    img, src = r.forward_raw(tp, 0.3, v)
    assert tuple(img.shape) == (3, h, w) and tuple(src.shape) == (h, w)
    gp = torch.full_like(tp, 9.0)
    gs = torch.full((1,), 4.0, device="cuda")
    r.vjp_raw(tp, src, torch.empty((3, h, w), device="cuda"), v, gp, gs, accumulate=accumulate)
    torch.cuda.synchronize()
    fill = 9.0 if accumulate else 0.0
    assert bool((gp == fill).all()) and float(gs[0]) == (4.0 if accumulate else 0.0)
    gp.fill_(9.0)
    gs.fill_(4.0)
    loss = torch.full((1,), 2.0, device="cuda")
    r.l2_step(tp, 0.3, v, torch.empty((3, h, w), device="cuda"), gplanes=gp, gsky=gs,
              loss=loss, accumulate=accumulate)
    torch.cuda.synchronize()
    assert bool((gp == fill).all())
    assert float(gs[0]) == (4.0 if accumulate else 0.0)
    assert float(loss[0]) == (2.0 if accumulate else 0.0)


def test_batches_refuse_empty_views(ctx, small):
    m = small[0]
    out = torch.zeros(64 * 64, dtype=torch.int32, device="cuda")
    with pytest.raises(FutspaceError, match="unsupported size"):
        ctx.render_batch([m], [_view(64, 64), _view(0, 64)], out, [0, 0])
    with pytest.raises(FutspaceError, match="no views"):
        ctx.render_batch([m], [], out, [])


def test_invalid_arguments_leave_a_message(ctx, small):
    """Bad shapes, a pitch below the width and null arguments fail with a
    message for futhark_context_get_error; the context stays usable."""
    import ctypes
    lib, c, m = ctx.lib, ctx.ptr, small[0]
    assert not lib.futhark_new_u32_2d(c, None, -1, 4)
    assert "invalid array shape" in ctx.error()
    assert not lib.futhark_new_i32_2d(c, None, 1 << 20, 1 << 20)
    assert "invalid array shape" in ctx.error()
    out = torch.zeros((8, 16), dtype=torch.int32, device="cuda")
    cv = _view(8, 16).to_c()
    assert lib.fsx_render(c, m.ptr, ctypes.byref(cv), ctypes.c_void_p(out.data_ptr()), 8) != 0
    assert "pitch" in ctx.error()
    assert lib.fsx_render(c, m.ptr, ctypes.byref(cv), None, 0) != 0
    assert "fsx_render: null argument" in ctx.error()
    arr = ctypes.c_void_p()
    assert lib.futhark_entry_render(c, ctypes.byref(arr), None) != 0
    assert "futhark_entry_render: null argument" in ctx.error()
    assert lib.fsx_batch_run(c, None) != 0
    assert "null argument" in ctx.error()
    ctx.render(m, _view(8, 16), out)
    ctx.sync()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ctx.render(m, _view(8, 16)))


def test_depth_slice_limit(ctx, small):
    """More than 2^24 depth slices -- given outright or implied by a tiny
    slice spacing -- is refused with a message by every render entry."""
    m, col, _ = small
    huge = _view(8, 16)
    huge.n_slices = (1 << 24) + 1
    implied = _view(8, 16)
    implied.d, implied.z0 = 1e-12, 0.0      # get_zs: about 4.5e7 slices to distance 300
    for v in (huge, implied):
        with pytest.raises(FutspaceError, match="depth slices"):
            ctx.render(m, v)
        r = DiffRenderer(ctx, m)
        with pytest.raises(FutspaceError, match="depth slices"):
            r.forward_raw(torch.from_numpy(planes_of(col)).cuda(), 0.3, v)
        out = torch.zeros(8 * 16, dtype=torch.int32, device="cuda")
        with pytest.raises(FutspaceError, match="depth slices"):
            ctx.render_batch([m], [v], out, [0])
    ok = _view(8, 16)
    ok.n_slices = 1 << 12
    assert ctx.render(m, ok).shape == (8, 16)
