"""fsx_render_f32_l2_step_noalloc against fsx_render_f32_l2_step (needs a
GPU), called through the C-ABI: the same image, visible slices and loss
gradient bit for bit, the same loss (its partials are summed in a fixed
order), and plane and sky gradients within 1e-5 (f32 atomics), on two
calls in a row, the second with a smaller frame, so the shared scratch is
reused."""
import ctypes

import numpy as np
import pytest

from futspace_amd import View
from futspace_amd.diff import DiffRenderer, planes_of
from futspace_amd import scene

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _step(ctx, r, name, planes, v, target, sky):
    """One fused step through the named entry on fresh, poisoned outputs."""
    img = torch.full((3, v.h, v.w), 7.0, dtype=torch.float32, device="cuda")
    src = torch.full((v.h, v.w), 5, dtype=torch.int32, device="cuda")
    gimg = torch.full_like(img, 3.0)
    gp = torch.full_like(planes, 9.0)
    gs = torch.full((1,), 4.0, device="cuda")
    loss = torch.full((1,), 2.0, device="cuda")
    cv = v.to_c()
    torch.cuda.synchronize()
    rc = getattr(ctx.lib, name)(ctx.ptr, r.map.ptr, ctypes.byref(cv), _ptr(planes), _ptr(r.sf),
                                r.hc, r.wc, ctypes.c_float(sky), _ptr(target), 0, _ptr(img),
                                _ptr(src), _ptr(gimg), _ptr(gp), _ptr(gs), _ptr(loss))
    ctx.check(rc, name)
    ctx.sync()
    return [t.cpu().numpy() for t in (img, src, gimg, gp, gs, loss)]


def test_noalloc_step_matches_the_allocating_step(ctx):
    col, hts = scene.scene_maps(11, 256)
    m = ctx.map(col, hts, shade_size=256)
    r = DiffRenderer(ctx, m)
    planes = torch.from_numpy(planes_of(col)).cuda()
    views = [
        View(x=120.3, y=77.7, height=280.0, angle=0.9, horizon=300.0, distance=400.0, fov=1.2,
             h=720, w=1280, d=0.001, z0=0.4, n_slices=900),
        View(x=64.3, y=100.7, height=260.0, angle=0.4, horizon=60.0, distance=300.0, fov=1.2,
             h=96, w=128, d=0.01, z0=0.5),
    ]
    rng = np.random.default_rng(77)
    for v in views:
        print(f"noalloc step vs allocating step: {v.h}x{v.w}", flush=True)
        target = torch.from_numpy(rng.random((3, v.h, v.w), dtype=np.float32)).cuda()
        want = _step(ctx, r, "fsx_render_f32_l2_step", planes, v, target, 0.37)
        got = _step(ctx, r, "fsx_render_f32_l2_step_noalloc", planes, v, target, 0.37)
        for k in (0, 1, 2):
            assert np.array_equal(got[k], want[k]), k
        assert got[5][0] == want[5][0], (got[5], want[5])
        gp_w, gp_g = want[3].astype(np.float64), got[3].astype(np.float64)
        assert np.abs(gp_g - gp_w).max() <= 1e-5 * np.abs(gp_w).max()
        assert abs(float(got[4][0]) - float(want[4][0])) <= 1e-5 * max(abs(float(want[4][0])), 1.0)


def test_noalloc_step_refuses_oversized_frames(ctx):
    """The same limit as the allocating step, with a message naming it."""
    col, hts = scene.scene_maps(11, 128)
    m = ctx.map(col, hts, shade_size=128)
    r = DiffRenderer(ctx, m)
    planes = torch.from_numpy(planes_of(col)).cuda()
    v = View(x=64.3, y=100.7, height=260.0, angle=0.4, horizon=60.0, distance=300.0, fov=1.2,
             h=4128, w=8192, d=0.01, z0=0.5)
    target = torch.zeros((3, v.h, v.w), dtype=torch.float32, device="cuda")
    with pytest.raises(Exception, match="16384"):
        _step(ctx, r, "fsx_render_f32_l2_step_noalloc", planes, v, target, 0.37)
