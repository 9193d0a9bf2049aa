"""Differentiable renderer (needs a GPU).

The HIP forward must reproduce the oracle's f32 relaxation (visible slice
per pixel exactly, colours bit for bit), the adjoint must match the
oracle's vjp within 1e-5 of the gradient's scale, and both must agree
with central finite differences of the loss.
"""
import numpy as np
import pytest

from futspace_amd import View, scene, sky_color
from oracle import fs_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from futspace_amd.diff import DiffRenderer, fit, planes_of  # noqa: E402

SUN_H, SUN_A = 0.1, 0.1


def _setup(ctx, seed=11, size=128, bits=8, sun=(SUN_H, SUN_A)):
    if bits == 8:
        col, hts = scene.scene_maps(seed, size)
    else:
        hts = scene.heightmap(seed, size, 0.6, 600)
        col = scene.colormap(seed, hts, 600)
    m = ctx.map(col, hts, height_bits=bits, shade_size=size, sun_height=sun[0], sun_ang=sun[1])
    r = DiffRenderer(ctx, m)
    alt = (np.asarray(hts, dtype=np.int32) & ((1 << bits) - 1)).astype(np.int32)
    counts = O.shadow_counts(alt, O.sun(*sun), size)
    sf = O.shade_factor(counts)
    return r, col, alt, sf


VIEWS = [
    View(x=64.3, y=100.7, height=260.0, angle=0.4, horizon=60.0, distance=300.0, fov=1.2,
         h=96, w=128, d=0.01, z0=0.5),
    View(x=64.0, y=64.0, height=200.0, angle=0.0, horizon=50.0, distance=200.0, fov=1.2,
         h=64, w=80, d=0.01, z0=0.5),                       # yaw 0: integral x, transparent
    View(x=10.5, y=3.25, height=300.0, angle=1.5707964, horizon=90.0, distance=400.0, fov=0.8,
         h=100, w=77, d=0.005, z0=1.0, scale=120.0),
    View(x=33.3, y=77.7, height=250.0, angle=5.0, horizon=300.0, distance=700.0, fov=1.2,
         h=257, w=129, d=0.003, z0=0.4, scale=64.0, n_slices=3000),
]


def test_shade_factor_matches_oracle(ctx):
    for sun in [(SUN_H, SUN_A), (0.6, 2.0), (-0.3, 0.7)]:
        r, _, _, sf = _setup(ctx, sun=sun)
        torch.cuda.synchronize()
        assert np.array_equal(r.sf.cpu().numpy().view(np.uint32), sf.view(np.uint32)), sun


@pytest.mark.parametrize("vi", range(len(VIEWS)))
def test_forward_matches_oracle(ctx, vi):
    r, col, alt, sf = _setup(ctx)
    v = VIEWS[vi]
    planes = planes_of(col)
    img, src = r.forward_raw(torch.from_numpy(planes).cuda(), 0.37, v)
    torch.cuda.synchronize()
    want_img, want_src = O.render_f32(v, planes, sf, alt, 0.37)
    got_src = src.cpu().numpy()
    cols = np.arange(v.w, dtype=np.int32)[None, :]
    got_src_o = np.where(got_src < 0, -1, got_src * v.w + cols)
    assert np.array_equal(got_src_o, want_src), f"{np.count_nonzero(got_src_o != want_src)} pixels"
    got = img.cpu().numpy()
    bad = np.count_nonzero(got.view(np.uint32) != want_img.view(np.uint32))
    assert bad == 0, f"{bad} channel values differ, max abs {np.abs(got - want_img).max()}"


def test_forward_sixteen_bit_heights(ctx):
    r, col, alt, sf = _setup(ctx, seed=13, bits=16)
    v = View(x=50.0, y=90.0, height=700.0, angle=1.0, horizon=120.0, distance=1000.0, fov=1.2,
             h=144, w=160, d=0.002, z0=0.5)
    planes = planes_of(col)
    img, src = r.forward_raw(torch.from_numpy(planes).cuda(), 0.1, v)
    want_img, want_src = O.render_f32(v, planes, sf, alt, 0.1)
    got_src = src.cpu().numpy()
    cols = np.arange(v.w, dtype=np.int32)[None, :]
    assert np.array_equal(np.where(got_src < 0, -1, got_src * v.w + cols), want_src)
    assert np.array_equal(img.cpu().numpy(), want_img)


@pytest.mark.parametrize("h,w", [(4320, 20), (9000, 7)])
def test_forward_tall_frame(ctx, h, w):
    """An 8K-tall (4320-row) frame takes 4-column slice-mode tiles, a
    9000-row one the column kernel: both bit-exact, and the fused step of
    the 9000-row one (its gradient cleared in a launch of its own) gives
    the separate calls' result."""
    r, col, alt, sf = _setup(ctx)
    v = View(x=64.3, y=100.7, height=260.0, angle=0.4, horizon=h * 0.4, distance=300.0, fov=1.2,
             h=h, w=w, d=0.01, z0=0.5, scale=h / 3.0)
    planes = planes_of(col)
    img, src = r.forward_raw(torch.from_numpy(planes).cuda(), 0.37, v)
    want_img, want_src = O.render_f32(v, planes, sf, alt, 0.37)
    got_src = src.cpu().numpy()
    cols = np.arange(v.w, dtype=np.int32)[None, :]
    assert np.array_equal(np.where(got_src < 0, -1, got_src * v.w + cols), want_src)
    assert np.array_equal(img.cpu().numpy().view(np.uint32), want_img.view(np.uint32))
    if h > 7680:
        tp = torch.from_numpy(planes).cuda()
        target = torch.from_numpy(np.random.default_rng(h).random((3, h, w), dtype=np.float32)).cuda()
        gimg, loss = r.l2_loss_grad(img, target)
        gp, gs = r.vjp_raw(tp, src, gimg, v)
        gp2 = torch.full_like(tp, 9.0)
        _, src2, gp2, gs2, loss2 = r.l2_step(tp, 0.37, v, target, gplanes=gp2)
        torch.cuda.synchronize()
        assert torch.equal(src2, src)
        assert abs(float(loss2[0]) - float(loss[0])) <= 1e-5 * float(loss[0])
        assert _rel_err(gp2.cpu().numpy(), gp.cpu().numpy()) <= 1e-5


def _rel_err(got, want):
    scale = max(float(np.abs(want).max()), 1e-30)
    return float(np.abs(got.astype(np.float64) - want.astype(np.float64)).max()) / scale


@pytest.mark.parametrize("vi", range(len(VIEWS)))
def test_vjp_matches_oracle(ctx, vi):
    """Colour-plane gradients within 1e-5 of the largest gradient; the sky
    gradient within 1e-5 of the L1 mass of its terms (f32 atomics sum it
    in no fixed order)."""
    r, col, alt, sf = _setup(ctx)
    v = VIEWS[vi]
    planes = planes_of(col)
    rng = np.random.default_rng(100 + vi)
    gimg = rng.standard_normal((3, v.h, v.w)).astype(np.float32)
    tp = torch.from_numpy(planes).cuda()
    _, src = r.forward_raw(tp, 0.37, v)
    gp, gs = r.vjp_raw(tp, src, torch.from_numpy(gimg).cuda(), v)
    torch.cuda.synchronize()
    _, want_src = O.render_f32(v, planes, sf, alt, 0.37)
    want_gp, want_gs = O.render_f32_vjp(v, planes, sf, want_src, gimg)
    assert _rel_err(gp.cpu().numpy(), want_gp) <= 1e-5
    sky_terms = np.abs(gimg[:, want_src < 0]).sum(axis=1) @ np.array([144, 144, 224]) / 255.0
    assert abs(float(gs.cpu()[0]) - want_gs) <= 1e-5 * max(sky_terms, 1e-3)


def test_vjp_accumulates(ctx):
    r, col, _, _ = _setup(ctx)
    tp = torch.from_numpy(planes_of(col)).cuda()
    g = torch.zeros_like(tp)
    gs = torch.zeros(1, device="cuda")
    total = torch.zeros_like(tp)
    for i, v in enumerate(VIEWS[:3]):
        _, src = r.forward_raw(tp, 0.2, v)
        gimg = torch.ones((3, v.h, v.w), device="cuda")
        one, _ = r.vjp_raw(tp, src, gimg, v)
        total += one
        r.vjp_raw(tp, src, gimg, v, g, gs, accumulate=i > 0)
    torch.cuda.synchronize()
    assert torch.allclose(g, total, rtol=1e-5, atol=1e-6)


def _loss64(r, planes, v, target, sky):
    img, _ = r.forward_raw(planes, sky, v)
    d = (img - target).double()
    return float((d * d).sum())


def test_autograd_function_and_finite_differences(ctx):
    """dL/dplanes through torch.autograd equals the adjoint, and agrees with
    central differences (h = 1e-2) on 64 texels, mostly visible ones."""
    r, col, alt, sf = _setup(ctx, seed=21, size=256)
    v = View(x=128.3, y=40.7, height=230.0, angle=2.7, horizon=110.0, distance=500.0, fov=1.2,
             h=180, w=240, d=0.002, z0=0.5, n_slices=1200)
    rng = np.random.default_rng(9)
    planes0 = planes_of(col)
    target = np.clip(planes0 + rng.normal(0.0, 8.0 / 255.0, planes0.shape), 0.0, 1.0)
    tgt_img, _ = r.forward_raw(torch.from_numpy(target.astype(np.float32)).cuda(), SUN_H, v)
    planes = torch.from_numpy(planes0).cuda().requires_grad_(True)
    shading = torch.tensor([0.15, SUN_A], dtype=torch.float32, requires_grad=True)
    img = r.render(planes, shading, v)
    loss = ((img - tgt_img) ** 2).sum()
    loss.backward()
    g = planes.grad.detach().clone()
    # the adjoint called directly
    _, src = r.forward_raw(planes.detach(), 0.15, v)
    g2, gs2 = r.vjp_raw(planes.detach(), src, 2.0 * (img.detach() - tgt_img), v)
    assert float((g - g2).abs().max()) <= 1e-5 * float(g2.abs().max())
    assert abs(float(shading.grad[0]) - float(gs2.cpu()[0])) <= 1e-4 * abs(float(gs2.cpu()[0]))
    assert float(shading.grad[1]) == 0.0
    # central differences
    # texels whose value stays positive under +-h (the channel mix is even
    # in each texel value, so FD across 0 would not be a derivative)
    h = 1e-2
    gh = g.cpu().numpy()
    ok = planes0 > 5 * h
    nz = np.argwhere((np.abs(gh) > 1e-3 * np.abs(gh).max()) & ok)
    pick = nz[rng.choice(len(nz), 48, replace=False)]
    allok = np.argwhere(ok)
    rand = allok[rng.choice(len(allok), 16, replace=False)]
    base = planes.detach().clone()
    for ch, y, x in list(map(tuple, pick)) + list(map(tuple, rand)):
        p = base.clone()
        p[ch, y, x] += h
        lp = _loss64(r, p, v, tgt_img, 0.15)
        p[ch, y, x] -= 2 * h
        lm = _loss64(r, p, v, tgt_img, 0.15)
        fd = (lp - lm) / (2 * h)
        an = float(gh[ch, y, x])
        assert abs(fd - an) <= 2e-2 * max(abs(an), abs(fd)) + 1e-3, (ch, y, x, fd, an)
    # sun_height through the sky pixels
    def loss_sky(s):
        return _loss64(r, base, v, tgt_img, s)
    fd = (loss_sky(0.15 + h) - loss_sky(0.15 - h)) / (2 * h)
    an = float(shading.grad[0])
    assert abs(fd - an) <= 1e-3 * max(abs(an), 1.0), (fd, an)


def test_fit_reduces_loss(ctx):
    r, col, _, _ = _setup(ctx, seed=31, size=128)
    rng = np.random.default_rng(8)
    planes0 = planes_of(col)
    target_planes = np.clip(planes0 + rng.normal(0.0, 8.0 / 255.0, planes0.shape), 0, 1)
    views = VIEWS[:3]
    tp = torch.from_numpy(target_planes.astype(np.float32)).cuda()
    targets = [r.forward_raw(tp, SUN_H, v)[0].clone() for v in views]
    planes = torch.from_numpy(planes0).cuda()
    losses = fit(r, planes, views, targets, SUN_H, steps=20, lr=2.0).cpu().numpy()
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0], losses
    p = planes.cpu().numpy()
    assert p.min() >= 0.0 and p.max() <= 1.0


def test_fit_at_the_workload_rate_does_not_diverge(ctx):
    """Adam at the fitting workload's lr 0.5 (one half u8 channel step):
    the loss falls over the first steps instead of blowing up, and the
    planes stay inside [0, 1]."""
    r, col, _, _ = _setup(ctx, seed=31, size=128)
    rng = np.random.default_rng(8)
    planes0 = planes_of(col)
    target_planes = np.clip(planes0 + rng.normal(0.0, 8.0 / 255.0, planes0.shape), 0, 1)
    views = VIEWS[:3]
    tp = torch.from_numpy(target_planes.astype(np.float32)).cuda()
    targets = [r.forward_raw(tp, SUN_H, v)[0].clone() for v in views]
    planes = torch.from_numpy(planes0).cuda()
    losses = fit(r, planes, views, targets, SUN_H, steps=12, lr=0.5).cpu().numpy()
    assert np.all(np.isfinite(losses))
    assert np.all(np.diff(losses) <= 1e-6 * losses[0]), losses
    assert losses[-1] < losses[0], losses
    p = planes.cpu().numpy()
    assert p.min() >= 0.0 and p.max() <= 1.0


def test_fused_l2_loss(ctx):
    r, _, _, _ = _setup(ctx)
    gen = torch.Generator(device="cuda").manual_seed(3)
    for n in (4, 7, 3 * 97 * 131, 3 * 720 * 1280):
        a = torch.rand(n, device="cuda", generator=gen)
        b = torch.rand(n, device="cuda", generator=gen)
        g, loss = r.l2_loss_grad(a, b)
        g2, loss2 = r.l2_loss_grad(a, b, loss=loss.clone(), accumulate=True)
        torch.cuda.synchronize()
        want = ((a.double() - b.double()) ** 2).sum()
        assert torch.equal(g, 2.0 * (a - b))
        assert abs(float(loss[0]) - float(want)) <= 1e-5 * float(want)
        assert abs(float(loss2[0]) - 2 * float(want)) <= 2e-5 * float(want)


@pytest.mark.parametrize("vi", range(len(VIEWS)))
def test_fused_step_matches_separate_calls(ctx, vi):
    """fsx_render_f32_l2_step = fsx_render_f32, fsx_l2_loss and
    fsx_render_f32_vjp: img, src and the loss gradient bit-exact, the loss
    within f32 summation order, the gradients within 1e-5 (f32 atomics);
    and against the oracle's adjoint of the same loss gradient."""
    r, col, alt, sf = _setup(ctx)
    v = VIEWS[vi]
    planes = planes_of(col)
    tp = torch.from_numpy(planes).cuda()
    rng = np.random.default_rng(200 + vi)
    target = torch.from_numpy(rng.random((3, v.h, v.w), dtype=np.float32)).cuda()
    img, src = r.forward_raw(tp, 0.37, v)
    gimg, loss = r.l2_loss_grad(img, target)
    gp, gs = r.vjp_raw(tp, src, gimg, v)
    # stale values in the outputs: the step must overwrite them
    f_img = torch.full_like(img, 7.0)
    f_src = torch.full_like(src, 5)
    f_gimg = torch.full_like(img, 3.0)
    f_gp = torch.full_like(tp, 9.0)
    f_gs = torch.full((1,), 4.0, device="cuda")
    f_loss = torch.full((1,), 2.0, device="cuda")
    r.l2_step(tp, 0.37, v, target, f_img, f_src, f_gimg, f_gp, f_gs, f_loss)
    torch.cuda.synchronize()
    assert torch.equal(f_img, img) and torch.equal(f_src, src) and torch.equal(f_gimg, gimg)
    assert abs(float(f_loss[0]) - float(loss[0])) <= 1e-5 * float(loss[0])
    assert _rel_err(f_gp.cpu().numpy(), gp.cpu().numpy()) <= 1e-5
    _, want_src = O.render_f32(v, planes, sf, alt, 0.37)
    g_np = gimg.cpu().numpy()
    want_gp, want_gs = O.render_f32_vjp(v, planes, sf, want_src, g_np)
    assert _rel_err(f_gp.cpu().numpy(), want_gp) <= 1e-5
    sky_terms = np.abs(g_np[:, want_src < 0]).sum(axis=1) @ np.array([144, 144, 224]) / 255.0
    assert abs(float(f_gs.cpu()[0]) - want_gs) <= 1e-5 * max(sky_terms, 1e-3)


@pytest.mark.parametrize("vi", [0, 3])
def test_fused_step_without_image_outputs(ctx, vi):
    """outputs=False (NULL img and dL/dimg) leaves those buffers untouched
    and gives the same src and loss bit for bit, the same gradients within
    the f32-atomic order, as the step that writes them."""
    r, col, _, _ = _setup(ctx)
    v = VIEWS[vi]
    tp = torch.from_numpy(planes_of(col)).cuda()
    rng = np.random.default_rng(300 + vi)
    target = torch.from_numpy(rng.random((3, v.h, v.w), dtype=np.float32)).cuda()
    img, src, gp, gs, loss = r.l2_step(tp, 0.37, v, target)
    src2 = torch.full_like(src, 5)
    gp2 = torch.full_like(tp, 9.0)
    gs2 = torch.full((1,), 4.0, device="cuda")
    loss2 = torch.full((1,), 2.0, device="cuda")
    out = r.l2_step(tp, 0.37, v, target, None, src2, None, gp2, gs2, loss2, outputs=False)
    torch.cuda.synchronize()
    assert out[0] is None
    assert torch.equal(src2, src)
    assert float(loss2[0]) == float(loss[0])
    assert _rel_err(gp2.cpu().numpy(), gp.cpu().numpy()) <= 1e-5
    assert abs(float(gs2[0]) - float(gs[0])) <= 1e-5 * max(abs(float(gs[0])), 1.0)


def test_fused_step_accumulates(ctx):
    """accumulate adds the loss and both gradients to the given values."""
    r, col, _, _ = _setup(ctx)
    tp = torch.from_numpy(planes_of(col)).cuda()
    rng = np.random.default_rng(7)
    g = torch.empty_like(tp)
    gs = torch.empty(1, device="cuda")
    loss = torch.empty(1, device="cuda")
    tot_g = torch.zeros_like(tp)
    tot_gs = torch.zeros(1, device="cuda")
    tot_loss = 0.0
    for i, v in enumerate(VIEWS[:3]):
        target = torch.from_numpy(rng.random((3, v.h, v.w), dtype=np.float32)).cuda()
        _, _, g1, gs1, l1 = r.l2_step(tp, 0.2, v, target)
        tot_g += g1
        tot_gs += gs1
        tot_loss += float(l1[0])
        r.l2_step(tp, 0.2, v, target, None, None, None, g, gs, loss, accumulate=i > 0)
    torch.cuda.synchronize()
    assert torch.allclose(g, tot_g, rtol=1e-5, atol=1e-6)
    assert abs(float(gs[0]) - float(tot_gs[0])) <= 1e-5 * max(abs(float(tot_gs[0])), 1.0)
    assert abs(float(loss[0]) - tot_loss) <= 1e-5 * tot_loss


@pytest.mark.parametrize("size", [99, 100])
def test_fused_step_odd_plane_sizes(ctx, size):
    """Planes that are not powers of two (floor-mod wrap), and 3*99*99
    values (not a multiple of 4: the gradient clear takes its own launch);
    a view wide enough to wrap around the map."""
    col, hts = scene.scene_maps(17, 128)
    col = np.ascontiguousarray(col[:size, :size])
    hts = np.ascontiguousarray(hts[:size, :size])
    m = ctx.map(col, hts, shade_size=size, sun_height=SUN_H, sun_ang=SUN_A)
    r = DiffRenderer(ctx, m)
    alt = np.asarray(hts, dtype=np.int32) & 255
    sf = O.shade_factor(O.shadow_counts(alt, O.sun(SUN_H, SUN_A), size))
    v = View(x=50.5, y=20.25, height=240.0, angle=2.0, horizon=80.0, distance=400.0, fov=1.4,
             h=72, w=150, d=0.004, z0=0.6, scale=90.0)
    planes = planes_of(col)
    tp = torch.from_numpy(planes).cuda()
    target = torch.from_numpy(np.random.default_rng(size).random((3, v.h, v.w),
                                                                 dtype=np.float32)).cuda()
    gp = torch.full_like(tp, 9.0)
    img, src, gp, gs, loss = r.l2_step(tp, 0.2, v, target, gplanes=gp,
                                       gsky=torch.zeros(1, device="cuda"),
                                       loss=torch.zeros(1, device="cuda"))
    torch.cuda.synchronize()
    want_img, want_src = O.render_f32(v, planes, sf, alt, 0.2)
    assert np.array_equal(img.cpu().numpy(), want_img)
    d = (want_img.astype(np.float64) - target.cpu().numpy())
    assert abs(float(loss[0]) - float((d * d).sum())) <= 1e-5 * float((d * d).sum())
    g_np = (2.0 * (img - target)).cpu().numpy()
    want_gp, _ = O.render_f32_vjp(v, planes, sf, want_src, g_np)
    assert _rel_err(gp.cpu().numpy(), want_gp) <= 1e-5


def test_forward_tall_frame(ctx):
    """The differentiable forward on a frame tall enough for 4-column
    slice-mode tiles."""
    r, col, alt, sf = _setup(ctx)
    v = View(x=64.3, y=100.7, height=260.0, angle=0.4, horizon=900.0, distance=300.0, fov=1.2,
             h=2200, w=40, d=0.01, z0=0.5, scale=700.0)
    planes = planes_of(col)
    img, src = r.forward_raw(torch.from_numpy(planes).cuda(), 0.37, v)
    torch.cuda.synchronize()
    want_img, want_src = O.render_f32(v, planes, sf, alt, 0.37)
    got_src = src.cpu().numpy()
    cols = np.arange(v.w, dtype=np.int32)[None, :]
    assert np.array_equal(np.where(got_src < 0, -1, got_src * v.w + cols), want_src)
    assert np.array_equal(img.cpu().numpy(), want_img)


@pytest.mark.parametrize("w", [128, 1600])
def test_fused_step_gradient_clear(ctx, w):
    """The fused step clears stale gradient planes either in spare
    workgroups of the slice render (a narrow view leaves most CUs idle) or
    in a launch of its own (a 1600-column view leaves too few): both give
    the separate calls' result."""
    r, col, _, _ = _setup(ctx)
    v = View(x=64.3, y=100.7, height=260.0, angle=0.4, horizon=0.0, distance=300.0, fov=1.2,
             h=40, w=w, d=0.01, z0=0.5, scale=64.0)
    tp = torch.from_numpy(planes_of(col)).cuda()
    target = torch.from_numpy(np.random.default_rng(w).random((3, v.h, v.w),
                                                              dtype=np.float32)).cuda()
    img, src = r.forward_raw(tp, 0.3, v)
    gimg, loss = r.l2_loss_grad(img, target)
    gp, gs = r.vjp_raw(tp, src, gimg, v)
    assert float(gp.abs().sum()) > 0
    for _ in range(2):
        f_gp = torch.full_like(tp, 9.0)
        out = r.l2_step(tp, 0.3, v, target, gplanes=f_gp, gsky=torch.full((1,), 4.0, device="cuda"),
                        loss=torch.full((1,), 2.0, device="cuda"))
        torch.cuda.synchronize()
        assert torch.equal(out[0], img) and torch.equal(out[1], src)
        assert _rel_err(f_gp.cpu().numpy(), gp.cpu().numpy()) <= 1e-5
        assert abs(float(out[3][0]) - float(gs[0])) <= 1e-5 * max(abs(float(gs[0])), 1e-3)
        assert abs(float(out[4][0]) - float(loss[0])) <= 1e-5 * float(loss[0])


def test_fused_step_of_a_4096x2160_frame(ctx):
    """A 4096 x 2160 frame (2^22 pixels or more: the adjoint's blocks are 64
    x 32 pixels, 4352 of them): the fused step still equals the three
    separate calls (image, src and loss gradient bit for bit, loss within
    f32 summation order, gradients within 1e-5)."""
    r, col, _, _ = _setup(ctx, seed=11, size=256)
    v = View(x=120.3, y=77.7, height=280.0, angle=0.9, horizon=700.0, distance=400.0, fov=1.2,
             h=2160, w=4096, d=0.001, z0=0.4, n_slices=900)
    planes = torch.from_numpy(planes_of(col)).cuda()
    rng = np.random.default_rng(41)
    target = torch.from_numpy(rng.random((3, v.h, v.w), dtype=np.float32)).cuda()
    img, src = r.forward_raw(planes, 0.37, v)
    gimg, loss = r.l2_loss_grad(img, target)
    gp, gs = r.vjp_raw(planes, src, gimg, v)
    f_gimg = torch.empty_like(gimg)
    f_img, f_src, f_gp, f_gs, f_loss = r.l2_step(planes, 0.37, v, target, gimg=f_gimg)
    torch.cuda.synchronize()
    assert torch.equal(f_img, img) and torch.equal(f_src, src) and torch.equal(f_gimg, gimg)
    assert abs(float(f_loss[0]) - float(loss[0])) <= 1e-5 * float(loss[0])
    assert _rel_err(f_gp.cpu().numpy(), gp.cpu().numpy()) <= 1e-5
    assert abs(float(f_gs[0]) - float(gs[0])) <= 1e-5 * max(abs(float(gs[0])), 1.0)


def test_fused_step_refuses_frames_beyond_its_partials(ctx):
    """8192 x 4128 needs 16512 loss partials (blocks of 64 x 32 pixels, more
    than 16384): the step is refused with a message naming the limit,
    before anything is enqueued (the outputs keep their contents)."""
    from futspace_amd.renderer import FutspaceError
    r, col, _, _ = _setup(ctx, seed=11, size=128)
    v = View(x=64.3, y=100.7, height=260.0, angle=0.4, horizon=60.0, distance=300.0, fov=1.2,
             h=4128, w=8192, d=0.01, z0=0.5)
    planes = torch.from_numpy(planes_of(col)).cuda()
    target = torch.zeros((3, v.h, v.w), dtype=torch.float32, device="cuda")
    img = torch.full_like(target, 7.0)
    gimg = torch.full_like(target, 3.0)
    src = torch.full((v.h, v.w), 5, dtype=torch.int32, device="cuda")
    gp = torch.full_like(planes, 9.0)
    gs = torch.full((1,), 4.0, device="cuda")
    loss = torch.full((1,), 2.0, device="cuda")
    with pytest.raises(FutspaceError, match="16384"):
        r.l2_step(planes, 0.37, v, target, img, src, gimg, gp, gs, loss)
    torch.cuda.synchronize()
    assert bool((img == 7.0).all()) and bool((gimg == 3.0).all()) and bool((src == 5).all())
    assert bool((gp == 9.0).all()) and float(gs[0]) == 4.0 and float(loss[0]) == 2.0


def test_adam_step_matches_torch_adam(ctx):
    """fsx_adam_step: three steps equal torch's Adam (default betas and
    eps, no weight decay) within the hardware sqrt / reciprocal's few ulps,
    then the clamp to [lo, hi]; a length that is not a multiple of 4 takes
    the scalar tail."""
    r, _, _, _ = _setup(ctx)
    rng = np.random.default_rng(31)
    for n in (4096, 1003):
        p0 = torch.from_numpy(rng.random(n, dtype=np.float32)).cuda()
        grads = [torch.from_numpy(rng.standard_normal(n).astype(np.float32)).cuda()
                 for _ in range(3)]
        p, m, v = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)
        q = p0.clone().requires_grad_(True)
        opt = torch.optim.Adam([q], lr=0.01, foreach=False)
        for t, g in enumerate(grads):
            r.adam_step(p, g, m, v, 0.01, t + 1, lo=-10.0, hi=10.0)
            q.grad = g.clone()
            opt.step()
        torch.cuda.synchronize()
        assert float((p - q.detach()).abs().max()) <= 1e-6
        st = opt.state[q]
        assert torch.allclose(m, st["exp_avg"], rtol=1e-6, atol=1e-7)
        assert torch.allclose(v, st["exp_avg_sq"], rtol=1e-6, atol=1e-7)
        r.adam_step(p, grads[0] * 1e4, m, v, 1.0, 4, lo=0.25, hi=0.75)
        torch.cuda.synchronize()
        assert float(p.min()) >= 0.25 and float(p.max()) <= 0.75
