"""Differentiable renderer: torch autograd over the library's f32-plane entries.

The forward pass renders a view from f32 colour planes planes[3][M][M]
(r, g, b in [0,1]) through fsx_render_f32, which also records the depth
slice visible at each pixel; the backward pass (fsx_render_f32_vjp)
scatter-adds dL/dimg into dL/dplanes with the visibility held fixed.  The
two shading parameters are (sun_height, sun_ang): sun_height scales the
sky colour (argb.scale 0xFF9090e0 sun_height, fut/interactive.fut:163),
so its gradient is the sky pixels' term; the shadow counts that sun_ang
and sun_height set are piecewise constant, so their gradient through the
shade factor is zero almost everywhere and is reported as zero.

Every device call runs on the renderer context's own stream, ordered
after torch's current stream with stream waits (no host synchronisation),
and torch's current stream waits for it afterwards.
"""
import contextlib
import ctypes

import numpy as np
import torch

from .renderer import FutspaceError, View


class DiffRenderer:
    """The map's heights (visibility) and shade factor, resident on the
    device; colour planes are supplied per call.

    The colour planes have the shade map's size, so create the map with
    shade_size equal to the colour map's side to differentiate with
    respect to the full-resolution colour map."""

    def __init__(self, ctx, m):
        self.ctx = ctx
        self.map = m
        n = m.shade_size
        self.hc = self.wc = n
        self._depth = 0          # nesting of stream()
        self._ext = None
        self.sf = torch.empty((n, n), dtype=torch.float32, device="cuda")
        with self.stream():
            ctx.check(ctx.lib.fsx_map_shade_factor(ctx.ptr, m.ptr, _ptr(self.sf)),
                      "fsx_map_shade_factor")

    @contextlib.contextmanager
    def stream(self):
        """The context's HIP stream, ordered after torch's current stream on
        every entry (so the renderer sees torch's earlier work, also inside
        an outer block) and before it at the outermost exit: torch work
        issued inside a block must not read this renderer's outputs before
        the outermost exit."""
        if self._depth:
            self._ext.wait_stream(torch.cuda.current_stream())
            self._depth += 1
            try:
                yield self._ext
            finally:
                self._depth -= 1
            return
        handle = self.ctx.lib.fsx_context_stream(self.ctx.ptr)
        if not handle:
            raise FutspaceError("the renderer context has no stream")
        ext = torch.cuda.ExternalStream(handle)
        cur = torch.cuda.current_stream()
        ext.wait_stream(cur)
        self._ext, self._depth = ext, 1
        try:
            yield ext
        finally:
            self._ext, self._depth = None, 0
            cur.wait_stream(ext)

    def _check_planes(self, planes):
        if (planes.dtype != torch.float32 or not planes.is_cuda or not planes.is_contiguous()
                or tuple(planes.shape) != (3, self.hc, self.wc)):
            raise ValueError(f"colour planes must be a contiguous CUDA float32 tensor of shape "
                             f"(3, {self.hc}, {self.wc}), got {tuple(planes.shape)} "
                             f"{planes.dtype}")

    def forward_raw(self, planes, sky_scale, view: View, img=None, src=None):
        """(img[3][h][w] f32, src[h][w] i32) for one view."""
        self._check_planes(planes)
        if img is None:
            img = torch.empty((3, view.h, view.w), dtype=torch.float32, device="cuda")
        if src is None:
            src = torch.empty((view.h, view.w), dtype=torch.int32, device="cuda")
        cv = view.to_c()
        with self.stream():
            self.ctx.check(self.ctx.lib.fsx_render_f32(
                self.ctx.ptr, self.map.ptr, ctypes.byref(cv), _ptr(planes), _ptr(self.sf),
                self.hc, self.wc, ctypes.c_float(sky_scale), _ptr(img), _ptr(src)),
                "fsx_render_f32")
        return img, src

    def vjp_raw(self, planes, src, gimg, view: View, gplanes=None, gsky=None, accumulate=False):
        """dL/dplanes and dL/dsky_scale (a 1-element tensor) for one view;
        with accumulate the gradients are added into gplanes / gsky."""
        self._check_planes(planes)
        gimg = gimg.contiguous()
        if tuple(gimg.shape) != (3, view.h, view.w) or tuple(src.shape) != (view.h, view.w):
            raise ValueError("gradient or src does not match the view")
        if gplanes is None:
            gplanes = torch.empty_like(planes)
            accumulate = False
        if gsky is None:
            gsky = torch.empty(1, dtype=torch.float32, device="cuda")
            accumulate = False
        cv = view.to_c()
        with self.stream():
            self.ctx.check(self.ctx.lib.fsx_render_f32_vjp(
                self.ctx.ptr, ctypes.byref(cv), _ptr(planes), _ptr(self.sf), self.hc, self.wc,
                _ptr(src), _ptr(gimg), 1 if accumulate else 0, _ptr(gplanes), _ptr(gsky)),
                "fsx_render_f32_vjp")
        return gplanes, gsky

    def l2_loss_grad(self, img, target, grad=None, loss=None, accumulate=False):
        """One fused pass: grad = 2 * (img - target) and loss (a 1-element
        tensor) = sum((img - target)^2), added into loss with accumulate."""
        if img.shape != target.shape or not img.is_contiguous() or not target.is_contiguous():
            raise ValueError("img and target must be contiguous tensors of one shape")
        if grad is None:
            grad = torch.empty_like(img)
        if loss is None:
            loss = torch.empty(1, dtype=torch.float32, device="cuda")
            accumulate = False
        with self.stream():
            self.ctx.check(self.ctx.lib.fsx_l2_loss(
                self.ctx.ptr, _ptr(img), _ptr(target), img.numel(), 1 if accumulate else 0,
                _ptr(grad), _ptr(loss)), "fsx_l2_loss")
        return grad, loss

    def l2_step(self, planes, sky_scale, view: View, target, img=None, src=None, gimg=None,
                gplanes=None, gsky=None, loss=None, accumulate=False, outputs=True):
        """One fitting step in one call: forward_raw, l2_loss_grad against
        target and vjp_raw with that gradient (fsx_render_f32_l2_step).
        Returns (img, src, gplanes, gsky, loss); with accumulate the loss and
        the gradients are added into the given tensors.  With outputs=False
        the image and its loss gradient are neither allocated nor written
        (img and gimg are ignored and None is returned for img)."""
        self._check_planes(planes)
        if tuple(target.shape) != (3, view.h, view.w) or not target.is_contiguous():
            raise ValueError("target must be a contiguous [3][h][w] tensor")
        if not outputs:
            img = gimg = None
        elif img is None:
            img = torch.empty((3, view.h, view.w), dtype=torch.float32, device="cuda")
        if src is None:
            src = torch.empty((view.h, view.w), dtype=torch.int32, device="cuda")
        if outputs and gimg is None:
            gimg = torch.empty_like(img)
        if gplanes is None or gsky is None or loss is None:
            accumulate = False
        if gplanes is None:
            gplanes = torch.empty_like(planes)
        if gsky is None:
            gsky = torch.empty(1, dtype=torch.float32, device="cuda")
        if loss is None:
            loss = torch.empty(1, dtype=torch.float32, device="cuda")
        cv = view.to_c()
        with self.stream():
            self.ctx.check(self.ctx.lib.fsx_render_f32_l2_step(
                self.ctx.ptr, self.map.ptr, ctypes.byref(cv), _ptr(planes), _ptr(self.sf),
                self.hc, self.wc, ctypes.c_float(sky_scale), _ptr(target),
                1 if accumulate else 0, _ptr_or_null(img), _ptr(src), _ptr_or_null(gimg),
                _ptr(gplanes), _ptr(gsky), _ptr(loss)), "fsx_render_f32_l2_step")
        return img, src, gplanes, gsky, loss

    def adam_step(self, param, grad, m, v, lr, step, betas=(0.9, 0.999), eps=1e-8, lo=0.0,
                  hi=1.0):
        """One Adam step (fsx_adam_step: bias-corrected, no weight decay,
        step >= 1) on param in place, its moments m and v updated in place,
        then param clamped to [lo, hi]; on the context stream."""
        for t in (param, grad, m, v):
            if (t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous()
                    or t.shape != param.shape):
                raise ValueError("param, grad, m and v must be contiguous CUDA float32 "
                                 "tensors of one shape")
        with self.stream():
            self.ctx.check(self.ctx.lib.fsx_adam_step(
                self.ctx.ptr, _ptr(param), _ptr(grad), _ptr(m), _ptr(v), param.numel(),
                float(lr), float(betas[0]), float(betas[1]), float(eps), int(step),
                ctypes.c_float(lo), ctypes.c_float(hi)),
                "fsx_adam_step")

    def render(self, planes, shading, view: View):
        """Differentiable render: img[3][h][w] as a function of planes and
        shading = (sun_height, sun_ang)."""
        return RenderF32.apply(planes, shading, self, view)


class RenderF32(torch.autograd.Function):
    @staticmethod
    def forward(fctx, planes, shading, renderer, view):
        sky_scale = float(shading.detach().reshape(-1)[0])
        img, src = renderer.forward_raw(planes.detach(), sky_scale, view)
        fctx.save_for_backward(planes, src)
        fctx.renderer = renderer
        fctx.view = view
        fctx.shading_meta = (shading.device, shading.dtype, tuple(shading.shape))
        fctx.mark_non_differentiable(src)
        return img

    @staticmethod
    def backward(fctx, gimg):
        planes, src = fctx.saved_tensors
        gplanes, gsky = fctx.renderer.vjp_raw(planes.detach(), src, gimg, fctx.view)
        device, dtype, shape = fctx.shading_meta
        gshade = torch.zeros(shape, dtype=dtype, device=device)
        gshade.view(-1)[0] = gsky.to(device=device, dtype=dtype)[0]
        return gplanes, gshade, None, None


def planes_of(color):
    """f32 colour planes [3][h][w] (r, g, b / 255) of a u32 ARGB map."""
    c = np.asarray(color, dtype=np.uint32)
    return np.stack([((c >> 16) & 255), ((c >> 8) & 255), (c & 255)]).astype(np.float32) / \
        np.float32(255.0)


COLOUR_UNIT = 255.0   # a colour-map channel step (u8) in plane units


def fit(renderer, planes, views, targets, sky_scale, steps, lr=0.5, on_step=None):
    """Multi-view colour-map fitting: `steps` Adam steps on planes (in
    place) against target frames targets[i] (3 x h x w f32 CUDA tensors),
    loss = sum over views of the squared L2 distance.  Gradients of all
    views accumulate into one buffer through the adjoint's scatter-add.

    lr is in the colour map's own unit, one u8 channel step (the planes hold
    channel / 255), so Adam moves each plane value by about lr / 255 per
    step; each step is one pass of the library's Adam kernel, which also
    clamps the planes back to [0, 1], the domain of the relaxation.
    Returns the per-step losses (a CUDA tensor, read once at the end)."""
    planes.requires_grad_(False)
    step_lr = lr / COLOUR_UNIT
    grad = torch.zeros_like(planes)
    m = torch.zeros_like(planes)
    v = torch.zeros_like(planes)
    gsky = torch.zeros(1, dtype=torch.float32, device="cuda")
    src = None
    losses = torch.zeros(steps, dtype=torch.float32, device="cuda")
    for it in range(steps):
        # one ordering of the renderer's stream against torch's per step:
        # the views' calls and the optimizer step queue back to back
        with renderer.stream():
            for i, (view, t) in enumerate(zip(views, targets)):
                if src is None or tuple(src.shape) != (view.h, view.w):
                    src = torch.empty((view.h, view.w), dtype=torch.int32, device="cuda")
                # the loop needs neither the image nor its loss gradient
                renderer.l2_step(planes, sky_scale, view, t, None, src, None, grad, gsky,
                                 losses[it:it + 1], accumulate=i > 0, outputs=False)
            renderer.adam_step(planes, grad, m, v, step_lr, it + 1)
        if on_step is not None:
            on_step(it)
    return losses


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _ptr_or_null(t):
    return ctypes.c_void_p(None) if t is None else ctypes.c_void_p(t.data_ptr())
