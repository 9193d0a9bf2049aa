#!/usr/bin/env python3
"""Benchmark of the futspace hot path on one MI355X.

Default workload (plain_2048): the 120-frame fly-through over the seeded
2048^2 map at 1920x1080 with 2048 depth slices; one step renders one frame
(slice table + render kernel) from maps already resident in HBM.  Prints
one JSON line with frames/s, Gpix/s, the render kernel's roofline and the
CPU baseline (the oracle restatement, OpenMP over columns, on this host).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("frames/s & Gpix/s at 1080p/2048² map, plain vs batched; "
          "fwd+bwd grad step ms")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--settle-s", type=float, default=0.5,
                    help="untimed rendering before the warmup steps, while GPU clocks ramp up")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU work for the baseline sample")
    ap.add_argument("--quick", action="store_true", help="few steps, for profiling runs")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 256-view leg")
    ap.add_argument("--batch-reps", type=int, default=20,
                    help="timed graph replays of the 256-view batch")
    ap.add_argument("--no-diff", action="store_true",
                    help="skip the differentiation legs (2048^2 grad step, 8192^2 fit)")
    ap.add_argument("--no-fit", action="store_true", help="skip the 8192^2 multi-view fit")
    ap.add_argument("--fit-steps", type=int, default=50)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 64 MiB device copy (known bytes) to calibrate HBM counters")
    return ap.parse_args()


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_render_plain_2048_1080p.md")
COUNTER_PROFILE = os.path.join(ROOT, "profiles", "render_counters.md")
# non-packed VALU lane-operations per second: 256 CUs x 64 lanes x 2.4 GHz
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9


def _profile_value(path, key):
    """A `key = value` line of a committed rocprofv3 summary; None if absent."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split("=", 1)[1])
    except OSError:
        pass
    return None


def measured_traffic():
    """HBM bytes per k_render launch on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (each scaled by what the
    counter reads for a 64 MiB copy); None if the summary is absent."""
    return _profile_value(TRAFFIC_PROFILE, "traffic_bytes_per_launch")


def valu_roofline(kern_ms):
    """The VALU side of the kernel: SQ_INSTS_VALU per launch (committed
    rocprofv3 pass) x 64 lanes against the non-packed lane-op peak."""
    insts = _profile_value(COUNTER_PROFILE, "valu_insts_per_launch")
    if insts is None:
        return None
    achieved = insts * 64 / (kern_ms * 1e-3)
    return {"insts_per_launch": insts, "lane_ops_per_launch": insts * 64,
            "achieved_lane_ops_per_s": achieved, "peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
            "frac": achieved / VALU_PEAK_LANE_OPS,
            "source": os.path.relpath(COUNTER_PROFILE, ROOT)}


def calib_copy():
    """A plain device-to-device copy of known size: 64 MiB read + 64 MiB
    written per launch, the yardstick for FETCH_SIZE / WRITE_SIZE."""
    import torch
    src = torch.ones(16 * 1024 * 1024, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def bench_batched(ctx, reps):
    """256 views of mixed resolution and depth over four resident maps:
    one persistent launch captured in a HIP graph (shape autotuned on this
    workload), against a loop of plain renders of the same views."""
    import numpy as np
    import torch

    from futspace_amd import autotune, scene

    data = scene.batch_maps()
    maps = [ctx.map(c, h) for c, h in data]
    views = scene.batched_views([h for _, h in data])
    offs, total = scene.batch_offsets(views)
    out = torch.zeros(total, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()

    def make_run():
        b = ctx.batch(maps, views, out, offs)
        return b.run

    best, timings = autotune.tune(ctx, make_run, "batched-256", reps=3, use_cache=False)
    b = ctx.batch(maps, views, out, offs)
    b.capture()
    for _ in range(3):
        b.replay()
    ctx.sync()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        b.replay()
    ctx.sync()
    t1 = time.perf_counter()
    graph_ms, launches = ctx.profile_read()
    ctx.profile(False)
    batch_ms = 1e3 * (t1 - t0) / reps

    # the same views as a loop of plain renders (automatic shape)
    ctx.set_tuning(0, 0, 0)
    slices = [out[o:o + v.h * v.w].view(v.h, v.w) for v, o in zip(views, offs)]
    loop_reps = max(2, reps // 4)
    for v, sl in zip(views, slices):
        ctx.render(maps[v.map], v, sl)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(loop_reps):
        for v, sl in zip(views, slices):
            ctx.render(maps[v.map], v, sl)
    ctx.sync()
    t1 = time.perf_counter()
    loop_ms = 1e3 * (t1 - t0) / loop_reps
    b.close()
    pixels = int(sum(v.h * v.w for v in views))
    alg = float(sum(scene.algorithmic_bytes(v, 1) for v in views))
    kern_ms = graph_ms / max(launches, 1)
    return {
        "views": len(views), "maps": list(scene.BATCH_MAP_SIZES), "pixels": pixels,
        "ms_per_batch": batch_ms, "graph_kernel_ms": kern_ms,
        "views_per_s": len(views) / (batch_ms * 1e-3),
        "gpix_per_s": pixels / (batch_ms * 1e-3) / 1e9,
        "achieved_gbs": alg / (kern_ms * 1e-3) / 1e9, "algorithmic_bytes": alg,
        "plain_loop_ms": loop_ms, "speedup_vs_plain_loop": loop_ms / batch_ms,
        "tuning": {"tile_columns": best[0], "wg_per_cu": best[1], "chunk": best[2]},
        "autotune_ms": {"x".join(map(str, k)): round(v, 4) for k, v in timings.items()},
    }


def bench_host_frames(ctx, m, views, n):
    """The PCIe-inclusive rate (not `value`): each fly-through frame
    rendered and copied to host memory by fsx_render_host, as a frame loop
    that reads every frame back (futhark_values) would do; into pinned and
    into pageable host buffers."""
    import ctypes

    import numpy as np
    import torch

    from futspace_amd import _native

    lib = _native.futspace()
    H, W = views[0].h, views[0].w
    pinned = torch.empty((H, W), dtype=torch.int32, pin_memory=True)
    pageable = np.empty((H, W), dtype=np.uint32)

    def rate(ptr):
        def one(v):
            cv = v.to_c()
            ctx.check(lib.fsx_render_host(ctx.ptr, m.ptr, ctypes.byref(cv), ptr),
                      "fsx_render_host")
        for v in views[:3]:
            one(v)
        t0 = time.perf_counter()
        for i in range(n):
            one(views[i % len(views)])
        return n / (time.perf_counter() - t0)

    return {"pinned_frames_per_s": rate(ctypes.c_void_p(pinned.data_ptr())),
            "pageable_frames_per_s": rate(ctypes.c_void_p(pageable.ctypes.data)),
            "frames": n, "frame_bytes": H * W * 4}


def _kernel_ms(ctx, fn, reps):
    ctx.sync()
    ctx.profile(True)
    for _ in range(reps):
        fn()
    ms, n = ctx.profile_read()
    ctx.profile(False)
    return ms / max(n, 1)


def bench_diff(ctx, reps):
    """One 1280x720 view of the 2048^2 map, 2048 slices: forward (f32
    planes) and forward + L2 loss + adjoint, the gradient w.r.t. the
    [3, 2048, 2048] colour planes and the shading parameters."""
    import ctypes

    import torch

    from futspace_amd import scene
    from futspace_amd.diff import DiffRenderer, planes_of

    col, hts = scene.diff_2048_maps()
    m = ctx.map(col, hts, shade_size=2048)
    r = DiffRenderer(ctx, m)
    v = scene.diff_2048_view(hts)
    planes_np = planes_of(col)
    planes = torch.from_numpy(planes_np).cuda()
    target_planes = torch.from_numpy(
        scene.perturbed_planes(planes_np, scene.DIFF_2048_TARGET_SEED)).cuda()
    target, _ = r.forward_raw(target_planes, scene.SUN_HEIGHT, v)
    target = target.clone()
    img = torch.empty((3, v.h, v.w), dtype=torch.float32, device="cuda")
    src = torch.empty((v.h, v.w), dtype=torch.int32, device="cuda")
    gplanes = torch.empty_like(planes)
    gsky = torch.empty(1, dtype=torch.float32, device="cuda")
    gimg = torch.empty_like(img)
    loss = torch.empty(1, dtype=torch.float32, device="cuda")

    def fwd():
        r.forward_raw(planes, scene.SUN_HEIGHT, v, img, src)

    def step():
        # forward + L2 loss + adjoint (gradient cleared first), one call
        r.l2_step(planes, scene.SUN_HEIGHT, v, target, img, src, gimg, gplanes, gsky, loss)
        return loss

    def step_separate():
        r.forward_raw(planes, scene.SUN_HEIGHT, v, img, src)
        r.l2_loss_grad(img, target, gimg, loss)
        r.vjp_raw(planes, src, gimg, v, gplanes, gsky)
        return loss

    def wall(fn):
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / reps

    for _ in range(5):
        step()
        step_separate()
    torch.cuda.synchronize()

    # GPU time per call as the plain render's render_kernel_ms measures it:
    # HIP events on the context stream bracketing reps back-to-back calls of
    # the C-ABI entry (no torch stream ordering in between)
    lib, cv, P = ctx.lib, v.to_c(), ctypes.c_void_p
    sky = ctypes.c_float(scene.SUN_HEIGHT)

    def c_fwd():
        ctx.check(lib.fsx_render_f32(ctx.ptr, m.ptr, ctypes.byref(cv), P(planes.data_ptr()),
                                     P(r.sf.data_ptr()), r.hc, r.wc, sky, P(img.data_ptr()),
                                     P(src.data_ptr())), "fsx_render_f32")

    def c_step():
        ctx.check(lib.fsx_render_f32_l2_step(
            ctx.ptr, m.ptr, ctypes.byref(cv), P(planes.data_ptr()), P(r.sf.data_ptr()), r.hc,
            r.wc, sky, P(target.data_ptr()), 0, P(img.data_ptr()), P(src.data_ptr()),
            P(gimg.data_ptr()), P(gplanes.data_ptr()), P(gsky.data_ptr()), P(loss.data_ptr())),
            "fsx_render_f32_l2_step")

    def c_step_fit():
        # the step as a fitting loop calls it: no image and no g image out
        ctx.check(lib.fsx_render_f32_l2_step(
            ctx.ptr, m.ptr, ctypes.byref(cv), P(planes.data_ptr()), P(r.sf.data_ptr()), r.hc,
            r.wc, sky, P(target.data_ptr()), 0, None, P(src.data_ptr()), None,
            P(gplanes.data_ptr()), P(gsky.data_ptr()), P(loss.data_ptr())),
            "fsx_render_f32_l2_step")

    def c_vjp():
        ctx.check(lib.fsx_render_f32_vjp(
            ctx.ptr, ctypes.byref(cv), P(planes.data_ptr()), P(r.sf.data_ptr()), r.hc, r.wc,
            P(src.data_ptr()), P(gimg.data_ptr()), 0, P(gplanes.data_ptr()), P(gsky.data_ptr())),
            "fsx_render_f32_vjp")

    ext = torch.cuda.ExternalStream(lib.fsx_context_stream(ctx.ptr))

    def region_ms(fn):
        torch.cuda.synchronize()
        fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(ext)
        for _ in range(reps):
            fn()
        e1.record(ext)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    fwd_kernel_ms = region_ms(c_fwd)
    step_kernel_ms = region_ms(c_step)
    step_fit_kernel_ms = region_ms(c_step_fit)
    vjp_kernel_ms = region_ms(c_vjp)          # the adjoint alone: clear + scatter-add
    # per-call event pairs (fsx_profile_*) around the torch-ordered calls:
    # each pair also times its own event packets and the stream ordering
    fwd_pair_ms = _kernel_ms(ctx, fwd, reps)
    step_pair_ms = _kernel_ms(ctx, step, reps)
    fwd_ms = wall(fwd)
    step_ms = wall(step)
    separate_ms = wall(step_separate)
    step()
    visible = int((src >= 0).sum())
    return {"workload": "1280x720 view, 2048^2 map, 2048 slices, grad w.r.t. f32 "
                        "[3,2048,2048] colour planes + (sun_height, sun_ang)",
            "fwd_ms": fwd_ms, "fwd_bwd_step_ms": step_ms, "step_over_fwd": step_ms / fwd_ms,
            "fwd_kernel_ms": fwd_kernel_ms, "step_kernel_ms": step_kernel_ms,
            "step_over_fwd_kernels": step_kernel_ms / fwd_kernel_ms,
            "step_no_image_out_kernel_ms": step_fit_kernel_ms,
            "fwd_kernel_ms_event_pairs": fwd_pair_ms, "step_kernel_ms_event_pairs": step_pair_ms,
            "separate_calls_step_ms": separate_ms, "vjp_kernel_ms": vjp_kernel_ms,
            "loss": float(loss[0]), "terrain_pixels": visible}


def bench_fit(ctx, steps):
    """64 views of 3840x2160 with 4096 slices over the 8192^2 map with
    16-bit heights: colour-map fitting with Adam (lr 0.5 in u8 channel
    units, planes clamped to [0, 1]: the library's fsx_adam_step), the
    views' gradients accumulated by the adjoint's scatter-add."""
    import numpy as np
    import torch

    from futspace_amd import scene
    from futspace_amd.diff import DiffRenderer, fit, planes_of

    t0 = time.perf_counter()
    col, hts = scene.diff_8192_maps()
    m = ctx.map(col, hts, height_bits=16, shade_size=8192)
    r = DiffRenderer(ctx, m)
    views = scene.diff_8192_views(hts)
    planes_np = planes_of(col)
    del col
    target_planes = torch.from_numpy(
        scene.perturbed_planes(planes_np, scene.DIFF_8192_SEED)).cuda()
    targets = [r.forward_raw(target_planes, scene.SUN_HEIGHT, v)[0] for v in views]
    del target_planes
    planes = torch.from_numpy(planes_np).cuda()
    del planes_np
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    fit(r, planes.clone(), views[:2], targets[:2], scene.SUN_HEIGHT, steps=1, lr=0.5)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = fit(r, planes, views, targets, scene.SUN_HEIGHT, steps=steps, lr=0.5)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ls = losses.cpu().numpy()
    return {"workload": "64 views 3840x2160, 4096 slices, 8192^2 u16 heights + colour, "
                        f"{steps} Adam steps lr 0.5 (u8 channel units)",
            "ms_per_step": 1e3 * dt / steps, "loss_first": float(ls[0]),
            "loss_last": float(ls[-1]), "loss_every_10": [float(x) for x in ls[::10]],
            "setup_s": setup_s}


def cpu_baseline(views, shaded, alt, budget_s):
    """The oracle painter (the futhark multicore stand-in) on this host's
    cores over a bounded sample of the same fly-through."""
    from oracle import fs_oracle as O
    threads = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    frames = 0
    while True:
        O.render_painter_mt(views[frames % len(views)], shaded, alt, threads)
        frames += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    # single-thread stand-in for `futhark c`, one frame
    t1 = time.perf_counter()
    O.render_painter_mt(views[0], shaded, alt, 1)
    st = time.perf_counter() - t1
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": frames / dt, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{frames} frames cycling through the 1920x1080 / 2048-slice "
                      f"fly-through ({dt:.1f} s), oracle painter, OpenMP over columns",
            "single_thread_frames_per_s": 1.0 / st, "cpu_model": cpu_model}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by view or column range, see the design notes)",
              file=sys.stderr)
        sys.exit(2)
    if args.quick:
        args.steps, args.warmup = min(args.steps, 30), min(args.warmup, 5)

    import numpy as np
    import torch

    from futspace_amd import Context, scene
    from futspace_amd import _native

    col, hts = scene.plain_2048_maps()
    views = scene.plain_2048_views(hts)
    ctx = Context(device=0)
    m = ctx.map(col, hts)                  # reference shading: 1024^2 shade map
    stream = torch.cuda.Stream()
    ctx.set_stream(stream)
    H, W = views[0].h, views[0].w
    out = torch.empty((H, W), dtype=torch.int32, device="cuda")

    with torch.cuda.stream(stream):
        # settle: a fresh box's GPU clocks ramp up under load for a few
        # hundred milliseconds; render untimed until that has passed, so
        # that the timed steps measure the steady state
        t_settle = time.perf_counter()
        while time.perf_counter() - t_settle < args.settle_s:
            for v in views:
                ctx.render(m, v, out)
            torch.cuda.synchronize()
        for i in range(args.warmup):
            ctx.render(m, views[i % len(views)], out)
        torch.cuda.synchronize()
        # the timed steps: nothing but the renders, bracketed by HIP events
        # on the context stream (the stream the kernel is launched on)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        # the timed steps stride through the whole fly-through whatever
        # their number: step i renders frame i * 120 // steps
        timed = [(i * len(views)) // args.steps for i in range(args.steps)]
        ev0.record(stream)
        t0 = time.perf_counter()
        for k in timed:
            ctx.render(m, views[k], out)
        ev1.record(stream)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        region_ms = ev0.elapsed_time(ev1)
        # per-launch event pairs (fsx_profile_*) in a separate pass over the
        # same frames: each pair also times its own event packets
        import ctypes
        lib = _native.futspace()
        ctx.check(lib.fsx_profile_enable(ctx.ptr, 1), "fsx_profile_enable")
        for i in range(min(args.steps, 2 * len(views))):
            ctx.render(m, views[i % len(views)], out)
        torch.cuda.synchronize()
        tot_ms, n = ctypes.c_double(), ctypes.c_int64()
        ctx.check(lib.fsx_profile_read(ctx.ptr, ctypes.byref(tot_ms), ctypes.byref(n)),
                  "fsx_profile_read")
        ctx.check(lib.fsx_profile_enable(ctx.ptr, 0), "fsx_profile_enable")

    elapsed = t1 - t0
    fps = args.steps / elapsed
    # the kernel's average launch duration over the timed region: the
    # launches run back to back (the GPU, not the host, sets the pace), so
    # the region's event time per launch is the kernel time plus the
    # dispatch gap between launches
    kern_ms = region_ms / args.steps
    launch_pair_ms = tot_ms.value / max(n.value, 1)
    bytes_per_frame = float(np.mean([scene.algorithmic_bytes(views[k], 1) for k in timed]))
    achieved = bytes_per_frame / (kern_ms * 1e-3) / 1e9
    rec = {
        "metric": METRIC,
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded diamond-square 2048^2 map, Catmull-Rom fly-through)",
        "config": {"workload": "plain_2048: 1920x1080 frames, 2048 depth slices, 2048^2 u8 "
                               "height + u32 colour map, 120-frame fly-through",
                   "frame": [H, W], "slices": 2048, "map": 2048,
                   "frames_timed": f"{len(set(timed))} distinct of {len(views)} "
                                   f"(step i renders frame i*{len(views)}//{args.steps})"},
        "gpix_per_s": fps * H * W / 1e9,
        "render_kernel_ms": kern_ms,
        "render_kernel_ms_event_pairs": launch_pair_ms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(),
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
                     "kernel": "k_render", "algorithmic_bytes_per_launch": bytes_per_frame,
                     "valu": valu_roofline(kern_ms)},
    }
    if not args.no_batch:
        ctx.set_stream(None)
        rec["batched"] = bench_batched(ctx, max(3, args.batch_reps if not args.quick else 3))
    ctx.set_stream(None)
    rec["host_frames"] = bench_host_frames(ctx, m, views, 5 if args.quick else 60)
    if not args.no_diff:
        rec["diff"] = bench_diff(ctx, 50 if not args.quick else 5)
    if not args.no_fit:
        rec["fit"] = bench_fit(ctx, args.fit_steps if not args.quick else 3)
    if args.calib_copy:
        calib_copy()
    if not args.no_cpu:
        from oracle import fs_oracle as O
        shaded, alt = O.update_map(col, hts)
        rec["cpu_baseline"] = cpu_baseline(views, shaded, alt, args.cpu_seconds)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
