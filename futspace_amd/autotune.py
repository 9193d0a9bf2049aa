"""Autotuner for the render kernels' shape.

The kernels take three knobs (fsx_set_tuning): frame columns per
workgroup tile (4 or 8), workgroups per CU (one of 1024 threads or two of
512, which overlap each other's stalls but need shorter depth chunks to
fit two tiles in LDS) and depth slices per chunk.  The tuner times
every supported combination on the caller's own workload with the
library's kernel events, keeps the fastest, and caches the choice in a
JSON file keyed by device and workload signature.
"""
import json
import os
import time

CANDIDATES = (
    (8, 2, 512),
    (8, 2, 256),
    (8, 1, 2048),
    (8, 1, 1024),
    (8, 1, 512),
    (4, 2, 2048),
    (4, 1, 2048),
    (4, 2, 1024),
    (4, 1, 1024),
)


def cache_path():
    return os.environ.get("FSX_TUNING_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "futspace_amd",
                                       "tuning.json"))


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:
        return "unknown"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(ctx, make_run, signature, reps=5, candidates=CANDIDATES, use_cache=True):
    """Pick the fastest kernel shape for a workload and set it on ctx.

    make_run() is called after each candidate shape is set and returns a
    callable that enqueues one instance of the workload on the context
    stream (a batch must be prepared after the shape is set, since it is
    planned for one shape).  Returns (best_shape, {shape: ms}).
    """
    key = f"{_device_name()}|{signature}"
    path = cache_path()
    cache = _load(path) if use_cache else {}
    if key in cache:
        best = tuple(cache[key]["shape"])
        ctx.set_tuning(*best)
        return best, {tuple(json.loads(k)): v for k, v in cache[key]["ms"].items()}
    timings = {}
    for cand in candidates:
        try:
            ctx.set_tuning(*cand)
            run = make_run()
            run()
            ctx.sync()
            ctx.profile(True)
            for _ in range(reps):
                run()
            ms, n = ctx.profile_read()
            ctx.profile(False)
            timings[cand] = ms / max(n, 1)
        except RuntimeError:
            ctx.profile(False)
            continue
    if not timings:
        ctx.set_tuning(0, 0, 0)
        return (0, 0, 0), {}
    best = min(timings, key=timings.get)
    ctx.set_tuning(*best)
    if use_cache:
        cache[key] = {"shape": list(best), "ms": {json.dumps(list(k)): v for k, v in timings.items()},
                      "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f, indent=1, sort_keys=True)
        except OSError:
            pass
    return best, timings
