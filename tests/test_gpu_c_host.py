"""The library from a plain C host process (needs a GPU).

tests/c_host/lys_loop.c is compiled with gcc against include/libfutspace.h
and run as a child process with no Python or torch in it: it drives the
generated entries the way c/interactive.c and the lys loop do (init,
load_map's new / update_map / free, resize, key events, step, render,
futhark_values_u32_2d, sync) on the reference's own C1W/D1 map crop, then
renders one explicit view through fsx_map_new / fsx_render_host.  Both
frames must equal the oracle's for the camera the state reports.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from futspace_amd import View
from oracle import fs_oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "futspace_amd")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_plain_c_host_runs_the_frame_loop(tmp_path):
    exe = tmp_path / "lys_loop"
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror",
           f"-I{os.path.join(ROOT, 'include')}", os.path.join(ROOT, "tests", "c_host", "lys_loop.c"),
           f"-L{LIBDIR}", "-lfutspace", "-L/opt/rocm/lib", "-Wl,--allow-shlib-undefined",
           f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

    g = np.load(os.path.join(ROOT, "tests", "golden", "maps.npz"))
    col, hts = g["c1w_colour"], g["d1_height"]
    mapf, outf = tmp_path / "map.bin", tmp_path / "out.bin"
    with open(mapf, "wb") as f:
        f.write(np.array(col.shape, dtype=np.int64).tobytes())
        f.write(np.ascontiguousarray(col, dtype=np.uint32).tobytes())
        f.write(np.ascontiguousarray(hts, dtype=np.int32).tobytes())
    h, w, frames = 180, 240, 6
    print(f"running the C host: {h}x{w}, {frames} frames", flush=True)
    r = subprocess.run([str(exe), str(mapf), str(outf), str(h), str(w), str(frames), "7"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = outf.read_bytes()
    hdr = np.frombuffer(raw[:12], dtype=np.int32)
    assert tuple(hdr) == (frames, h, w)
    x, y, ang, hgt, hor, dist, sun_h, sun_a, fov = np.frombuffer(raw[12:48], dtype=np.float32)
    npx = h * w
    state_frame = np.frombuffer(raw[48:48 + 4 * npx], dtype=np.uint32).reshape(h, w)
    view_frame = np.frombuffer(raw[48 + 4 * npx:48 + 8 * npx], dtype=np.uint32).reshape(h, w)
    assert len(raw) == 48 + 8 * npx

    # the keys moved and turned the camera and raised the horizon
    assert (x, y) != (0.0, 0.0) and hor > 0.0
    shaded, alt = O.update_map(col, hts, sun_h, sun_a)
    state_view = View(x=float(x), y=float(y), height=float(hgt), angle=float(ang),
                      horizon=float(hor), distance=float(dist), fov=float(fov),
                      sky_color=O.lib().fso_sky(sun_h), h=h, w=w, d=0.001, z0=0.0, scale=0.0,
                      n_slices=0)
    want = O.render_state(state_view, "png", False, shaded, alt, O.sun(sun_h, sun_a))
    assert np.array_equal(state_frame, want)

    rows, cols = col.shape
    v = View(x=float(np.float32(0.3) * np.float32(cols)),
             y=float(np.float32(0.45) * np.float32(rows)), height=160.0, angle=0.8,
             horizon=float(np.float32(h) / np.float32(3.0)), distance=400.0, fov=1.2,
             sky_color=0xFF9090E0, h=h, w=w, d=0.002, z0=0.5)
    shaded0, alt0 = O.update_map(col, hts)
    want = O.render_off(v, shaded0, alt0)
    assert np.array_equal(view_frame, want)
    assert len(np.unique(view_frame)) > 100       # a real frame, not a uniform one
