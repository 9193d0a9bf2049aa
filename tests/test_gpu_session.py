"""The generated-library entries driven like c/interactive.c and lys do.

init -> update_map -> key/step -> render, compared with the oracle on the
camera the state reports through text_content.
"""
import numpy as np
import pytest

from futspace_amd import Session, View, scene
from oracle import fs_oracle as O

pytestmark = pytest.mark.gpu


def _state_view(sess, sky, h=1024, w=1024):
    x, y, ang, hgt, hor, dist, sun_h, sun_a, fov = sess.text_content()
    return View(x=x, y=y, height=hgt, angle=ang, horizon=hor, distance=dist, fov=fov,
                sky_color=sky, h=h, w=w, d=0.001, z0=0.0, scale=0.0, n_slices=0)


def test_render_before_update_map(ctx):
    """init's dummy landscape is [[0],[0]] everywhere (fut/interactive.fut:38-43)."""
    s = Session(ctx)
    s.resize(64, 96)
    got = s.render()
    z2 = np.zeros((2, 1), np.uint32)
    want = O.render_off(_state_view(s, 0, 64, 96), z2, z2.astype(np.int32))
    assert np.array_equal(got, want)
    s.close()


def test_update_map_then_render_is_uniform_at_z0_zero(ctx):
    """With the reference's z_0 = 0 the first slice sits at z = 0, inv_z is
    +inf and, under x86 casts, every column is covered by the sample under
    the camera (fut/voxel_renderer.fut:103, :217, :225)."""
    col, hts = scene.plain_1024_maps()
    s = Session(ctx)
    s.resize(200, 300)
    s.update_map(col, hts)
    s.step()
    got = s.render()
    shaded, alt = O.update_map(col, hts)
    want = O.render_off(_state_view(s, O.lib().fso_sky(0.1), 200, 300), shaded, alt)
    assert np.array_equal(got, want)
    assert len(np.unique(got)) <= 2
    s.close()


def test_keys_move_camera_like_process_inputs(ctx):
    col, hts = scene.plain_1024_maps()
    s = Session(ctx)
    s.update_map(col, hts)
    f = np.float32
    x0, y0, ang0, h0, hor0, dist0, sh0, sa0, fov0 = s.text_content()
    s.key(True, "w")
    s.key(True, "a")
    s.key(True, "r")
    s.key(True, "q")
    s.key(True, "UP")
    s.key(True, "o")
    s.step()
    x1, y1, ang1, h1, hor1, dist1, sh1, sa1, fov1 = s.text_content()
    assert f(x1) == f(f(x0) - f(f(3.0) * f(np.sin(f(ang0), dtype=np.float32))))
    assert f(ang1) == f(f(ang0) + f(0.10))
    assert f(hor1) == f(f(hor0) + f(20.0))
    assert f(dist1) == f(f(dist0) + f(30.0))
    assert f(fov1) == f(f(fov0) + f(0.1))
    # height: +10, then terrain_collision keeps it at or above the ground
    ground = hts[int(np.trunc(y1)) % 1024, int(np.trunc(x1)) % 1024] & 0xFF
    assert f(h1) == max(f(f(h0) + f(10.0)), f(ground))
    s.key(False, "w")
    s.key(False, "a")
    s.step()
    x2, y2, ang2, *_ = s.text_content()
    assert x2 == x1 and y2 == y1 and ang2 == ang1
    s.close()


def test_sun_keys_reshade_with_previous_sun(ctx):
    """U held: sun_height += 0.005 while the shadow map is recomputed from
    the incoming sun (fut/interactive.fut:118-145)."""
    col, hts = scene.scene_maps(21, 256)
    s = Session(ctx)
    s.resize(90, 120)
    s.update_map(col, hts)
    s.key(True, "u")
    s.step()
    s.step()
    s.key(False, "u")
    s.key(True, "d")
    s.step()
    s.key(False, "d")
    # the second step reshaded with sun_height 0.105 (the value after step 1)
    _, _, _, _, _, _, sun_h, sun_a, _ = s.text_content()
    assert np.float32(sun_h) == np.float32(np.float32(np.float32(0.1) + np.float32(0.005))
                                           + np.float32(0.005))
    shaded, alt = O.update_map(col, hts, np.float32(np.float32(0.1) + np.float32(0.005)), 0.1)
    sky = O.lib().fso_sky(np.float32(sun_h))
    got = s.render()
    want = O.render_off(_state_view(s, sky, 90, 120), shaded, alt)
    assert np.array_equal(got, want)
    s.close()


def test_mismatched_update_map_reports_error(ctx):
    s = Session(ctx)
    with pytest.raises(RuntimeError, match="differ"):
        s.update_map(np.zeros((4, 4), np.uint32), np.zeros((4, 5), np.int32))
    s.close()


def _toggle(s, key):
    s.key(True, key)
    s.step()
    s.key(False, key)


@pytest.mark.parametrize("keys,mode,smooth", [(("2",), "png", True), (("1",), "math", False),
                                              (("1", "2"), "math", True)])
def test_math_mode_and_smoothing_match_the_reference_render(ctx, keys, mode, smooth):
    """Key 1 toggles #png/#math, key 2 the smoothing (fut/interactive.fut:146-159);
    render follows fut/interactive.fut:173-183 for every combination."""
    col, hts = scene.scene_maps(21, 256)
    s = Session(ctx)
    s.resize(70, 90)
    s.update_map(col, hts)
    for k in keys:
        _toggle(s, k)
    s.key(True, "w")          # move, so the camera is not on a grid line of #math
    s.step()
    s.key(False, "w")
    _, _, _, _, _, _, sun_h, sun_a, _ = s.text_content()
    sky = O.lib().fso_sky(np.float32(sun_h))
    shaded, alt = O.update_map(col, hts, size=1024)
    view = _state_view(s, sky, 70, 90)
    want = O.render_state(view, mode, smooth, shaded, alt, O.sun(sun_h, sun_a))
    got = s.render()
    assert np.array_equal(got, want)
    s.close()
