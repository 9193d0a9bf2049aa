/*
 * Compile-and-link check of include/libfutspace.h against the call shapes
 * of the reference host (c/interactive.c:50-56, :64, :99) and of the lys
 * frame loop that drives the state entries (resize, key, mouse, wheel,
 * step, render, then futhark_values_u32_2d of the frame).  Built with
 * -Werror by tests/test_abi.py and linked against libfutspace.so; never run.
 */
#include <stdint.h>
#include <stdlib.h>

#include "libfutspace.h"

int drive(struct futhark_context_config *cfg, const uint32_t *colour, const int32_t *height,
          int64_t rows, int64_t cols, uint32_t seed, uint32_t *frame_out)
{
    futhark_context_config_set_debugging(cfg, 0);
    futhark_context_config_set_profiling(cfg, 0);
    futhark_context_config_set_logging(cfg, 0);
    futhark_context_config_set_device(cfg, "#0");
    struct futhark_context *ctx = futhark_context_new(cfg);
    if (!ctx)
        return 1;

    struct futhark_opaque_state *state = NULL;
    if (futhark_entry_init(ctx, &state, seed) != 0)
        goto fail;

    /* load_map: arrays are copied, so the host buffers can go at once */
    struct futhark_u32_2d *cmap = futhark_new_u32_2d(ctx, colour, rows, cols);
    struct futhark_i32_2d *hmap = futhark_new_i32_2d(ctx, height, rows, cols);
    if (futhark_entry_update_map(ctx, &state, cmap, hmap, state) != 0)
        goto fail;
    futhark_free_u32_2d(ctx, cmap);
    futhark_free_i32_2d(ctx, hmap);

    struct futhark_opaque_state *next = NULL;
    if (futhark_entry_resize(ctx, &next, 600, 800, state) != 0)
        goto fail;
    futhark_free_opaque_state(ctx, state);
    state = next;
    if (futhark_entry_key(ctx, &next, 0, 119, state) != 0)
        goto fail;
    futhark_free_opaque_state(ctx, state);
    state = next;
    if (futhark_entry_mouse(ctx, &next, 1, 10, 20, state) != 0)
        goto fail;
    futhark_free_opaque_state(ctx, state);
    state = next;
    if (futhark_entry_wheel(ctx, &next, 0, 1, state) != 0)
        goto fail;
    futhark_free_opaque_state(ctx, state);
    state = next;
    if (futhark_entry_step(ctx, &next, 0.016f, state) != 0)
        goto fail;
    futhark_free_opaque_state(ctx, state);
    state = next;

    float x, y, angle, h, horizon, distance, sun_height, sun_ang, fov;
    if (futhark_entry_text_content(ctx, &x, &y, &angle, &h, &horizon, &distance, &sun_height,
                                   &sun_ang, &fov, state) != 0)
        goto fail;

    struct futhark_u32_2d *frame = NULL;
    if (futhark_entry_render(ctx, &frame, state) != 0)
        goto fail;
    const int64_t *shape = futhark_shape_u32_2d(ctx, frame);
    (void)shape;
    if (futhark_values_u32_2d(ctx, frame, frame_out) != 0 || futhark_context_sync(ctx) != 0)
        goto fail;
    futhark_free_u32_2d(ctx, frame);
    futhark_free_opaque_state(ctx, state);
    futhark_context_clear_caches(ctx);
    futhark_context_pause_profiling(ctx);
    futhark_context_unpause_profiling(ctx);
    futhark_debugging_report(ctx);
    futhark_context_free(ctx);
    return 0;
fail:
    free(futhark_context_get_error(ctx));
    futhark_context_free(ctx);
    return 1;
}

int main(void)
{
    struct futhark_context_config *cfg = futhark_context_config_new();
    futhark_context_config_select_device_interactively(cfg);
    int rc = drive(cfg, NULL, NULL, 0, 0, 1u, NULL);
    futhark_context_config_free(cfg);
    return rc;
}
