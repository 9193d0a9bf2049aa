/*
 * A plain C host of the generated-library entries, driven the way the
 * reference host does: c/interactive.c:89-116 (init, then load_map, whose
 * futhark_new_* / update_map / futhark_free_* sequence is
 * c/interactive.c:25-57) and the lys frame loop around it (window resize,
 * key events, step, render, futhark_values_u32_2d of the frame, sync,
 * free).  After the loop it renders one explicit view of the same map
 * through the extension entries (fsx_map_new, fsx_render_host), which go
 * through the depth-sliced kernel rather than the state's z0 = 0 frame.
 * It links against libfutspace.so through include/libfutspace.h only, in
 * a process of its own.
 *
 *   lys_loop MAP OUT H W FRAMES SEED
 *
 * MAP holds int64 rows, int64 cols, then rows*cols u32 colours and
 * rows*cols i32 heights (native byte order).  OUT receives int32 frames,
 * int32 h, int32 w, the nine text_content floats after the last frame,
 * then the last frame's h*w u32 pixels, then the explicit view's h*w u32
 * pixels (the view: camera (0.3 cols, 0.45 rows), height 160, angle 0.8,
 * horizon h/3, distance 400, fov 1.2, sky 0xFF9090E0, d = 0.002,
 * z0 = 0.5, default scale and slice count).  Exit status 0 on success.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "libfutspace.h"

#define SDLK_a 97
#define SDLK_w 119
#define SDLK_UP 1073741906
#define KEY_DOWN 0
#define KEY_UP 1

static int fail(struct futhark_context *ctx, const char *what)
{
    char *e = ctx ? futhark_context_get_error(ctx) : NULL;
    fprintf(stderr, "lys_loop: %s failed%s%s\n", what, e ? ": " : "", e ? e : "");
    free(e);
    return 1;
}

/* The state entries return a new state; the old one is the caller's to free. */
#define ADVANCE(call)                                                               \
    do {                                                                            \
        struct futhark_opaque_state *next_ = NULL;                                  \
        if ((call) != 0)                                                            \
            return fail(ctx, #call);                                                \
        futhark_free_opaque_state(ctx, state);                                      \
        state = next_;                                                              \
    } while (0)

static int run(struct futhark_context *ctx, const uint32_t *colour, const int32_t *height,
               int64_t rows, int64_t cols, int32_t h, int32_t w, int32_t frames, uint32_t seed,
               FILE *out)
{
    struct futhark_opaque_state *state = NULL;
    if (futhark_entry_init(ctx, &state, seed) != 0)
        return fail(ctx, "futhark_entry_init");

    /* load_map */
    struct futhark_u32_2d *cmap = futhark_new_u32_2d(ctx, colour, rows, cols);
    struct futhark_i32_2d *hmap = futhark_new_i32_2d(ctx, height, rows, cols);
    if (!cmap || !hmap)
        return fail(ctx, "futhark_new_*_2d");
    ADVANCE(futhark_entry_update_map(ctx, &next_, cmap, hmap, state));
    futhark_free_u32_2d(ctx, cmap);
    futhark_free_i32_2d(ctx, hmap);

    /* the window opens at its size */
    ADVANCE(futhark_entry_resize(ctx, &next_, h, w, state));

    uint32_t *pixels = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)h * (size_t)w);
    if (!pixels)
        return fail(ctx, "malloc");
    for (int32_t i = 0; i < frames; i++) {
        /* events of this frame, as lys hands them over before stepping */
        if (i == 0)
            ADVANCE(futhark_entry_key(ctx, &next_, KEY_DOWN, SDLK_w, state));
        if (i == 1) {
            ADVANCE(futhark_entry_key(ctx, &next_, KEY_DOWN, SDLK_a, state));
            ADVANCE(futhark_entry_key(ctx, &next_, KEY_DOWN, SDLK_UP, state));
        }
        if (i == 3) {
            ADVANCE(futhark_entry_key(ctx, &next_, KEY_UP, SDLK_a, state));
            ADVANCE(futhark_entry_key(ctx, &next_, KEY_UP, SDLK_UP, state));
            ADVANCE(futhark_entry_mouse(ctx, &next_, 0, w / 2, h / 2, state));
            ADVANCE(futhark_entry_wheel(ctx, &next_, 0, 1, state));
        }
        ADVANCE(futhark_entry_step(ctx, &next_, 1.0f / 60.0f, state));

        struct futhark_u32_2d *frame = NULL;
        if (futhark_entry_render(ctx, &frame, state) != 0)
            return fail(ctx, "futhark_entry_render");
        const int64_t *shape = futhark_shape_u32_2d(ctx, frame);
        if (!shape || shape[0] != h || shape[1] != w)
            return fail(ctx, "futhark_shape_u32_2d");
        if (futhark_values_u32_2d(ctx, frame, pixels) != 0 || futhark_context_sync(ctx) != 0)
            return fail(ctx, "futhark_values_u32_2d");
        futhark_free_u32_2d(ctx, frame);
    }

    float tc[9];
    if (futhark_entry_text_content(ctx, &tc[0], &tc[1], &tc[2], &tc[3], &tc[4], &tc[5], &tc[6],
                                   &tc[7], &tc[8], state) != 0)
        return fail(ctx, "futhark_entry_text_content");
    futhark_free_opaque_state(ctx, state);
    const int32_t hdr[3] = {frames, h, w};
    int ok = fwrite(hdr, sizeof hdr, 1, out) == 1 && fwrite(tc, sizeof tc, 1, out) == 1 &&
             fwrite(pixels, sizeof(uint32_t) * (size_t)h * (size_t)w, 1, out) == 1;

    /* one explicit view through the extension entries */
    struct fsx_map *map = NULL;
    if (fsx_map_new(ctx, &map, colour, height, rows, cols, 8, 0, 0.1f, 0.1f) != 0)
        return fail(ctx, "fsx_map_new");
    struct fsx_view v = {0};
    v.x = 0.3f * (float)cols;
    v.y = 0.45f * (float)rows;
    v.height = 160.0f;
    v.angle = 0.8f;
    v.horizon = (float)h / 3.0f;
    v.distance = 400.0f;
    v.fov = 1.2f;
    v.sky_color = 0xFF9090E0u;
    v.h = h;
    v.w = w;
    v.d = 0.002f;
    v.z0 = 0.5f;
    if (fsx_render_host(ctx, map, &v, pixels) != 0)
        return fail(ctx, "fsx_render_host");
    ok = ok && fwrite(pixels, sizeof(uint32_t) * (size_t)h * (size_t)w, 1, out) == 1;
    fsx_map_free(ctx, map);
    free(pixels);
    return ok ? 0 : fail(ctx, "writing the output");
}

int main(int argc, char **argv)
{
    if (argc != 7) {
        fprintf(stderr, "usage: %s MAP OUT H W FRAMES SEED\n", argv[0]);
        return 2;
    }
    const int32_t h = atoi(argv[3]), w = atoi(argv[4]), frames = atoi(argv[5]);
    const uint32_t seed = (uint32_t)strtoul(argv[6], NULL, 10);
    FILE *in = fopen(argv[1], "rb");
    int64_t dims[2];
    if (!in || fread(dims, sizeof dims, 1, in) != 1 || dims[0] <= 0 || dims[1] <= 0 ||
        h <= 0 || w <= 0 || frames <= 0) {
        fprintf(stderr, "lys_loop: bad input\n");
        return 2;
    }
    const size_t n = (size_t)dims[0] * (size_t)dims[1];
    uint32_t *colour = (uint32_t *)malloc(n * sizeof(uint32_t));
    int32_t *height = (int32_t *)malloc(n * sizeof(int32_t));
    if (!colour || !height || fread(colour, sizeof(uint32_t), n, in) != n ||
        fread(height, sizeof(int32_t), n, in) != n) {
        fprintf(stderr, "lys_loop: short map file\n");
        return 2;
    }
    fclose(in);

    struct futhark_context_config *cfg = futhark_context_config_new();
    struct futhark_context *ctx = cfg ? futhark_context_new(cfg) : NULL;
    if (!ctx) {
        fprintf(stderr, "lys_loop: no context\n");
        return 1;
    }
    FILE *out = fopen(argv[2], "wb");
    int rc = out ? run(ctx, colour, height, dims[0], dims[1], h, w, frames, seed, out) : 1;
    if (out && fclose(out) != 0)
        rc = 1;
    free(colour);
    free(height);
    futhark_context_free(ctx);
    futhark_context_config_free(cfg);
    return rc;
}
