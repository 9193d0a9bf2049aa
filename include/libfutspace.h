/*
 * libfutspace.h -- C-ABI of the MI355X (gfx950) futspace renderer.
 *
 * Part 1 declares exactly the symbols `futhark opencl --library
 * fut/interactive_entrypoints.fut` generates for the reference
 * (Makefile:24-25), with the same prototypes, so the reference host
 * c/interactive.c and the lys frame loop link against this library
 * unchanged (the header is passed as PROGHEADER, Makefile:19).
 * Part 2 declares extension entries with no reference counterpart:
 * explicit-camera, batched and differentiable renders over resident maps,
 * used by tests, benchmarks and the PyTorch harness.
 *
 * Conventions (those of the generated library):
 *  - every entry returns 0 on success; on failure it returns nonzero and
 *    leaves a message that futhark_context_get_error() hands over (a
 *    malloc'd string the caller frees);
 *  - out-parameters come first; every returned array or state is a new
 *    object owned by the caller and freed with futhark_free_*; inputs are
 *    borrowed;
 *  - one context per thread; entries are not reentrant on a context.
 * All device work of a context runs on one HIP stream.
 */
#ifndef LIBFUTSPACE_H
#define LIBFUTSPACE_H

#include <stdint.h>
#include <stddef.h>
#include <stdbool.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ==================================================================== */
/* Part 1: the generated library's interface                             */
/* ==================================================================== */

/* ---- Initialisation (generated for every backend) ---- */
struct futhark_context_config;
struct futhark_context_config *futhark_context_config_new(void);
void futhark_context_config_free(struct futhark_context_config *cfg);
void futhark_context_config_set_debugging(struct futhark_context_config *cfg, int flag);
void futhark_context_config_set_profiling(struct futhark_context_config *cfg, int flag);
void futhark_context_config_set_logging(struct futhark_context_config *cfg, int flag);
/* GPU backends: select a device by "#<index>" or by a substring of its
 * name (called by lys' context setup when -d is given, c/interactive.c:127-129). */
void futhark_context_config_set_device(struct futhark_context_config *cfg, const char *s);
/* OpenCL backend: list devices (c/interactive.c:130-132 -i); here it lists
 * the HIP devices on stderr and keeps the current choice. */
void futhark_context_config_select_device_interactively(struct futhark_context_config *cfg);

struct futhark_context;
struct futhark_context *futhark_context_new(struct futhark_context_config *cfg);
void futhark_context_free(struct futhark_context *ctx);
int futhark_context_sync(struct futhark_context *ctx);
char *futhark_context_get_error(struct futhark_context *ctx);
int futhark_context_clear_caches(struct futhark_context *ctx);
void futhark_context_pause_profiling(struct futhark_context *ctx);
void futhark_context_unpause_profiling(struct futhark_context *ctx);
void futhark_debugging_report(struct futhark_context *ctx);

/* ---- Arrays (c/interactive.c:50-56) ---- */
struct futhark_i32_2d;
struct futhark_i32_2d *futhark_new_i32_2d(struct futhark_context *ctx, const int32_t *data,
                                          int64_t dim0, int64_t dim1);
int futhark_free_i32_2d(struct futhark_context *ctx, struct futhark_i32_2d *arr);
int futhark_values_i32_2d(struct futhark_context *ctx, struct futhark_i32_2d *arr, int32_t *data);
const int64_t *futhark_shape_i32_2d(struct futhark_context *ctx, struct futhark_i32_2d *arr);

struct futhark_u32_2d;
struct futhark_u32_2d *futhark_new_u32_2d(struct futhark_context *ctx, const uint32_t *data,
                                          int64_t dim0, int64_t dim1);
int futhark_free_u32_2d(struct futhark_context *ctx, struct futhark_u32_2d *arr);
int futhark_values_u32_2d(struct futhark_context *ctx, struct futhark_u32_2d *arr, uint32_t *data);
const int64_t *futhark_shape_u32_2d(struct futhark_context *ctx, struct futhark_u32_2d *arr);

/* ---- Opaque values: `state`, fut/interactive_entrypoints.fut:4 ---- */
struct futhark_opaque_state;
int futhark_free_opaque_state(struct futhark_context *ctx, struct futhark_opaque_state *obj);

/* ---- Entry points, fut/interactive_entrypoints.fut:6-32 ---- */
/* entry resize (h: i32) (w: i32) (s: state): state        (:6-7)  */
int futhark_entry_resize(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                         const int32_t in0, const int32_t in1,
                         const struct futhark_opaque_state *in2);
/* entry key (e: i32) (key: i32) (s: state): state          (:9-11) */
int futhark_entry_key(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                      const int32_t in0, const int32_t in1,
                      const struct futhark_opaque_state *in2);
/* entry mouse (buttons: i32) (x: i32) (y: i32) (s: state)  (:13-14) */
int futhark_entry_mouse(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                        const int32_t in0, const int32_t in1, const int32_t in2,
                        const struct futhark_opaque_state *in3);
/* entry wheel (dx: i32) (dy: i32) (s: state)                (:16-17) */
int futhark_entry_wheel(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                        const int32_t in0, const int32_t in1,
                        const struct futhark_opaque_state *in2);
/* entry render (s: state): [h][w]u32 -- the hot path        (:19-20) */
int futhark_entry_render(struct futhark_context *ctx, struct futhark_u32_2d **out0,
                         const struct futhark_opaque_state *in0);
/* entry init (seed: u32): state                             (:22-23) */
int futhark_entry_init(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                       const uint32_t in0);
/* entry text_content (s: state): 9 x f32                    (:25-26) */
int futhark_entry_text_content(struct futhark_context *ctx, float *out0, float *out1,
                               float *out2, float *out3, float *out4, float *out5,
                               float *out6, float *out7, float *out8,
                               const struct futhark_opaque_state *in0);
/* entry step (td: f32) (s: state): state                    (:28-29) */
int futhark_entry_step(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                       const float in0, const struct futhark_opaque_state *in1);
/* entry update_map [h][w] (color_map: [h][w]u32) (height_map: [h][w]i32)
 *   (s: state): state                                        (:31-32) */
int futhark_entry_update_map(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                             const struct futhark_u32_2d *in0,
                             const struct futhark_i32_2d *in1,
                             const struct futhark_opaque_state *in2);

/* ==================================================================== */
/* Part 2: extension entries (no counterpart in the reference)           */
/* ==================================================================== */

/* One view: the reference camera record (fut/voxel_renderer.fut:5-12),
 * the frame size, and the depth/projection parameters the reference
 * hard-codes -- d (invz_param1, fut/voxel_renderer.fut:104), z0
 * (invz_param2, :103), scale (magic_number, :217; <= 0 means f32(w/2)),
 * n_slices (<= 0 means get_zs' count from distance, :28-34; a render
 * refuses more than 2^24 slices, and frames of more than 2^22 rows;
 * frames of more than 7680 rows take a slower one-thread-per-column
 * kernel, and a batched view has at most 7680 rows).
 * `map` indexes the map list of a batched render. */
struct fsx_view {
    float x, y, height, angle, horizon, distance, fov;
    uint32_t sky_color;
    int32_t h, w;
    float d, z0, scale;
    int32_t n_slices;
    int32_t map;
    int32_t reserved;
};

/* A device-resident map pair: heights masked to `height_bits` (8 is the
 * reference's `& 0xFF`, fut/interactive.fut:189; 16 keeps u16 heights)
 * and the shaded colour map of generate_shadowmap_accumulated for the
 * given sun (fut/effects.fut:108-125) on a shade_size^2 grid (0 selects
 * the reference's fixed 1024). */
struct fsx_map;
int fsx_map_new(struct futhark_context *ctx, struct fsx_map **out,
                const uint32_t *color, const int32_t *height, int64_t rows, int64_t cols,
                int32_t height_bits, int32_t shade_size, float sun_height, float sun_ang);
int fsx_map_free(struct futhark_context *ctx, struct fsx_map *map);
/* Copy the shaded colour map (shade_size^2 u32) to host memory. */
int fsx_map_shaded(struct futhark_context *ctx, const struct fsx_map *map, uint32_t *host_out);
int32_t fsx_map_shade_size(const struct fsx_map *map);

/* Render one view into device memory `dev_out` ([h][w] u32, row pitch
 * `pitch` elements; 0 means w).  Asynchronous on the context stream. */
int fsx_render(struct futhark_context *ctx, const struct fsx_map *map,
               const struct fsx_view *view, uint32_t *dev_out, int64_t pitch);
/* Same, into host memory (synchronous). */
int fsx_render_host(struct futhark_context *ctx, const struct fsx_map *map,
                    const struct fsx_view *view, uint32_t *host_out);

/* render2 (fut/voxel_renderer.fut:257-305), the reference's ray-march
 * renderer, which no generated entry point calls: pixel (y, x) marches
 * dist = 0 ... i32(distance / 4) along angle + (x / w) * 1.5 and takes the
 * bilinear unshaded colour of the first point whose bilinear height is
 * above height - dist * (horizon / 50 + (y / h) * 0.5), else 0.  Uses the
 * view's x, y, height, angle, horizon, distance and frame size; distance / 4
 * at most 2^16 (the march's length per pixel).  dev_out [h][w] u32 (row pitch in elements, 0 = w),
 * asynchronous on the context stream after a short synchronous upload of
 * the per-column cos / sin. */
int fsx_render2(struct futhark_context *ctx, const struct fsx_map *map,
                const struct fsx_view *view, uint32_t *dev_out, int64_t pitch);

/* Batched render: views[i] draws maps[views[i].map] into dev_out +
 * offsets[i] (u32 elements, row pitch views[i].w).  One persistent
 * launch covers all views; per-XCD work queues hand out column tiles. */
int fsx_render_batch(struct futhark_context *ctx, const struct fsx_map *const *maps,
                     int32_t n_maps, const struct fsx_view *views, int32_t n_views,
                     uint32_t *dev_out, const int64_t *offsets);

/* A prepared batch (same arguments as fsx_render_batch), launched any
 * number of times.  fsx_batch_run enqueues it on the context stream;
 * fsx_batch_capture records that launch into a HIP graph once and
 * fsx_batch_replay relaunches the graph.  The batch holds references to
 * its maps; the output buffer must outlive it. */
struct fsx_batch;
int fsx_batch_new(struct futhark_context *ctx, const struct fsx_map *const *maps, int32_t n_maps,
                  const struct fsx_view *views, int32_t n_views, uint32_t *dev_out,
                  const int64_t *offsets, struct fsx_batch **out);
int fsx_batch_run(struct futhark_context *ctx, struct fsx_batch *batch);
int fsx_batch_capture(struct futhark_context *ctx, struct fsx_batch *batch);
int fsx_batch_replay(struct futhark_context *ctx, struct fsx_batch *batch);
int fsx_batch_free(struct futhark_context *ctx, struct fsx_batch *batch);

/* The stateless program of tests/futspace.fut:47-145 (`render`): nearest
 * sampling (i32.f32 truncation, floor-mod) of i32 colours and unmasked i32
 * heights, both [q][r] device arrays; z_0 = 1, delta = 0.005, inv_z =
 * (1/z) * 240 as the program fixes them; view supplies the camera
 * (x, y, height, angle, horizon, distance, fov), the frame size (l = h,
 * m = w) and the sky colour, which is also the fill's `empty` colour.
 * Writes dev_out[h][w] i32, asynchronously on the context stream.  The
 * program's `main` is this with camera (512, 800, 78, 0, 100, 800, fov 1),
 * sky 0xFF9090e0 and a 400 x 800 frame (:127-145). */
int fsx_render_stateless(struct futhark_context *ctx, const int32_t *dev_color,
                         const int32_t *dev_height, int32_t q, int32_t r,
                         const struct fsx_view *view, int32_t *dev_out);

/* Differentiable renderer: the f32 relaxation of the png render path
 * (fut/interactive.fut:179-182 over fut/voxel_renderer.fut:214-251 with
 * png_color_filtered, fut/render_functions.fut:95-105), with the shadow
 * counts of fut/effects.fut:108-125 and visibility held fixed.
 *
 * fsx_map_shade_factor writes the map's per-texel shade factor
 * sf = sqrt(1 / (4*count + 1)) ([shade_size][shade_size] f32, device).
 *
 * fsx_render_f32 renders `view` from f32 colour planes
 * dev_planes[3][hc][wc] (r, g, b in [0,1]) times dev_sf[hc][wc]:
 * dev_img[3][h][w] f32 and dev_src[h][w], the depth slice whose sample
 * shows at each pixel (-1: sky, whose colour is (0x90,0x90,0xE0)/255 *
 * sky_scale).  A sample at integral x or y is transparent, as argb.mix
 * makes it in the u32 path.  Visibility comes from the map's heights,
 * exactly as in fsx_render.
 *
 * fsx_render_f32_vjp adds dL/d planes into dev_gplanes[3][hc][wc] and
 * dL/d sky_scale into *dev_gsky, given dev_gimg = dL/d img and the src of
 * the forward pass; accumulate = 0 clears both first.  f32 atomics: the
 * summation order, hence the last bits, varies from run to run.  Its
 * per-workgroup partial sums limit the frame to ceil(w/64) * ceil(h/b)
 * <= 32768 blocks, b = 16 rows, or 32 in frames of 2^22 pixels or more
 * (every frame up to 8192 x 8192); larger frames are refused before
 * anything is enqueued.
 *
 * All three are asynchronous on the context stream. */
int fsx_map_shade_factor(struct futhark_context *ctx, const struct fsx_map *map, float *dev_sf);
int fsx_render_f32(struct futhark_context *ctx, const struct fsx_map *map,
                   const struct fsx_view *view, const float *dev_planes, const float *dev_sf,
                   int32_t hc, int32_t wc, float sky_scale, float *dev_img, int32_t *dev_src);
int fsx_render_f32_vjp(struct futhark_context *ctx, const struct fsx_view *view,
                       const float *dev_planes, const float *dev_sf, int32_t hc, int32_t wc,
                       const int32_t *dev_src, const float *dev_gimg, int32_t accumulate,
                       float *dev_gplanes, float *dev_gsky);

/* Squared-L2 loss of the fitting harness in one pass over n f32 values
 * (16-byte aligned device buffers): dev_grad = 2 * (img - target) and
 * *dev_loss = sum((img - target)^2), added to the old value when
 * accumulate is nonzero.  Asynchronous on the context stream. */
int fsx_l2_loss(struct futhark_context *ctx, const float *dev_img, const float *dev_target,
                int64_t n, int32_t accumulate, float *dev_grad, float *dev_loss);

/* One Adam step of a fitting loop (no weight decay, bias-corrected
 * moments, step >= 1) on dev_param[n] with gradient dev_grad and moments
 * dev_m, dev_v (zero before the first step), then dev_param clamped to
 * [lo, hi] -- e.g. the colour planes' [0, 1] -- in one pass over the four
 * arrays (16-byte aligned device buffers).  The hyper-parameters are
 * doubles: 1 - beta and the bias corrections are formed in double and
 * rounded to f32 once, as torch's Adam forms them from Python floats.
 * Asynchronous on the context stream. */
int fsx_adam_step(struct futhark_context *ctx, float *dev_param, const float *dev_grad,
                  float *dev_m, float *dev_v, int64_t n, double lr, double beta1,
                  double beta2, double eps, int32_t step, float lo, float hi);

/* One fitting step of the differentiable renderer in one call: the same
 * results as fsx_render_f32, then fsx_l2_loss against dev_target[3][h][w]
 * (dev_gimg = 2 (img - target), *dev_loss), then fsx_render_f32_vjp with
 * that gradient.  dev_img and dev_gimg may be NULL when the caller does not
 * need them (a fitting loop needs neither); dev_src is always written.
 * accumulate = 0 overwrites *dev_loss, dev_gplanes and *dev_gsky, else
 * adds to them.  Three launches: the visibility render (which also clears
 * the gradient on CUs its tiles leave idle, else a clear launch precedes
 * the adjoint), the adjoint, which evaluates each visible sample's colour
 * once and forms the image, the loss terms and the gradient from it, and a
 * one-workgroup sum of the loss and sky partials.  The loss is summed per
 * workgroup in a fixed order, so the loss of a step is reproducible; the
 * gradients carry the f32-atomic summation order of fsx_render_f32_vjp.
 * The loss and sky partial sums (one each per block of 64 columns by
 * b = 16 rows, or 32 in frames of 2^22 pixels or more) limit the frame to
 * ceil(w/64) * ceil(h/b) <= 16384 blocks (every frame up to 8192 x 4096);
 * larger frames are refused before anything is enqueued.  Asynchronous
 * on the context stream. */
int fsx_render_f32_l2_step(struct futhark_context *ctx, const struct fsx_map *map,
                           const struct fsx_view *view, const float *dev_planes,
                           const float *dev_sf, int32_t hc, int32_t wc, float sky_scale,
                           const float *dev_target, int32_t accumulate, float *dev_img,
                           int32_t *dev_src, float *dev_gimg, float *dev_gplanes,
                           float *dev_gsky, float *dev_loss);

/* fsx_render_f32_l2_step that allocates nothing: no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving or training loops).  Same
 * arguments, results and frame limit (16384 blocks, 8192 x 4096);
 * the per-workgroup partial sums live in a fixed scratch array of 32768
 * floats (128 KiB) per device, set up when the first context on the
 * device is created and held until the process ends, instead of the
 * context buffer fsx_render_f32_l2_step allocates on first use.  Every
 * call on a device shares that device's array, so calls that overlap on
 * one GPU (from several contexts or streams) race; calls ordered on one
 * stream do not. */
int fsx_render_f32_l2_step_noalloc(struct futhark_context *ctx, const struct fsx_map *map,
                                   const struct fsx_view *view, const float *dev_planes,
                                   const float *dev_sf, int32_t hc, int32_t wc, float sky_scale,
                                   const float *dev_target, int32_t accumulate, float *dev_img,
                                   int32_t *dev_src, float *dev_gimg, float *dev_gplanes,
                                   float *dev_gsky, float *dev_loss);

/* Render kernel shape, as chosen by the autotuner: columns per workgroup
 * tile (4 or 8), workgroups per CU (1 of 1024 threads, or 2 of 512 whose
 * depth chunks are shortened until two tiles fit in LDS) and depth slices
 * per chunk (256-2048); 0 leaves a knob automatic.  The setting belongs
 * to this context: it applies to its later renders and to batches it
 * prepares afterwards, not to other contexts.  A new context starts from
 * the FSX_RENDER_TC, FSX_RENDER_WG_PER_CU and FSX_RENDER_CHUNK environment
 * variables (automatic where unset). */
int fsx_set_tuning(struct futhark_context *ctx, int32_t tile_columns, int32_t wg_per_cu,
                   int32_t chunk);
/* The context's current tuning, as (tile_columns, wg_per_cu, chunk). */
int fsx_get_tuning(struct futhark_context *ctx, int32_t out[3]);

/* Use `stream` (a hipStream_t, or NULL for the context's own stream) for
 * all later device work of the context. */
int fsx_context_set_stream(struct futhark_context *ctx, void *stream);
/* The hipStream_t the context currently enqueues on. */
void *fsx_context_stream(struct futhark_context *ctx);
/* HIP device index the context runs on. */
int fsx_context_device(struct futhark_context *ctx);

/* Kernel timing: while enabled, every render-kernel launch is bracketed
 * by HIP events on its own stream (up to 4096 launches are kept).
 * fsx_profile_read synchronises, returns the summed kernel time in ms and
 * the number of launches, and clears the record. */
int fsx_profile_enable(struct futhark_context *ctx, int enable);
int fsx_profile_read(struct futhark_context *ctx, double *total_ms, int64_t *launches);

/* The sky colour `step` sets for a sun height: argb.scale 0xFF9090e0
 * sun_height (fut/interactive.fut:163). */
uint32_t fsx_sky(float sun_height);

/* Run device self-test `which` of the kernels' arithmetic shortcuts
 * (0: byte/255, 1: sqrt on [0,1], 2: channel quantisation, 3: screen-row
 * clamp, 4: the table quantisation of mixed channels, 5: the slice
 * table's reciprocal, 6: its divisions by the frame width, 7: sqrt of the
 * relaxed colour mix) and report how many inputs disagree with the plain formula and
 * the smallest such input's bit pattern (-1 when none; may be NULL). */
int fsx_selftest(struct futhark_context *ctx, int which, int64_t *mismatches,
                 int64_t *first_bad);

/* Version string of this library. */
const char *fsx_version(void);

#ifdef __cplusplus
}
#endif
#endif
