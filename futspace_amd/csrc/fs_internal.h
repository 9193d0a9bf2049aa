/*
 * fs_internal.h -- interface between the C host library (fs_host.c) and
 * the HIP kernel launchers (fs_kernels.hip).  Not installed; no torch or
 * HIP C++ types cross it.
 */
#ifndef FS_INTERNAL_H
#define FS_INTERNAL_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Device view of a resident map pair.
 *   hq: per texel (r,c) the bilinear quad of masked heights
 *       h(r,c), h(r,c+1), h(r+1,c), h(r+1,c+1) (indices wrap), packed as
 *       4 x u8 in a u32 (hbits 8) or 4 x u16 in a uint2 (hbits 16);
 *   cq: the same quad of shaded colours, one uint4 per shade texel. */
typedef struct {
    const void *hq;
    int32_t hrows, hcols, hbits;
    const void *cq;
    int32_t crows, ccols;
} fsk_map;

/* Per-frame scalars, computed on the host exactly as `futhark c` does:
 *   lx = Lx*z, ly = Ly*z, rx = Rx*z, ry = Ry*z  (fut/voxel_renderer.fut:47-50)
 * with Lx = (-cos a) - sin a*fov, Ly = sin a - cos a*fov,
 *      Rx = cos a - sin a*fov,   Ry = (-sin a) - cos a*fov. */
typedef struct {
    float cam_x, cam_y, cam_h, horizon;
    float Lx, Ly, Rx, Ry;
    float d, z0, scale, fw;
    float rfw;           /* 1 / fw, correctly rounded (the host's f32 division) */
    int32_t n, rows, cols;
    uint32_t sky;
    int32_t tame;        /* fsk_frame_tame(this); set by the launchers */
} fsk_frame;

/* 1 when every sample of the frame stays where the render kernel may skip
 * the out-of-range branches of the x86 casts: finite scalars, slice depths
 * increasing from z0 > 0, every sample coordinate below 2^22 and every
 * projected row below 2^30 in magnitude for any 16-bit height. */
int32_t fsk_frame_tame(const fsk_frame *f);

/* Once per device, before its first render: uploads the kernels' constant
 * tables (the byte quantisation table of the colour path).  Thread-safe;
 * later calls on the same device return at once. */
int fsk_device_init(void);

/* The device address of `device`'s fixed scratch array (set up by
 * fsk_device_init; NULL before) and its size in floats.  Shared by every
 * caller on that device. */
float *fsk_scratch(int32_t device, int64_t *n_floats);

/* render2 (ray march) of one frame into out[rows][pitch]: d_cs/d_sn[cols]
 * the host's cos / sin of angle + horz_ang per column, d_color the map's
 * unshaded colours [hrows][hcols], n = i32(distance / 4). */
int fsk_render2(const fsk_map *m, const fsk_frame *f, const float *d_cs, const float *d_sn,
                const uint32_t *d_color, int32_t n, uint32_t *d_out, int64_t pitch, hipStream_t s);

/* Tuning knobs of the render kernels (0 = automatic). */
typedef struct {
    int32_t tc;          /* frame columns per workgroup tile: 4 or 8 */
    int32_t wg_per_cu;   /* 1 or 2 workgroups per CU */
    int32_t chunk;       /* depth slices per chunk: 256..2048 */
} fsk_tuning;

/* The tuning a new context starts from: the FSX_RENDER_TC,
 * FSX_RENDER_WG_PER_CU and FSX_RENDER_CHUNK environment variables (unset:
 * 0, automatic). */
void fsk_default_tuning(fsk_tuning *t);

/* Plain render of one frame into out[rows][pitch] in the kernel shape
 * that t selects; the kernel builds its own slice table per depth chunk. */
int fsk_render(const fsk_map *m, const fsk_frame *f, const fsk_tuning *t, uint32_t *d_out,
               int64_t pitch, hipStream_t s);
/* Frame rows a render takes (frames taller than a tile's limit take a
 * slower one-thread-per-column kernel), and the rows of a batched view. */
int32_t fsk_render_max_rows(void);
int32_t fsk_batch_max_rows(void);
/* Pixel rows of one adjoint workgroup's 64-column block in a w x h frame
 * (16, or 32 in frames of 2^22 pixels or more); one partial sum per block. */
int32_t fsk_grad_block_rows(int32_t w, int32_t h);

/* One view of a batched render, as the persistent kernel reads it. */
typedef struct {
    fsk_map m;
    fsk_frame f;
    int64_t offset;      /* first output element of the view */
    int32_t ntiles;
} fsk_batch_view;

/* A prepared batch: the device arrays stay valid while it is launched. */
typedef struct {
    const fsk_batch_view *views;   /* device */
    const uint32_t *items;         /* device: view << 16 | tile, grouped by queue */
    const uint32_t *qmeta;         /* device: qstart[nqueues], qlen[nqueues] */
    uint32_t *counters;            /* device: nqueues claim counters */
    uint32_t *out;
    int32_t nqueues, max_rows, hbits, pow2;
    int32_t grid;                  /* persistent workgroups */
    fsk_tuning tune;
} fsk_batch;

/* Columns per tile the kernels use for frames of max_rows rows. */
int32_t fsk_tile_columns(int32_t max_rows, const fsk_tuning *t);
/* Persistent workgroups for a batch (CUs x workgroups per CU). */
int32_t fsk_batch_grid(int32_t max_rows, const fsk_tuning *t);
/* Reset the claim counters and run the persistent batched render. */
int fsk_batch_run(const fsk_batch *b, hipStream_t s);

/* Device self-test `which` (0-7) of the arithmetic shortcuts; adds the
 * number of mismatching inputs to *d_bad. */
int fsk_selftest(int which, unsigned long long *d_bad, hipStream_t s);

/* Map preparation. */
int fsk_height_quads(const int32_t *d_alt, int32_t rows, int32_t cols, int32_t hbits,
                     void *d_hq, uint16_t *d_hplain, hipStream_t s);
int fsk_shade(const uint16_t *d_hplain, int32_t hrows, int32_t hcols,
              const uint32_t *d_color, int32_t crows, int32_t ccols,
              float sun0, float sun1, float sun2, int32_t size,
              uint32_t *d_shaded, hipStream_t s);
int fsk_color_quads(const uint32_t *d_col, int32_t rows, int32_t cols, void *d_cq,
                    hipStream_t s);

/* Differentiable renderer (f32 colour planes, visibility held fixed).
 * fsk_render_slices writes, per pixel, the depth slice whose sample shows
 * there (-1: sky); fsk_diff_color turns it into img[3][rows][cols];
 * fsk_diff_grad adds dL/dcp into d_gcp[3][hc][wc] and dL/dsky_scale into
 * *d_gsky (both accumulate). */
int fsk_render_slices(const fsk_map *m, const fsk_frame *f, const fsk_tuning *t, int32_t *d_src,
                      int64_t pitch, hipStream_t s);
/* fsk_render_slices that also clears d_zero[zero_n] (float4-aligned) in
 * extra workgroups on the CUs its tiles leave idle, when enough are; sets
 * *cleared to 1 if it did (else the caller clears). */
int fsk_render_slices_clear(const fsk_map *m, const fsk_frame *f, const fsk_tuning *t,
                            int32_t *d_src, int64_t pitch, float *d_zero, size_t zero_n,
                            int32_t *cleared, hipStream_t s);
/* The stateless program of tests/futspace.fut: m.hq = i32 heights and
 * m.cq = i32 colours, both [hrows][hcols], nearest sampling; f.sky is the
 * program's sky (also its `empty` colour). */
int fsk_render_nearest(const fsk_map *m, const fsk_frame *f, const fsk_tuning *t,
                       uint32_t *d_out, int64_t pitch, hipStream_t s);
int fsk_diff_color(const fsk_frame *f, const float *d_cp, const float *d_sf, int32_t hc,
                   int32_t wc, float sky_scale, const int32_t *d_src, float *d_img, hipStream_t s);
int fsk_diff_grad(const fsk_frame *f, const float *d_cp, const float *d_sf, int32_t hc,
                  int32_t wc, const int32_t *d_src, const float *d_gimg, float *d_gcp,
                  float *d_gsky, int32_t sky_accumulate, float *d_partials, int32_t n_partials,
                  hipStream_t s);
/* The fused L2 step after the slice render: for each pixel the colour px
 * (written to d_img if not null), g = 2 (px - d_tgt) (written to d_g if not
 * null), the loss sum (px - tgt)^2 into *d_loss and dL/dcp into d_gcp (which
 * accumulates; the caller clears it first), dL/dsky_scale into *d_gsky.  The
 * loss and sky terms go through per-workgroup partials d_loss_part and
 * d_sky_part [n_partials each], summed in a fixed order. */
int fsk_diff_step_l2(const fsk_frame *f, const float *d_cp, const float *d_sf, int32_t hc,
                     int32_t wc, float sky_scale, const int32_t *d_src, const float *d_tgt,
                     float *d_img, float *d_g, float *d_gcp, float *d_gsky,
                     int32_t sky_accumulate, float *d_loss, int32_t loss_accumulate,
                     float *d_loss_part, float *d_sky_part, int32_t n_partials, hipStream_t s);
int fsk_zero_f32(float *d_p, size_t n, hipStream_t s);
/* One bias-corrected Adam step (step t >= 1) on d_p[n] with gradient d_g
 * and moments d_m, d_v, then d_p clamped to [lo, hi]. */
int fsk_adam(float *d_p, const float *d_g, float *d_m, float *d_v, size_t n, double lr,
             double beta1, double beta2, double eps, int32_t step, float lo, float hi,
             hipStream_t s);
int fsk_fill_u32(uint32_t *d_p, size_t n, uint32_t v, hipStream_t s);
/* Reductions write one partial sum per workgroup into d_partials (at most
 * n_partials) and a one-workgroup pass adds their sum, in a fixed order,
 * to the scalar: no atomics on the scalar, reproducible results. */
int fsk_l2_loss(const float *d_img, const float *d_tgt, size_t n, float *d_g, float *d_loss,
                int32_t accumulate, float *d_partials, int32_t n_partials, hipStream_t s);
int fsk_shade_factor(const uint16_t *d_hplain, int32_t hrows, int32_t hcols, float sun0,
                     float sun1, float sun2, int32_t size, float *d_sf, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
