/*
 * fs_oracle.h -- CPU restatement of the futspace Voxel Space renderer.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * renderer and the timed CPU baseline; nothing in the product links it.
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.
 *
 * It restates, in plain C11 with -ffp-contract=off, the f32 operation
 * order of the Futhark reference as `futhark c` would execute it on
 * x86-64:
 *   - i32.f32 / u32.f32 are C casts (cvttss2si: NaN, +-inf and out of
 *     range give INT_MIN; u32 goes through a 64-bit conversion);
 *   - integer % is floor-mod;
 *   - f32.sqrt, floor, ceil, sin, cos and ** are glibc sqrtf, floorf,
 *     ceilf, sinf, cosf and powf;
 *   - scan is a left fold from the neutral element, scatter drops
 *     out-of-range indices.
 *
 * Pinning status of each function (what checks it independently of this
 * file) is listed in tests/test_oracle.py and summarised here:
 *   fso_get_zs            closed form z_i = i*z0 + d*i*(i-1)/2 and the
 *                         s-macke stepping of tests/javascript_zs.py
 *   fso_get_h_line        yaw 0 / yaw pi/2 special cases, brute force
 *   fso_png_height_*      constant and linear-ramp maps, integer points
 *   fso_render_off        independent painter restatement (fso_render_
 *                         painter), flat-map horizon row, z0=0 uniform
 *                         frame, invariants
 *   fso_shadowmap         closed-form sun vector, flat map (count 0),
 *                         single spike counted by hand
 *   fso_render_nearest    painter cross-check (tests/futspace.fut)
 *   fso_render2           camera inside terrain (uniform frame of the
 *                         bilinear colour under it), flat-map hit rows in
 *                         closed form (vert_ang, inclusive march end)
 *   fso_mix/from/to/scale PARITY UNPINNED: matte 0.1.2 (colour.fut) is
 *                         not vendored in the reference; the formula is
 *                         the Gloss mixColors restatement, checked only
 *                         against its own algebra (mix of equal colours,
 *                         weight 0/1 limits), not against matte itself.
 */
#ifndef FS_ORACLE_H
#define FS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* camera record, fut/voxel_renderer.fut:5-12 */
typedef struct {
    float x, y, height, angle, horizon, distance, fov;
    uint32_t sky_color;
} fso_camera;

/* Depth-slice and projection parameters.  The reference fixes d = 0.001
 * and z_0 = 0 (fut/voxel_renderer.fut:103-104) and the projection scale
 * to f32(w/2) (fut/voxel_renderer.fut:217); README.md:37-41 names them
 * invz_param1 (d), invz_param2 (z_0) and magic_number (scale).
 * scale <= 0 selects f32(w/2); n_slices <= 0 derives the count from
 * get_zs, n_slices > 0 keeps get_zs' term formula for i = 1..n_slices. */
typedef struct {
    float d, z0, scale;
    int32_t n_slices;
} fso_depth;

/* ---- Futhark scalar semantics (runtime, x86-64 `futhark c`) ---- */
int32_t  fso_i32_f32(float v);
uint32_t fso_u32_f32(float v);
int32_t  fso_mod_i32(int32_t a, int32_t m);

/* ---- matte argb (external dependency, parity unpinned) ---- */
void     fso_to_rgba(uint32_t c, float out[4]);
uint32_t fso_from_rgba(float r, float g, float b, float a);
uint32_t fso_mix(float m1, uint32_t c1, float m2, uint32_t c2);
uint32_t fso_scale(uint32_t c, float s);
uint32_t fso_black(void);

/* ---- depth slices and lines ---- */
int32_t fso_get_zs_count(float c, float d, float z0);
int32_t fso_get_zs(float c, float d, float z0, float *out, int32_t cap);
int32_t fso_slices(const fso_camera *cam, const fso_depth *dp, float *out, int32_t cap);
void    fso_get_h_line(float z, const fso_camera *cam, int32_t w, float out[4]);

/* ---- samplers (fut/render_functions.fut) ---- */
float    fso_png_height(const int32_t *alt, int32_t h, int32_t w, float x, float y);
float    fso_png_height_filtered(const int32_t *alt, int32_t h, int32_t w, float x, float y);
uint32_t fso_png_color(const uint32_t *col, int32_t h, int32_t w, float x, float y);
uint32_t fso_png_color_filtered(const uint32_t *col, int32_t h, int32_t w, float x, float y);

/* ---- renderer, #off branch (fut/voxel_renderer.fut:214-251) ---- */
/* render2, fut/voxel_renderer.fut:257-305 (ray march over the altitude and
 * unshaded colour maps [r][s]) into out[h][w]. */
int fso_render2(const fso_camera *cam, const uint32_t *col, const int32_t *alt, int32_t r,
                int32_t s, int32_t h, int32_t w, uint32_t *out);

int fso_render_off(const fso_camera *cam, const fso_depth *dp,
                   const uint32_t *col, int32_t hc, int32_t wc,
                   const int32_t *alt, int32_t ha, int32_t wa,
                   int32_t h, int32_t w, uint32_t *out);

/* Same frame through an independent front-to-back painter with a
 * y-buffer; also reports per pixel the (slice, column) sample that
 * coloured it as slice*w + column, or -1 for sky. */
int fso_render_painter(const fso_camera *cam, const fso_depth *dp,
                       const uint32_t *col, int32_t hc, int32_t wc,
                       const int32_t *alt, int32_t ha, int32_t wa,
                       int32_t h, int32_t w, uint32_t *out, int32_t *src);

/* Number of visible samples (strict y-buffer lowerings) per frame. */
int64_t fso_count_visible(const fso_camera *cam, const fso_depth *dp,
                          const int32_t *alt, int32_t ha, int32_t wa,
                          int32_t h, int32_t w);

/* Multi-threaded (OpenMP over columns) painter, used as the timed
 * `futhark multicore` stand-in.  threads <= 0 uses all cores. */
int fso_render_painter_mt(const fso_camera *cam, const fso_depth *dp,
                          const uint32_t *col, int32_t hc, int32_t wc,
                          const int32_t *alt, int32_t ha, int32_t wa,
                          int32_t h, int32_t w, uint32_t *out, int threads);
int fso_max_threads(void);

/* ---- shading (fut/effects.fut:6-25, :108-125) ---- */
void fso_vec3_rotate(int axis, float angle, const float v[3], float out[3]);
void fso_sun(float sun_height, float sun_ang, float out[3]);
int  fso_shadowmap(const uint32_t *col, int32_t hc, int32_t wc,
                   const int32_t *alt, int32_t ha, int32_t wa,
                   const float sun[3], int32_t size, uint32_t *out);
/* The per-texel count of shadowing sun-ray steps that fso_shadowmap mixes
 * with black (fut/effects.fut:116-122). */
int  fso_shadow_counts(const int32_t *alt, int32_t ha, int32_t wa, const float sun[3],
                       int32_t size, int32_t *counts);
uint32_t fso_sky(float sun_height);

/* ---- #math samplers and the state's render (both smoothing modes) ---- */
float    fso_expo2(float x, float y);
uint32_t fso_function_coloring(float x, float y);
int      fso_math_shadowmap(const float sun[3], uint32_t *out);   /* [1024][1024] */
/* The state's render entry: mode 0 #math, 1 #png; z_0 = 0, d = 0.001,
 * scale f32(w/2) as in fut/voxel_renderer.fut:101-104. */
int      fso_render_state(const fso_camera *cam, int32_t mode, int32_t smoothing_on,
                          const uint32_t *shaded, int32_t hc, int32_t wc,
                          const int32_t *alt, int32_t ha, int32_t wa, const float sun[3],
                          int32_t h, int32_t w, uint32_t *out);

/* ---- stateless benchmark program (tests/futspace.fut:47-145) ---- */
int fso_render_nearest(const fso_camera *cam, const int32_t *col, const int32_t *alt,
                       int32_t q, int32_t r, int32_t l, int32_t m, int32_t sky,
                       int32_t *out);
int fso_render_nearest_painter(const fso_camera *cam, const int32_t *col, const int32_t *alt,
                               int32_t q, int32_t r, int32_t l, int32_t m, int32_t sky,
                               int32_t *out);

/* ---- f32-plane relaxation used by the differentiable renderer ----
 * Colour planes cp[3][hc][wc] in [0,1] and a per-texel shade factor
 * sf[hc][wc] = sqrt(1/(4*count+1)) (the relaxed argb.mix with black of
 * fut/effects.fut:123); visibility comes from the integer renderer.
 * Output img[3][h][w].  The channel mix is the Gloss formula without
 * u8 quantisation. */
int fso_render_f32(const fso_camera *cam, const fso_depth *dp,
                   const float *cp, const float *sf, int32_t hc, int32_t wc,
                   const int32_t *alt, int32_t ha, int32_t wa,
                   int32_t h, int32_t w, float sky_scale, float *img, int32_t *src);
/* Adjoint of fso_render_f32 w.r.t. the colour planes and sky_scale. */
int fso_render_f32_vjp(const fso_camera *cam, const fso_depth *dp,
                       const float *cp, const float *sf, int32_t hc, int32_t wc,
                       int32_t h, int32_t w, const int32_t *src,
                       const float *gimg, float *gcp, float *gsky);

#ifdef __cplusplus
}
#endif
#endif
