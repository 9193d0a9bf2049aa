/*
 * fs_oracle.c -- CPU restatement of the futspace renderer (test
 * infrastructure only; see fs_oracle.h for scope and pinning status).
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math [-fopenmp]
 * Every expression keeps the reference's left-to-right f32 operation
 * order; the parenthesisation below is the Futhark parse made explicit.
 */
#include "fs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Futhark scalar semantics under `futhark c` on x86-64                */
/* ------------------------------------------------------------------ */

/* i32.f32: a C cast, which gcc emits as cvttss2si.  NaN, +-inf and values
 * outside [-2^31, 2^31) give the "integer indefinite" INT_MIN. */
int32_t fso_i32_f32(float v)
{
    if (!(v >= -2147483648.0f && v < 2147483648.0f))
        return INT32_MIN;
    return (int32_t)v;
}

/* u32.f32: gcc converts through a 64-bit cvttss2si and keeps the low
 * 32 bits; NaN and |v| >= 2^63 give 0x8000000000000000, i.e. 0. */
uint32_t fso_u32_f32(float v)
{
    if (!(v >= -9223372036854775808.0f && v < 9223372036854775808.0f))
        return 0u;
    return (uint32_t)(uint64_t)(int64_t)v;
}

/* Futhark `%` on signed integers rounds the quotient toward -infinity. */
int32_t fso_mod_i32(int32_t a, int32_t m)
{
    int32_t r = a % m;
    if (r != 0 && ((r < 0) != (m < 0)))
        r += m;
    return r;
}

/* ------------------------------------------------------------------ */
/* matte 0.1.2 argb (github.com/athas/matte, futhark.pkg:3).            */
/* Not vendored by the reference: this is the Gloss-derived formula     */
/* (mixColors: normalised weights, RGB mixed as root of weighted        */
/* squares, alpha mixed linearly) with channels stored as ARGB bytes.   */
/* PARITY UNPINNED against matte itself.                                */
/* ------------------------------------------------------------------ */

static float chan(uint32_t c, int shift)
{
    return (float)((c >> shift) & 0xFFu) / 255.0f;
}

void fso_to_rgba(uint32_t c, float out[4])
{
    out[0] = chan(c, 16);
    out[1] = chan(c, 8);
    out[2] = chan(c, 0);
    out[3] = chan(c, 24);
}

static float clamp_channel(float x)
{
    if (x < 0.0f)
        return 0.0f;
    else if (x > 1.0f)
        return 1.0f;
    return x;
}

uint32_t fso_from_rgba(float r, float g, float b, float a)
{
    return (fso_u32_f32(clamp_channel(a) * 255.0f) << 24)
         | (fso_u32_f32(clamp_channel(r) * 255.0f) << 16)
         | (fso_u32_f32(clamp_channel(g) * 255.0f) << 8)
         |  fso_u32_f32(clamp_channel(b) * 255.0f);
}

uint32_t fso_mix(float m1, uint32_t c1, float m2, uint32_t c2)
{
    float p[4], q[4];
    fso_to_rgba(c1, p);
    fso_to_rgba(c2, q);
    float m12 = m1 + m2;
    float m1n = m1 / m12;
    float m2n = m2 / m12;
    float r1s = p[0] * p[0], r2s = q[0] * q[0];
    float g1s = p[1] * p[1], g2s = q[1] * q[1];
    float b1s = p[2] * p[2], b2s = q[2] * q[2];
    return fso_from_rgba(sqrtf((m1n * r1s) + (m2n * r2s)),
                         sqrtf((m1n * g1s) + (m2n * g2s)),
                         sqrtf((m1n * b1s) + (m2n * b2s)),
                         ((m1 * p[3]) + (m2 * q[3])) / m12);
}

uint32_t fso_scale(uint32_t c, float s)
{
    float p[4];
    fso_to_rgba(c, p);
    return fso_from_rgba(p[0] * s, p[1] * s, p[2] * s, p[3]);
}

uint32_t fso_black(void)
{
    return fso_from_rgba(0.0f, 0.0f, 0.0f, 1.0f);
}

/* ------------------------------------------------------------------ */
/* Depth slices: get_zs, fut/voxel_renderer.fut:28-34                  */
/* ------------------------------------------------------------------ */

int32_t fso_get_zs_count(float c, float d, float z0)
{
    float sq = sqrtf(powf(d - 2.0f * z0, 2.0f) + (8.0f * d) * c);
    float num = (sq - 2.0f * z0) + d;
    float div = 2.0f * d;
    int32_t n = fso_i32_f32(floorf(num / div));
    return n < 0 ? 0 : n;
}

static float zs_term(int32_t i, float d, float z0)
{
    float fi = (float)i;
    return (fi / 2.0f) * ((2.0f * z0) + ((fi - 1.0f) * d));
}

int32_t fso_get_zs(float c, float d, float z0, float *out, int32_t cap)
{
    int32_t n = fso_get_zs_count(c, d, z0);
    for (int32_t i = 1; i <= n && i <= cap; i++)
        out[i - 1] = zs_term(i, d, z0);
    return n;
}

/* Slice depths for a render: the reference's get_zs on (cam.distance, d,
 * z0) (fut/voxel_renderer.fut:108), or an explicit count keeping the
 * same term formula. */
int32_t fso_slices(const fso_camera *cam, const fso_depth *dp, float *out, int32_t cap)
{
    int32_t n = dp->n_slices > 0 ? dp->n_slices
                                 : fso_get_zs_count(cam->distance, dp->d, dp->z0);
    if (out)
        for (int32_t i = 1; i <= n && i <= cap; i++)
            out[i - 1] = zs_term(i, dp->d, dp->z0);
    return n;
}

/* get_h_line, fut/voxel_renderer.fut:43-60.  out = {start_x, start_y,
 * dx, dy}. */
void fso_get_h_line(float z, const fso_camera *cam, int32_t w, float out[4])
{
    float sin_ang = sinf(cam->angle);
    float cos_ang = cosf(cam->angle);
    float view = cam->fov;
    float left_x = ((-cos_ang) - (sin_ang * view)) * z;
    float left_y = (sin_ang - (cos_ang * view)) * z;
    float right_x = (cos_ang - (sin_ang * view)) * z;
    float right_y = ((-sin_ang) - (cos_ang * view)) * z;
    float dx = (right_x - left_x) / (float)w;
    float dy = (right_y - left_y) / (float)w;
    out[0] = left_x + cam->x;
    out[1] = left_y + cam->y;
    out[2] = dx;
    out[3] = dy;
}

/* ------------------------------------------------------------------ */
/* Samplers, fut/render_functions.fut:63-77 and :91-105                */
/* ------------------------------------------------------------------ */

static inline size_t idx2(int32_t r, int32_t c, int32_t h, int32_t w)
{
    return (size_t)fso_mod_i32(r, h) * (size_t)w + (size_t)fso_mod_i32(c, w);
}

/* png_height, fut/render_functions.fut:63-64 (nearest, truncating). */
float fso_png_height(const int32_t *alt, int32_t h, int32_t w, float x, float y)
{
    return (float)alt[idx2(fso_i32_f32(y), fso_i32_f32(x), h, w)];
}

/* png_height_filtered, fut/render_functions.fut:67-77. */
float fso_png_height_filtered(const int32_t *alt, int32_t h, int32_t w, float x, float y)
{
    float floor_x = floorf(x), ceil_x = ceilf(x);
    float floor_y = floorf(y), ceil_y = ceilf(y);
    int32_t ify = fso_i32_f32(floor_y), icy = fso_i32_f32(ceil_y);
    int32_t ifx = fso_i32_f32(floor_x), icx = fso_i32_f32(ceil_x);
    float x_interpolated1 = ((ceil_x - x) * (float)alt[idx2(ify, ifx, h, w)])
                          + ((x - floor_x) * (float)alt[idx2(ify, icx, h, w)]);
    float x_interpolated2 = ((ceil_x - x) * (float)alt[idx2(icy, ifx, h, w)])
                          + ((x - floor_x) * (float)alt[idx2(icy, icx, h, w)]);
    return ((ceil_y - y) * x_interpolated1) + ((y - floor_y) * x_interpolated2);
}

/* png_color, fut/render_functions.fut:91-92. */
uint32_t fso_png_color(const uint32_t *col, int32_t h, int32_t w, float x, float y)
{
    return col[idx2(fso_i32_f32(y), fso_i32_f32(x), h, w)];
}

/* png_color_filtered, fut/render_functions.fut:95-105. */
uint32_t fso_png_color_filtered(const uint32_t *col, int32_t h, int32_t w, float x, float y)
{
    float floor_x = floorf(x), ceil_x = ceilf(x);
    float floor_y = floorf(y), ceil_y = ceilf(y);
    int32_t ify = fso_i32_f32(floor_y), icy = fso_i32_f32(ceil_y);
    int32_t ifx = fso_i32_f32(floor_x), icx = fso_i32_f32(ceil_x);
    uint32_t x_interpolated1 = fso_mix(ceil_x - x, col[idx2(ify, ifx, h, w)],
                                       x - floor_x, col[idx2(ify, icx, h, w)]);
    uint32_t x_interpolated2 = fso_mix(ceil_x - x, col[idx2(icy, ifx, h, w)],
                                       x - floor_x, col[idx2(icy, icx, h, w)]);
    return fso_mix(ceil_y - y, x_interpolated1, y - floor_y, x_interpolated2);
}

/* render2, fut/voxel_renderer.fut:257-305: the ray-march renderer, which
 * no entry point calls.  Pixel (y, x): vert_ang = horizon/50 +
 * (y / (h - 0)) * (1.5 - 1.0) (:292), horz_ang = (x / w) * (2.5 - 1.0)
 * (:295); for every dist in 0 ... i32(distance / 4), both ends included
 * (:296-300), the point (x + dist cos(angle + horz_ang), y + dist
 * sin(angle + horz_ang)) holds the bilinear height and colour of the
 * landscape's altitude and (unshaded) colour maps (render_map, :262-290);
 * it is (colour, dist) when height - dist * vert_ang < that height, else
 * (0, i32(distance + 1)); the pixel is the reduce with redstuff
 * (:254-255) from (0, 100000).  Restated literally: every dist is
 * evaluated, colour included, and folded in order. */
int fso_render2(const fso_camera *cam, const uint32_t *col, const int32_t *alt, int32_t r,
                int32_t s, int32_t h, int32_t w, uint32_t *out)
{
    if (r <= 0 || s <= 0 || h < 0 || w < 0)
        return 1;
    const int32_t n = fso_i32_f32(cam->distance / 4.0f);
    const int32_t miss = fso_i32_f32(cam->distance + 1.0f);
    for (int32_t y = 0; y < h; y++) {
        const float vert_ang = (cam->horizon / 50.0f) + (((float)y / ((float)h - 0.0f)) * (1.5f - 1.0f));
        for (int32_t x = 0; x < w; x++) {
            const float horz_ang = ((float)x / (float)w) * (2.5f - 1.0f);
            const float ca = cosf(cam->angle + horz_ang), sa = sinf(cam->angle + horz_ang);
            uint32_t best_c = 0u;
            int32_t best_d = 100000;
            for (int32_t dist = 0; dist <= n; dist++) {
                const float fd = (float)dist;
                const float px = cam->x + (fd * ca), py = cam->y + (fd * sa);
                const float height = fso_png_height_filtered(alt, r, s, px, py);
                const uint32_t c = fso_png_color_filtered(col, r, s, px, py);
                const int hit = (cam->height - (fd * vert_ang)) < height;
                const uint32_t ec = hit ? c : 0u;
                const int32_t ed = hit ? dist : miss;
                /* redstuff acc elem: acc if acc.dist < elem.dist, else elem */
                if (!(best_d < ed)) {
                    best_c = ec;
                    best_d = ed;
                }
            }
            out[(size_t)y * (size_t)w + (size_t)x] = best_c;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Renderer, #off branch                                               */
/* ------------------------------------------------------------------ */

/* Per-slice projection factor, fut/voxel_renderer.fut:217. */
static float inv_z_of(float z, const fso_depth *dp, int32_t w)
{
    float scale = dp->scale > 0.0f ? dp->scale : (float)(w / 2);
    return (1.0f / z) * scale;
}

/* Projection, fut/voxel_renderer.fut:223-225. */
static int32_t project(const fso_camera *cam, float map_height, float inv_z)
{
    float height_diff = cam->height - map_height;
    float relative_height = (height_diff * inv_z) + cam->horizon;
    int32_t a = fso_i32_f32(relative_height);
    return a > 0 ? a : 0;
}

/* render #off, fut/voxel_renderer.fut:214-251, restated literally: the
 * [n][w] (colour, screen-row) map, a scan with `occlude` from (0, h) per
 * column, scatter into a zero column, scan with `fill_vline` from 0, then
 * 0 -> sky. */
int fso_render_off(const fso_camera *cam, const fso_depth *dp,
                   const uint32_t *col, int32_t hc, int32_t wc,
                   const int32_t *alt, int32_t ha, int32_t wa,
                   int32_t h, int32_t w, uint32_t *out)
{
    int32_t n = fso_slices(cam, dp, NULL, 0);
    float *zs = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    uint32_t *pc = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1) * (size_t)w);
    int32_t *ph = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1) * (size_t)w);
    uint32_t *vline = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(h > 0 ? h : 1));
    if (!zs || !pc || !ph || !vline) {
        free(zs); free(pc); free(ph); free(vline);
        return 1;
    }
    fso_slices(cam, dp, zs, n);
    for (int32_t k = 0; k < n; k++) {
        float line[4];
        fso_get_h_line(zs[k], cam, w, line);
        float inv_z = inv_z_of(zs[k], dp, w);
        for (int32_t i = 0; i < w; i++) {
            float x = line[0] + ((float)i * line[2]);
            float y = line[1] + ((float)i * line[3]);
            uint32_t c = fso_png_color_filtered(col, hc, wc, x, y);
            float hgt = fso_png_height_filtered(alt, ha, wa, x, y);
            pc[(size_t)k * w + i] = c;
            ph[(size_t)k * w + i] = project(cam, hgt, inv_z);
        }
    }
    for (int32_t i = 0; i < w; i++) {
        for (int32_t r = 0; r < h; r++)
            vline[r] = 0;
        uint32_t acc_c = 0;
        int32_t acc_h = h;
        for (int32_t k = 0; k < n; k++) {
            uint32_t c2 = pc[(size_t)k * w + i];
            int32_t h2 = ph[(size_t)k * w + i];
            if (!(acc_h <= h2)) {
                acc_c = c2;
                acc_h = h2;
            }
            if (acc_h >= 0 && acc_h < h)
                vline[acc_h] = acc_c;
        }
        uint32_t fill = 0;
        for (int32_t r = 0; r < h; r++) {
            if (vline[r] != 0)
                fill = vline[r];
            out[(size_t)r * w + i] = fill == 0 ? cam->sky_color : fill;
        }
    }
    free(zs); free(pc); free(ph); free(vline);
    return 0;
}

/* One column of the front-to-back painter: a y-buffer `ymin`, the bottom
 * of the still-unpainted part of the column, spans painted as samples
 * rise strictly above the y-buffer; colour 0 leaves its span to the next
 * visible sample; the rest is sky. */
static void painter_column(const fso_camera *cam, const fso_depth *dp,
                           const float *zs, int32_t n, const float (*lines)[4],
                           const float *invz,
                           const uint32_t *col, int32_t hc, int32_t wc,
                           const int32_t *alt, int32_t ha, int32_t wa,
                           int32_t h, int32_t w, int32_t i,
                           uint32_t *out, int32_t *src)
{
    (void)dp; (void)zs;
    int32_t ymin = h, bottom = h;
    for (int32_t k = 0; k < n && ymin > 0; k++) {
        float x = lines[k][0] + ((float)i * lines[k][2]);
        float y = lines[k][1] + ((float)i * lines[k][3]);
        float hgt = fso_png_height_filtered(alt, ha, wa, x, y);
        int32_t row = project(cam, hgt, invz[k]);
        if (row < ymin) {
            uint32_t c = fso_png_color_filtered(col, hc, wc, x, y);
            if (c != 0) {
                for (int32_t r = row; r < bottom; r++) {
                    out[(size_t)r * w + i] = c;
                    if (src)
                        src[(size_t)r * w + i] = k * w + i;
                }
                bottom = row;
            }
            ymin = row;
        }
    }
    for (int32_t r = 0; r < bottom; r++) {
        out[(size_t)r * w + i] = cam->sky_color;
        if (src)
            src[(size_t)r * w + i] = -1;
    }
}

static int painter_setup(const fso_camera *cam, const fso_depth *dp, int32_t w,
                         int32_t *n_out, float **zs, float (**lines)[4], float **invz)
{
    int32_t n = fso_slices(cam, dp, NULL, 0);
    *zs = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    *lines = (float (*)[4])malloc(sizeof(float[4]) * (size_t)(n > 0 ? n : 1));
    *invz = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    if (!*zs || !*lines || !*invz)
        return 1;
    fso_slices(cam, dp, *zs, n);
    for (int32_t k = 0; k < n; k++) {
        fso_get_h_line((*zs)[k], cam, w, (*lines)[k]);
        (*invz)[k] = inv_z_of((*zs)[k], dp, w);
    }
    *n_out = n;
    return 0;
}

int fso_render_painter(const fso_camera *cam, const fso_depth *dp,
                       const uint32_t *col, int32_t hc, int32_t wc,
                       const int32_t *alt, int32_t ha, int32_t wa,
                       int32_t h, int32_t w, uint32_t *out, int32_t *src)
{
    int32_t n;
    float *zs = NULL, *invz = NULL;
    float (*lines)[4] = NULL;
    if (painter_setup(cam, dp, w, &n, &zs, &lines, &invz)) {
        free(zs); free(lines); free(invz);
        return 1;
    }
    for (int32_t i = 0; i < w; i++)
        painter_column(cam, dp, zs, n, (const float (*)[4])lines, invz, col, hc, wc,
                       alt, ha, wa, h, w, i, out, src);
    free(zs); free(lines); free(invz);
    return 0;
}

int fso_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int fso_render_painter_mt(const fso_camera *cam, const fso_depth *dp,
                          const uint32_t *col, int32_t hc, int32_t wc,
                          const int32_t *alt, int32_t ha, int32_t wa,
                          int32_t h, int32_t w, uint32_t *out, int threads)
{
    int32_t n;
    float *zs = NULL, *invz = NULL;
    float (*lines)[4] = NULL;
    if (painter_setup(cam, dp, w, &n, &zs, &lines, &invz)) {
        free(zs); free(lines); free(invz);
        return 1;
    }
#ifdef _OPENMP
    if (threads <= 0)
        threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads)
#else
    (void)threads;
#endif
    for (int32_t i = 0; i < w; i++)
        painter_column(cam, dp, zs, n, (const float (*)[4])lines, invz, col, hc, wc,
                       alt, ha, wa, h, w, i, out, NULL);
    free(zs); free(lines); free(invz);
    return 0;
}

int64_t fso_count_visible(const fso_camera *cam, const fso_depth *dp,
                          const int32_t *alt, int32_t ha, int32_t wa,
                          int32_t h, int32_t w)
{
    int32_t n;
    float *zs = NULL, *invz = NULL;
    float (*lines)[4] = NULL;
    if (painter_setup(cam, dp, w, &n, &zs, &lines, &invz)) {
        free(zs); free(lines); free(invz);
        return -1;
    }
    int64_t total = 0;
    for (int32_t i = 0; i < w; i++) {
        int32_t ymin = h;
        for (int32_t k = 0; k < n && ymin > 0; k++) {
            float x = lines[k][0] + ((float)i * lines[k][2]);
            float y = lines[k][1] + ((float)i * lines[k][3]);
            int32_t row = project(cam, fso_png_height_filtered(alt, ha, wa, x, y), invz[k]);
            if (row < ymin) {
                ymin = row;
                total++;
            }
        }
    }
    free(zs); free(lines); free(invz);
    return total;
}

/* ------------------------------------------------------------------ */
/* Shading, fut/effects.fut                                            */
/* ------------------------------------------------------------------ */

/* vec3_rotate, fut/effects.fut:6-25: a 3x3 rotation applied with
 * linalg's matvecmul_row, each row a dot product folded left from 0. */
void fso_vec3_rotate(int axis, float angle, const float v[3], float out[3])
{
    float c = cosf(angle), s = sinf(angle);
    float R[3][3];
    if (axis == 0) {
        float m[3][3] = {{1, 0, 0}, {0, c, -s}, {0, s, c}};
        memcpy(R, m, sizeof R);
    } else if (axis == 1) {
        float m[3][3] = {{c, 0, s}, {0, 1, 0}, {-s, 0, c}};
        memcpy(R, m, sizeof R);
    } else {
        float m[3][3] = {{c, -s, 0}, {s, c, 0}, {0, 0, 1}};
        memcpy(R, m, sizeof R);
    }
    for (int i = 0; i < 3; i++) {
        float acc = 0.0f;
        for (int j = 0; j < 3; j++)
            acc = acc + (R[i][j] * v[j]);
        out[i] = acc;
    }
}

/* sun = vec3_rotate #y sun_ang (vec3_rotate #z sun_height [0,1,0]),
 * fut/interactive.fut:196 with the initial sun of fut/interactive.fut:56. */
void fso_sun(float sun_height, float sun_ang, float out[3])
{
    float up[3] = {0.0f, 1.0f / sqrtf(1.0f), 0.0f};
    float t[3];
    fso_vec3_rotate(2, sun_height, up, t);
    fso_vec3_rotate(1, sun_ang, t, out);
}

/* generate_shadowmap_accumulated, fut/effects.fut:108-125, with the
 * output side `size` (1024 in the reference). */
#define SHADOW_MAX_DIST 1024
#define SHADOW_STEPS 256

/* Sun-ray steps from texel (x, y) that pass more than half a unit below
 * the terrain, fut/effects.fut:116-122. */
static int32_t shadow_count(const int32_t *alt, int32_t ha, int32_t wa, const float sun[3],
                            int32_t x, int32_t y)
{
    const float step_size = (float)(SHADOW_MAX_DIST / SHADOW_STEPS);
    float fx = (float)x, fy = (float)y;
    float height = fso_png_height(alt, ha, wa, fx, fy);
    int32_t count = 0;
    for (int32_t dist = 1; dist < SHADOW_STEPS; dist++) {
        float fdist = (float)dist;
        float t = fdist * step_size;
        float probe = fso_png_height(alt, ha, wa, fx + (t * sun[0]), fy + (t * sun[2]));
        if (((height + (t * sun[1])) - probe) < -0.5f)
            count += 1;
    }
    return count;
}

int fso_shadow_counts(const int32_t *alt, int32_t ha, int32_t wa, const float sun[3],
                      int32_t size, int32_t *counts)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t y = 0; y < size; y++)
        for (int32_t x = 0; x < size; x++)
            counts[(size_t)y * size + x] = shadow_count(alt, ha, wa, sun, x, y);
    return 0;
}

int fso_shadowmap(const uint32_t *col, int32_t hc, int32_t wc,
                  const int32_t *alt, int32_t ha, int32_t wa,
                  const float sun[3], int32_t size, uint32_t *out)
{
    const float step_size = (float)(SHADOW_MAX_DIST / SHADOW_STEPS);
    const uint32_t black = fso_black();
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t y = 0; y < size; y++) {
        for (int32_t x = 0; x < size; x++) {
            float fx = (float)x, fy = (float)y;
            float amount = (float)shadow_count(alt, ha, wa, sun, x, y);
            out[(size_t)y * size + x] =
                fso_mix(step_size * amount, black, 1.0f, fso_png_color(col, hc, wc, fx, fy));
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* #math mode samplers, fut/render_functions.fut:23-42                  */
/* ------------------------------------------------------------------ */

/* expo2: (x**2 - y**2) / 1024 inside (-1024, 1024)^2, else 0 (`**` is
 * powf, as futhark c emits it). */
float fso_expo2(float x, float y)
{
    float output = (powf(x, 2.0f) - powf(y, 2.0f)) / 1024.0f;
    if (x > -1024.0f && x < 1024.0f && y > -1024.0f && y < 1024.0f)
        return output;
    return 0.0f;
}

/* function_coloring: black on the 64-texel grid lines, else white. */
uint32_t fso_function_coloring(float x, float y)
{
    if (fso_mod_i32(fso_i32_f32(x), 64) == 0 || fso_mod_i32(fso_i32_f32(y), 64) == 0)
        return fso_black();
    return fso_from_rgba(1.0f, 1.0f, 1.0f, 1.0f);
}

/* generate_shadowmap_accumulated function_coloring expo2 sun
 * (fut/effects.fut:108-125 with the #math samplers), 1024^2. */
int fso_math_shadowmap(const float sun[3], uint32_t *out)
{
    const float step_size = (float)(SHADOW_MAX_DIST / SHADOW_STEPS);
    const uint32_t black = fso_black();
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t y = 0; y < 1024; y++) {
        for (int32_t x = 0; x < 1024; x++) {
            float fx = (float)x, fy = (float)y;
            float height = fso_expo2(fx, fy);
            int32_t count = 0;
            for (int32_t dist = 1; dist < SHADOW_STEPS; dist++) {
                float t = (float)dist * step_size;
                float probe = fso_expo2(fx + (t * sun[0]), fy + (t * sun[2]));
                if (((height + (t * sun[1])) - probe) < -0.5f)
                    count += 1;
            }
            out[(size_t)y * 1024 + x] =
                fso_mix(step_size * (float)count, black, 1.0f, fso_function_coloring(fx, fy));
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* The state's render, fut/interactive.fut:173-183 over                 */
/* fut/voxel_renderer.fut:101-251, both smoothing branches              */
/* ------------------------------------------------------------------ */

typedef struct {
    int32_t mode;                 /* 0 #math, 1 #png */
    const uint32_t *col;          /* #png: shadowed colours; #math: the math shadow map */
    int32_t hc, wc;
    const int32_t *alt;
    int32_t ha, wa;
} state_samplers;

static uint32_t state_colour(const state_samplers *sm, float x, float y)
{
    if (sm->mode == 1)
        return fso_png_color_filtered(sm->col, sm->hc, sm->wc, x, y);
    return fso_png_color(sm->col, 1024, 1024, x, y);
}

static float state_height(const state_samplers *sm, float x, float y)
{
    if (sm->mode == 1)
        return fso_png_height_filtered(sm->alt, sm->ha, sm->wa, x, y);
    return fso_expo2(x, y);
}

typedef struct {
    uint32_t c0, c1;
    int32_t h0, h1, i0, i1;
} smooth_entry;

static int entry_is_neutral(const smooth_entry *e)
{
    return e->c0 == 0 && e->c1 == 0 && e->h0 == 0 && e->h1 == 0 && e->i0 == 1000 &&
           e->i1 == 1000;
}

/* render with z_0 = 0, d = 0.001 and inv_z = (1/z) * f32(w/2)
 * (fut/voxel_renderer.fut:101-104, :217).  #off: scan occlude, scatter,
 * scan fill_vline, sky (:214-251).  #on (:174-213): scan occlude2 over
 * (colour, row, slice), the tuple of each slice with its predecessor's
 * (rotate (-1), so slice 0 pairs with the last slice), scatter into a
 * column of (0,0,0,0,1000,1000) -- the last write to a row wins, as in
 * futhark c's sequential scatter -- scan fill_vline3, then rows whose
 * entry pairs slice i with slice i-1 are argb.mix-ed between the two
 * colours across the span; finally 0 -> sky. */
int fso_render_state(const fso_camera *cam, int32_t mode, int32_t smoothing_on,
                     const uint32_t *shaded, int32_t hc, int32_t wc,
                     const int32_t *alt, int32_t ha, int32_t wa, const float sun[3],
                     int32_t h, int32_t w, uint32_t *out)
{
    if (h < 0 || w < 0)
        return 1;
    state_samplers sm = {mode, shaded, hc, wc, alt, ha, wa};
    uint32_t *mathmap = NULL;
    if (mode == 0) {
        mathmap = (uint32_t *)malloc(sizeof(uint32_t) * 1024 * 1024);
        if (!mathmap)
            return 1;
        fso_math_shadowmap(sun, mathmap);
        sm.col = mathmap;
    }
    const float d = 0.001f, z_0 = 0.0f;
    int32_t n = fso_get_zs_count(cam->distance, d, z_0);
    if (n < 0)
        n = 0;
    float *zs = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    uint32_t *pc = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1) * (size_t)(w > 0 ? w : 1));
    int32_t *ph = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1) * (size_t)(w > 0 ? w : 1));
    smooth_entry *vl = (smooth_entry *)malloc(sizeof(smooth_entry) * (size_t)(h > 0 ? h : 1));
    smooth_entry *cx = (smooth_entry *)malloc(sizeof(smooth_entry) * (size_t)(n > 0 ? n : 1));
    uint32_t *colors = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *heights = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *idxs = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int rc = 1;
    if (!zs || !pc || !ph || !vl || !cx || !colors || !heights || !idxs)
        goto done;
    fso_get_zs(cam->distance, d, z_0, zs, n);
    for (int32_t k = 0; k < n; k++) {
        float line[4];
        fso_get_h_line(zs[k], cam, w, line);
        float inv_z = (1.0f / zs[k]) * (float)(w / 2);
        for (int32_t i = 0; i < w; i++) {
            float x = line[0] + ((float)i * line[2]);
            float y = line[1] + ((float)i * line[3]);
            pc[(size_t)k * w + i] = state_colour(&sm, x, y);
            ph[(size_t)k * w + i] = project(cam, state_height(&sm, x, y), inv_z);
        }
    }
    for (int32_t i = 0; i < w; i++) {
        if (!smoothing_on) {
            uint32_t *vline = (uint32_t *)vl;       /* h u32 fit in h entries */
            for (int32_t r = 0; r < h; r++)
                vline[r] = 0;
            uint32_t acc_c = 0;
            int32_t acc_h = h;
            for (int32_t k = 0; k < n; k++) {
                uint32_t c2 = pc[(size_t)k * w + i];
                int32_t h2 = ph[(size_t)k * w + i];
                if (!(acc_h <= h2)) {
                    acc_c = c2;
                    acc_h = h2;
                }
                if (acc_h >= 0 && acc_h < h)
                    vline[acc_h] = acc_c;
            }
            uint32_t fill = 0;
            for (int32_t r = 0; r < h; r++) {
                if (vline[r] != 0)
                    fill = vline[r];
                out[(size_t)r * w + i] = fill == 0 ? cam->sky_color : fill;
            }
            continue;
        }
        uint32_t acc_c = 0;
        int32_t acc_h = h, acc_i = 0;
        for (int32_t k = 0; k < n; k++) {
            uint32_t c2 = pc[(size_t)k * w + i];
            int32_t h2 = ph[(size_t)k * w + i];
            if (!(acc_h <= h2)) {
                acc_c = c2;
                acc_h = h2;
                acc_i = k;
            }
            colors[k] = acc_c;
            heights[k] = acc_h;
            idxs[k] = acc_i;
        }
        for (int32_t k = 0; k < n; k++) {
            int32_t p = (k - 1 + n) % n;
            smooth_entry e = {colors[k], colors[p], heights[k], heights[p], idxs[k], idxs[p]};
            cx[k] = e;
        }
        const smooth_entry neutral = {0, 0, 0, 0, 1000, 1000};
        for (int32_t r = 0; r < h; r++)
            vl[r] = neutral;
        for (int32_t k = 0; k < n; k++)
            if (heights[k] >= 0 && heights[k] < h)
                vl[heights[k]] = cx[k];
        smooth_entry acc = neutral;
        for (int32_t r = 0; r < h; r++) {
            if (!entry_is_neutral(&vl[r]))
                acc = vl[r];
            float range = fmaxf(1.0f, (float)(acc.h1 - acc.h0));
            float delta1 = fabsf((float)r - (float)acc.h1) / range;
            float delta2 = fabsf((float)acc.h0 - (float)r) / range;
            uint32_t colour = (acc.i0 - acc.i1 == 1) ? fso_mix(delta2, acc.c1, delta1, acc.c0)
                                                      : acc.c0;
            out[(size_t)r * w + i] = colour == 0 ? cam->sky_color : colour;
        }
    }
    rc = 0;
done:
    free(zs); free(pc); free(ph); free(vl); free(cx); free(colors); free(heights); free(idxs);
    free(mathmap);
    return rc;
}

/* Sky colour set by `step`, fut/interactive.fut:163. *//* Sky colour set by `step`, fut/interactive.fut:163. */
uint32_t fso_sky(float sun_height)
{
    return fso_scale(0xFF9090E0u, sun_height);
}

/* ------------------------------------------------------------------ */
/* Stateless benchmark program, tests/futspace.fut:47-145              */
/* ------------------------------------------------------------------ */

static int32_t nearest_slices(const fso_camera *cam, float **zs_out)
{
    const float z_0 = 1.0f, delta = 0.005f;
    float sq = sqrtf(powf(delta - 2.0f * z_0, 2.0f) + (8.0f * delta) * cam->distance);
    float num = (sq - 2.0f * z_0) + delta;
    float div = 2.0f * delta;
    int32_t n = fso_i32_f32(num / div);
    if (n < 0)
        n = 0;
    float *zs = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 1; i <= n; i++)
        zs[i - 1] = zs_term(i, delta, z_0);
    *zs_out = zs;
    return n;
}

static void nearest_sample(const fso_camera *cam, const float line[4], int32_t j,
                           const int32_t *col, const int32_t *alt, int32_t q, int32_t r,
                           float inv_z, int32_t *c_out, int32_t *h_out)
{
    int32_t x = fso_i32_f32(line[0] + ((float)j * line[2]));
    int32_t y = fso_i32_f32(line[1] + ((float)j * line[3]));
    size_t id = (size_t)fso_mod_i32(y, q) * (size_t)r + (size_t)fso_mod_i32(x, r);
    float height_diff = cam->height - (float)alt[id];
    float rel = (height_diff * inv_z) + cam->horizon;
    int32_t a = fso_i32_f32(rel);
    *c_out = col[id];
    *h_out = a > 0 ? a : 0;
}

int fso_render_nearest(const fso_camera *cam, const int32_t *col, const int32_t *alt,
                       int32_t q, int32_t r, int32_t l, int32_t m, int32_t sky,
                       int32_t *out)
{
    float *zs;
    int32_t n = nearest_slices(cam, &zs);
    int32_t *pc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1) * (size_t)m);
    int32_t *ph = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1) * (size_t)m);
    int32_t *colv = (int32_t *)malloc(sizeof(int32_t) * (size_t)(l > 0 ? l : 1));
    for (int32_t k = 0; k < n; k++) {
        float line[4];
        fso_get_h_line(zs[k], cam, m, line);
        float inv_z = (1.0f / zs[k]) * 240.0f;
        for (int32_t j = 0; j < m; j++)
            nearest_sample(cam, line, j, col, alt, q, r, inv_z,
                           &pc[(size_t)k * m + j], &ph[(size_t)k * m + j]);
    }
    for (int32_t j = 0; j < m; j++) {
        for (int32_t y = 0; y < l; y++)
            colv[y] = sky;
        int32_t acc_c = 0, acc_h = l;
        for (int32_t k = 0; k < n; k++) {
            int32_t c2 = pc[(size_t)k * m + j], h2 = ph[(size_t)k * m + j];
            if (!(acc_h <= h2)) {
                acc_c = c2;
                acc_h = h2;
            }
            if (acc_h >= 0 && acc_h < l)
                colv[acc_h] = acc_c;
        }
        int32_t fill = sky;
        for (int32_t y = 0; y < l; y++) {
            if (colv[y] != sky)
                fill = colv[y];
            out[(size_t)y * m + j] = fill;
        }
    }
    free(zs); free(pc); free(ph); free(colv);
    return 0;
}

int fso_render_nearest_painter(const fso_camera *cam, const int32_t *col, const int32_t *alt,
                               int32_t q, int32_t r, int32_t l, int32_t m, int32_t sky,
                               int32_t *out)
{
    float *zs;
    int32_t n = nearest_slices(cam, &zs);
    for (int32_t j = 0; j < m; j++) {
        int32_t ymin = l, bottom = l;
        for (int32_t k = 0; k < n && ymin > 0; k++) {
            float line[4];
            fso_get_h_line(zs[k], cam, m, line);
            float inv_z = (1.0f / zs[k]) * 240.0f;
            int32_t c, row;
            nearest_sample(cam, line, j, col, alt, q, r, inv_z, &c, &row);
            if (row < ymin) {
                if (c != sky) {
                    for (int32_t y = row; y < bottom; y++)
                        out[(size_t)y * m + j] = c;
                    bottom = row;
                }
                ymin = row;
            }
        }
        for (int32_t y = 0; y < bottom; y++)
            out[(size_t)y * m + j] = sky;
    }
    free(zs);
    return 0;
}

/* ------------------------------------------------------------------ */
/* f32-plane relaxation (differentiable renderer)                      */
/* ------------------------------------------------------------------ */

static const float sky_base[3] = {144.0f / 255.0f, 144.0f / 255.0f, 224.0f / 255.0f};

typedef struct {
    size_t t[4];            /* texel offsets: (fy,fx) (fy,cx) (cy,fx) (cy,cx) */
    float wx0, wx1, wy0, wy1;
} bilin;

static int bilin_of(float x, float y, int32_t hc, int32_t wc, bilin *b)
{
    float floor_x = floorf(x), ceil_x = ceilf(x);
    float floor_y = floorf(y), ceil_y = ceilf(y);
    int32_t ify = fso_i32_f32(floor_y), icy = fso_i32_f32(ceil_y);
    int32_t ifx = fso_i32_f32(floor_x), icx = fso_i32_f32(ceil_x);
    b->t[0] = idx2(ify, ifx, hc, wc);
    b->t[1] = idx2(ify, icx, hc, wc);
    b->t[2] = idx2(icy, ifx, hc, wc);
    b->t[3] = idx2(icy, icx, hc, wc);
    b->wx0 = ceil_x - x;
    b->wx1 = x - floor_x;
    b->wy0 = ceil_y - y;
    b->wy1 = y - floor_y;
    /* the u32 path's argb.mix divides by a zero weight sum here: colour 0,
     * i.e. a transparent sample */
    return (floor_x == ceil_x) || (floor_y == ceil_y);
}

static float mixf(float m1, float a, float m2, float b)
{
    float m12 = m1 + m2;
    float m1n = m1 / m12, m2n = m2 / m12;
    return sqrtf((m1n * (a * a)) + (m2n * (b * b)));
}

int fso_render_f32(const fso_camera *cam, const fso_depth *dp,
                   const float *cp, const float *sf, int32_t hc, int32_t wc,
                   const int32_t *alt, int32_t ha, int32_t wa,
                   int32_t h, int32_t w, float sky_scale, float *img, int32_t *src)
{
    int32_t n;
    float *zs = NULL, *invz = NULL;
    float (*lines)[4] = NULL;
    if (painter_setup(cam, dp, w, &n, &zs, &lines, &invz)) {
        free(zs); free(lines); free(invz);
        return 1;
    }
    size_t plane = (size_t)hc * (size_t)wc, fplane = (size_t)h * (size_t)w;
    for (int32_t i = 0; i < w; i++) {
        int32_t ymin = h, bottom = h;
        for (int32_t k = 0; k < n && ymin > 0; k++) {
            float x = lines[k][0] + ((float)i * lines[k][2]);
            float y = lines[k][1] + ((float)i * lines[k][3]);
            int32_t row = project(cam, fso_png_height_filtered(alt, ha, wa, x, y), invz[k]);
            if (row < ymin) {
                bilin b;
                if (!bilin_of(x, y, hc, wc, &b)) {
                    float px[3];
                    for (int ch = 0; ch < 3; ch++) {
                        const float *P = cp + (size_t)ch * plane;
                        float s0 = P[b.t[0]] * sf[b.t[0]], s1 = P[b.t[1]] * sf[b.t[1]];
                        float s2 = P[b.t[2]] * sf[b.t[2]], s3 = P[b.t[3]] * sf[b.t[3]];
                        float X1 = mixf(b.wx0, s0, b.wx1, s1);
                        float X2 = mixf(b.wx0, s2, b.wx1, s3);
                        px[ch] = mixf(b.wy0, X1, b.wy1, X2);
                    }
                    for (int32_t r = row; r < bottom; r++) {
                        for (int ch = 0; ch < 3; ch++)
                            img[(size_t)ch * fplane + (size_t)r * w + i] = px[ch];
                        src[(size_t)r * w + i] = k * w + i;
                    }
                    bottom = row;
                }
                ymin = row;
            }
        }
        for (int32_t r = 0; r < bottom; r++) {
            for (int ch = 0; ch < 3; ch++)
                img[(size_t)ch * fplane + (size_t)r * w + i] = sky_base[ch] * sky_scale;
            src[(size_t)r * w + i] = -1;
        }
    }
    free(zs); free(lines); free(invz);
    return 0;
}

/* d sqrt(a'*p^2 + b'*q^2) / dp = a'*p / out (0 where out == 0). */
static void mixf_grad(float m1, float a, float m2, float b, double g, double *ga, double *gb)
{
    float m12 = m1 + m2;
    double m1n = (double)(m1 / m12), m2n = (double)(m2 / m12);
    double o = sqrt(m1n * (double)a * (double)a + m2n * (double)b * (double)b);
    if (o > 0.0) {
        *ga = g * m1n * (double)a / o;
        *gb = g * m2n * (double)b / o;
    } else {
        *ga = 0.0;
        *gb = 0.0;
    }
}

int fso_render_f32_vjp(const fso_camera *cam, const fso_depth *dp,
                       const float *cp, const float *sf, int32_t hc, int32_t wc,
                       int32_t h, int32_t w, const int32_t *src,
                       const float *gimg, float *gcp, float *gsky)
{
    int32_t n;
    float *zs = NULL, *invz = NULL;
    float (*lines)[4] = NULL;
    if (painter_setup(cam, dp, w, &n, &zs, &lines, &invz)) {
        free(zs); free(lines); free(invz);
        return 1;
    }
    size_t plane = (size_t)hc * (size_t)wc, fplane = (size_t)h * (size_t)w;
    double *acc = (double *)calloc(3 * plane, sizeof(double));
    double sky_acc = 0.0;
    for (size_t p = 0; p < fplane; p++) {
        int32_t s = src[p];
        if (s < 0) {
            for (int ch = 0; ch < 3; ch++)
                sky_acc += (double)gimg[(size_t)ch * fplane + p] * (double)sky_base[ch];
            continue;
        }
        int32_t k = s / w, i = s % w;
        float x = lines[k][0] + ((float)i * lines[k][2]);
        float y = lines[k][1] + ((float)i * lines[k][3]);
        bilin b;
        bilin_of(x, y, hc, wc, &b);
        for (int ch = 0; ch < 3; ch++) {
            const float *P = cp + (size_t)ch * plane;
            float sv[4];
            for (int t = 0; t < 4; t++)
                sv[t] = P[b.t[t]] * sf[b.t[t]];
            float X1 = mixf(b.wx0, sv[0], b.wx1, sv[1]);
            float X2 = mixf(b.wx0, sv[2], b.wx1, sv[3]);
            double g = (double)gimg[(size_t)ch * fplane + p];
            double gX1, gX2, g0, g1, g2, g3;
            mixf_grad(b.wy0, X1, b.wy1, X2, g, &gX1, &gX2);
            mixf_grad(b.wx0, sv[0], b.wx1, sv[1], gX1, &g0, &g1);
            mixf_grad(b.wx0, sv[2], b.wx1, sv[3], gX2, &g2, &g3);
            double gs[4] = {g0, g1, g2, g3};
            for (int t = 0; t < 4; t++)
                acc[(size_t)ch * plane + b.t[t]] += gs[t] * (double)sf[b.t[t]];
        }
    }
    for (size_t q = 0; q < 3 * plane; q++)
        gcp[q] = (float)acc[q];
    *gsky = (float)sky_acc;
    free(acc);
    free(zs); free(lines); free(invz);
    return 0;
}
