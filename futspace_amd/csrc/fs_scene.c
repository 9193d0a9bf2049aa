/* fs_scene.c -- seeded synthetic scenes (see fs_scene.h). */
#include "fs_scene.h"

#include <math.h>
#include <stdlib.h>

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* PCG32 (XSH-RR), seeded from SplitMix64 so nearby seeds decorrelate. */
void fss_rng_seed(fss_rng *r, uint64_t seed, uint64_t stream)
{
    uint64_t s = seed ^ (stream * 0xD1B54A32D192ED03ull);
    r->state = 0;
    r->inc = (splitmix64(&s) << 1) | 1u;
    fss_rng_u32(r);
    r->state += splitmix64(&s);
    fss_rng_u32(r);
}

uint32_t fss_rng_u32(fss_rng *r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ull + r->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

double fss_rng_unit(fss_rng *r)
{
    return (double)fss_rng_u32(r) * (1.0 / 4294967296.0);
}

int32_t fss_rng_int(fss_rng *r, int32_t lo, int32_t hi)
{
    uint32_t span = (uint32_t)(hi - lo) + 1u;
    if (span == 0)
        return (int32_t)fss_rng_u32(r);
    /* Lemire-style rejection keeps the draw unbiased */
    uint32_t lim = (uint32_t)(-span) % span;
    for (;;) {
        uint32_t v = fss_rng_u32(r);
        uint64_t m = (uint64_t)v * span;
        if ((uint32_t)m >= lim)
            return lo + (int32_t)(m >> 32);
    }
}

int fss_heightmap(uint64_t seed, int32_t size, double roughness, int32_t max_value,
                  int32_t *out)
{
    if (size < 2 || (size & (size - 1)) != 0 || max_value < 1)
        return 1;
    size_t n = (size_t)size;
    double *g = (double *)malloc(sizeof(double) * n * n);
    if (!g)
        return 1;
    fss_rng r;
    fss_rng_seed(&r, seed, 1);
    const size_t mask = n - 1;
#define G(yy, xx) g[(((size_t)(yy)) & mask) * n + (((size_t)(xx)) & mask)]
    G(0, 0) = 0.0;
    double amp = 1.0;
    for (size_t step = n; step > 1; step /= 2) {
        size_t half = step / 2;
        /* diamond step: centres of squares */
        for (size_t y = 0; y < n; y += step)
            for (size_t x = 0; x < n; x += step) {
                double avg = (G(y, x) + G(y, x + step) + G(y + step, x) + G(y + step, x + step)) * 0.25;
                G(y + half, x + half) = avg + (2.0 * fss_rng_unit(&r) - 1.0) * amp;
            }
        /* square step: edge midpoints */
        for (size_t y = 0; y < n; y += half) {
            size_t x0 = ((y / half) % 2 == 0) ? half : 0;
            for (size_t x = x0; x < n; x += step) {
                double avg = (G(y, x + n - half) + G(y, x + half) + G(y + n - half, x) + G(y + half, x)) * 0.25;
                G(y, x) = avg + (2.0 * fss_rng_unit(&r) - 1.0) * amp;
            }
        }
        amp *= roughness;
    }
#undef G
    double lo = g[0], hi = g[0];
    for (size_t i = 0; i < n * n; i++) {
        if (g[i] < lo) lo = g[i];
        if (g[i] > hi) hi = g[i];
    }
    double span = hi > lo ? hi - lo : 1.0;
    for (size_t i = 0; i < n * n; i++) {
        double v = (g[i] - lo) / span * (double)max_value;
        int32_t q = (int32_t)floor(v + 0.5);
        out[i] = q < 0 ? 0 : (q > max_value ? max_value : q);
    }
    free(g);
    return 0;
}

int fss_colormap(uint64_t seed, const int32_t *heights, int32_t h, int32_t w,
                 int32_t max_value, int32_t noise, uint32_t *out)
{
    static const int32_t band_top[5] = {77, 92, 153, 204, 256};   /* of 256 */
    static const int32_t pal[5][3] = {
        {36, 74, 148},    /* water */
        {200, 184, 134},  /* sand */
        {72, 136, 58},    /* grass */
        {116, 104, 94},   /* rock */
        {238, 238, 244},  /* snow */
    };
    fss_rng r;
    fss_rng_seed(&r, seed, 2);
    size_t n = (size_t)h * (size_t)w;
    for (size_t i = 0; i < n; i++) {
        int64_t t = (int64_t)heights[i] * 256 / (int64_t)(max_value + 1);
        int b = 0;
        while (b < 4 && t >= band_top[b])
            b++;
        uint32_t c = 0xFF000000u;
        for (int ch = 0; ch < 3; ch++) {
            int32_t v = pal[b][ch] + fss_rng_int(&r, -noise, noise);
            v = v < 0 ? 0 : (v > 255 ? 255 : v);
            c |= (uint32_t)v << (16 - 8 * ch);
        }
        out[i] = c;
    }
    return 0;
}

static double catmull(double p0, double p1, double p2, double p3, double t)
{
    double t2 = t * t, t3 = t2 * t;
    return 0.5 * ((2.0 * p1) + (-p0 + p2) * t + (2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3) * t2
                  + (-p0 + 3.0 * p1 - 3.0 * p2 + p3) * t3);
}

int fss_flythrough(uint64_t seed, int32_t n, int32_t n_ctrl, int32_t map_size,
                   float alt_lo, float alt_hi, float horizon_mid, float horizon_amp,
                   fss_pose *out)
{
    if (n < 1 || n_ctrl < 4)
        return 1;
    fss_rng r;
    fss_rng_seed(&r, seed, 3);
    double *cx = (double *)malloc(sizeof(double) * (size_t)n_ctrl);
    double *cy = (double *)malloc(sizeof(double) * (size_t)n_ctrl);
    double *ch = (double *)malloc(sizeof(double) * (size_t)n_ctrl);
    if (!cx || !cy || !ch) {
        free(cx); free(cy); free(ch);
        return 1;
    }
    for (int32_t i = 0; i < n_ctrl; i++) {
        cx[i] = fss_rng_unit(&r) * map_size;
        cy[i] = fss_rng_unit(&r) * map_size;
        ch[i] = alt_lo + fss_rng_unit(&r) * (alt_hi - alt_lo);
    }
    double phase = fss_rng_unit(&r) * 6.283185307179586;
    for (int32_t f = 0; f < n; f++) {
        double u = (double)f / (double)n;           /* loop parameter in [0,1) */
        double s = u * n_ctrl;
        int32_t seg = (int32_t)floor(s);
        double t = s - seg;
        int32_t i0 = (seg - 1 + n_ctrl) % n_ctrl, i1 = seg % n_ctrl;
        int32_t i2 = (seg + 1) % n_ctrl, i3 = (seg + 2) % n_ctrl;
        double hgt = catmull(ch[i0], ch[i1], ch[i2], ch[i3], t);
        if (hgt < alt_lo) hgt = alt_lo;
        if (hgt > alt_hi) hgt = alt_hi;
        out[f].x = (float)catmull(cx[i0], cx[i1], cx[i2], cx[i3], t);
        out[f].y = (float)catmull(cy[i0], cy[i1], cy[i2], cy[i3], t);
        out[f].height = (float)hgt;
        out[f].angle = (float)(u * 6.283185307179586);
        out[f].horizon = (float)(horizon_mid + horizon_amp * sin(2.0 * 6.283185307179586 * u + phase));
    }
    free(cx); free(cy); free(ch);
    return 0;
}
