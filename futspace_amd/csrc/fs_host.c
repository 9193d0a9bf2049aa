/*
 * fs_host.c -- host side of libfutspace for MI355X.
 *
 * Implements the C-ABI that `futhark opencl --library` generates for
 * fut/interactive_entrypoints.fut (context, arrays, opaque state, nine
 * entries), plus the fsx_* extension entries.  The state machine of
 * fut/interactive.fut (camera keys, sun, shading refresh, terrain
 * collision) is scalar host C; all per-pixel and per-texel work runs in
 * the HIP kernels of fs_kernels.hip.  Per-frame scalars (sin/cos of the
 * yaw, the get_zs slice count) are computed here with glibc, as the
 * `futhark c` build of the reference does.
 *
 * Build with -std=c11 -ffp-contract=off -fno-fast-math: the f32 host
 * arithmetic must round exactly like the reference's.
 */
#define _POSIX_C_SOURCE 200809L
#include "../../include/libfutspace.h"
#include "fs_internal.h"

#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#define FS_VERSION "futspace-mi355x 0.1"

/* ------------------------------------------------------------------ */
/* Context                                                             */
/* ------------------------------------------------------------------ */

struct futhark_context_config {
    int device;
    int debugging, profiling, logging;
    char device_name[256];
};

struct fs_terrain;
struct fs_shade;

#define FS_POOL_SLOTS 16

struct futhark_context {
    int device;
    int debugging, logging;
    hipStream_t own_stream;
    hipStream_t stream;
    char *error;
    struct {
        void *ptr;
        size_t bytes;
    } pool[FS_POOL_SLOTS];    /* freed output buffers kept for reuse */
    struct fs_terrain *empty_terrain;
    struct fs_shade *empty_shade;
    fsk_tuning tune;          /* render kernel shape (0 = automatic) */
    int prof_on;              /* bracket render kernels with events */
    int prof_n;
    hipEvent_t *prof_ev;      /* FS_PROF_MAX pairs */
    float *d_partials;        /* per-workgroup partial sums of the reductions */
    float *d_trig;            /* render2's per-column cos / sin */
    size_t trig_cap;          /* floats in d_trig */
};

#define FS_PARTIALS 32768     /* floats in d_partials */
#define FS_STEP_BLOCK_COLS 64 /* pixel blocks of the adjoint (and of the fused step) */

/* Pixel blocks of a w x h frame in the differentiable passes (64 columns
 * by fsk_grad_block_rows rows): one partial sum per block. */
static int64_t step_blocks(int32_t w, int32_t h)
{
    const int32_t br = fsk_grad_block_rows(w, h);
    return (int64_t)((w + FS_STEP_BLOCK_COLS - 1) / FS_STEP_BLOCK_COLS) *
           (int64_t)((h + br - 1) / br);
}

#define FS_PROF_MAX 4096

static void set_error(struct futhark_context *ctx, const char *fmt, ...)
{
    if (!ctx)
        return;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx->error) {
        /* keep the first error, like the generated library's error chain */
        size_t l0 = strlen(ctx->error), l1 = strlen(buf);
        char *e = (char *)realloc(ctx->error, l0 + l1 + 2);
        if (e) {
            e[l0] = '\n';
            memcpy(e + l0 + 1, buf, l1 + 1);
            ctx->error = e;
        }
        return;
    }
    ctx->error = strdup(buf);
}

/* A NULL context fails silently (there is nowhere to put a message);
 * any other missing argument leaves one. */
#define REQUIRE_ARGS(ctx, ok, fn)                                                   \
    do {                                                                            \
        if (!(ctx))                                                                 \
            return 1;                                                               \
        if (!(ok)) {                                                                \
            set_error((ctx), "%s: null argument", (fn));                            \
            return 1;                                                               \
        }                                                                           \
    } while (0)

#define HIPCHECK(ctx, call)                                                         \
    do {                                                                            \
        hipError_t e_ = (hipError_t)(call);                                         \
        if (e_ != hipSuccess) {                                                     \
            set_error((ctx), "%s:%d: %s: %s", __FILE__, __LINE__, #call,            \
                      hipGetErrorString(e_));                                       \
            return 1;                                                               \
        }                                                                           \
    } while (0)

static int dev_alloc(struct futhark_context *ctx, void **p, size_t bytes)
{
    *p = NULL;
    if (bytes == 0)
        bytes = 4;
    for (int i = 0; i < FS_POOL_SLOTS; i++) {
        if (ctx->pool[i].ptr && ctx->pool[i].bytes == bytes) {
            *p = ctx->pool[i].ptr;
            ctx->pool[i].ptr = NULL;
            return 0;
        }
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        /* release cached buffers and retry once */
        for (int i = 0; i < FS_POOL_SLOTS; i++)
            if (ctx->pool[i].ptr) {
                hipFree(ctx->pool[i].ptr);
                ctx->pool[i].ptr = NULL;
            }
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        *p = NULL;
        set_error(ctx, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        return 1;
    }
    return 0;
}

/* Return a buffer to the reuse pool (stream order keeps reuse safe: every
 * later writer runs after every earlier reader on the one stream). */
static void dev_release(struct futhark_context *ctx, void *p, size_t bytes)
{
    if (!p)
        return;
    if (bytes == 0)
        bytes = 4;
    for (int i = 0; i < FS_POOL_SLOTS; i++) {
        if (!ctx->pool[i].ptr) {
            ctx->pool[i].ptr = p;
            ctx->pool[i].bytes = bytes;
            return;
        }
    }
    hipStreamSynchronize(ctx->stream);
    hipFree(p);
}

struct futhark_context_config *futhark_context_config_new(void)
{
    struct futhark_context_config *cfg = (struct futhark_context_config *)calloc(1, sizeof *cfg);
    if (cfg)
        cfg->device = -1;
    return cfg;
}

void futhark_context_config_free(struct futhark_context_config *cfg)
{
    free(cfg);
}

void futhark_context_config_set_debugging(struct futhark_context_config *cfg, int flag)
{
    if (cfg)
        cfg->debugging = flag;
}

void futhark_context_config_set_profiling(struct futhark_context_config *cfg, int flag)
{
    if (cfg)
        cfg->profiling = flag;
}

void futhark_context_config_set_logging(struct futhark_context_config *cfg, int flag)
{
    if (cfg)
        cfg->logging = flag;
}

void futhark_context_config_set_device(struct futhark_context_config *cfg, const char *s)
{
    if (!cfg || !s)
        return;
    if (s[0] == '#') {
        cfg->device = atoi(s + 1);
        cfg->device_name[0] = 0;
    } else {
        cfg->device = -1;
        strncpy(cfg->device_name, s, sizeof cfg->device_name - 1);
    }
}

void futhark_context_config_select_device_interactively(struct futhark_context_config *cfg)
{
    (void)cfg;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        n = 0;
    for (int i = 0; i < n; i++) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, i) == hipSuccess)
            fprintf(stderr, "[%d]: %s (%s)\n", i, p.name, p.gcnArchName);
    }
}

static int pick_device(const struct futhark_context_config *cfg)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return -1;
    if (cfg && cfg->device >= 0)
        return cfg->device < n ? cfg->device : -1;
    if (cfg && cfg->device_name[0]) {
        for (int i = 0; i < n; i++) {
            hipDeviceProp_t p;
            if (hipGetDeviceProperties(&p, i) == hipSuccess &&
                (strstr(p.name, cfg->device_name) || strstr(p.gcnArchName, cfg->device_name)))
                return i;
        }
        return -1;
    }
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess)
        cur = 0;
    return cur;
}

static struct fs_terrain *terrain_zero(struct futhark_context *ctx);
static struct fs_shade *shade_zero(struct futhark_context *ctx);
static void terrain_unref(struct futhark_context *ctx, struct fs_terrain *t);
static void shade_unref(struct futhark_context *ctx, struct fs_shade *s);

struct futhark_context *futhark_context_new(struct futhark_context_config *cfg)
{
    struct futhark_context *ctx = (struct futhark_context *)calloc(1, sizeof *ctx);
    if (!ctx)
        return NULL;
    ctx->device = pick_device(cfg);
    if (ctx->device < 0) {
        fprintf(stderr, "futspace: no usable HIP device\n");
        free(ctx);
        return NULL;
    }
    if (cfg) {
        ctx->debugging = cfg->debugging;
        ctx->logging = cfg->logging;
    }
    if (hipSetDevice(ctx->device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        free(ctx);
        return NULL;
    }
    ctx->stream = ctx->own_stream;
    fsk_default_tuning(&ctx->tune);
    if (fsk_device_init() != hipSuccess) {
        fprintf(stderr, "futspace: device table upload failed\n");
        hipStreamDestroy(ctx->own_stream);
        free(ctx);
        return NULL;
    }
    ctx->empty_terrain = terrain_zero(ctx);
    ctx->empty_shade = shade_zero(ctx);
    if (!ctx->empty_terrain || !ctx->empty_shade) {
        futhark_context_free(ctx);
        return NULL;
    }
    return ctx;
}

void futhark_context_free(struct futhark_context *ctx)
{
    if (!ctx)
        return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    if (ctx->empty_terrain)
        terrain_unref(ctx, ctx->empty_terrain);
    if (ctx->empty_shade)
        shade_unref(ctx, ctx->empty_shade);
    for (int i = 0; i < FS_POOL_SLOTS; i++)
        if (ctx->pool[i].ptr)
            hipFree(ctx->pool[i].ptr);
    if (ctx->prof_ev) {
        for (int i = 0; i < 2 * FS_PROF_MAX; i++)
            if (ctx->prof_ev[i])
                hipEventDestroy(ctx->prof_ev[i]);
        free(ctx->prof_ev);
    }
    if (ctx->d_partials)
        hipFree(ctx->d_partials);
    if (ctx->d_trig)
        hipFree(ctx->d_trig);
    if (ctx->own_stream)
        hipStreamDestroy(ctx->own_stream);
    free(ctx->error);
    free(ctx);
}

int futhark_context_sync(struct futhark_context *ctx)
{
    if (!ctx)
        return 1;
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHECK(ctx, hipGetLastError());
    return 0;
}

char *futhark_context_get_error(struct futhark_context *ctx)
{
    if (!ctx)
        return NULL;
    char *e = ctx->error;
    ctx->error = NULL;
    return e;
}

int futhark_context_clear_caches(struct futhark_context *ctx)
{
    if (!ctx)
        return 1;
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < FS_POOL_SLOTS; i++)
        if (ctx->pool[i].ptr) {
            hipFree(ctx->pool[i].ptr);
            ctx->pool[i].ptr = NULL;
        }
    return 0;
}

void futhark_context_pause_profiling(struct futhark_context *ctx)
{
    (void)ctx;
}

void futhark_context_unpause_profiling(struct futhark_context *ctx)
{
    (void)ctx;
}

void futhark_debugging_report(struct futhark_context *ctx)
{
    if (ctx && ctx->debugging)
        fprintf(stderr, "futspace: device %d, %s\n", ctx->device, FS_VERSION);
}

int fsx_context_set_stream(struct futhark_context *ctx, void *stream)
{
    if (!ctx)
        return 1;
    hipStream_t next = stream ? (hipStream_t)stream : ctx->own_stream;
    if (next == ctx->stream)
        return 0;
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stream = next;
    return 0;
}

void *fsx_context_stream(struct futhark_context *ctx)
{
    return ctx ? (void *)ctx->stream : NULL;
}

int fsx_context_device(struct futhark_context *ctx)
{
    return ctx ? ctx->device : -1;
}

const char *fsx_version(void)
{
    return FS_VERSION;
}

int fsx_selftest(struct futhark_context *ctx, int which, int64_t *mismatches,
                 int64_t *first_bad)
{
    REQUIRE_ARGS(ctx, mismatches, "fsx_selftest");
    unsigned long long *d = NULL, h[2] = {0, 0};
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    if (dev_alloc(ctx, (void **)&d, sizeof h))
        return 1;
    /* count = 0, smallest failing input = all ones */
    hipError_t e = hipMemsetAsync(d, 0, sizeof h[0], ctx->stream);
    if (e == hipSuccess)
        e = hipMemsetAsync(d + 1, 0xFF, sizeof h[1], ctx->stream);
    if (e == hipSuccess)
        e = (hipError_t)fsk_selftest(which, d, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    dev_release(ctx, d, sizeof h);
    if (e != hipSuccess) {
        set_error(ctx, "fsx_selftest(%d): %s", which, hipGetErrorString(e));
        return 1;
    }
    *mismatches = (int64_t)h[0];
    if (first_bad)
        *first_bad = h[0] ? (int64_t)h[1] : -1;
    return 0;
}

int fsx_profile_enable(struct futhark_context *ctx, int enable)
{
    if (!ctx)
        return 1;
    if (enable && !ctx->prof_ev) {
        ctx->prof_ev = (hipEvent_t *)calloc(2 * FS_PROF_MAX, sizeof(hipEvent_t));
        if (!ctx->prof_ev) {
            set_error(ctx, "fsx_profile_enable: out of host memory");
            return 1;
        }
        for (int i = 0; i < 2 * FS_PROF_MAX; i++)
            HIPCHECK(ctx, hipEventCreate(&ctx->prof_ev[i]));
    }
    ctx->prof_on = enable ? 1 : 0;
    ctx->prof_n = 0;
    return 0;
}

int fsx_profile_read(struct futhark_context *ctx, double *total_ms, int64_t *launches)
{
    REQUIRE_ARGS(ctx, total_ms && launches, "fsx_profile_read");
    double sum = 0.0;
    if (ctx->prof_n > 0) {
        HIPCHECK(ctx, hipEventSynchronize(ctx->prof_ev[2 * (ctx->prof_n - 1) + 1]));
        for (int i = 0; i < ctx->prof_n; i++) {
            float ms = 0.0f;
            HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->prof_ev[2 * i], ctx->prof_ev[2 * i + 1]));
            sum += ms;
        }
    }
    *total_ms = sum;
    *launches = ctx->prof_n;
    ctx->prof_n = 0;
    return 0;
}

/* Event pair around one kernel launch on the context stream. */
static int prof_begin(struct futhark_context *ctx)
{
    if (!ctx->prof_on || ctx->prof_n >= FS_PROF_MAX)
        return -1;
    int i = ctx->prof_n++;
    if (hipEventRecord(ctx->prof_ev[2 * i], ctx->stream) != hipSuccess)
        return -1;
    return i;
}

static void prof_end(struct futhark_context *ctx, int i)
{
    if (i >= 0)
        hipEventRecord(ctx->prof_ev[2 * i + 1], ctx->stream);
}

/* ------------------------------------------------------------------ */
/* Arrays                                                              */
/* ------------------------------------------------------------------ */

struct futhark_u32_2d {
    int64_t shape[2];
    uint32_t *d;
};

struct futhark_i32_2d {
    int64_t shape[2];
    int32_t *d;
};

static int dims_ok(struct futhark_context *ctx, int64_t d0, int64_t d1)
{
    if (d0 < 0 || d1 < 0 || (d0 > 0 && d1 > INT32_MAX / d0)) {
        set_error(ctx, "invalid array shape [%lld][%lld]", (long long)d0, (long long)d1);
        return 0;
    }
    return 1;
}

struct futhark_u32_2d *futhark_new_u32_2d(struct futhark_context *ctx, const uint32_t *data,
                                          int64_t dim0, int64_t dim1)
{
    if (!ctx || !dims_ok(ctx, dim0, dim1))
        return NULL;
    struct futhark_u32_2d *a = (struct futhark_u32_2d *)calloc(1, sizeof *a);
    if (!a)
        return NULL;
    a->shape[0] = dim0;
    a->shape[1] = dim1;
    size_t bytes = (size_t)(dim0 * dim1) * sizeof(uint32_t);
    if (hipSetDevice(ctx->device) != hipSuccess || dev_alloc(ctx, (void **)&a->d, bytes)) {
        free(a);
        return NULL;
    }
    if (bytes && data) {
        hipError_t e = hipMemcpyAsync(a->d, data, bytes, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            set_error(ctx, "futhark_new_u32_2d: %s", hipGetErrorString(e));
            hipFree(a->d);
            free(a);
            return NULL;
        }
    }
    return a;
}

int futhark_free_u32_2d(struct futhark_context *ctx, struct futhark_u32_2d *arr)
{
    if (!ctx)
        return 1;
    if (!arr)
        return 0;
    dev_release(ctx, arr->d, (size_t)(arr->shape[0] * arr->shape[1]) * sizeof(uint32_t));
    free(arr);
    return 0;
}

int futhark_values_u32_2d(struct futhark_context *ctx, struct futhark_u32_2d *arr, uint32_t *data)
{
    REQUIRE_ARGS(ctx, arr, "futhark_values_u32_2d");
    size_t bytes = (size_t)(arr->shape[0] * arr->shape[1]) * sizeof(uint32_t);
    if (bytes) {
        HIPCHECK(ctx, hipMemcpyAsync(data, arr->d, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return 0;
}

const int64_t *futhark_shape_u32_2d(struct futhark_context *ctx, struct futhark_u32_2d *arr)
{
    (void)ctx;
    return arr ? arr->shape : NULL;
}

struct futhark_i32_2d *futhark_new_i32_2d(struct futhark_context *ctx, const int32_t *data,
                                          int64_t dim0, int64_t dim1)
{
    if (!ctx || !dims_ok(ctx, dim0, dim1))
        return NULL;
    struct futhark_i32_2d *a = (struct futhark_i32_2d *)calloc(1, sizeof *a);
    if (!a)
        return NULL;
    a->shape[0] = dim0;
    a->shape[1] = dim1;
    size_t bytes = (size_t)(dim0 * dim1) * sizeof(int32_t);
    if (hipSetDevice(ctx->device) != hipSuccess || dev_alloc(ctx, (void **)&a->d, bytes)) {
        free(a);
        return NULL;
    }
    if (bytes && data) {
        hipError_t e = hipMemcpyAsync(a->d, data, bytes, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            set_error(ctx, "futhark_new_i32_2d: %s", hipGetErrorString(e));
            hipFree(a->d);
            free(a);
            return NULL;
        }
    }
    return a;
}

int futhark_free_i32_2d(struct futhark_context *ctx, struct futhark_i32_2d *arr)
{
    if (!ctx)
        return 1;
    if (!arr)
        return 0;
    dev_release(ctx, arr->d, (size_t)(arr->shape[0] * arr->shape[1]) * sizeof(int32_t));
    free(arr);
    return 0;
}

int futhark_values_i32_2d(struct futhark_context *ctx, struct futhark_i32_2d *arr, int32_t *data)
{
    REQUIRE_ARGS(ctx, arr, "futhark_values_i32_2d");
    size_t bytes = (size_t)(arr->shape[0] * arr->shape[1]) * sizeof(int32_t);
    if (bytes) {
        HIPCHECK(ctx, hipMemcpyAsync(data, arr->d, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return 0;
}

const int64_t *futhark_shape_i32_2d(struct futhark_context *ctx, struct futhark_i32_2d *arr)
{
    (void)ctx;
    return arr ? arr->shape : NULL;
}

/* ------------------------------------------------------------------ */
/* Resident maps: terrain (heights + raw colours) and shade (shadowed   */
/* colours), both immutable and reference-counted so states share them */
/* ------------------------------------------------------------------ */

struct fs_terrain {
    int refs;
    int32_t rows, cols, hbits;
    void *d_hq;            /* bilinear height quads */
    uint16_t *d_hp;        /* masked heights, one per texel */
    uint32_t *d_color;     /* raw colour map */
    int32_t *h_alt;        /* host copy of the masked heights */
};

struct fs_shade {
    int refs;
    int32_t rows, cols;
    float sun[3];          /* the sun vector it was shaded with */
    uint32_t *d_shaded;
    void *d_cq;            /* bilinear colour quads (uint4 per texel) */
};

struct fsx_map {
    struct fs_terrain *t;
    struct fs_shade *s;
    int32_t shade_size;
};

static void terrain_unref(struct futhark_context *ctx, struct fs_terrain *t)
{
    if (!t || --t->refs > 0)
        return;
    hipStreamSynchronize(ctx->stream);
    hipFree(t->d_hq);
    hipFree(t->d_hp);
    hipFree(t->d_color);
    free(t->h_alt);
    free(t);
}

static void shade_unref(struct futhark_context *ctx, struct fs_shade *s)
{
    if (!s || --s->refs > 0)
        return;
    hipStreamSynchronize(ctx->stream);
    hipFree(s->d_shaded);
    hipFree(s->d_cq);
    free(s);
}

/* Build a terrain from device colours and device i32 heights (both
 * [rows][cols]); heights are masked to hbits. */
static struct fs_terrain *terrain_new(struct futhark_context *ctx, const uint32_t *d_color_src,
                                      const int32_t *d_alt, int32_t rows, int32_t cols,
                                      int32_t hbits)
{
    if ((int64_t)rows * (int64_t)cols > ((int64_t)1 << 28)) {
        set_error(ctx, "map of %d x %d texels exceeds the supported 2^28", rows, cols);
        return NULL;
    }
    struct fs_terrain *t = (struct fs_terrain *)calloc(1, sizeof *t);
    if (!t)
        return NULL;
    t->refs = 1;
    t->rows = rows;
    t->cols = cols;
    t->hbits = hbits;
    size_t n = (size_t)rows * (size_t)cols;
    size_t qbytes = n * (hbits == 8 ? 4 : 8);
    hipError_t e = hipMalloc(&t->d_hq, qbytes ? qbytes : 4);
    if (e == hipSuccess) e = hipMalloc((void **)&t->d_hp, n ? n * 2 : 4);
    if (e == hipSuccess) e = hipMalloc((void **)&t->d_color, n ? n * 4 : 4);
    t->h_alt = (int32_t *)malloc(n ? n * 4 : 4);
    if (e != hipSuccess || !t->h_alt) {
        set_error(ctx, "terrain allocation (%d x %d) failed: %s", rows, cols, hipGetErrorString(e));
        hipFree(t->d_hq); hipFree(t->d_hp); hipFree(t->d_color); free(t->h_alt); free(t);
        return NULL;
    }
    if (n) {
        e = hipMemcpyAsync(t->d_color, d_color_src, n * 4, hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess)
            e = (hipError_t)fsk_height_quads(d_alt, rows, cols, hbits, t->d_hq, t->d_hp, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(t->h_alt, d_alt, n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            set_error(ctx, "terrain upload failed: %s", hipGetErrorString(e));
            t->refs = 1;
            terrain_unref(ctx, t);
            return NULL;
        }
        uint32_t mk = hbits == 8 ? 0xFFu : 0xFFFFu;
        for (size_t i = 0; i < n; i++)
            t->h_alt[i] = (int32_t)((uint32_t)t->h_alt[i] & mk);
    }
    return t;
}

/* A shade of side `size` for terrain t under the sun of (sun_height,
 * sun_ang): generate_shadowmap_accumulated (fut/effects.fut:108-125) with
 * sun = R_y(sun_ang) R_z(sun_height) [0,1,0] (fut/interactive.fut:196),
 * whose closed form is (cos a * -sin h, cos h, -sin a * -sin h). */
/* vec3_rotate #y sun_ang (vec3_rotate #z sun_height [0, 1, 0])
 * (fut/effects.fut:6-25): matvecmul rows folded left from 0; the zero
 * terms do not change the non-zero sums, so the rotations reduce to
 * (cos a * (-sin h), cos h, (-sin a) * (-sin h)) bit for bit. */
static void sun_vector(float sun_height, float sun_ang, float out[3])
{
    float sh = sinf(sun_height), chh = cosf(sun_height);
    float sa = sinf(sun_ang), ca = cosf(sun_ang);
    float v0 = -sh, v1 = chh;
    out[0] = ca * v0;
    out[1] = v1;
    out[2] = (-sa) * v0;
}

static struct fs_shade *shade_new(struct futhark_context *ctx, const struct fs_terrain *t,
                                  int32_t size, float sun_height, float sun_ang)
{
    struct fs_shade *s = (struct fs_shade *)calloc(1, sizeof *s);
    if (!s)
        return NULL;
    s->refs = 1;
    s->rows = size;
    s->cols = size;
    size_t n = (size_t)size * (size_t)size;
    hipError_t e = hipMalloc((void **)&s->d_shaded, n * 4);
    if (e == hipSuccess)
        e = hipMalloc(&s->d_cq, n * 16);
    if (e != hipSuccess) {
        set_error(ctx, "shade allocation (%d^2) failed: %s", size, hipGetErrorString(e));
        hipFree(s->d_shaded); hipFree(s->d_cq); free(s);
        return NULL;
    }
    sun_vector(sun_height, sun_ang, s->sun);
    e = (hipError_t)fsk_shade(t->d_hp, t->rows, t->cols, t->d_color, t->rows, t->cols,
                              s->sun[0], s->sun[1], s->sun[2], size, s->d_shaded, ctx->stream);
    if (e == hipSuccess)
        e = (hipError_t)fsk_color_quads(s->d_shaded, size, size, s->d_cq, ctx->stream);
    if (e != hipSuccess) {
        set_error(ctx, "shading failed: %s", hipGetErrorString(e));
        shade_unref(ctx, s);
        return NULL;
    }
    return s;
}

/* The dummy landscape of `init`: altitude, color and shadowed_color all
 * [[0],[0]] (fut/interactive.fut:38-43). */
static struct fs_terrain *terrain_zero(struct futhark_context *ctx)
{
    int32_t *d_alt = NULL;
    uint32_t *d_col = NULL;
    if (hipMalloc((void **)&d_alt, 8) != hipSuccess || hipMalloc((void **)&d_col, 8) != hipSuccess ||
        hipMemsetAsync(d_alt, 0, 8, ctx->stream) != hipSuccess ||
        hipMemsetAsync(d_col, 0, 8, ctx->stream) != hipSuccess) {
        hipFree(d_alt); hipFree(d_col);
        return NULL;
    }
    struct fs_terrain *t = terrain_new(ctx, d_col, d_alt, 2, 1, 8);
    hipStreamSynchronize(ctx->stream);
    hipFree(d_alt);
    hipFree(d_col);
    return t;
}

static struct fs_shade *shade_zero(struct futhark_context *ctx)
{
    struct fs_shade *s = (struct fs_shade *)calloc(1, sizeof *s);
    if (!s)
        return NULL;
    s->refs = 1;
    s->rows = 2;
    s->cols = 1;
    if (hipMalloc((void **)&s->d_shaded, 8) != hipSuccess || hipMalloc(&s->d_cq, 32) != hipSuccess ||
        hipMemsetAsync(s->d_shaded, 0, 8, ctx->stream) != hipSuccess ||
        hipMemsetAsync(s->d_cq, 0, 32, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        hipFree(s->d_shaded); hipFree(s->d_cq); free(s);
        return NULL;
    }
    return s;
}

/* ------------------------------------------------------------------ */
/* Frame rendering                                                     */
/* ------------------------------------------------------------------ */

/* get_zs' slice count, fut/voxel_renderer.fut:29-32, x86 `futhark c`
 * semantics for the i32 conversion. */
static int32_t zs_count(float c, float d, float z0)
{
    float sq = sqrtf(powf(d - 2.0f * z0, 2.0f) + (8.0f * d) * c);
    float num = (sq - 2.0f * z0) + d;
    float div = 2.0f * d;
    float q = floorf(num / div);
    if (!(q >= -2147483648.0f && q < 2147483648.0f))
        return 0;
    int32_t n = (int32_t)q;
    return n < 0 ? 0 : n;
}

/* Depth slices one render may march: 4096 times the largest workload
 * measured here; a longer march would keep each tile busy for seconds. */
#define FS_MAX_SLICES (1 << 24)

static int slices_ok(struct futhark_context *ctx, const char *what, int32_t n)
{
    if (n > FS_MAX_SLICES) {
        set_error(ctx, "%s: %d depth slices, above the supported %d", what, n, FS_MAX_SLICES);
        return 0;
    }
    return 1;
}

static void frame_of_view(const struct fsx_view *v, fsk_frame *f)
{
    float sin_ang = sinf(v->angle), cos_ang = cosf(v->angle), view = v->fov;
    f->cam_x = v->x;
    f->cam_y = v->y;
    f->cam_h = v->height;
    f->horizon = v->horizon;
    f->Lx = (-cos_ang) - (sin_ang * view);
    f->Ly = sin_ang - (cos_ang * view);
    f->Rx = cos_ang - (sin_ang * view);
    f->Ry = (-sin_ang) - (cos_ang * view);
    f->d = v->d;
    f->z0 = v->z0;
    f->scale = v->scale > 0.0f ? v->scale : (float)(v->w / 2);
    f->fw = (float)v->w;
    f->rfw = 1.0f / f->fw;
    f->n = v->n_slices > 0 ? v->n_slices : zs_count(v->distance, v->d, v->z0);
    f->rows = v->h;
    f->cols = v->w;
    f->sky = v->sky_color;
    f->tame = 0;
}

static int render_into(struct futhark_context *ctx, const struct fs_terrain *t,
                       const struct fs_shade *s, const struct fsx_view *v,
                       uint32_t *d_out, int64_t pitch)
{
    if (v->h < 0 || v->w < 0) {
        set_error(ctx, "render: negative frame size %d x %d", v->h, v->w);
        return 1;
    }
    if (v->h > fsk_render_max_rows()) {
        set_error(ctx, "render: %d rows exceed the supported %d", v->h, fsk_render_max_rows());
        return 1;
    }
    if (v->h == 0 || v->w == 0)
        return 0;
    fsk_frame f;
    frame_of_view(v, &f);
    if (!slices_ok(ctx, "render", f.n))
        return 1;
    fsk_map m;
    m.hq = t->d_hq;
    m.hrows = t->rows;
    m.hcols = t->cols;
    m.hbits = t->hbits;
    m.cq = s->d_cq;
    m.crows = s->rows;
    m.ccols = s->cols;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int pe = prof_begin(ctx);
    HIPCHECK(ctx, fsk_render(&m, &f, &ctx->tune, d_out, pitch > 0 ? pitch : v->w, ctx->stream));
    prof_end(ctx, pe);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Opaque state: sized_state of fut/interactive.fut:9-24               */
/* ------------------------------------------------------------------ */

enum { MODE_MATH = 0, MODE_PNG = 1 };
enum { SMOOTH_ON = 0, SMOOTH_OFF = 1 };

/* input_state slots, fut/interactive_input.fut:6-48 */
enum {
    K_A = 0, K_Z = 25, K_UP = 26, K_LEFT = 27, K_DOWN = 28, K_RIGHT = 29,
    K_1 = 30, K_9 = 38, K_0 = 39, K_COUNT = 40
};

struct futhark_opaque_state {
    int32_t mode, smoothing;
    float cam_x, cam_y, cam_height, cam_angle, cam_horizon, cam_distance, cam_fov;
    uint32_t sky_color;
    struct fs_terrain *terrain;
    struct fs_shade *shade;
    int32_t height, width;
    int8_t inputs[K_COUNT];
    float random, sun_height, sun_ang;
    float sun[3];
};

static struct futhark_opaque_state *state_copy(const struct futhark_opaque_state *s)
{
    struct futhark_opaque_state *o = (struct futhark_opaque_state *)malloc(sizeof *o);
    if (!o)
        return NULL;
    *o = *s;
    o->terrain->refs++;
    o->shade->refs++;
    return o;
}

int futhark_free_opaque_state(struct futhark_context *ctx, struct futhark_opaque_state *obj)
{
    if (!ctx)
        return 1;
    if (!obj)
        return 0;
    terrain_unref(ctx, obj->terrain);
    shade_unref(ctx, obj->shade);
    free(obj);
    return 0;
}

/* init, fut/interactive.fut:28-57 */
int futhark_entry_init(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                       const uint32_t in0)
{
    (void)in0;   /* the seed is unused by the reference */
    REQUIRE_ARGS(ctx, out0, "futhark_entry_init");
    struct futhark_opaque_state *s = (struct futhark_opaque_state *)calloc(1, sizeof *s);
    if (!s) {
        set_error(ctx, "init: out of host memory");
        return 1;
    }
    s->mode = MODE_PNG;
    s->smoothing = SMOOTH_OFF;
    s->cam_x = 0.98f;
    s->cam_y = 0.6f;
    s->cam_height = 58.0f;
    s->cam_angle = 2.2f;
    s->cam_horizon = 200.0f;
    s->cam_distance = 800.0f;
    s->cam_fov = 1.2f;
    s->sky_color = 0;
    s->terrain = ctx->empty_terrain;
    s->terrain->refs++;
    s->shade = ctx->empty_shade;
    s->shade->refs++;
    s->height = 1024;
    s->width = 1024;
    s->random = 1.0f;
    s->sun_height = 0.1f;
    s->sun_ang = 0.1f;
    s->sun[0] = 0.0f;
    s->sun[1] = 1.0f / sqrtf(1.0f);
    s->sun[2] = 0.0f;
    *out0 = s;
    return 0;
}

/* resize, fut/interactive.fut:59-61 */
int futhark_entry_resize(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                         const int32_t in0, const int32_t in1,
                         const struct futhark_opaque_state *in2)
{
    REQUIRE_ARGS(ctx, out0 && in2, "futhark_entry_resize");
    struct futhark_opaque_state *o = state_copy(in2);
    if (!o)
        return 1;
    o->height = in0;
    o->width = in1;
    *out0 = o;
    return 0;
}

/* SDL2 keycodes the reference binds (fut/interactive_input.fut:50-134). */
static int key_slot(int32_t key)
{
    if (key >= 'a' && key <= 'z')
        return K_A + (key - 'a');
    if (key >= '1' && key <= '9')
        return K_1 + (key - '1');
    if (key == '0')
        return K_0;
    switch (key) {
    case 1073741906: return K_UP;      /* SDLK_UP */
    case 1073741904: return K_LEFT;    /* SDLK_LEFT */
    case 1073741905: return K_DOWN;    /* SDLK_DOWN */
    case 1073741903: return K_RIGHT;   /* SDLK_RIGHT */
    default: return -1;
    }
}

/* key, fut/interactive_entrypoints.fut:9-11: e == 0 is keydown. */
int futhark_entry_key(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                      const int32_t in0, const int32_t in1,
                      const struct futhark_opaque_state *in2)
{
    REQUIRE_ARGS(ctx, out0 && in2, "futhark_entry_key");
    struct futhark_opaque_state *o = state_copy(in2);
    if (!o)
        return 1;
    int slot = key_slot(in1);
    if (slot >= 0)
        o->inputs[slot] = in0 == 0 ? 1 : 0;
    *out0 = o;
    return 0;
}

/* mouse and wheel events fall through `event` unchanged
 * (fut/interactive.fut:166-171). */
int futhark_entry_mouse(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                        const int32_t in0, const int32_t in1, const int32_t in2,
                        const struct futhark_opaque_state *in3)
{
    (void)in0; (void)in1; (void)in2;
    REQUIRE_ARGS(ctx, out0 && in3, "futhark_entry_mouse");
    *out0 = state_copy(in3);
    return *out0 ? 0 : 1;
}

int futhark_entry_wheel(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                        const int32_t in0, const int32_t in1,
                        const struct futhark_opaque_state *in2)
{
    (void)in0; (void)in1;
    REQUIRE_ARGS(ctx, out0 && in2, "futhark_entry_wheel");
    *out0 = state_copy(in2);
    return *out0 ? 0 : 1;
}

/* squares, fut/render_functions.fut:4-11 (terrain of #math mode); f32
 * `%` is C fmodf. */
static float squares(float x, float y)
{
    x = fmodf(x, 1024.0f);
    y = fmodf(y, 1024.0f);
    return (x > 256.0f && fabsf(x) < 768.0f && fabsf(y) > 256.0f && fabsf(y) < 768.0f) ? 200.0f : 0.0f;
}

static int32_t mod_i32(int32_t a, int32_t m)
{
    int32_t r = a % m;
    return (r != 0 && ((r < 0) != (m < 0))) ? r + m : r;
}

static int32_t i32_f32(float v)
{
    if (!(v >= -2147483648.0f && v < 2147483648.0f))
        return INT32_MIN;
    return (int32_t)v;
}

/* argb.scale 0xFF9090e0 sun_height (fut/interactive.fut:163) with the
 * matte argb restatement: channels c/255 scaled, alpha kept. */
static uint32_t u8_of(float v)
{
    float c = (v < 0.0f) ? 0.0f : ((v > 1.0f) ? 1.0f : v);
    float s = c * 255.0f;
    if (!(s >= -9223372036854775808.0f && s < 9223372036854775808.0f))
        return 0u;
    return (uint32_t)(uint64_t)(int64_t)s;
}

/* argb.mix with the Gloss rule: normalised weights, channels as the
 * root of the weighted squares, alpha mixed linearly. */
static uint32_t argb_mix(float m1, uint32_t c1, float m2, uint32_t c2)
{
    float p[4], q[4];
    for (int k = 0; k < 4; k++) {
        int sh = k == 3 ? 24 : 16 - 8 * k;
        p[k] = (float)((c1 >> sh) & 255u) / 255.0f;
        q[k] = (float)((c2 >> sh) & 255u) / 255.0f;
    }
    float m12 = m1 + m2;
    float m1n = m1 / m12, m2n = m2 / m12;
    uint32_t out = u8_of(((m1 * p[3]) + (m2 * q[3])) / m12) << 24;
    for (int k = 0; k < 3; k++)
        out |= u8_of(sqrtf((m1n * (p[k] * p[k])) + (m2n * (q[k] * q[k])))) << (16 - 8 * k);
    return out;
}

/* expo2, fut/render_functions.fut:23-28 (heights of #math mode). */
static float expo2(float x, float y)
{
    float output = (powf(x, 2.0f) - powf(y, 2.0f)) / 1024.0f;
    return (x > -1024.0f && x < 1024.0f && y > -1024.0f && y < 1024.0f) ? output : 0.0f;
}

/* function_coloring, fut/render_functions.fut:37-41 (colours of #math mode). */
static uint32_t function_coloring(float x, float y)
{
    return (mod_i32(i32_f32(x), 64) == 0 || mod_i32(i32_f32(y), 64) == 0) ? 0xFF000000u
                                                                           : 0xFFFFFFFFu;
}

/* The texel png_color reads at (x, y) from the #math shadow map
 * generate_shadowmap_accumulated function_coloring expo2 sun
 * (fut/interactive.fut:176, fut/effects.fut:108-125): only the one texel
 * is evaluated, with the map's own loop. */
static uint32_t math_texel(float x, float y, const float sun[3])
{
    int32_t tx = mod_i32(i32_f32(x), 1024), ty = mod_i32(i32_f32(y), 1024);
    const float step_size = 4.0f;
    float fx = (float)tx, fy = (float)ty;
    float height = expo2(fx, fy);
    int32_t count = 0;
    for (int32_t dist = 1; dist < 256; dist++) {
        float t = (float)dist * step_size;
        float probe = expo2(fx + (t * sun[0]), fy + (t * sun[2]));
        if (((height + (t * sun[1])) - probe) < -0.5f)
            count += 1;
    }
    return argb_mix(step_size * (float)count, 0xFF000000u, 1.0f, function_coloring(fx, fy));
}

static uint32_t argb_scale(uint32_t col, float s)
{
    float r = (float)((col >> 16) & 255u) / 255.0f;
    float g = (float)((col >> 8) & 255u) / 255.0f;
    float b = (float)(col & 255u) / 255.0f;
    float a = (float)((col >> 24) & 255u) / 255.0f;
    return (u8_of(a) << 24) | (u8_of(r * s) << 16) | (u8_of(g * s) << 8) | u8_of(b * s);
}

/* step, fut/interactive.fut:161-164: sky and random first, then
 * process_inputs (:89-159, every field computed from the incoming state),
 * then terrain_collision (:67-87). */
int futhark_entry_step(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                       const float in0, const struct futhark_opaque_state *in1)
{
    (void)in0;   /* the frame time is unused by the reference */
    REQUIRE_ARGS(ctx, out0 && in1, "futhark_entry_step");
    struct futhark_opaque_state s = *in1;
    s.random = s.random + 0.005f;
    s.sky_color = argb_scale(0xFF9090E0u, s.sun_height);
    const int8_t *k = s.inputs;
    float sin_a = sinf(s.cam_angle), cos_a = cosf(s.cam_angle);

    struct futhark_opaque_state o = s;
    o.cam_x = k['w' - 'a'] == 1 ? s.cam_x - 3.0f * sin_a
            : k['s' - 'a'] == 1 ? s.cam_x + 3.0f * sin_a : s.cam_x;
    o.cam_y = k['w' - 'a'] == 1 ? s.cam_y - 3.0f * cos_a
            : k['s' - 'a'] == 1 ? s.cam_y + 3.0f * cos_a : s.cam_y;
    o.cam_angle = k['d' - 'a'] == 1 ? s.cam_angle - 0.10f
                : k['a' - 'a'] == 1 ? s.cam_angle + 0.10f : s.cam_angle;
    o.cam_horizon = k['e' - 'a'] == 1 ? s.cam_horizon - 20.0f
                  : k['q' - 'a'] == 1 ? s.cam_horizon + 20.0f : s.cam_horizon;
    o.cam_height = k['r' - 'a'] == 1 ? s.cam_height + 10.0f
                 : k['f' - 'a'] == 1 ? s.cam_height - 10.0f : s.cam_height;
    o.cam_distance = k[K_UP] == 1 ? s.cam_distance + 30.0f
                   : k[K_DOWN] == 1 ? s.cam_distance - 30.0f : s.cam_distance;
    o.cam_fov = k['o' - 'a'] == 1 ? s.cam_fov + 0.1f
              : k['l' - 'a'] == 1 ? s.cam_fov - 0.1f : s.cam_fov;
    o.sun_height = k['u' - 'a'] == 1 ? s.sun_height + 0.005f
                 : k['j' - 'a'] == 1 ? s.sun_height - 0.005f : s.sun_height;
    o.sun_ang = k['n' - 'a'] == 1 ? s.sun_ang + 0.005f
              : k['m' - 'a'] == 1 ? s.sun_ang - 0.005f : s.sun_ang;
    o.mode = k[K_1] == 1 ? (s.mode == MODE_PNG ? MODE_MATH : MODE_PNG) : s.mode;
    o.smoothing = k[K_1 + 1] == 1 ? (s.smoothing == SMOOTH_OFF ? SMOOTH_ON : SMOOTH_OFF)
                                  : s.smoothing;
    struct fs_shade *sh = s.shade;
    int reshade = k['u' - 'a'] == 1 || k['j' - 'a'] == 1 || k['n' - 'a'] == 1 || k['m' - 'a'] == 1;
    if (reshade) {
        /* the refresh uses the incoming sun_height/sun_ang */
        sh = shade_new(ctx, s.terrain, 1024, s.sun_height, s.sun_ang);
        if (!sh)
            return 1;
    } else {
        sh->refs++;
    }
    o.shade = sh;
    o.terrain->refs++;

    /* terrain_collision */
    if (o.mode == MODE_MATH) {
        float th = squares(o.cam_x, o.cam_y);
        o.cam_height = o.cam_height <= th ? th : o.cam_height;
    } else {
        const struct fs_terrain *t = o.terrain;
        int32_t xi = i32_f32(o.cam_x), yi = i32_f32(o.cam_y);
        float th = (float)t->h_alt[(size_t)mod_i32(yi, t->rows) * t->cols + mod_i32(xi, t->cols)];
        o.cam_height = o.cam_height <= th ? th : o.cam_height;
    }
    struct futhark_opaque_state *res = (struct futhark_opaque_state *)malloc(sizeof *res);
    if (!res) {
        terrain_unref(ctx, o.terrain);
        shade_unref(ctx, o.shade);
        set_error(ctx, "step: out of host memory");
        return 1;
    }
    *res = o;
    *out0 = res;
    return 0;
}

/* text_content, fut/interactive.fut:185-186 */
int futhark_entry_text_content(struct futhark_context *ctx, float *out0, float *out1,
                               float *out2, float *out3, float *out4, float *out5,
                               float *out6, float *out7, float *out8,
                               const struct futhark_opaque_state *in0)
{
    REQUIRE_ARGS(ctx, in0, "futhark_entry_text_content");
    *out0 = in0->cam_x;
    *out1 = in0->cam_y;
    *out2 = in0->cam_angle;
    *out3 = in0->cam_height;
    *out4 = in0->cam_horizon;
    *out5 = in0->cam_distance;
    *out6 = in0->sun_height;
    *out7 = in0->sun_ang;
    *out8 = in0->cam_fov;
    return 0;
}

/* update_map, fut/interactive.fut:188-198 */
int futhark_entry_update_map(struct futhark_context *ctx, struct futhark_opaque_state **out0,
                             const struct futhark_u32_2d *in0,
                             const struct futhark_i32_2d *in1,
                             const struct futhark_opaque_state *in2)
{
    REQUIRE_ARGS(ctx, out0 && in0 && in1 && in2, "futhark_entry_update_map");
    if (in0->shape[0] != in1->shape[0] || in0->shape[1] != in1->shape[1]) {
        set_error(ctx, "update_map: colour map [%lld][%lld] and height map [%lld][%lld] differ",
                  (long long)in0->shape[0], (long long)in0->shape[1],
                  (long long)in1->shape[0], (long long)in1->shape[1]);
        return 1;
    }
    if (in0->shape[0] == 0 || in0->shape[1] == 0) {
        set_error(ctx, "update_map: empty maps cannot be indexed");
        return 1;
    }
    if (hipSetDevice(ctx->device) != hipSuccess)
        return 1;
    struct fs_terrain *t = terrain_new(ctx, in0->d, in1->d, (int32_t)in0->shape[0],
                                       (int32_t)in0->shape[1], 8);
    if (!t)
        return 1;
    struct fs_shade *s = shade_new(ctx, t, 1024, in2->sun_height, in2->sun_ang);
    if (!s) {
        terrain_unref(ctx, t);
        return 1;
    }
    struct futhark_opaque_state *o = (struct futhark_opaque_state *)malloc(sizeof *o);
    if (!o) {
        terrain_unref(ctx, t);
        shade_unref(ctx, s);
        return 1;
    }
    *o = *in2;
    o->terrain = t;
    o->shade = s;
    *out0 = o;
    return 0;
}

/* render, fut/interactive.fut:173-183 -> voxel_renderer.render with
 * z_0 = 0, d = 0.001 and scale f32(w/2) (fut/voxel_renderer.fut:101-251).
 *
 * render fixes z_0 = 0 for every mode, so its first depth slice lies at
 * z = 0: inv_z = (1/0) * f32(w/2) is +inf (NaN for w = 1), every column's
 * first sample sits at (cam.x, cam.y) and projects, through the x86 cast of
 * +-inf or NaN, to row 0; no later slice can rise above row 0, so each
 * column shows that one sample on every row -- or the sky if its colour is
 * 0.  The #on smoothing branch (:174-213) reaches the same frame: every
 * slice's running minimum is slice 0's, so the entry that survives the
 * scatter pairs slice 0 with itself and is not blended.  #png frames are
 * rendered by the HIP kernel like any other view (which also covers the
 * general case); a #math frame (:175-178) is the colour of the #math
 * shadow map under the camera, filled on the device. */
int futhark_entry_render(struct futhark_context *ctx, struct futhark_u32_2d **out0,
                         const struct futhark_opaque_state *in0)
{
    REQUIRE_ARGS(ctx, out0 && in0, "futhark_entry_render");
    if (in0->height < 0 || in0->width < 0) {
        set_error(ctx, "render: negative frame size");
        return 1;
    }
    struct fsx_view v;
    memset(&v, 0, sizeof v);
    v.x = in0->cam_x;
    v.y = in0->cam_y;
    v.height = in0->cam_height;
    v.angle = in0->cam_angle;
    v.horizon = in0->cam_horizon;
    v.distance = in0->cam_distance;
    v.fov = in0->cam_fov;
    v.sky_color = in0->sky_color;
    v.h = in0->height;
    v.w = in0->width;
    v.d = 0.001f;
    v.z0 = 0.0f;
    v.scale = 0.0f;
    v.n_slices = 0;
    struct futhark_u32_2d *a = (struct futhark_u32_2d *)calloc(1, sizeof *a);
    if (!a)
        return 1;
    a->shape[0] = v.h;
    a->shape[1] = v.w;
    if (hipSetDevice(ctx->device) != hipSuccess ||
        dev_alloc(ctx, (void **)&a->d, (size_t)v.h * (size_t)v.w * 4)) {
        free(a);
        return 1;
    }
    if (in0->mode == MODE_MATH) {
        size_t npx = (size_t)v.h * (size_t)v.w;
        uint32_t c = 0;
        if (npx > 0 && zs_count(v.distance, v.d, v.z0) > 0) {
            /* the first slice's sample of column 0, as get_h_line and
             * get_segment form it at z = 0 (every column shares it) */
            fsk_frame f;
            frame_of_view(&v, &f);
            const float z = 0.0f;
            float lx = f.Lx * z, ly = f.Ly * z, rx = f.Rx * z, ry = f.Ry * z;
            float dx = (rx - lx) / f.fw, dy = (ry - ly) / f.fw;
            float x0 = (lx + f.cam_x) + (0.0f * dx);
            float y0 = (ly + f.cam_y) + (0.0f * dy);
            float sun[3];
            sun_vector(in0->sun_height, in0->sun_ang, sun);
            c = math_texel(x0, y0, sun);
        }
        hipError_t e = (hipError_t)fsk_fill_u32(a->d, npx, c ? c : v.sky_color, ctx->stream);
        if (e != hipSuccess) {
            set_error(ctx, "render: %s", hipGetErrorString(e));
            futhark_free_u32_2d(ctx, a);
            return 1;
        }
    } else if (render_into(ctx, in0->terrain, in0->shade, &v, a->d, v.w)) {
        futhark_free_u32_2d(ctx, a);
        return 1;
    }
    *out0 = a;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Extension entries                                                   */
/* ------------------------------------------------------------------ */

uint32_t fsx_sky(float sun_height)
{
    return argb_scale(0xFF9090E0u, sun_height);
}

int fsx_map_new(struct futhark_context *ctx, struct fsx_map **out,
                const uint32_t *color, const int32_t *height, int64_t rows, int64_t cols,
                int32_t height_bits, int32_t shade_size, float sun_height, float sun_ang)
{
    REQUIRE_ARGS(ctx, out && color && height, "fsx_map_new");
    if (rows <= 0 || cols <= 0 || rows > 65536 || cols > 65536) {
        set_error(ctx, "fsx_map_new: invalid map size %lld x %lld", (long long)rows, (long long)cols);
        return 1;
    }
    if (height_bits != 8 && height_bits != 16) {
        set_error(ctx, "fsx_map_new: height_bits must be 8 or 16");
        return 1;
    }
    if (shade_size < 0 || shade_size > 65536) {
        set_error(ctx, "fsx_map_new: invalid shade size %d", shade_size);
        return 1;
    }
    struct futhark_u32_2d *c = futhark_new_u32_2d(ctx, color, rows, cols);
    struct futhark_i32_2d *h = futhark_new_i32_2d(ctx, height, rows, cols);
    if (!c || !h) {
        futhark_free_u32_2d(ctx, c);
        futhark_free_i32_2d(ctx, h);
        return 1;
    }
    struct fsx_map *m = (struct fsx_map *)calloc(1, sizeof *m);
    int rc = 1;
    if (m) {
        m->shade_size = shade_size == 0 ? 1024 : shade_size;
        m->t = terrain_new(ctx, c->d, h->d, (int32_t)rows, (int32_t)cols, height_bits);
        if (m->t)
            m->s = shade_new(ctx, m->t, m->shade_size, sun_height, sun_ang);
        if (m->t && m->s) {
            rc = 0;
        } else {
            terrain_unref(ctx, m->t);
            free(m);
            m = NULL;
        }
    }
    hipStreamSynchronize(ctx->stream);
    futhark_free_u32_2d(ctx, c);
    futhark_free_i32_2d(ctx, h);
    *out = m;
    return rc;
}

int fsx_map_free(struct futhark_context *ctx, struct fsx_map *map)
{
    if (!ctx)
        return 1;
    if (!map)
        return 0;
    terrain_unref(ctx, map->t);
    shade_unref(ctx, map->s);
    free(map);
    return 0;
}

int32_t fsx_map_shade_size(const struct fsx_map *map)
{
    return map ? map->shade_size : 0;
}

int fsx_map_shaded(struct futhark_context *ctx, const struct fsx_map *map, uint32_t *host_out)
{
    REQUIRE_ARGS(ctx, map && host_out, "fsx_map_shaded");
    size_t n = (size_t)map->s->rows * (size_t)map->s->cols;
    HIPCHECK(ctx, hipMemcpyAsync(host_out, map->s->d_shaded, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsx_render(struct futhark_context *ctx, const struct fsx_map *map,
               const struct fsx_view *view, uint32_t *dev_out, int64_t pitch)
{
    REQUIRE_ARGS(ctx, map && view && (dev_out || view->h <= 0 || view->w <= 0), "fsx_render");
    if (pitch != 0 && pitch < view->w) {
        set_error(ctx, "fsx_render: pitch %lld below width %d", (long long)pitch, view->w);
        return 1;
    }
    return render_into(ctx, map->t, map->s, view, dev_out, pitch);
}

/* render2 (fut/voxel_renderer.fut:257-305), the reference's ray-march
 * renderer, over the map's masked heights and unshaded colours.  The
 * per-column cos / sin of angle + horz_ang are computed here with the C
 * library, as `futhark c` computes them, and copied to the device after
 * the context stream has drained (the buffer is reused between calls). */
int fsx_render2(struct futhark_context *ctx, const struct fsx_map *map,
                const struct fsx_view *view, uint32_t *dev_out, int64_t pitch)
{
    REQUIRE_ARGS(ctx, map && view, "fsx_render2");
    if (view->h < 0 || view->w < 0 || (pitch != 0 && pitch < view->w)) {
        set_error(ctx, "fsx_render2: bad frame %d x %d (pitch %lld)", view->h, view->w,
                  (long long)pitch);
        return 1;
    }
    if (view->h == 0 || view->w == 0)
        return 0;
    if (!dev_out) {
        set_error(ctx, "fsx_render2: null output");
        return 1;
    }
    const int32_t n = i32_f32(view->distance / 4.0f);
    const int32_t miss = i32_f32(view->distance + 1.0f);
    /* the march stops at the first hit, which is the reference's result
     * while a miss (distance i32(distance + 1)) cannot beat a hit */
    if (n > (1 << 16) || (n >= 0 && miss <= n)) {
        set_error(ctx, "fsx_render2: distance %g outside the supported range (distance / 4 "
                       "at most 2^16)", (double)view->distance);
        return 1;
    }
    const size_t need = (size_t)2 * (size_t)view->w;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (need > ctx->trig_cap) {
        if (ctx->d_trig)
            HIPCHECK(ctx, hipFree(ctx->d_trig));
        ctx->d_trig = NULL;
        ctx->trig_cap = 0;
        HIPCHECK(ctx, hipMalloc((void **)&ctx->d_trig, sizeof(float) * need));
        ctx->trig_cap = need;
    }
    float *h = (float *)malloc(sizeof(float) * need);
    if (!h) {
        set_error(ctx, "fsx_render2: out of host memory");
        return 1;
    }
    for (int32_t x = 0; x < view->w; x++) {
        const float horz_ang = ((float)x / (float)view->w) * (2.5f - 1.0f);
        h[x] = cosf(view->angle + horz_ang);
        h[view->w + x] = sinf(view->angle + horz_ang);
    }
    hipError_t e = hipMemcpy(ctx->d_trig, h, sizeof(float) * need, hipMemcpyHostToDevice);
    free(h);
    HIPCHECK(ctx, e);
    fsk_frame f;
    frame_of_view(view, &f);
    fsk_map m;
    m.hq = map->t->d_hq;
    m.hrows = map->t->rows;
    m.hcols = map->t->cols;
    m.hbits = map->t->hbits;
    m.cq = NULL;
    m.crows = 0;
    m.ccols = 0;
    int pe = prof_begin(ctx);
    HIPCHECK(ctx, fsk_render2(&m, &f, ctx->d_trig, ctx->d_trig + view->w, map->t->d_color, n,
                              dev_out, pitch > 0 ? pitch : view->w, ctx->stream));
    prof_end(ctx, pe);
    return 0;
}

int fsx_render_host(struct futhark_context *ctx, const struct fsx_map *map,
                    const struct fsx_view *view, uint32_t *host_out)
{
    REQUIRE_ARGS(ctx, map && view && host_out, "fsx_render_host");
    size_t bytes = (size_t)(view->h > 0 ? view->h : 0) * (size_t)(view->w > 0 ? view->w : 0) * 4;
    if (bytes == 0)
        return render_into(ctx, map->t, map->s, view, NULL, 0);
    uint32_t *d = NULL;
    if (dev_alloc(ctx, (void **)&d, bytes))
        return 1;
    int rc = render_into(ctx, map->t, map->s, view, d, view->w);
    if (rc == 0) {
        hipError_t e = hipMemcpyAsync(host_out, d, bytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            set_error(ctx, "fsx_render_host: %s", hipGetErrorString(e));
            rc = 1;
        }
    }
    dev_release(ctx, d, bytes);
    return rc;
}

/* ------------------------------------------------------------------ */
/* The stateless program of tests/futspace.fut                          */
/* ------------------------------------------------------------------ */

/* get_zs delta cam.distance z_0 with delta = 0.005, z_0 = 1
 * (tests/futspace.fut:47-54, :84-86): n = i32.f32 (num / div). */
static int32_t stateless_zs_count(float d)
{
    const float delta = 0.005f, z_0 = 1.0f;
    float sq = sqrtf(powf(delta - 2.0f * z_0, 2.0f) + (8.0f * delta) * d);
    float num = (sq - 2.0f * z_0) + delta;
    return i32_f32(num / (2.0f * delta));
}

int fsx_render_stateless(struct futhark_context *ctx, const int32_t *dev_color,
                         const int32_t *dev_height, int32_t q, int32_t r,
                         const struct fsx_view *view, int32_t *dev_out)
{
    REQUIRE_ARGS(ctx, view, "fsx_render_stateless");
    if (view->h < 0 || view->w < 0 || view->h > fsk_render_max_rows() || q <= 0 || r <= 0 ||
        (int64_t)q * r > ((int64_t)1 << 30)) {
        set_error(ctx, "fsx_render_stateless: unsupported sizes (frame %d x %d, map %d x %d)",
                  view->h, view->w, q, r);
        return 1;
    }
    if (view->h == 0 || view->w == 0)
        return 0;
    if (!dev_color || !dev_height || !dev_out) {
        set_error(ctx, "fsx_render_stateless: null buffer");
        return 1;
    }
    fsk_frame f;
    frame_of_view(view, &f);
    f.d = 0.005f;
    f.z0 = 1.0f;
    f.scale = 240.0f;
    f.n = stateless_zs_count(view->distance);
    if (!slices_ok(ctx, "fsx_render_stateless", f.n))
        return 1;
    fsk_map m;
    m.hq = dev_height;
    m.hrows = q;
    m.hcols = r;
    m.hbits = 32;
    m.cq = dev_color;
    m.crows = q;
    m.ccols = r;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int pe = prof_begin(ctx);
    HIPCHECK(ctx, fsk_render_nearest(&m, &f, &ctx->tune, (uint32_t *)dev_out, view->w,
                                     ctx->stream));
    prof_end(ctx, pe);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Differentiable renderer                                             */
/* ------------------------------------------------------------------ */

int fsx_map_shade_factor(struct futhark_context *ctx, const struct fsx_map *map, float *dev_sf)
{
    REQUIRE_ARGS(ctx, map && dev_sf, "fsx_map_shade_factor");
    const struct fs_terrain *t = map->t;
    const struct fs_shade *s = map->s;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    HIPCHECK(ctx, fsk_shade_factor(t->d_hp, t->rows, t->cols, s->sun[0], s->sun[1], s->sun[2],
                                   s->rows, dev_sf, ctx->stream));
    return 0;
}

/* The context's scratch for per-workgroup partial sums (allocated once;
 * stream-ordered reuse). */
static int partials(struct futhark_context *ctx, float **p)
{
    if (!ctx->d_partials) {
        hipError_t e = hipMalloc((void **)&ctx->d_partials, sizeof(float) * FS_PARTIALS);
        if (e != hipSuccess) {
            ctx->d_partials = NULL;
            set_error(ctx, "scratch allocation failed: %s", hipGetErrorString(e));
            return 1;
        }
    }
    *p = ctx->d_partials;
    return 0;
}

static int diff_frame(struct futhark_context *ctx, const char *what, const struct fsx_view *v,
                      int32_t hc, int32_t wc, fsk_frame *f)
{
    if (v->h < 0 || v->w < 0 || v->h > fsk_render_max_rows()) {
        set_error(ctx, "%s: unsupported frame size %d x %d", what, v->h, v->w);
        return 1;
    }
    if (hc <= 0 || wc <= 0 || (int64_t)hc * wc > ((int64_t)1 << 31)) {
        set_error(ctx, "%s: colour planes of %d x %d (at most 2^31 texels)", what, hc, wc);
        return 1;
    }
    frame_of_view(v, f);
    return slices_ok(ctx, what, f->n) ? 0 : 1;
}

int fsx_render_f32(struct futhark_context *ctx, const struct fsx_map *map,
                   const struct fsx_view *view, const float *dev_planes, const float *dev_sf,
                   int32_t hc, int32_t wc, float sky_scale, float *dev_img, int32_t *dev_src)
{
    REQUIRE_ARGS(ctx, map && view, "fsx_render_f32");
    fsk_frame f;
    if (diff_frame(ctx, "fsx_render_f32", view, hc, wc, &f))
        return 1;
    if (view->h == 0 || view->w == 0)
        return 0;
    if (!dev_planes || !dev_sf || !dev_img || !dev_src) {
        set_error(ctx, "fsx_render_f32: null buffer");
        return 1;
    }
    fsk_map m;
    m.hq = map->t->d_hq;
    m.hrows = map->t->rows;
    m.hcols = map->t->cols;
    m.hbits = map->t->hbits;
    m.cq = NULL;
    m.crows = hc;
    m.ccols = wc;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int pe = prof_begin(ctx);
    HIPCHECK(ctx, fsk_render_slices(&m, &f, &ctx->tune, dev_src, view->w, ctx->stream));
    HIPCHECK(ctx, fsk_diff_color(&f, dev_planes, dev_sf, hc, wc, sky_scale, dev_src, dev_img,
                                 ctx->stream));
    prof_end(ctx, pe);
    return 0;
}

int fsx_render_f32_vjp(struct futhark_context *ctx, const struct fsx_view *view,
                       const float *dev_planes, const float *dev_sf, int32_t hc, int32_t wc,
                       const int32_t *dev_src, const float *dev_gimg, int32_t accumulate,
                       float *dev_gplanes, float *dev_gsky)
{
    REQUIRE_ARGS(ctx, view, "fsx_render_f32_vjp");
    fsk_frame f;
    if (diff_frame(ctx, "fsx_render_f32_vjp", view, hc, wc, &f))
        return 1;
    if (!dev_planes || !dev_sf || !dev_gplanes || !dev_gsky ||
        ((!dev_src || !dev_gimg) && view->h > 0 && view->w > 0)) {
        set_error(ctx, "fsx_render_f32_vjp: null buffer");
        return 1;
    }
    if (step_blocks(view->w, view->h) > FS_PARTIALS) {
        set_error(ctx, "fsx_render_f32_vjp: a %d x %d frame has %lld blocks of %d x %d pixels, "
                  "more than the %d partial sums of the adjoint", view->w, view->h,
                  (long long)step_blocks(view->w, view->h), FS_STEP_BLOCK_COLS,
                  fsk_grad_block_rows(view->w, view->h), FS_PARTIALS);
        return 1;
    }
    float *part;
    if (partials(ctx, &part))
        return 1;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int pe = prof_begin(ctx);
    if (!accumulate)
        HIPCHECK(ctx, fsk_zero_f32(dev_gplanes, (size_t)3 * (size_t)hc * (size_t)wc, ctx->stream));
    if (view->h > 0 && view->w > 0)
        HIPCHECK(ctx, fsk_diff_grad(&f, dev_planes, dev_sf, hc, wc, dev_src, dev_gimg,
                                    dev_gplanes, dev_gsky, accumulate != 0, part, FS_PARTIALS,
                                    ctx->stream));
    else if (!accumulate)
        HIPCHECK(ctx, fsk_zero_f32(dev_gsky, 1, ctx->stream));
    prof_end(ctx, pe);
    return 0;
}

/* The fused step behind both entries; `noalloc` takes the partial sums
 * from the device's fixed scratch array instead of the context's buffer
 * (which the first use allocates). */
static int l2_step(struct futhark_context *ctx, const char *what, int noalloc,
                   const struct fsx_map *map, const struct fsx_view *view,
                   const float *dev_planes, const float *dev_sf, int32_t hc, int32_t wc,
                   float sky_scale, const float *dev_target, int32_t accumulate, float *dev_img,
                   int32_t *dev_src, float *dev_gimg, float *dev_gplanes, float *dev_gsky,
                   float *dev_loss)
{
    REQUIRE_ARGS(ctx, map && view, what);
    fsk_frame f;
    if (diff_frame(ctx, what, view, hc, wc, &f))
        return 1;
    if (!dev_planes || !dev_sf || !dev_gplanes || !dev_gsky || !dev_loss ||
        ((!dev_target || !dev_src) && view->h > 0 && view->w > 0)) {
        set_error(ctx, "%s: null buffer", what);
        return 1;
    }
    const size_t nplanes = (size_t)3 * (size_t)hc * (size_t)wc;
    /* the loss and sky partials take half of the scratch each */
    if (step_blocks(view->w, view->h) > FS_PARTIALS / 2) {
        set_error(ctx, "%s: a %d x %d frame has %lld blocks of %d x %d pixels, more than the "
                  "%d loss partial sums of the fused step", what, view->w, view->h,
                  (long long)step_blocks(view->w, view->h), FS_STEP_BLOCK_COLS,
                  fsk_grad_block_rows(view->w, view->h), FS_PARTIALS / 2);
        return 1;
    }
    float *part = NULL;
    if (noalloc) {
        int64_t cap = 0;
        part = fsk_scratch(ctx->device, &cap);
        if (!part || cap < FS_PARTIALS) {
            set_error(ctx, "%s: the device scratch of %lld floats is not set up", what,
                      (long long)cap);
            return 1;
        }
    }
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int pe = prof_begin(ctx);
    if (view->h == 0 || view->w == 0) {
        if (!accumulate) {
            HIPCHECK(ctx, fsk_zero_f32(dev_gplanes, nplanes, ctx->stream));
            HIPCHECK(ctx, fsk_zero_f32(dev_gsky, 1, ctx->stream));
            HIPCHECK(ctx, fsk_zero_f32(dev_loss, 1, ctx->stream));
        }
        prof_end(ctx, pe);
        return 0;
    }
    if (!noalloc && partials(ctx, &part))
        return 1;
    fsk_map m;
    m.hq = map->t->d_hq;
    m.hrows = map->t->rows;
    m.hcols = map->t->cols;
    m.hbits = map->t->hbits;
    m.cq = NULL;
    m.crows = hc;
    m.ccols = wc;
    /* unless accumulating, the gradient planes are cleared by the slice
     * render's spare workgroups (CUs its tiles leave idle) or else by a
     * launch of their own */
    int32_t cleared = 0;
    HIPCHECK(ctx, fsk_render_slices_clear(&m, &f, &ctx->tune, dev_src, view->w,
                                          accumulate ? NULL : dev_gplanes, nplanes, &cleared,
                                          ctx->stream));
    if (!accumulate && !cleared)
        HIPCHECK(ctx, fsk_zero_f32(dev_gplanes, nplanes, ctx->stream));
    /* colour, loss terms and adjoint in one launch; the loss and sky
     * partials take the scratch's two halves */
    const int32_t q = FS_PARTIALS / 2;
    HIPCHECK(ctx, fsk_diff_step_l2(&f, dev_planes, dev_sf, hc, wc, sky_scale, dev_src, dev_target,
                                   dev_img, dev_gimg, dev_gplanes, dev_gsky, accumulate != 0,
                                   dev_loss, accumulate != 0, part, part + q, q, ctx->stream));
    prof_end(ctx, pe);
    return 0;
}

int fsx_render_f32_l2_step(struct futhark_context *ctx, const struct fsx_map *map,
                           const struct fsx_view *view, const float *dev_planes,
                           const float *dev_sf, int32_t hc, int32_t wc, float sky_scale,
                           const float *dev_target, int32_t accumulate, float *dev_img,
                           int32_t *dev_src, float *dev_gimg, float *dev_gplanes,
                           float *dev_gsky, float *dev_loss)
{
    return l2_step(ctx, "fsx_render_f32_l2_step", 0, map, view, dev_planes, dev_sf, hc, wc,
                   sky_scale, dev_target, accumulate, dev_img, dev_src, dev_gimg, dev_gplanes,
                   dev_gsky, dev_loss);
}

int fsx_render_f32_l2_step_noalloc(struct futhark_context *ctx, const struct fsx_map *map,
                                   const struct fsx_view *view, const float *dev_planes,
                                   const float *dev_sf, int32_t hc, int32_t wc, float sky_scale,
                                   const float *dev_target, int32_t accumulate, float *dev_img,
                                   int32_t *dev_src, float *dev_gimg, float *dev_gplanes,
                                   float *dev_gsky, float *dev_loss)
{
    return l2_step(ctx, "fsx_render_f32_l2_step_noalloc", 1, map, view, dev_planes, dev_sf, hc,
                   wc, sky_scale, dev_target, accumulate, dev_img, dev_src, dev_gimg,
                   dev_gplanes, dev_gsky, dev_loss);
}

int fsx_adam_step(struct futhark_context *ctx, float *dev_param, const float *dev_grad,
                  float *dev_m, float *dev_v, int64_t n, double lr, double beta1,
                  double beta2, double eps, int32_t step, float lo, float hi)
{
    if (!ctx)
        return 1;
    if (n < 0 || step < 1 || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) ||
        ((!dev_param || !dev_grad || !dev_m || !dev_v) && n > 0) ||
        (((uintptr_t)dev_param | (uintptr_t)dev_grad | (uintptr_t)dev_m | (uintptr_t)dev_v) &
         15u)) {
        set_error(ctx, "fsx_adam_step: invalid arguments (null or misaligned buffer, step < 1 "
                       "or a beta outside [0, 1))");
        return 1;
    }
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    HIPCHECK(ctx, fsk_adam(dev_param, dev_grad, dev_m, dev_v, (size_t)n, lr, beta1, beta2, eps,
                           step, lo, hi, ctx->stream));
    return 0;
}

int fsx_l2_loss(struct futhark_context *ctx, const float *dev_img, const float *dev_target,
                int64_t n, int32_t accumulate, float *dev_grad, float *dev_loss)
{
    if (!ctx)
        return 1;
    if (n < 0 || ((!dev_img || !dev_target || !dev_grad) && n > 0) || !dev_loss ||
        (((uintptr_t)dev_img | (uintptr_t)dev_target | (uintptr_t)dev_grad) & 15u)) {
        set_error(ctx, "fsx_l2_loss: null or misaligned buffer");
        return 1;
    }
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    float *part;
    if (partials(ctx, &part))
        return 1;
    HIPCHECK(ctx, fsk_l2_loss(dev_img, dev_target, (size_t)n, dev_grad, dev_loss, accumulate, part,
                              FS_PARTIALS, ctx->stream));
    return 0;
}

/* ------------------------------------------------------------------ */
/* Batched renders                                                     */
/* ------------------------------------------------------------------ */

#define FS_QUEUES 8   /* one work queue per XCD */

struct fsx_batch {
    fsk_batch b;
    void *d_mem;                   /* views | items | qmeta | counters */
    int32_t n_views, n_maps;
    struct fs_terrain **terrains;  /* references held while the batch lives */
    struct fs_shade **shades;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

int fsx_set_tuning(struct futhark_context *ctx, int32_t tile_columns, int32_t wg_per_cu,
                   int32_t chunk)
{
    if (!ctx)
        return 1;
    if ((tile_columns != 0 && tile_columns != 4 && tile_columns != 8) ||
        (wg_per_cu != 0 && wg_per_cu != 1 && wg_per_cu != 2) ||
        (chunk != 0 && (chunk < 256 || chunk > 2048))) {
        set_error(ctx, "fsx_set_tuning: unsupported tuning (%d, %d, %d)", tile_columns, wg_per_cu,
                  chunk);
        return 1;
    }
    fsk_tuning t = {tile_columns, wg_per_cu, chunk};
    ctx->tune = t;
    return 0;
}

int fsx_get_tuning(struct futhark_context *ctx, int32_t out[3])
{
    REQUIRE_ARGS(ctx, out, "fsx_get_tuning");
    out[0] = ctx->tune.tc;
    out[1] = ctx->tune.wg_per_cu;
    out[2] = ctx->tune.chunk;
    return 0;
}

static int is_pow2(int32_t v)
{
    return v > 0 && (v & (v - 1)) == 0;
}

/* Deal the views' tiles to the per-XCD queues: costlier views first, each
 * view's tiles cut into FS_QUEUES contiguous blocks so that neighbouring
 * tiles share an XCD. */
static void deal_items(const fsk_batch_view *bv, int32_t n_views, uint32_t *items,
                       uint32_t *qmeta)
{
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_views > 0 ? n_views : 1));
    for (int32_t i = 0; i < n_views; i++)
        order[i] = i;
    /* insertion sort by descending work (slices x columns + pixels) */
    for (int32_t i = 1; i < n_views; i++) {
        int32_t k = order[i];
        double wk = (double)bv[k].f.n * bv[k].f.cols + (double)bv[k].f.rows * bv[k].f.cols;
        int32_t j = i - 1;
        while (j >= 0) {
            int32_t o = order[j];
            double wo = (double)bv[o].f.n * bv[o].f.cols + (double)bv[o].f.rows * bv[o].f.cols;
            if (wo >= wk)
                break;
            order[j + 1] = o;
            j--;
        }
        order[j + 1] = k;
    }
    uint32_t len[FS_QUEUES] = {0};
    for (int32_t i = 0; i < n_views; i++) {
        int32_t nt = bv[order[i]].ntiles;
        for (int q = 0; q < FS_QUEUES; q++)
            len[q] += (uint32_t)((int64_t)nt * (q + 1) / FS_QUEUES - (int64_t)nt * q / FS_QUEUES);
    }
    uint32_t start = 0, fill[FS_QUEUES];
    for (int q = 0; q < FS_QUEUES; q++) {
        qmeta[q] = start;
        qmeta[FS_QUEUES + q] = len[q];
        fill[q] = start;
        start += len[q];
    }
    for (int32_t i = 0; i < n_views; i++) {
        int32_t v = order[i], nt = bv[v].ntiles;
        for (int q = 0; q < FS_QUEUES; q++) {
            int32_t t0 = (int32_t)((int64_t)nt * q / FS_QUEUES);
            int32_t t1 = (int32_t)((int64_t)nt * (q + 1) / FS_QUEUES);
            for (int32_t t = t0; t < t1; t++)
                items[fill[q]++] = ((uint32_t)v << 16) | (uint32_t)t;
        }
    }
    free(order);
}

int fsx_batch_new(struct futhark_context *ctx, const struct fsx_map *const *maps, int32_t n_maps,
                  const struct fsx_view *views, int32_t n_views, uint32_t *dev_out,
                  const int64_t *offsets, struct fsx_batch **out)
{
    if (!ctx)
        return 1;
    if (!maps || !views || !dev_out || !offsets || !out || n_views < 1 || n_maps < 1) {
        set_error(ctx, "fsx_batch_new: null argument, or no views or maps");
        return 1;
    }
    if (n_views > 65535) {
        set_error(ctx, "fsx_batch_new: at most 65535 views per batch");
        return 1;
    }
    int32_t max_rows = 0, hbits = -1, pow2 = 1;
    for (int32_t i = 0; i < n_views; i++) {
        const struct fsx_view *v = &views[i];
        if (v->map < 0 || v->map >= n_maps || !maps[v->map]) {
            set_error(ctx, "fsx_batch_new: view %d names map %d of %d", i, v->map, n_maps);
            return 1;
        }
        if (v->h <= 0 || v->w <= 0 || v->h > fsk_batch_max_rows() || v->w > 65535 * 4) {
            set_error(ctx, "fsx_batch_new: view %d has unsupported size %d x %d", i, v->h, v->w);
            return 1;
        }
        const struct fsx_map *m = maps[v->map];
        if (hbits < 0)
            hbits = m->t->hbits;
        if (m->t->hbits != hbits) {
            set_error(ctx, "fsx_batch_new: maps with 8- and 16-bit heights in one batch");
            return 1;
        }
        pow2 &= is_pow2(m->t->rows) && is_pow2(m->t->cols) && is_pow2(m->s->rows) &&
                is_pow2(m->s->cols);
        if (v->h > max_rows)
            max_rows = v->h;
    }
    struct fsx_batch *bt = (struct fsx_batch *)calloc(1, sizeof *bt);
    fsk_batch_view *hv = (fsk_batch_view *)calloc((size_t)n_views, sizeof *hv);
    if (!bt || !hv) {
        free(bt);
        free(hv);
        set_error(ctx, "fsx_batch_new: out of host memory");
        return 1;
    }
    bt->b.tune = ctx->tune;
    int32_t tc = fsk_tile_columns(max_rows, &bt->b.tune);
    int64_t n_items = 0;
    for (int32_t i = 0; i < n_views; i++) {
        const struct fsx_map *m = maps[views[i].map];
        hv[i].m.hq = m->t->d_hq;
        hv[i].m.hrows = m->t->rows;
        hv[i].m.hcols = m->t->cols;
        hv[i].m.hbits = m->t->hbits;
        hv[i].m.cq = m->s->d_cq;
        hv[i].m.crows = m->s->rows;
        hv[i].m.ccols = m->s->cols;
        frame_of_view(&views[i], &hv[i].f);
        if (hv[i].f.n > FS_MAX_SLICES) {
            const int32_t n = hv[i].f.n;
            free(bt);
            free(hv);
            set_error(ctx, "fsx_batch_new: view %d has %d depth slices, above the supported %d", i,
                      n, FS_MAX_SLICES);
            return 1;
        }
        hv[i].f.tame = fsk_frame_tame(&hv[i].f);
        hv[i].offset = offsets[i];
        hv[i].ntiles = (views[i].w + tc - 1) / tc;
        n_items += hv[i].ntiles;
    }
    size_t vb = sizeof(fsk_batch_view) * (size_t)n_views;
    size_t ib = sizeof(uint32_t) * (size_t)n_items;
    size_t mb = sizeof(uint32_t) * 2 * FS_QUEUES;
    size_t cb = sizeof(uint32_t) * FS_QUEUES;
    uint32_t *hitems = (uint32_t *)malloc(ib + mb);
    int rc = 1;
    if (hitems && hipSetDevice(ctx->device) == hipSuccess &&
        hipMalloc(&bt->d_mem, vb + ib + mb + cb) == hipSuccess) {
        deal_items(hv, n_views, hitems, hitems + n_items);
        char *base = (char *)bt->d_mem;
        hipError_t e = hipMemcpy(base, hv, vb, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(base + vb, hitems, ib + mb, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            bt->b.views = (const fsk_batch_view *)base;
            bt->b.items = (const uint32_t *)(base + vb);
            bt->b.qmeta = (const uint32_t *)(base + vb + ib);
            bt->b.counters = (uint32_t *)(base + vb + ib + mb);
            bt->b.out = dev_out;
            bt->b.nqueues = FS_QUEUES;
            bt->b.max_rows = max_rows;
            bt->b.hbits = hbits;
            bt->b.pow2 = pow2;
            bt->b.grid = fsk_batch_grid(max_rows, &bt->b.tune);
            rc = 0;
        } else {
            set_error(ctx, "fsx_batch_new: upload failed: %s", hipGetErrorString(e));
        }
    } else {
        set_error(ctx, "fsx_batch_new: allocation failed");
    }
    free(hitems);
    free(hv);
    if (rc == 0) {
        bt->n_views = n_views;
        bt->n_maps = n_maps;
        bt->terrains = (struct fs_terrain **)calloc((size_t)n_maps, sizeof *bt->terrains);
        bt->shades = (struct fs_shade **)calloc((size_t)n_maps, sizeof *bt->shades);
        if (!bt->terrains || !bt->shades)
            rc = 1;
        for (int32_t i = 0; rc == 0 && i < n_maps; i++) {
            if (!maps[i])
                continue;
            bt->terrains[i] = maps[i]->t;
            bt->shades[i] = maps[i]->s;
            maps[i]->t->refs++;
            maps[i]->s->refs++;
        }
    }
    if (rc) {
        if (bt->d_mem)
            hipFree(bt->d_mem);
        free(bt->terrains);
        free(bt->shades);
        free(bt);
        return 1;
    }
    *out = bt;
    return 0;
}

int fsx_batch_run(struct futhark_context *ctx, struct fsx_batch *batch)
{
    REQUIRE_ARGS(ctx, batch, "fsx_batch_run");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int pe = prof_begin(ctx);
    HIPCHECK(ctx, fsk_batch_run(&batch->b, ctx->stream));
    prof_end(ctx, pe);
    return 0;
}

int fsx_batch_capture(struct futhark_context *ctx, struct fsx_batch *batch)
{
    REQUIRE_ARGS(ctx, batch, "fsx_batch_capture");
    if (ctx->stream == NULL) {
        set_error(ctx, "fsx_batch_capture: the context stream cannot be captured");
        return 1;
    }
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    if (batch->exec) {
        hipGraphExecDestroy(batch->exec);
        batch->exec = NULL;
    }
    if (batch->graph) {
        hipGraphDestroy(batch->graph);
        batch->graph = NULL;
    }
    HIPCHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = (hipError_t)fsk_batch_run(&batch->b, ctx->stream);
    hipGraph_t g = NULL;
    hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g)
            hipGraphDestroy(g);
        set_error(ctx, "fsx_batch_capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
        return 1;
    }
    batch->graph = g;
    HIPCHECK(ctx, hipGraphInstantiate(&batch->exec, g, NULL, NULL, 0));
    return 0;
}

int fsx_batch_replay(struct futhark_context *ctx, struct fsx_batch *batch)
{
    if (!ctx || !batch || !batch->exec) {
        if (ctx)
            set_error(ctx, "fsx_batch_replay: the batch has not been captured");
        return 1;
    }
    int pe = prof_begin(ctx);
    HIPCHECK(ctx, hipGraphLaunch(batch->exec, ctx->stream));
    prof_end(ctx, pe);
    return 0;
}

int fsx_batch_free(struct futhark_context *ctx, struct fsx_batch *batch)
{
    if (!ctx)
        return 1;
    if (!batch)
        return 0;
    hipStreamSynchronize(ctx->stream);
    if (batch->exec)
        hipGraphExecDestroy(batch->exec);
    if (batch->graph)
        hipGraphDestroy(batch->graph);
    hipFree(batch->d_mem);
    for (int32_t i = 0; i < batch->n_maps; i++) {
        terrain_unref(ctx, batch->terrains[i]);
        shade_unref(ctx, batch->shades[i]);
    }
    free(batch->terrains);
    free(batch->shades);
    free(batch);
    return 0;
}

int fsx_render_batch(struct futhark_context *ctx, const struct fsx_map *const *maps,
                     int32_t n_maps, const struct fsx_view *views, int32_t n_views,
                     uint32_t *dev_out, const int64_t *offsets)
{
    struct fsx_batch *b = NULL;
    if (fsx_batch_new(ctx, maps, n_maps, views, n_views, dev_out, offsets, &b))
        return 1;
    int rc = fsx_batch_run(ctx, b);
    fsx_batch_free(ctx, b);
    return rc;
}
