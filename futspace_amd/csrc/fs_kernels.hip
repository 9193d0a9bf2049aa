/*
 * fs_kernels.hip -- gfx950 kernels of the futspace renderer.
 *
 * Compiled with -ffp-contract=off and correctly rounded f32 divide/sqrt:
 * every f32 expression below is evaluated in the reference's order
 * (fut/voxel_renderer.fut, fut/render_functions.fut, fut/effects.fut and
 * the matte argb operations they call), so frames are bit-exact with the
 * `futhark c` build of the reference on x86-64.
 *
 * Hot path (k_render): one workgroup owns a stripe of TC adjacent frame
 * columns for the whole depth range.  Lane (slice group g, column c) walks
 * KR consecutive slices per chunk; a prefix-min across slice groups gives
 * each lane the y-buffer value in front of its slices, so a lane knows
 * which of its samples rise strictly above everything nearer (the
 * reference's `scan occlude` + `scatter`, fut/voxel_renderer.fut:231,244).
 * Only those samples get the expensive bilinear argb colour.  The stripe
 * is assembled in LDS, filled top-down (`scan fill_vline` + sky,
 * :246-249) and written once, row-major.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fs_internal.h"

namespace {

/* ------------------------------------------------------------------ */
/* Scalar semantics                                                    */
/* ------------------------------------------------------------------ */

/* Texel index from a floor/ceil value.  gfx950's v_cvt_i32_f32 saturates
 * (NaN -> 0); indices only have to be exact where |v| < 2^24, because at
 * larger magnitudes every f32 is an integer and both bilinear weights are
 * zero, so the texel value is multiplied by 0. */
__device__ __forceinline__ int32_t idx_of(float v)
{
    return (int32_t)fminf(fmaxf(v, -1073741824.0f), 1073741824.0f);
}

/* i32.f32 as `futhark c` runs it on x86-64 (cvttss2si): NaN and values
 * outside [-2^31, 2^31) give INT_MIN. */
__device__ __forceinline__ int32_t i32_f32_x86(float v)
{
    bool ok = (v >= -2147483648.0f) && (v < 2147483648.0f);
    return ok ? (int32_t)fminf(fmaxf(v, -2147483648.0f), 2147483520.0f) : INT32_MIN;
}

/* i32.max 0 (i32.f32 rel), fut/voxel_renderer.fut:225, x86 semantics:
 * NaN, >= 2^31 and < -2^31 all give 0. */
__device__ __forceinline__ int32_t screen_row(float rel)
{
    return (rel < 2147483648.0f) ? max(0, (int32_t)fmaxf(rel, -1.0f)) : 0;
}

/* Futhark's signed `%` (floor-mod). */
__device__ __forceinline__ int32_t fmod_i32(int32_t a, int32_t m)
{
    int32_t r = a % m;
    return (r != 0 && ((r < 0) != (m < 0))) ? r + m : r;
}

template <bool POW2>
__device__ __forceinline__ int32_t wrap(int32_t v, int32_t m)
{
    return POW2 ? (v & (m - 1)) : fmod_i32(v, m);
}

/* ------------------------------------------------------------------ */
/* matte argb, as restated for the reference (Gloss mixColors)         */
/* ------------------------------------------------------------------ */

/* f32.u32 b / 255 for a byte b, correctly rounded: one FMA correction
 * of the reciprocal product (exhaustively equal to the division for all
 * 256 bytes; checked on the device by the test suite). */
__device__ __forceinline__ float unit8(uint32_t b)
{
    const float inv = 1.0f / 255.0f;
    float fb = (float)b;
    float q0 = fb * inv;
    float r = __builtin_fmaf(-q0, 255.0f, fb);
    return __builtin_fmaf(r, inv, q0);
}

/* u32.f32 (clamp_channel v * 255); clamp passes NaN, u32 of NaN is 0. */
__device__ __forceinline__ uint32_t chan_u8(float v)
{
    float c = (v < 0.0f) ? 0.0f : ((v > 1.0f) ? 1.0f : v);
    float s = c * 255.0f;
    return (s == s) ? (uint32_t)s : 0u;
}

__device__ __forceinline__ uint32_t from_rgba(float r, float g, float b, float a)
{
    return (chan_u8(a) << 24) | (chan_u8(r) << 16) | (chan_u8(g) << 8) | chan_u8(b);
}

/* argb.mix m1 c1 m2 c2.  When m1 + m2 rounds to exactly 1 the
 * normalisation divides by 1, which is exact, so the quotient is the
 * dividend; that skips three correctly rounded divisions on the common
 * bilinear path (weights (c - x) and (x - f) sum to 1 for |x| >= 1). */
__device__ __forceinline__ uint32_t argb_mix(float m1, uint32_t c1, float m2, uint32_t c2)
{
    float r1 = unit8((c1 >> 16) & 255u), g1 = unit8((c1 >> 8) & 255u);
    float b1 = unit8(c1 & 255u), a1 = unit8(c1 >> 24);
    float r2 = unit8((c2 >> 16) & 255u), g2 = unit8((c2 >> 8) & 255u);
    float b2 = unit8(c2 & 255u), a2 = unit8(c2 >> 24);
    float m12 = m1 + m2;
    float m1n = m1, m2n = m2;
    float al = (m1 * a1) + (m2 * a2);
    if (m12 != 1.0f) {
        m1n = m1 / m12;
        m2n = m2 / m12;
        al = al / m12;
    }
    float r = __builtin_sqrtf((m1n * (r1 * r1)) + (m2n * (r2 * r2)));
    float g = __builtin_sqrtf((m1n * (g1 * g1)) + (m2n * (g2 * g2)));
    float b = __builtin_sqrtf((m1n * (b1 * b1)) + (m2n * (b2 * b2)));
    return from_rgba(r, g, b, al);
}

/* ------------------------------------------------------------------ */
/* Samplers                                                            */
/* ------------------------------------------------------------------ */

/* png_height_filtered, fut/render_functions.fut:67-77, from the height
 * quad at (floor y, floor x). */
template <int HBITS, bool HP2>
__device__ __forceinline__ float height_bilinear(const fsk_map &m, float x, float y)
{
    float fx = floorf(x), cx = ceilf(x), fy = floorf(y), cy = ceilf(y);
    int32_t ix = wrap<HP2>(idx_of(fx), m.hcols);
    int32_t iy = wrap<HP2>(idx_of(fy), m.hrows);
    size_t t = (size_t)iy * (size_t)m.hcols + (size_t)ix;
    float h00, h01, h10, h11;
    if (HBITS == 8) {
        uint32_t q = ((const uint32_t *)m.hq)[t];
        h00 = (float)(q & 255u);
        h01 = (float)((q >> 8) & 255u);
        h10 = (float)((q >> 16) & 255u);
        h11 = (float)(q >> 24);
    } else {
        uint2 q = ((const uint2 *)m.hq)[t];
        h00 = (float)(q.x & 65535u);
        h01 = (float)(q.x >> 16);
        h10 = (float)(q.y & 65535u);
        h11 = (float)(q.y >> 16);
    }
    float wx0 = cx - x, wx1 = x - fx;
    float x_interpolated1 = (wx0 * h00) + (wx1 * h01);
    float x_interpolated2 = (wx0 * h10) + (wx1 * h11);
    return ((cy - y) * x_interpolated1) + ((y - fy) * x_interpolated2);
}

/* png_color_filtered, fut/render_functions.fut:95-105. */
template <bool CP2>
__device__ __forceinline__ uint32_t color_bilinear(const fsk_map &m, float x, float y)
{
    float fx = floorf(x), cx = ceilf(x), fy = floorf(y), cy = ceilf(y);
    int32_t ix = wrap<CP2>(idx_of(fx), m.ccols);
    int32_t iy = wrap<CP2>(idx_of(fy), m.crows);
    uint4 q = ((const uint4 *)m.cq)[(size_t)iy * (size_t)m.ccols + (size_t)ix];
    float wx0 = cx - x, wx1 = x - fx;
    uint32_t x_interpolated1 = argb_mix(wx0, q.x, wx1, q.y);
    uint32_t x_interpolated2 = argb_mix(wx0, q.z, wx1, q.w);
    return argb_mix(cy - y, x_interpolated1, y - fy, x_interpolated2);
}

/* ------------------------------------------------------------------ */
/* Slice table: get_zs terms (fut/voxel_renderer.fut:34), get_h_line   */
/* (:43-60) and inv_z (:217), one thread per slice                     */
/* ------------------------------------------------------------------ */

__global__ void k_slices(fsk_frame f, float4 *__restrict__ lines, float *__restrict__ invz)
{
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= f.n)
        return;
    float i = (float)(k + 1);
    float z = (i / 2.0f) * ((2.0f * f.z0) + ((i - 1.0f) * f.d));
    float left_x = f.Lx * z, left_y = f.Ly * z;
    float right_x = f.Rx * z, right_y = f.Ry * z;
    float dx = (right_x - left_x) / f.fw;
    float dy = (right_y - left_y) / f.fw;
    lines[k] = make_float4(left_x + f.cam_x, left_y + f.cam_y, dx, dy);
    invz[k] = (1.0f / z) * f.scale;
}

/* ------------------------------------------------------------------ */
/* The render kernel                                                   */
/* ------------------------------------------------------------------ */

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kTC = 4;                  /* columns per workgroup */
constexpr int kSG = kThreads / kTC;     /* slice groups per workgroup */
constexpr int kKR = 8;                  /* slices per lane per chunk */
constexpr int kChunk = kSG * kKR;       /* slices per chunk */
constexpr int kMaxRows = 4096;          /* queue entries keep the row in 16 bits */

/* blockIdx -> column tile, keeping adjacent tiles on one XCD (workgroups
 * are dispatched round-robin over the 8 XCDs) so neighbouring stripes share
 * an L2 for height texels and for merging the 16-byte row writes. */
__device__ __forceinline__ int32_t xcd_tile(int32_t b, int32_t n)
{
    const int32_t X = 8;
    int32_t q = n / X, r = n % X, x = b % X, s = b / X;
    return (x < r) ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

struct RenderArgs {
    fsk_map m;
    const float4 *lines;
    const float *invz;
    uint32_t *out;
    int64_t pitch;
    int32_t n, rows, cols, ntiles;
    float cam_h, horizon;
    uint32_t sky;
};

template <int HBITS, bool HP2, bool CP2>
__global__ __launch_bounds__(kThreads) void k_render(RenderArgs a)
{
    extern __shared__ uint32_t lds[];
    uint32_t *stripe = lds;                                  /* [rows][kTC] */
    uint32_t *queue = lds + (size_t)a.rows * kTC;            /* [rows*kTC]  */
    __shared__ int32_t wtot[2][kWaves][kTC];
    __shared__ uint32_t wfill[kWaves][kTC];
    __shared__ uint32_t qn;

    const int32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t tile = xcd_tile(blockIdx.x, a.ntiles);
    const int32_t col0 = tile * kTC;
    const int32_t c = tid % kTC, g = tid / kTC;
    const int32_t col = col0 + c;
    const bool colok = col < a.cols;
    const float fcol = (float)col;
    const int32_t H = a.rows;

    for (int32_t i = tid; i < H * kTC; i += kThreads)
        stripe[i] = 0u;
    if (tid == 0)
        qn = 0u;
    __syncthreads();

    int32_t carry[kTC];
#pragma unroll
    for (int j = 0; j < kTC; j++)
        carry[j] = H;

    int32_t it = 0;
    for (int32_t base = 0; base < a.n; base += kChunk, it ^= 1) {
        /* phase 1: screen rows of this lane's KR slices */
        int32_t ah[kKR];
        int32_t lmin = H;
#pragma unroll
        for (int j = 0; j < kKR; j++) {
            int32_t k = base + g * kKR + j;
            int32_t r = H;
            if (colok && k < a.n) {
                float4 L = a.lines[k];
                float iz = a.invz[k];
                float x = L.x + (fcol * L.z);
                float y = L.y + (fcol * L.w);
                float hgt = height_bilinear<HBITS, HP2>(a.m, x, y);
                float rel = ((a.cam_h - hgt) * iz) + a.horizon;
                r = min(screen_row(rel), H);
            }
            ah[j] = r;
            lmin = min(lmin, r);
        }
        /* phase 2: exclusive prefix-min over slice groups of my column */
        int32_t v = lmin;
#pragma unroll
        for (int off = kTC; off < 64; off <<= 1) {
            int32_t o = __shfl_up(v, off);
            if (lane >= off)
                v = min(v, o);
        }
        int32_t excl = __shfl_up(v, kTC);
        if (lane < kTC)
            excl = H;
        if (lane >= 64 - kTC)
            wtot[it][wave][c] = v;
        __syncthreads();
        int32_t run = carry[c];
        for (int w2 = 0; w2 < wave; w2++)
            run = min(run, wtot[it][w2][c]);
        run = min(run, excl);
        /* phase 3: samples strictly above the y-buffer are visible */
#pragma unroll
        for (int j = 0; j < kKR; j++) {
            bool vis = ah[j] < run;
            if (vis)
                run = ah[j];
            unsigned long long mask = __ballot(vis);
            if (mask) {
                uint32_t cnt = (uint32_t)__popcll(mask);
                uint32_t b0 = 0;
                if (lane == 0)
                    b0 = atomicAdd(&qn, cnt);
                b0 = __shfl(b0, 0);
                if (vis) {
                    uint32_t off = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
                    int32_t k = base + g * kKR + j;
                    queue[b0 + off] = ((uint32_t)c << 16) | (uint32_t)ah[j];
                    stripe[ah[j] * kTC + c] = (uint32_t)k + 1u;
                }
            }
        }
        bool done = true;
#pragma unroll
        for (int j = 0; j < kTC; j++) {
            for (int w2 = 0; w2 < kWaves; w2++)
                carry[j] = min(carry[j], wtot[it][w2][j]);
            done = done && (carry[j] == 0);
        }
        if (done)
            break;
    }
    __syncthreads();

    /* phase 4: colour of every visible sample */
    const uint32_t nq = qn;
    for (uint32_t i = tid; i < nq; i += kThreads) {
        uint32_t e = queue[i];
        int32_t cc = (int32_t)(e >> 16), row = (int32_t)(e & 0xFFFFu);
        int32_t k = (int32_t)stripe[row * kTC + cc] - 1;
        float4 L = a.lines[k];
        float fc = (float)(col0 + cc);
        float x = L.x + (fc * L.z);
        float y = L.y + (fc * L.w);
        stripe[row * kTC + cc] = color_bilinear<CP2>(a.m, x, y);
    }
    __syncthreads();

    /* phase 5: top-down fill (last non-zero at or above each row), sky,
     * and the row-major write-out */
    const int32_t per = (H + kThreads - 1) / kThreads;
    const int32_t r0 = min(tid * per, H), r1 = min(r0 + per, H);
    uint32_t last[kTC];
#pragma unroll
    for (int j = 0; j < kTC; j++)
        last[j] = 0u;
    for (int32_t r = r0; r < r1; r++) {
#pragma unroll
        for (int j = 0; j < kTC; j++) {
            uint32_t s = stripe[r * kTC + j];
            last[j] = s ? s : last[j];
        }
    }
    uint32_t pre[kTC];
#pragma unroll
    for (int j = 0; j < kTC; j++) {
        uint32_t v = last[j];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t o = __shfl_up(v, off);
            if (lane >= off)
                v = v ? v : o;
        }
        uint32_t ex = __shfl_up(v, 1);
        if (lane == 0)
            ex = 0u;
        if (lane == 63)
            wfill[wave][j] = v;
        pre[j] = ex;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kTC; j++) {
        uint32_t p = 0u;
        for (int w2 = 0; w2 < wave; w2++)
            p = wfill[w2][j] ? wfill[w2][j] : p;
        pre[j] = pre[j] ? pre[j] : p;
    }
    const bool vec = (col0 + kTC <= a.cols) && ((a.pitch & 3) == 0) &&
                     ((((uintptr_t)a.out) & 15u) == 0);
    for (int32_t r = r0; r < r1; r++) {
        uint32_t px[kTC];
#pragma unroll
        for (int j = 0; j < kTC; j++) {
            uint32_t s = stripe[r * kTC + j];
            pre[j] = s ? s : pre[j];
            px[j] = pre[j] ? pre[j] : a.sky;
        }
        uint32_t *dst = a.out + (size_t)r * (size_t)a.pitch + col0;
        if (vec) {
            *(uint4 *)dst = make_uint4(px[0], px[1], px[2], px[3]);
        } else {
#pragma unroll
            for (int j = 0; j < kTC; j++)
                if (col0 + j < a.cols)
                    dst[j] = px[j];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Map preparation kernels                                             */
/* ------------------------------------------------------------------ */

/* Ingest: `elem & mask` (fut/interactive.fut:189) and the bilinear quad
 * of each texel (right and lower neighbours wrap, as floor-mod does). */
template <int HBITS>
__global__ void k_height_quads(const int32_t *__restrict__ alt, int32_t rows, int32_t cols,
                               void *__restrict__ hq, uint16_t *__restrict__ hplain)
{
    int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t y = blockIdx.y;
    if (x >= cols)
        return;
    const uint32_t mk = HBITS == 8 ? 0xFFu : 0xFFFFu;
    int32_t x1 = x + 1 == cols ? 0 : x + 1;
    int32_t y1 = y + 1 == rows ? 0 : y + 1;
    uint32_t h00 = (uint32_t)alt[(size_t)y * cols + x] & mk;
    uint32_t h01 = (uint32_t)alt[(size_t)y * cols + x1] & mk;
    uint32_t h10 = (uint32_t)alt[(size_t)y1 * cols + x] & mk;
    uint32_t h11 = (uint32_t)alt[(size_t)y1 * cols + x1] & mk;
    size_t t = (size_t)y * cols + x;
    if (HBITS == 8)
        ((uint32_t *)hq)[t] = h00 | (h01 << 8) | (h10 << 16) | (h11 << 24);
    else
        ((uint2 *)hq)[t] = make_uint2(h00 | (h01 << 16), h10 | (h11 << 16));
    hplain[t] = (uint16_t)h00;
}

__global__ void k_color_quads(const uint32_t *__restrict__ c, int32_t rows, int32_t cols,
                              uint4 *__restrict__ cq)
{
    int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t y = blockIdx.y;
    if (x >= cols)
        return;
    int32_t x1 = x + 1 == cols ? 0 : x + 1;
    int32_t y1 = y + 1 == rows ? 0 : y + 1;
    cq[(size_t)y * cols + x] = make_uint4(c[(size_t)y * cols + x], c[(size_t)y * cols + x1],
                                          c[(size_t)y1 * cols + x], c[(size_t)y1 * cols + x1]);
}

/* generate_shadowmap_accumulated, fut/effects.fut:108-125: for each
 * texel of the size^2 grid count the 255 sun-ray steps that pass below
 * the terrain by more than half a unit (nearest `png_height`,
 * fut/render_functions.fut:63-64), then darken with argb.mix. */
__global__ void k_shade(const uint16_t *__restrict__ hp, int32_t hrows, int32_t hcols,
                        const uint32_t *__restrict__ col, int32_t crows, int32_t ccols,
                        float s0, float s1, float s2, int32_t size, uint32_t *__restrict__ out)
{
    int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t y = blockIdx.y;
    if (x >= size)
        return;
    const float step_size = 4.0f;
    float fx = (float)x, fy = (float)y;
    float height = (float)hp[(size_t)fmod_i32(y, hrows) * hcols + fmod_i32(x, hcols)];
    int32_t count = 0;
    for (int32_t dist = 1; dist < 256; dist++) {
        float t = (float)dist * step_size;
        float px = fx + (t * s0);
        float py = fy + (t * s2);
        int32_t ix = fmod_i32(i32_f32_x86(px), hcols);
        int32_t iy = fmod_i32(i32_f32_x86(py), hrows);
        float probe = (float)hp[(size_t)iy * hcols + ix];
        count += (((height + (t * s1)) - probe) < -0.5f) ? 1 : 0;
    }
    float amount = (float)count;
    uint32_t c = col[(size_t)fmod_i32(y, crows) * ccols + fmod_i32(x, ccols)];
    out[(size_t)y * size + x] = argb_mix(step_size * amount, 0xFF000000u, 1.0f, c);
}

template <int HBITS, bool HP2, bool CP2>
hipError_t launch_render(const RenderArgs &a, size_t lds, hipStream_t s)
{
    auto kern = k_render<HBITS, HP2, CP2>;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)kern,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(2 * kMaxRows * kTC * sizeof(uint32_t)));
        if (e != hipSuccess)
            return e;
        attr_set = true;
    }
    hipLaunchKernelGGL(kern, dim3(a.ntiles), dim3(kThreads), lds, s, a);
    return hipGetLastError();
}

inline bool pow2(int32_t v)
{
    return v > 0 && (v & (v - 1)) == 0;
}

}  // namespace

extern "C" {

size_t fsk_render_lds_bytes(int32_t rows)
{
    return (size_t)2 * (size_t)rows * kTC * sizeof(uint32_t);
}

int32_t fsk_render_max_rows(void)
{
    return kMaxRows;
}

int fsk_slices(const fsk_frame *f, float *d_tab, hipStream_t s)
{
    if (f->n <= 0)
        return hipSuccess;
    float4 *lines = (float4 *)d_tab;
    float *invz = d_tab + 4 * (size_t)f->n;
    hipLaunchKernelGGL(k_slices, dim3((f->n + 255) / 256), dim3(256), 0, s, *f, lines, invz);
    return hipGetLastError();
}

int fsk_render(const fsk_map *m, const fsk_frame *f, const float *d_tab,
               uint32_t *d_out, int64_t pitch, hipStream_t s)
{
    if (f->rows <= 0 || f->cols <= 0)
        return hipSuccess;
    if (f->rows > kMaxRows)
        return hipErrorInvalidValue;
    RenderArgs a;
    a.m = *m;
    a.lines = (const float4 *)d_tab;
    a.invz = d_tab + 4 * (size_t)(f->n > 0 ? f->n : 0);
    a.out = d_out;
    a.pitch = pitch;
    a.n = f->n > 0 ? f->n : 0;
    a.rows = f->rows;
    a.cols = f->cols;
    a.ntiles = (f->cols + kTC - 1) / kTC;
    a.cam_h = f->cam_h;
    a.horizon = f->horizon;
    a.sky = f->sky;
    size_t lds = fsk_render_lds_bytes(f->rows);
    bool hp2 = pow2(m->hrows) && pow2(m->hcols);
    bool cp2 = pow2(m->crows) && pow2(m->ccols);
    hipError_t e;
    if (m->hbits == 8) {
        if (hp2 && cp2) e = launch_render<8, true, true>(a, lds, s);
        else if (hp2) e = launch_render<8, true, false>(a, lds, s);
        else if (cp2) e = launch_render<8, false, true>(a, lds, s);
        else e = launch_render<8, false, false>(a, lds, s);
    } else {
        if (hp2 && cp2) e = launch_render<16, true, true>(a, lds, s);
        else if (hp2) e = launch_render<16, true, false>(a, lds, s);
        else if (cp2) e = launch_render<16, false, true>(a, lds, s);
        else e = launch_render<16, false, false>(a, lds, s);
    }
    return e;
}

int fsk_height_quads(const int32_t *d_alt, int32_t rows, int32_t cols, int32_t hbits,
                     void *d_hq, uint16_t *d_hplain, hipStream_t s)
{
    dim3 grid((cols + 255) / 256, rows);
    if (hbits == 8)
        hipLaunchKernelGGL(k_height_quads<8>, grid, dim3(256), 0, s, d_alt, rows, cols, d_hq, d_hplain);
    else
        hipLaunchKernelGGL(k_height_quads<16>, grid, dim3(256), 0, s, d_alt, rows, cols, d_hq, d_hplain);
    return hipGetLastError();
}

int fsk_shade(const uint16_t *d_hplain, int32_t hrows, int32_t hcols,
              const uint32_t *d_color, int32_t crows, int32_t ccols,
              float sun0, float sun1, float sun2, int32_t size,
              uint32_t *d_shaded, hipStream_t s)
{
    dim3 grid((size + 255) / 256, size);
    hipLaunchKernelGGL(k_shade, grid, dim3(256), 0, s, d_hplain, hrows, hcols, d_color, crows,
                       ccols, sun0, sun1, sun2, size, d_shaded);
    return hipGetLastError();
}

int fsk_color_quads(const uint32_t *d_col, int32_t rows, int32_t cols, void *d_cq, hipStream_t s)
{
    dim3 grid((cols + 255) / 256, rows);
    hipLaunchKernelGGL(k_color_quads, grid, dim3(256), 0, s, d_col, rows, cols, (uint4 *)d_cq);
    return hipGetLastError();
}

}  // extern "C"
