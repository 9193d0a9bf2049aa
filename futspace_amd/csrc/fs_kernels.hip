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
#include <stdlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>

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

/* min(H, screen_row(rel)) with H = rows as float: truncating the median of
 * (rel, 0, H) equals clamping the truncation for every rel below 2^31;
 * rel >= 2^31 and NaN take the select's 0. */
__device__ __forceinline__ int32_t screen_row_clamped(float rel, float Hf)
{
    int32_t r = (int32_t)__builtin_amdgcn_fmed3f(rel, 0.0f, Hf);
    return (rel < 2147483648.0f) ? r : 0;
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
    /* the clamp is a median of three; a NaN v is caught by the final
     * select, so whatever the median does with it is discarded */
    float c = __builtin_amdgcn_fmed3f(v, 0.0f, 1.0f);
    uint32_t u = (uint32_t)(c * 255.0f);
    return (v == v) ? u : 0u;
}

/* Correctly rounded sqrt for t in [0, 1] or NaN: v_sqrt_f32 is within one
 * ulp, and the two FMA residuals pick the rounded root among its
 * neighbours.  Inputs are never negative or infinite here, which drops the
 * general sequence's special cases; inputs below 2^-96 (only reachable
 * within ~2^-90 of the map origin) take the general sqrt on a wave-uniform
 * branch, which skips it everywhere else.  The device self-test compares
 * this with the general sqrt on every f32 in [0, 1]. */
__device__ __forceinline__ float sqrt_unit(float t)
{
    float s = __builtin_amdgcn_sqrtf(t);
    float sm = __int_as_float(__float_as_int(s) - 1);
    float sp = __int_as_float(__float_as_int(s) + 1);
    float rm = __builtin_fmaf(-sm, s, t);
    float rp = __builtin_fmaf(-sp, s, t);
    s = (rm <= 0.0f) ? sm : s;
    s = (rp > 0.0f) ? sp : s;
    bool tiny = (t < 0x1p-96f) && (t > 0.0f);
    if (__ballot(tiny)) {
        /* the empty asm keeps the general sqrt inside the branch */
        float tt = t;
        asm volatile("" : "+v"(tt));
        if (tiny)
            s = __builtin_sqrtf(tt);
    }
    return s;
}

/* Correctly rounded sqrt of t >= 0 (or NaN): sqrt_unit's neighbour
 * correction holds for every non-negative f32; device self-test 7 checks
 * it against the correctly rounded sqrt on all of them. */
__device__ __forceinline__ float sqrt_rn(float t)
{
    return sqrt_unit(t);
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
/* The maps are device memory.  Their gathers read through pointers of the
 * global address space: a map pointer the kernel read from memory (a
 * batch's view records) is generic, and loads through it would be flat
 * loads, which also count against the LDS wait counter, so every LDS wait
 * of the tile would wait for the gathers in flight too. */
#define FS_GLOBAL __attribute__((address_space(1)))
typedef uint32_t fs_u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t fs_u32x4 __attribute__((ext_vector_type(4)));

template <int HBITS, bool HP2, bool TAME = false>
__device__ __forceinline__ uint2 height_quad(const fsk_map &m, float x, float y)
{
    /* TAME: |x|, |y| < 2^22, so floor converts exactly without a clamp */
    int32_t ix = wrap<HP2>(TAME ? (int32_t)floorf(x) : idx_of(floorf(x)), m.hcols);
    int32_t iy = wrap<HP2>(TAME ? (int32_t)floorf(y) : idx_of(floorf(y)), m.hrows);
    /* maps are at most 2^28 texels (host-checked), so 32-bit offsets; the
     * wrapped index is always inside the map, whatever x and y are */
    uint32_t t = HP2 ? ((uint32_t)iy << __builtin_ctz((uint32_t)m.hcols)) + (uint32_t)ix
                     : (uint32_t)iy * (uint32_t)m.hcols + (uint32_t)ix;
    const FS_GLOBAL char *hq = (const FS_GLOBAL char *)m.hq;
    if (HBITS == 8)
        return make_uint2(*(const FS_GLOBAL uint32_t *)(hq + t * 4u), 0u);
    const fs_u32x2 v = *(const FS_GLOBAL fs_u32x2 *)(hq + t * 8u);
    return make_uint2(v.x, v.y);
}

template <int HBITS>
__device__ __forceinline__ float height_interp(uint2 q, float x, float y)
{
    float fx = floorf(x), cx = ceilf(x), fy = floorf(y), cy = ceilf(y);
    float h00, h01, h10, h11;
    if (HBITS == 8) {
        h00 = (float)(q.x & 255u);
        h01 = (float)((q.x >> 8) & 255u);
        h10 = (float)((q.x >> 16) & 255u);
        h11 = (float)(q.x >> 24);
    } else {
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

/* ------------------------------------------------------------------ */
/* Slice parameters: get_zs term (fut/voxel_renderer.fut:34), get_h_line */
/* (:43-60) and inv_z (:217) of slice k (0-based)                      */
/* ------------------------------------------------------------------ */

struct Slice {
    float sx, sy, dx, dy, iz;
};

__device__ __forceinline__ Slice slice_params(const fsk_frame &f, int32_t k)
{
    float i = (float)(k + 1);
    float z = (i / 2.0f) * ((2.0f * f.z0) + ((i - 1.0f) * f.d));
    float left_x = f.Lx * z, left_y = f.Ly * z;
    float right_x = f.Rx * z, right_y = f.Ry * z;
    Slice s;
    s.dx = (right_x - left_x) / f.fw;
    s.dy = (right_y - left_y) / f.fw;
    s.sx = left_x + f.cam_x;
    s.sy = left_y + f.cam_y;
    s.iz = (1.0f / z) * f.scale;
    return s;
}

/* RN(1 / z) for z > 0 whose reciprocal is a normal number: the hardware
 * reciprocal (within an ulp) and one Newton step with FMA.  Device self-test
 * 5 compares it with the correctly rounded division for every such z. */
__device__ __forceinline__ float rcp_rn(float z)
{
    float y = __builtin_amdgcn_rcpf(z);
    float e = __builtin_fmaf(-z, y, 1.0f);
    return __builtin_fmaf(e, y, y);
}

/* RN(a / b) from rb = RN(1 / b), for finite a and b >= 1 (Markstein's
 * correction: q = RN(a * rb) is within an ulp of a / b, the residual
 * a - q * b is exact in one FMA, and q + residual * rb rounds to the
 * nearest quotient).  The residual is not exact for quotients near the
 * subnormal range, and the signs of zero differ, so quotients below 2^-100
 * in magnitude take the division itself, on a wave-uniform branch (a plain
 * select would have every lane compute it).  Device self-test 6 compares it
 * with the correctly rounded division for every finite f32 a at a set of
 * frame widths b. */
__device__ __forceinline__ float div_rn(float a, float b, float rb)
{
    float q = a * rb;
    float r = __builtin_fmaf(-q, b, a);
    q = __builtin_fmaf(r, rb, q);
    const bool small = !(fabsf(q) >= 0x1p-100f);
    if (__ballot(small)) {
        /* the empty asm keeps the compiler from hoisting the division out
         * of the branch (it may otherwise evaluate it on every wave) */
        float aa = a;
        asm volatile("" : "+v"(aa));
        if (small)
            q = aa / b;
    }
    return q;
}

/* slice_params of a tame frame (z >= z0 > 0, finite scalars) with the two
 * divisions by the frame width and the reciprocal of z taken by rcp_rn and
 * div_rn (f.rfw = RN(1 / fw)).  Bit-identical to slice_params there. */
__device__ __forceinline__ Slice slice_params_tame(const fsk_frame &f, int32_t k)
{
    const float rfw = f.rfw;
    float i = (float)(k + 1);
    float z = (i / 2.0f) * ((2.0f * f.z0) + ((i - 1.0f) * f.d));
    float left_x = f.Lx * z, left_y = f.Ly * z;
    float right_x = f.Rx * z, right_y = f.Ry * z;
    Slice s;
    s.dx = div_rn(right_x - left_x, f.fw, rfw);
    s.dy = div_rn(right_y - left_y, f.fw, rfw);
    s.sx = left_x + f.cam_x;
    s.sy = left_y + f.cam_y;
    s.iz = rcp_rn(z) * f.scale;
    return s;
}

/* ------------------------------------------------------------------ */
/* The render kernel                                                   */
/* ------------------------------------------------------------------ */

/* Inclusive prefix sum over the 64 lanes of a wave, in registers: DPP
 * row shifts by 1, 2, 4 and 8 scan each row of 16 lanes (lanes shifted in
 * from outside the row read 0), then the last lane of row 0 is added to
 * rows 1 and 3 (row_bcast:15; rows 0 and 2 add the old value 0) and lane
 * 31 to rows 2 and 3 (row_bcast:31). */
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);  /* row_shr:1 */
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);  /* row_shr:2 */
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);  /* row_shr:4 */
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);  /* row_shr:8 */
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false); /* row_bcast:15 */
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false); /* row_bcast:31 */
    return x;
}

/* Per-column minima over the slice groups of a wave of an 8-column tile
 * (lane = 8 * group + column, rows at most H): excl = the minimum over
 * the groups before the lane's (H for the first group), total = over all
 * eight.  In registers: DPP folds the two groups of each 16-lane row
 * (row_shr:8, then row_ror:8 to give the row's minimum to all its lanes),
 * and the permlane swaps hand every row the minima of all four rows with
 * the columns aligned (permlane32_swap exchanges rows 2-3 of its first
 * operand with rows 0-1 of its second; permlane16_swap the odd rows of the
 * first with the even rows of the second). */
__device__ __forceinline__ void wave_column_scan8(int32_t x, int32_t H, int32_t lane,
                                                  int32_t &excl, int32_t &total)
{
    /* the even group of the row in the odd group's lanes (row_shr:8; the
     * even group's lanes keep H) */
    const int32_t e = __builtin_amdgcn_update_dpp(H, x, 0x118, 0xf, 0xf, false);
    const int32_t i = min(x, e);
    const int32_t t = min(i, __builtin_amdgcn_update_dpp(H, i, 0x128, 0xf, 0xf, false));
    const auto y = __builtin_amdgcn_permlane32_swap(t, t, false, false);   /* T0 T1 T0 T1 | T2 T3 T2 T3 */
    const auto a = __builtin_amdgcn_permlane16_swap(y[0], y[0], false, false);   /* T0 | T1 */
    const auto b = __builtin_amdgcn_permlane16_swap(y[1], y[1], false, false);   /* T2 | T3 */
    const int32_t t0 = (int32_t)a[0], m01 = min(t0, (int32_t)a[1]);
    const int32_t m012 = min(m01, (int32_t)b[0]);
    total = min(m012, (int32_t)b[1]);
    const int32_t row = lane >> 4;
    const int32_t before = row == 0 ? H : row == 1 ? t0 : row == 2 ? m01 : m012;
    excl = (lane & 8) ? min(before, e) : before;
}

constexpr int kMaxWaves = 16;                    /* 1024-thread workgroups */
constexpr int kChunk = 2048;                     /* slices per chunk, at most */

/* Slice-table slots for chunks of up to `chunk` slices: the table is slot
 * major, slot j * SG + g holding slice g * kr + j of the chunk (slice group
 * g, its j-th slice), so that the slice groups of a wave read adjacent
 * records at an offset fixed per j.  Chunks are whole multiples of SG * 4
 * slots (the host rounds them), which covers the gather groups' reads past
 * kr. */
__host__ __device__ constexpr int tab_stride(int chunk)
{
    return chunk;
}
/* A queued visible sample: its slice group g, its slice j within the
 * group, its column c in the tile and its projected row, packed as
 * g << 22 | j << 17 | c << 13 | row (g < 256, j < 32, c < 16, row < 8192).
 * The long-span list packs c << 26 | first row << 13 | end row. */
constexpr uint32_t kQRowBits = 13;
constexpr uint32_t kQRowMask = (1u << kQRowBits) - 1u;
constexpr uint32_t kQColShift = 13, kQSliceShift = 17, kQGroupShift = 22;
constexpr uint32_t kQLaneMask = (0xFFu << kQGroupShift) | (15u << kQColShift);  /* g and c */
/* Frames up to kMaxRows rows: the row fields allow 8191, and a 4-column
 * tile of that height still fits in LDS. */
constexpr int kMaxRows = 7680;
constexpr int kChunkAlign = 4;                   /* chunks are multiples of SG * kChunkAlign */

/* blockIdx -> column tile.  Workgroups are dispatched round-robin over
 * the 8 XCDs (workgroup i runs on XCD i mod 8), so tile ranges are dealt
 * out per XCD: neighbouring stripes share one L2 for their height texels
 * and for merging their partial-line row writes. */
__device__ __forceinline__ int32_t xcd_tile(int32_t b, int32_t n)
{
    const int32_t X = 8;
    int32_t q = n / X, r = n % X, x = b % X, s = b / X;
    return (x < r) ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

struct RenderArgs {
    fsk_map m;
    fsk_frame f;
    uint32_t *out;
    int64_t pitch;
    int32_t ntiles;
    int32_t chunk_max;     /* slices per chunk, at most kChunk */
    /* slice mode: workgroups ntiles .. ntiles + nzero - 1 clear zero[zero_n4]
     * (the fused step's gradient) on CUs the tiles leave idle */
    float4 *zero;
    size_t zero_n4;
    int32_t nzero;
};

/* Fewest idle workgroup slots worth a gradient clear inside the slice
 * render: with fewer, each would clear too much of the buffer and
 * outlast the tiles. */
constexpr int32_t kMinZeroBlocks = 64;

#ifndef FS_SQ_LUT
#define FS_SQ_LUT 1
#endif

/* RN(RN(b/255)^2) for every byte: the squared channel of to_rgba. */
__device__ __forceinline__ void fill_square_lut(float *lut, int32_t tid)
{
    if (tid < 256) {
        float u = unit8((uint32_t)tid);
        lut[tid] = u * u;
    }
}

__device__ __forceinline__ float sq8(const float *lut, uint32_t b)
{
#if FS_SQ_LUT
    return lut[b];
#else
    (void)lut;
    float u = unit8(b);
    return u * u;
#endif
}

/* One channel of argb.mix with the weights already normalised:
 * u32.f32(clamp(sqrt(m1' * c1^2 + m2' * c2^2)) * 255). */
__device__ __forceinline__ uint32_t mix_channel(float m1n, float s1, float m2n, float s2)
{
    return chan_u8(sqrt_unit((m1n * s1) + (m2n * s2)));
}

/* The common case of png_color_filtered: |x|, |y| >= 1 and four opaque
 * texels.  Then the weights (c - x), (x - f) are exact and sum to exactly
 * 1 (or are both 0 at an integer coordinate), so every normalisation is
 * the identity, every sqrt argument lies in {0} U [2^-39, 1] (weights are
 * multiples of 2^-23, squares at least 255^-2), the clamp never acts and
 * no NaN arises; alpha is 255 unless a weight sum is 0. */
__device__ __forceinline__ uint32_t mix_channel_exact(float m1, float s1, float m2, float s2)
{
    float t = (m1 * s1) + (m2 * s2);
    float s = __builtin_amdgcn_sqrtf(t);
    float sm = __int_as_float(__float_as_int(s) - 1);
    float sp = __int_as_float(__float_as_int(s) + 1);
    float rm = __builtin_fmaf(-sm, s, t);
    float rp = __builtin_fmaf(-sp, s, t);
    s = (rm <= 0.0f) ? sm : s;
    s = (rp > 0.0f) ? sp : s;
    return (uint32_t)(s * 255.0f);
}

/* The byte quantisation of a mixed channel, u32.f32(sqrt(t) * 255), by
 * table: the map t -> byte is non-decreasing on [0, 1], so the float bits
 * of t (clamped below at kQLo, where the byte is still 0) fall into one of
 * kQBuckets buckets of 2^16 consecutive bit patterns, and no bucket holds
 * more than one step of the byte.  Entry i of the table, with
 * u = max(bits(t) - kQLo, 0) in bucket i = u >> 16, is
 *     (base << 17) + 2^17 - step - (i << 16)   (mod 2^32)
 * (base: the byte at the bucket's first pattern; step: the offset of the
 * pattern where the byte becomes base + 1, or 2^16 if none), so that
 * (entry + u) >> 17 is the byte for every t of the bucket.  The table is
 * built on the host with the C library's correctly rounded sqrtf and is
 * checked against the sqrt sequence for every f32 in [0, 1] by the device
 * self-test. */
/* Scratch of the allocation-free fused step (fsx_render_f32_l2_step_noalloc):
 * per-workgroup partial sums, kScratchFloats floats per device, held until
 * the process ends.  Host code reaches it only through the address
 * hipGetSymbolAddress returns for the current device (fsk_scratch). */
constexpr int kScratchFloats = 32768;
static __device__ __align__(256) float scratch[kScratchFloats];

constexpr uint32_t kQLo = 0x37810000u;           /* below sqrt(t) * 255 >= 1 */
constexpr int kQBuckets = 2048;                  /* (bits(1.0f) - kQLo) >> 16 = 2047 */
__device__ uint32_t g_qtab[kQBuckets];

/* (qt entry + u) for t = m1 * s1 + m2 * s2: the channel byte sits in bits
 * 17..24.  Needs t in [0, 1]. */
__device__ __forceinline__ uint32_t mix_q(const uint32_t *qt, float m1, float s1, float m2,
                                          float s2)
{
    float t = (m1 * s1) + (m2 * s2);
    uint32_t u = __builtin_elementwise_sub_sat(__float_as_uint(t), kQLo);
    return qt[u >> 16] + u;
}

/* color_fast with the table quantisation (qt: the bucket table in LDS). */
__device__ __forceinline__ uint32_t color_fast_q(uint4 q, const float *sq, const uint32_t *qt,
                                                 float wx0, float wx1, float wy0, float wy1)
{
    uint32_t out = 0;
#pragma unroll
    for (int sh = 16; sh >= 0; sh -= 8) {
        uint32_t X1 = mix_q(qt, wx0, sq8(sq, (q.x >> sh) & 255u), wx1,
                            sq8(sq, (q.y >> sh) & 255u)) >> 17;
        uint32_t X2 = mix_q(qt, wx0, sq8(sq, (q.z >> sh) & 255u), wx1,
                            sq8(sq, (q.w >> sh) & 255u)) >> 17;
        out |= (mix_q(qt, wy0, sq8(sq, X1), wy1, sq8(sq, X2)) >> 17) << sh;
    }
    bool opaque_sum = (wx0 + wx1 != 0.0f) && (wy0 + wy1 != 0.0f);
    return out | (opaque_sum ? 0xFF000000u : 0u);
}

__device__ __forceinline__ uint32_t color_fast(uint4 q, const float *sq, float wx0, float wx1,
                                               float wy0, float wy1)
{
    uint32_t out = 0;
#pragma unroll
    for (int sh = 16; sh >= 0; sh -= 8) {
        uint32_t X1 = mix_channel_exact(wx0, sq8(sq, (q.x >> sh) & 255u), wx1,
                                        sq8(sq, (q.y >> sh) & 255u));
        uint32_t X2 = mix_channel_exact(wx0, sq8(sq, (q.z >> sh) & 255u), wx1,
                                        sq8(sq, (q.w >> sh) & 255u));
        out |= mix_channel_exact(wy0, sq8(sq, X1), wy1, sq8(sq, X2)) << sh;
    }
    bool opaque_sum = (wx0 + wx1 != 0.0f) && (wy0 + wy1 != 0.0f);
    return out | (opaque_sum ? 0xFF000000u : 0u);
}

/* png_color_filtered (fut/render_functions.fut:95-105) evaluated channel
 * by channel: argb.mix is separable per channel once its weights are
 * normalised, so the two row mixes feed the column mix as bytes.
 *
 * argb.mix divides the weights by their sum m12.  A sum of exactly 1
 * makes the divisions exact no-ops; a sum of 0 (integer coordinate) makes
 * the reference's channels 0/0 = NaN, which quantise to byte 0 exactly as
 * the undivided zero weights do.  Bilinear weights sum to 0 or 1 whenever
 * |x|, |y| >= 1, so the divisions run only on waves where some lane has
 * another sum (a coordinate within 1 of the map origin); such waves, and
 * waves with translucent texels, take the general path below. */
template <bool CP2>
__device__ __forceinline__ uint4 color_quad(const fsk_map &m, float x, float y)
{
    int32_t ix = wrap<CP2>(idx_of(floorf(x)), m.ccols);
    int32_t iy = wrap<CP2>(idx_of(floorf(y)), m.crows);
    /* colour maps have at most 65536^2 texels (host-checked), so the quad
     * index fits 32 bits; the byte offset is formed in 64 */
    uint32_t t = CP2 ? ((uint32_t)iy << __builtin_ctz((uint32_t)m.ccols)) + (uint32_t)ix
                     : (uint32_t)iy * (uint32_t)m.ccols + (uint32_t)ix;
    const fs_u32x4 v = ((const FS_GLOBAL fs_u32x4 *)m.cq)[t];
    return make_uint4(v.x, v.y, v.z, v.w);
}

/* The colour of a sample at (x, y) from its texel quad q; with qt (the
 * quantisation table in LDS) the fast path quantises by table. */
template <bool QT = false>
__device__ __forceinline__ uint32_t color_of_quad(uint4 q, const float *sq, float x, float y,
                                                  const uint32_t *qt = nullptr)
{
    float fx = floorf(x), cx = ceilf(x), fy = floorf(y), cy = ceilf(y);
    float wx0 = cx - x, wx1 = x - fx, wy0 = cy - y, wy1 = y - fy;
    bool easy = (fabsf(x) >= 1.0f) && (fabsf(y) >= 1.0f) &&
                (((q.x & q.y & q.z & q.w) >> 24) == 255u);
    if (__ballot(!easy) == 0)
        return QT ? color_fast_q(q, sq, qt, wx0, wx1, wy0, wy1)
                  : color_fast(q, sq, wx0, wx1, wy0, wy1);
    float mx = wx0 + wx1, my = wy0 + wy1;
    bool odd = !((mx == 1.0f || mx == 0.0f) && (my == 1.0f || my == 0.0f));
    float nx0 = wx0, nx1 = wx1, ny0 = wy0, ny1 = wy1;
    if (__ballot(odd)) {
        if (odd) {
            nx0 = wx0 / mx;
            nx1 = wx1 / mx;
            ny0 = wy0 / my;
            ny1 = wy1 / my;
        }
    }
    uint32_t out = 0;
#pragma unroll
    for (int sh = 16; sh >= 0; sh -= 8) {
        uint32_t X1 = mix_channel(nx0, sq8(sq, (q.x >> sh) & 255u), nx1, sq8(sq, (q.y >> sh) & 255u));
        uint32_t X2 = mix_channel(nx0, sq8(sq, (q.z >> sh) & 255u), nx1, sq8(sq, (q.w >> sh) & 255u));
        out |= mix_channel(ny0, sq8(sq, X1), ny1, sq8(sq, X2)) << sh;
    }
    /* alpha: (m1*a1 + m2*a2) / m12, mixed linearly */
    float ax1 = (wx0 * unit8(q.x >> 24)) + (wx1 * unit8(q.y >> 24));
    float ax2 = (wx0 * unit8(q.z >> 24)) + (wx1 * unit8(q.w >> 24));
    if (__ballot(odd)) {
        if (odd) {
            ax1 = ax1 / mx;
            ax2 = ax2 / mx;
        }
    }
    float ay = (wy0 * unit8(chan_u8(ax1))) + (wy1 * unit8(chan_u8(ax2)));
    if (__ballot(odd)) {
        if (odd)
            ay = ay / my;
    }
    return out | (chan_u8(ay) << 24);
}

constexpr int kShortSpan = 40;   /* span rows a lane paints itself */
constexpr int kGather = 4;       /* height gathers a lane keeps in flight */
constexpr int kColourBatch = 1;  /* visible samples a lane colours per pass */

/* Dynamic LDS of a tile render of tc columns by nt threads: slice table
 * (float4 + inv_z per slot), the tc x rows colour stripe, the
 * visible-sample queue (at most min(rows, chunk) samples per column and
 * chunk), the y-buffer in front of each lane's slices (u16 per thread) and
 * the long-span list. */
inline size_t render_lds_bytes(int tc, int nt, int32_t rows, int32_t chunk)
{
    size_t hp = (size_t)((rows + 3) & ~3);
    size_t q = (size_t)tc * (size_t)(rows < chunk ? rows : chunk);
    size_t longs = (size_t)tc * (size_t)(rows / kShortSpan + 1);
    return sizeof(float) * 5 * (size_t)tab_stride(chunk) +
           sizeof(uint32_t) * (hp * tc + q + (size_t)nt / 2 + 2 * longs);
}

/* One workgroup renders a stripe of TC adjacent columns over all slices,
 * chunk by chunk:
 *   0. the chunk's slice table (position, step, inv_z) into LDS;
 *   1. lane (slice group g, column c) projects its kr consecutive slices;
 *   2. prefix-min over slice groups (wave scan, then each lane takes the
 *      minimum of the earlier waves' totals and the chunk's carry-in)
 *      gives the y-buffer in front of each lane's slices;
 *   3. samples strictly above it are visible (scan occlude + scatter,
 *      fut/voxel_renderer.fut:231, :244): they are queued with the bottom
 *      of their span;
 *   4. each queued sample gets its bilinear argb colour and paints its span
 *      down to the next visible row below (the reference's top-down
 *      `scan fill_vline`, :246-249, seen from the sample's side).
 * Finally the stripe is written row by row with 0 replaced by the sky.  A
 * sample whose colour is 0 is transparent for fill_vline; if one occurs the
 * stripe is instead filled with a top-down scan over all rows. */
/* Workgroup-shared scalars of one tile render. */
template <int TC>
struct TileShared {
    float sq[256];                 /* squares table, filled once per workgroup */
    uint32_t qt[kQBuckets];        /* byte quantisation table (g_qtab) */
    int32_t wtot[kMaxWaves][TC];   /* wave totals of the prefix-min */
    int32_t wpre[kMaxWaves][TC];   /* wave prefixes of the transparent fill */
    int32_t carry[TC];
    uint32_t qn, nlong, transparent;
};

/* The colour tables of a tile workgroup: squares and the byte
 * quantisation table, copied from g_qtab in two halves -- load() issues
 * this thread's loads, store() writes them to LDS -- so that the loads'
 * latency passes under other work of the tile's first phase. */
template <int TC, int NT>
struct TileTables {
    static_assert(kQBuckets % NT == 0, "table split");
    uint32_t qv[kQBuckets / NT];
    __device__ __forceinline__ void load(int32_t tid)
    {
#pragma unroll
        for (int k = 0; k < kQBuckets / NT; k++)
            qv[k] = g_qtab[tid + k * NT];
    }
    __device__ __forceinline__ void store(TileShared<TC> &sh, int32_t tid) const
    {
        fill_square_lut(sh.sq, tid);
#pragma unroll
        for (int k = 0; k < kQBuckets / NT; k++)
            sh.qt[tid + k * NT] = qv[k];
    }
};

/* No colour tables (tile modes without the argb colour path). */
struct NoTables {
    template <int TC>
    __device__ __forceinline__ void store(TileShared<TC> &, int32_t) const {}
};

/* What a tile writes per pixel: the u32 colour (the reference frame) or,
 * for the differentiable renderer, the index of the depth slice whose
 * sample shows there (-1 for sky).  The slice-index mode keeps the same
 * visibility and fill rules; a sample is transparent there iff its map
 * coordinate is integral in x or y, which is where argb.mix's weight sum
 * vanishes. */
enum { kOutColour = 0, kOutSlice = 1, kOutNearest = 2 };

/* kOutNearest is the stateless program of tests/futspace.fut:47-145: the
 * map is plain i32 colours (m.cq) and unmasked i32 heights (m.hq), sampled
 * at the truncated coordinate (floor-mod wrapped), and the column starts
 * as sky with sky as the `empty` colour of its fill (:110-121) -- colour 0
 * is an ordinary colour there.  The stripe holds colour ^ sky, so that its
 * 0 (unpainted) is the sky and a sky-coloured sample is the transparent
 * one. */
template <int MODE>
__device__ __forceinline__ uint32_t tile_value(uint32_t v, uint32_t sky)
{
    if (MODE == kOutColour)
        return v ? v : sky;
    if (MODE == kOutNearest)
        return v ^ sky;
    return v ? v - 1u : 0xFFFFFFFFu;
}

/* tile_value of a cell that is either painted with a non-zero value (v) or,
 * above its column's final y-buffer, unpainted (sky). */
template <int MODE>
__device__ __forceinline__ uint32_t painted_value(uint32_t v, bool sky_cell, uint32_t sky)
{
    if (MODE == kOutColour)
        return sky_cell ? sky : v;
    if (MODE == kOutNearest)
        return sky_cell ? sky : v ^ sky;
    return sky_cell ? 0xFFFFFFFFu : v - 1u;
}

/* Texel index of the stateless program's sample at (x, y): i32.f32 of
 * each coordinate (x86 truncation), floor-mod by the map size. */
template <bool P2>
__device__ __forceinline__ uint32_t nearest_index(const fsk_map &m, float x, float y)
{
    int32_t ix = wrap<P2>(i32_f32_x86(x), m.hcols);
    int32_t iy = wrap<P2>(i32_f32_x86(y), m.hrows);
    return (uint32_t)iy * (uint32_t)m.hcols + (uint32_t)ix;
}

/* chunk_max must be a multiple of NT/TC (the slice groups) and at most
 * CMAX; the LDS layout is sized for it. */
template <int TC, int NT, int CMAX, int HBITS, bool HP2, bool CP2, int MODE = kOutColour,
          int GATHER = kGather, int CBATCH = kColourBatch, class Tables = NoTables>
__device__ __forceinline__ void render_tile(const fsk_map &m, const fsk_frame &f, uint32_t *out,
                                            int64_t pitch, int32_t tile, int32_t chunk_max,
                                            uint32_t *lds, TileShared<TC> &sh,
                                            const Tables &tables = Tables())
{
    constexpr int NW = NT / 64;                  /* waves */
    constexpr int SG = NT / TC;                  /* slice groups */
    constexpr int KR = CMAX / SG;                /* slices per lane per chunk, at most */
    static_assert(NW <= kMaxWaves && KR % GATHER == 0 && GATHER % 2 == 0 && KR <= 32 &&
                      SG <= 256 && kChunkAlign % GATHER == 0,
                  "tile shape");
    const int32_t tabs = tab_stride(chunk_max);
    float4 *tl = (float4 *)lds;                                  /* [tabs] */
    float *tz = (float *)(lds + 4 * tabs);                       /* [tabs] */
    const int32_t H = f.rows, n = f.n;
    const int32_t Hp = (H + 3) & ~3;
    uint32_t *stripe = lds + 5 * tabs;                           /* [TC][Hp] */
    const int32_t qcap = TC * min(H, chunk_max);                 /* visible rows differ */
    uint32_t *queue = stripe + (size_t)Hp * TC;                  /* [qcap] */
    uint16_t *lanein = (uint16_t *)(queue + qcap);               /* [NT] y-buffer per lane */
    uint32_t *longs = queue + qcap + NT / 2;                     /* [2][TC*(H/kShortSpan+1)] */

    const int32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t col0 = tile * TC;
    const int32_t c = tid % TC, g = tid / TC;
    const int32_t col = col0 + c;
    const bool colok = col < f.cols;
    const float fcol = (float)col;
    const float Hf = (float)H;

    /* The stripe is not cleared: the spans of a column's visible samples
     * tile the rows from its final y-buffer down to H (transparent samples
     * paint 0), and the rows above it are sky, which the write phase tells
     * apart by the y-buffer. */
    if (tid < TC)
        sh.carry[tid] = H;
    if (tid == 0) {
        sh.nlong = 0u;
        sh.transparent = 0u;
    }

    const int32_t nch = (n + chunk_max - 1) / chunk_max;
    const int32_t kr = nch > 0 ? (n + nch * SG - 1) / (nch * SG) : 0;
    const int32_t csz = kr * SG;

    for (int32_t base = 0; base < n; base += csz) {
        const int32_t cn = min(csz, n - base);
        /* phase 0: the chunk's slice records in slot order (slot j*SG + g
         * holds slice g*kr + j; slots past the chunk repeat its last
         * slice).  Each wave computes the slots of its own slice groups --
         * the only ones its projection reads -- so no workgroup barrier
         * separates this phase from the next: a wave starts projecting as
         * soon as its own records are in LDS.  Computing them per tile is
         * cheaper than copying a precomputed per-view table, even in
         * batches. */
        {
            constexpr int GPW = 64 / TC;             /* slice groups per wave */
            const int32_t nj = ((kr + GATHER - 1) / GATHER) * GATHER;
            auto table = [&](auto tame_tag) {
                constexpr bool TAMEP = decltype(tame_tag)::value;
                for (int32_t q = lane; q < nj * GPW; q += 64) {
                    const int32_t gq = wave * GPW + q % GPW, j = q / GPW;
                    const int32_t o = j * SG + gq;
                    const int32_t s = min(gq * kr + j, cn - 1);
                    Slice sl = TAMEP ? slice_params_tame(f, base + s) : slice_params(f, base + s);
                    tl[o] = make_float4(sl.sx, sl.sy, sl.dx, sl.dy);
                    tz[o] = sl.iz;
                }
            };
            if (f.tame)
                table(std::true_type{});
            else
                table(std::false_type{});
        }
        if (tid == 0)
            sh.qn = 0u;
        /* the wave reads back what its own lanes wrote: LDS operations of a
         * wave complete in order; the fence keeps the compiler from moving
         * the reads above the writes */
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        /* phase 1: rows are < 2^16, two per register.  Branch-free in
         * groups of GATHER samples: all of a group's height gathers are
         * issued before the first is used.  Samples past the chunk, the
         * slice group or the frame read a defined table slot and an in-map
         * texel; the fix-up below raises them to row H. */
        uint32_t ah[KR / 2];
        int32_t lmin = H;
        const float4 *tlg = tl + g;
        const float *tzg = tz + g;
        auto project = [&](auto tame_tag, auto full_tag) {
            constexpr bool TAME = decltype(tame_tag)::value;
            constexpr bool FULL = decltype(full_tag)::value;
#pragma unroll
            for (int j0 = 0; j0 < KR; j0 += GATHER) {
                /* kr is uniform over the workgroup: whole groups past it
                 * are skipped, their rows stay H.  A full chunk (kr == KR)
                 * has no such group, and no branch between the groups, so
                 * the gathers of later groups may be issued early */
                if (!FULL && j0 >= kr) {
#pragma unroll
                    for (int u = 0; u < GATHER; u += 2)
                        ah[(j0 + u) / 2] = (uint32_t)H | ((uint32_t)H << 16);
                    continue;
                }
                float xs[GATHER], ys[GATHER], izs[GATHER];
                uint2 qs[GATHER];
#pragma unroll
                for (int u = 0; u < GATHER; u++) {
                    float4 L = tlg[(j0 + u) * SG];
                    xs[u] = L.x + (fcol * L.z);
                    ys[u] = L.y + (fcol * L.w);
                    izs[u] = tzg[(j0 + u) * SG];
                    if (MODE == kOutNearest)
                        qs[u] = make_uint2(
                            ((const uint32_t *)m.hq)[nearest_index<HP2>(m, xs[u], ys[u])], 0u);
                    else
                        qs[u] = height_quad<HBITS, HP2, TAME>(m, xs[u], ys[u]);
                }
#pragma unroll
                for (int u = 0; u < GATHER; u++) {
                    const int j = j0 + u;
                    float hgt = MODE == kOutNearest ? (float)(int32_t)qs[u].x
                                                    : height_interp<HBITS>(qs[u], xs[u], ys[u]);
                    float rel = ((f.cam_h - hgt) * izs[u]) + f.horizon;
                    /* a tame frame keeps |rel| < 2^30: the cast of the
                     * clamped value is the whole x86 semantics */
                    int32_t r = TAME ? (int32_t)__builtin_amdgcn_fmed3f(rel, 0.0f, Hf)
                                     : screen_row_clamped(rel, Hf);
                    if (j & 1)
                        ah[j / 2] |= (uint32_t)r << 16;
                    else
                        ah[j / 2] = (uint32_t)r;
                    lmin = min(lmin, r);
                }
            }
        };
        if (MODE != kOutNearest && f.tame && kr == KR)
            project(std::true_type{}, std::true_type{});
        else if (MODE != kOutNearest && f.tame)
            project(std::true_type{}, std::false_type{});
        else
            project(std::false_type{}, std::false_type{});
        /* lanes holding samples past the chunk, past kr (a last gather
         * group kr does not fill) or right of the frame: those samples
         * become row H */
        if (!colok || g * kr + kr > cn || (kr % GATHER) != 0) {
            lmin = H;
#pragma unroll
            for (int j = 0; j < KR; j++) {
                const bool ok = j < kr && g * kr + j < cn && colok;
                const uint32_t sh16 = (uint32_t)(j & 1) * 16u;
                uint32_t r = ok ? (ah[j / 2] >> sh16) & 0xFFFFu : (uint32_t)H;
                ah[j / 2] = (ah[j / 2] & ~(0xFFFFu << sh16)) | (r << sh16);
                lmin = min(lmin, (int32_t)r);
            }
        }

        /* phase 2 */
        int32_t excl;
        if constexpr (TC == 8) {
            int32_t total;
            wave_column_scan8(lmin, H, lane, excl, total);
            if (lane >= 64 - TC)
                sh.wtot[wave][c] = total;
        } else {
            int32_t v = lmin;
#pragma unroll
            for (int off = TC; off < 64; off <<= 1) {
                int32_t o = __shfl_up(v, off);
                if (lane >= off)
                    v = min(v, o);
            }
            excl = __shfl_up(v, TC);
            if (lane < TC)
                excl = H;
            if (lane >= 64 - TC)
                sh.wtot[wave][c] = v;
        }
        /* the chunk's carry-in is read before the barrier: the last slice
         * group overwrites it in phase 3 without a barrier in between (the
         * first chunk's is H) */
        const int32_t carry_in = base == 0 ? H : sh.carry[c];
        if (base == 0)
            tables.store(sh, tid);
        __syncthreads();
        /* exclusive prefix over the waves: each lane reads the totals of
         * the waves before its own (same address across the wave's lanes
         * of one column: broadcast reads) */
        int32_t run = min(carry_in, excl);
        for (int w = 0; w < wave; w++)
            run = min(run, sh.wtot[w][c]);

        /* phase 3: mark this lane's visible samples, reserve queue slots
         * with one wave-level prefix sum, then queue them in slice order.
         * A lane's samples take consecutive slots, so the bottom of a
         * sample's span (the y-buffer just before it: the next visible row
         * below in its column) is the row of the entry before it, or, for
         * the lane's first, the y-buffer in front of the lane (lanein) */
        const int32_t run_in = run;
        uint32_t vbits = 0u;
#pragma unroll
        for (int j = 0; j < KR; j++) {
            int32_t aj = (int32_t)((ah[j / 2] >> ((j & 1) * 16)) & 0xFFFFu);
            bool vis = aj < run;
            run = vis ? aj : run;
            vbits |= vis ? (1u << j) : 0u;
        }
        {
            uint32_t cnt = (uint32_t)__builtin_popcount(vbits);
            uint32_t incl = wave_incl_sum(cnt);
            uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            uint32_t b0 = 0;
            if (lane == 0 && total)
                b0 = atomicAdd(&sh.qn, total);
            b0 = __builtin_amdgcn_readfirstlane(b0);
            uint32_t slot = b0 + incl - cnt;
#pragma unroll
            for (int j = 0; j < KR; j++) {
                if (vbits & (1u << j)) {
                    int32_t aj = (int32_t)((ah[j / 2] >> ((j & 1) * 16)) & 0xFFFFu);
                    queue[slot] = ((uint32_t)g << kQGroupShift) | ((uint32_t)j << kQSliceShift) |
                                  ((uint32_t)c << kQColShift) | (uint32_t)aj;
                    slot++;
                }
            }
            lanein[tid] = (uint16_t)run_in;
        }
        /* the last slice group of each column holds the running minimum of
         * the whole chunk: the sh.carry into the next chunk */
        if (g == SG - 1)
            sh.carry[c] = run;
        __syncthreads();
        bool done = true;
#pragma unroll
        for (int j = 0; j < TC; j++)
            done = done && sh.carry[j] == 0;

        /* phase 4: colour and span, CBATCH queue entries per thread
         * at a time: their colour-quad gathers are all issued before the
         * first colour is computed */
        const uint32_t nq = sh.qn;
        for (uint32_t i0 = 0; i0 < nq; i0 += NT * CBATCH) {
            uint32_t es[CBATCH];
            float xs[CBATCH], ys[CBATCH];
            uint4 cqs[CBATCH];
#pragma unroll
            for (int u = 0; u < CBATCH; u++) {
                uint32_t e = queue[min(i0 + (uint32_t)(u * NT + tid), nq - 1u)];
                int32_t cc = (int32_t)((e >> kQColShift) & 15u);
                float4 L = tl[(int32_t)((e >> kQSliceShift) & 31u) * SG +
                              (int32_t)(e >> kQGroupShift)];
                float fc = (float)(col0 + cc);
                es[u] = e;
                xs[u] = L.x + (fc * L.z);
                ys[u] = L.y + (fc * L.w);
                if (MODE == kOutColour)
                    cqs[u] = color_quad<CP2>(m, xs[u], ys[u]);
            }
#pragma unroll
            for (int u = 0; u < CBATCH; u++) {
                const uint32_t i = i0 + (uint32_t)(u * NT + tid);
                if (i >= nq)
                    break;
                const uint32_t e = es[u];
                int32_t s = (int32_t)(e >> kQGroupShift) * kr + (int32_t)((e >> kQSliceShift) & 31u);
                int32_t cc = (int32_t)((e >> kQColShift) & 15u);
                int32_t row = (int32_t)(e & kQRowMask);
                const float x = xs[u], y = ys[u];
                uint32_t colour;
                if (MODE == kOutColour)
                    colour = color_of_quad<true>(cqs[u], sh.sq, x, y, sh.qt);
                else if (MODE == kOutNearest)
                    colour = ((const uint32_t *)m.cq)[nearest_index<HP2>(m, x, y)] ^ f.sky;
                else
                    colour = (floorf(x) == ceilf(x) || floorf(y) == ceilf(y))
                                 ? 0u : (uint32_t)(base + s + 1);
                /* a transparent sample paints 0 (fill_vline takes the
                 * colour above it) */
                if (colour == 0u)
                    sh.transparent = 1u;
                /* the entry before, if it is this lane's (same g and c),
                 * holds the row where this sample's span ends */
                const uint32_t ep = queue[i > 0 ? i - 1u : 0u];
                const bool first = i == 0 || ((e ^ ep) & kQLaneMask) != 0;
                int32_t bottom = first ? (int32_t)lanein[(int32_t)(e >> kQGroupShift) * TC + cc]
                                       : (int32_t)(ep & kQRowMask);
                uint32_t *dst = stripe + cc * Hp;
                int32_t stop = min(bottom, row + kShortSpan);
                for (int32_t r = row; r < stop; r++)
                    dst[r] = colour;
                if (bottom > stop) {
                    uint32_t li = atomicAdd(&sh.nlong, 1u);
                    longs[2 * li] = ((uint32_t)cc << 26) | ((uint32_t)stop << kQRowBits) |
                                    (uint32_t)bottom;
                    longs[2 * li + 1] = colour;
                }
            }
        }
        __syncthreads();
        if (done)
            break;
    }

    /* no chunk ran (no slices): publish the initial y-buffer (H, all sky) */
    if (n <= 0)
        __syncthreads();

    /* long spans: each wave paints every NW-th one with its 64 lanes */
    const uint32_t nl = sh.nlong;
    for (uint32_t li = (uint32_t)wave; li < nl; li += NW) {
        uint32_t d = longs[2 * li], colour = longs[2 * li + 1];
        int32_t cc = (int32_t)(d >> 26), r0 = (int32_t)((d >> kQRowBits) & kQRowMask);
        int32_t r1 = (int32_t)(d & kQRowMask);
        for (int32_t r = r0 + lane; r < r1; r += 64)
            stripe[cc * Hp + r] = colour;
    }
    if (nl)
        __syncthreads();

    const bool vec = (col0 + TC <= f.cols) && ((pitch & 3) == 0) &&
                     ((((uintptr_t)out) & 15u) == 0);
    /* a cell above its column's final y-buffer was never painted: 0 */
    auto cell = [&](int32_t cc, int32_t r, int32_t ybuf) -> uint32_t {
        return r < ybuf ? 0u : stripe[cc * Hp + r];
    };
    if (sh.transparent == 0u) {
        /* phase 5: every painted cell is final, the rest is sky.  TC/4
         * adjacent lanes share a row, so each store instruction writes
         * whole TC*4-byte row segments */
        constexpr int PARTS = TC / 4;
        const int32_t part = tid % PARTS;
        int32_t yb[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
            yb[j] = sh.carry[part * 4 + j];
        /* the loop is written twice so that the whole-segment one keeps
         * its single 16-byte store: with the edge test inside one loop the
         * compiler shares the last element's store between the two forms
         * and splits the segment into a 12-byte and a 4-byte store */
        auto rows = [&](auto vec_tag) {
            constexpr bool VEC = decltype(vec_tag)::value;
            for (int32_t r = tid / PARTS; r < H; r += NT / PARTS) {
                uint32_t px[4];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    px[j] = painted_value<MODE>(stripe[(part * 4 + j) * Hp + r], r < yb[j], f.sky);
                uint32_t *dst = out + (size_t)r * (size_t)pitch + col0 + part * 4;
                if (VEC) {
                    *(uint4 *)dst = make_uint4(px[0], px[1], px[2], px[3]);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (col0 + part * 4 + j < f.cols)
                            dst[j] = px[j];
                }
            }
        };
        if (vec)
            rows(std::true_type{});
        else
            rows(std::false_type{});
        return;
    }

    /* phase 5 with sh.transparent samples: a zero cell takes the last
     * non-zero colour above it (fill_vline), else the sky; each thread
     * owns a block of consecutive rows */
    const int32_t per = (H + NT - 1) / NT;
    const int32_t r0 = min(tid * per, H), r1 = min(r0 + per, H);
    uint32_t pre[TC];
    int32_t yb[TC];
#pragma unroll
    for (int j = 0; j < TC; j++) {
        yb[j] = sh.carry[j];
        uint32_t p = 0u;
        for (int32_t r = r0; r < r1; r++) {
            uint32_t s = cell(j, r, yb[j]);
            p = s ? s : p;
        }
        pre[j] = p;
    }
    uint32_t *wfill = (uint32_t *)&sh.wtot[0][0];      /* [NW][TC] */
    uint32_t *wfpre = (uint32_t *)&sh.wpre[0][0];
#pragma unroll
    for (int j = 0; j < TC; j++) {
        uint32_t v = pre[j];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t o = __shfl_up(v, off);
            if (lane >= off)
                v = v ? v : o;
        }
        uint32_t ex = __shfl_up(v, 1);
        pre[j] = lane == 0 ? 0u : ex;
        if (lane == 63)
            wfill[wave * TC + j] = v;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int half = 0; half < TC / 4; half++) {
            int32_t cc = (lane & 3) + 4 * half;
            uint32_t t = (lane >> 2) < NW ? wfill[(lane >> 2) * TC + cc] : 0u;
#pragma unroll
            for (int off = 4; off < 64; off <<= 1) {
                uint32_t o = __shfl_up(t, off);
                if (lane >= off)
                    t = t ? t : o;
            }
            uint32_t ex = __shfl_up(t, 4);
            if ((lane >> 2) < NW)
                wfpre[(lane >> 2) * TC + cc] = lane < 4 ? 0u : ex;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < TC; j++) {
        uint32_t p = wfpre[wave * TC + j];
        pre[j] = pre[j] ? pre[j] : p;
    }
    /* written twice for the same reason as the loop above */
    auto rows = [&](auto vec_tag) {
        constexpr bool VEC = decltype(vec_tag)::value;
        for (int32_t r = r0; r < r1; r++) {
            uint32_t px[TC];
#pragma unroll
            for (int j = 0; j < TC; j++) {
                uint32_t s = cell(j, r, yb[j]);
                pre[j] = s ? s : pre[j];
                px[j] = tile_value<MODE>(pre[j], f.sky);
            }
            uint32_t *dst = out + (size_t)r * (size_t)pitch + col0;
            if (VEC) {
#pragma unroll
                for (int j = 0; j < TC; j += 4)
                    *(uint4 *)(dst + j) = make_uint4(px[j], px[j + 1], px[j + 2], px[j + 3]);
            } else {
#pragma unroll
                for (int j = 0; j < TC; j++)
                    if (col0 + j < f.cols)
                        dst[j] = px[j];
            }
        }
    };
    if (vec)
        rows(std::true_type{});
    else
        rows(std::false_type{});
}

template <int TC, int NT, int CMAX, int HBITS, bool HP2, bool CP2, int WPE, int MODE>
__global__ __launch_bounds__(NT, WPE) void k_render(RenderArgs a)
{
    extern __shared__ __align__(16) uint32_t lds[];  /* float4 slice table at its start */
    __shared__ TileShared<TC> sh;
    if constexpr (MODE == kOutSlice) {
        if ((int32_t)blockIdx.x >= a.ntiles) {
            const size_t stride = (size_t)a.nzero * NT;
            for (size_t i = (size_t)(blockIdx.x - a.ntiles) * NT + threadIdx.x; i < a.zero_n4;
                 i += stride)
                a.zero[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            return;
        }
    }
    if constexpr (MODE == kOutColour) {
        TileTables<TC, NT> tables;
        tables.load(threadIdx.x);
        render_tile<TC, NT, CMAX, HBITS, HP2, CP2, MODE, kGather, kColourBatch,
                    TileTables<TC, NT>>(a.m, a.f, a.out, a.pitch, xcd_tile(blockIdx.x, a.ntiles),
                                        a.chunk_max, lds, sh, tables);
    } else {
        render_tile<TC, NT, CMAX, HBITS, HP2, CP2, MODE>(a.m, a.f, a.out, a.pitch,
                                                         xcd_tile(blockIdx.x, a.ntiles),
                                                         a.chunk_max, lds, sh);
    }
}

/* Frames taller than the tile's kMaxRows: one thread per column walks the
 * slices front to back and paints each visible sample's span straight into
 * the frame (the reference's scan occlude, scatter and fill_vline,
 * fut/voxel_renderer.fut:231-249, seen from the sample's side, as the tile
 * does), then writes the sky above the final y-buffer and, if a
 * transparent sample occurred, fills the column top-down.  Slow -- one
 * thread per column, spans stored one cell at a time -- but without a row
 * limit.  Cells of transparent samples hold the mode's `empty` value until
 * the fill: 0 in colour mode (a frame colour is never 0), the sky in the
 * stateless program's mode (its fill's own test), 0xFFFFFFFE in slice
 * mode (a slice index or -1 otherwise). */
constexpr int kMaxFrameRows = 1 << 22;

template <int HBITS, bool HP2, bool CP2, int MODE>
__global__ __launch_bounds__(64) void k_render_column(RenderArgs a)
{
    __shared__ float sq[256];
    if (MODE == kOutColour) {
        for (int32_t i = threadIdx.x; i < 256; i += 64)
            fill_square_lut(sq, i);
        __syncthreads();
    }
    const fsk_frame &f = a.f;
    const fsk_map &m = a.m;
    const int32_t col = blockIdx.x * 64 + threadIdx.x;
    if (col >= f.cols)
        return;
    const int32_t H = f.rows;
    const float Hf = (float)H, fcol = (float)col;
    const uint32_t sky = MODE == kOutSlice ? 0xFFFFFFFFu : f.sky;
    const uint32_t empty = MODE == kOutColour ? 0u : (MODE == kOutNearest ? f.sky : 0xFFFFFFFEu);
    uint32_t *dst = a.out + col;
    const size_t pitch = (size_t)a.pitch;
    int32_t ybuf = H;
    bool transp = false;
    for (int32_t k = 0; k < f.n && ybuf > 0; k++) {
        const Slice sl = slice_params(f, k);
        const float x = sl.sx + (fcol * sl.dx);
        const float y = sl.sy + (fcol * sl.dy);
        float hgt;
        if (MODE == kOutNearest)
            hgt = (float)((const int32_t *)m.hq)[nearest_index<HP2>(m, x, y)];
        else
            hgt = height_interp<HBITS>(height_quad<HBITS, HP2>(m, x, y), x, y);
        const float rel = ((f.cam_h - hgt) * sl.iz) + f.horizon;
        const int32_t r = screen_row_clamped(rel, Hf);
        if (r >= ybuf)
            continue;
        uint32_t v;
        if (MODE == kOutColour)
            v = color_of_quad(color_quad<CP2>(m, x, y), sq, x, y);
        else if (MODE == kOutNearest)
            v = ((const uint32_t *)m.cq)[nearest_index<HP2>(m, x, y)];
        else
            v = (floorf(x) == ceilf(x) || floorf(y) == ceilf(y)) ? empty : (uint32_t)k;
        transp = transp || v == empty;
        for (int32_t rr = r; rr < ybuf; rr++)
            dst[(size_t)rr * pitch] = v;
        ybuf = r;
    }
    for (int32_t rr = 0; rr < ybuf; rr++)
        dst[(size_t)rr * pitch] = sky;
    if (transp) {
        uint32_t last = sky;
        for (int32_t rr = 0; rr < H; rr++) {
            const uint32_t v = dst[(size_t)rr * pitch];
            if (v == empty)
                dst[(size_t)rr * pitch] = last;
            else
                last = v;
        }
    }
}

/* render2 (fut/voxel_renderer.fut:257-305), the reference's ray-march
 * renderer, one thread per pixel.  The reference evaluates every
 * dist = 0 ... n and keeps the hit with the smallest dist (redstuff,
 * :254-255); dists grow along the ray, so that is the first hit, and the
 * march stops there (hits are the only elements that can win while the
 * miss distance i32(distance + 1) exceeds n, which the host checks).  The
 * colour is bilinear over the map's unshaded colours, evaluated for the
 * hit only.  cs/sn hold cos and sin of angle + horz_ang per column,
 * computed on the host with the C library as `futhark c` does. */
struct Render2Args {
    fsk_map m;
    const uint32_t *color;     /* unshaded colours [hrows][hcols] */
    const float *cs, *sn;      /* [cols] */
    uint32_t *out;
    int64_t pitch;
    float cam_x, cam_y, cam_h, horizon;
    int32_t n, rows, cols;
};

template <bool P2>
__device__ __forceinline__ uint4 raw_color_quad(const uint32_t *col, int32_t rows, int32_t cols,
                                                float x, float y)
{
    const uint32_t fx = (uint32_t)wrap<P2>(idx_of(floorf(x)), cols);
    const uint32_t cx = (uint32_t)wrap<P2>(idx_of(ceilf(x)), cols);
    const uint32_t fy = (uint32_t)wrap<P2>(idx_of(floorf(y)), rows);
    const uint32_t cy = (uint32_t)wrap<P2>(idx_of(ceilf(y)), rows);
    const uint32_t w = (uint32_t)cols;
    return make_uint4(col[fy * w + fx], col[fy * w + cx], col[cy * w + fx], col[cy * w + cx]);
}

template <int HBITS, bool HP2>
__global__ __launch_bounds__(256) void k_render2(Render2Args a)
{
    __shared__ float sq[256];
    fill_square_lut(sq, threadIdx.x);
    __syncthreads();
    const int32_t x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int32_t y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= a.cols || y >= a.rows)
        return;
    const float vert_ang = (a.horizon / 50.0f) + (((float)y / ((float)a.rows - 0.0f)) * (1.5f - 1.0f));
    const float ca = a.cs[x], sa = a.sn[x];
    uint32_t c = 0u;
    for (int32_t dist = 0; dist <= a.n; dist++) {
        const float fd = (float)dist;
        const float px = a.cam_x + (fd * ca), py = a.cam_y + (fd * sa);
        const float hgt = height_interp<HBITS>(height_quad<HBITS, HP2>(a.m, px, py), px, py);
        if ((a.cam_h - (fd * vert_ang)) < hgt) {
            c = color_of_quad(raw_color_quad<HP2>(a.color, a.m.hrows, a.m.hcols, px, py), sq, px,
                              py);
            break;
        }
    }
    a.out[(size_t)y * (size_t)a.pitch + (size_t)x] = c;
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

/* ------------------------------------------------------------------ */
/* Device self-tests of the arithmetic shortcuts                       */
/* ------------------------------------------------------------------ */

/* 0: unit8(b) == b / 255 (correctly rounded) for all bytes.
 * 1: sqrt_unit(t) == correctly rounded sqrt(t) for every f32 in [0, 1]
 *    and for NaNs.
 * 2: chan_u8(v) == u32.f32(clamp_channel(v) * 255) with the if-chain
 *    clamp, for every f32 bit pattern.
 * 3: screen_row_clamped(rel, H) == min(screen_row(rel), H) for every f32
 *    bit pattern at H = 1080.
 * 4: the quantisation table (g_qtab) gives u32.f32(sqrt(t) * 255), with
 *    the correctly rounded sqrt, for every f32 t in [0, 1].
 * 5: rcp_rn(z) == 1 / z (correctly rounded) for every f32 z in
 *    [2^-126, 2^126).
 * 6: div_rn(a, w, RN(1 / w)) == a / w for every finite f32 a and every
 *    frame width w of kSelftestWidths.
 * 7: sqrt_rn(t) == correctly rounded sqrt(t) for every f32 t >= 0
 *    (+inf included) and NaN stays NaN. */
constexpr int32_t kSelftestWidths[] = {1920, 1280, 800, 640, 320, 3840, 128, 96, 1024,
                                       2048, 4096, 240, 1, 3, 7, 1000, 1366};
constexpr int kNSelftestWidths = sizeof kSelftestWidths / sizeof kSelftestWidths[0];
__global__ void k_selftest(int which, unsigned long long *bad)
{
    uint64_t n = which == 0 ? 256ull
                 : (which == 1 || which == 4) ? 0x3F800001ull
                 : which == 5 ? 0x7E000000ull
                 : which == 7 ? 0x7F800001ull
                 : which == 6 ? (uint64_t)kNSelftestWidths << 32
                              : 0x100000000ull;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long local = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t u = (uint32_t)i;
        bool ok = true;
        if (which == 7) {
            const float t = __uint_as_float(u);
            ok = __float_as_uint(sqrt_rn(t)) == __float_as_uint(__builtin_sqrtf(t));
            const float nan = __uint_as_float(0x7FC00000u | (u & 0x3FFFFFu));
            ok = ok && (sqrt_rn(nan) != sqrt_rn(nan));
        } else if (which == 5) {
            u += 0x00800000u;
            const float z = __uint_as_float(u);
            ok = __float_as_uint(rcp_rn(z)) == __float_as_uint(1.0f / z);
        } else if (which == 6) {
            const float w = (float)kSelftestWidths[i >> 32];
            const float a = __uint_as_float(u);
            if ((u & 0x7F800000u) != 0x7F800000u)
                ok = __float_as_uint(div_rn(a, w, 1.0f / w)) == __float_as_uint(a / w);
        } else if (which == 0) {
            ok = __float_as_uint(unit8(u)) == __float_as_uint((float)u / 255.0f);
        } else if (which == 1) {
            float t = __uint_as_float(u);
            ok = __float_as_uint(sqrt_unit(t)) == __float_as_uint(__builtin_sqrtf(t));
            float nan = __uint_as_float(0x7FC00000u | (u & 0x3FFFFFu));
            ok = ok && (sqrt_unit(nan) != sqrt_unit(nan));
        } else if (which == 2) {
            float v = __uint_as_float(u);
            float c = (v < 0.0f) ? 0.0f : ((v > 1.0f) ? 1.0f : v);
            float sc = c * 255.0f;
            uint32_t want = (sc == sc) ? (uint32_t)sc : 0u;
            ok = chan_u8(v) == want;
        } else if (which == 3) {
            float rel = __uint_as_float(u);
            ok = screen_row_clamped(rel, 1080.0f) == min(screen_row(rel), 1080);
        } else {
            float t = __uint_as_float(u);
            const uint32_t want = (uint32_t)(__builtin_sqrtf(t) * 255.0f);
            ok = (mix_q(g_qtab, 1.0f, t, 0.0f, 0.0f) >> 17) == want;
        }
        if (!ok) {
            local += 1;
            atomicMin(bad + 1, (unsigned long long)u);
        }
    }
    if (local)
        atomicAdd(bad, local);
}

/* Per-device caches of device properties and kernel set-up (a process may
 * drive several devices; each has its own). */
constexpr int kMaxDevices = 64;

inline int current_device()
{
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxDevices)
        d = 0;
    return d;
}

inline size_t lds_per_cu()
{
    static std::atomic<int> v[kMaxDevices];
    const int dev = current_device();
    int x = v[dev].load(std::memory_order_relaxed);
    if (x == 0) {
        if (hipDeviceGetAttribute(&x, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
                hipSuccess ||
            x <= 0)
            x = 65536;
        v[dev].store(x, std::memory_order_relaxed);
    }
    return (size_t)x;
}

inline int cu_count()
{
    static std::atomic<int> v[kMaxDevices];
    const int dev = current_device();
    int x = v[dev].load(std::memory_order_relaxed);
    if (x == 0) {
        if (hipDeviceGetAttribute(&x, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            x <= 0)
            x = 256;
        v[dev].store(x, std::memory_order_relaxed);
    }
    return x;
}

/* Lets `kern` take all of a CU's LDS but its static shared memory, once per
 * device; returns that dynamic maximum (0 on failure, with *err set). */
inline int dyn_lds_max(const void *kern, std::atomic<int> *cache, hipError_t *err)
{
    const int dev = current_device();
    int x = cache[dev].load(std::memory_order_relaxed);
    if (x > 0)
        return x;
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, kern);
    if (e == hipSuccess) {
        x = (int)lds_per_cu() - (int)fa.sharedSizeBytes;
        e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, x);
    }
    if (e != hipSuccess) {
        *err = e;
        return 0;
    }
    cache[dev].store(x, std::memory_order_relaxed);
    return x;
}

template <int TC, int NT, int CMAX, int HBITS, bool HP2, bool CP2, int WPE, int MODE>
hipError_t launch_render_cfg(const RenderArgs &a, hipStream_t s)
{
    auto kern = k_render<TC, NT, CMAX, HBITS, HP2, CP2, WPE, MODE>;
    static std::atomic<int> dyn_cache[kMaxDevices];
    hipError_t e = hipSuccess;
    const int dyn_max = dyn_lds_max((const void *)kern, dyn_cache, &e);
    if (!dyn_max)
        return e;
    size_t need = render_lds_bytes(TC, NT, a.f.rows, a.chunk_max);
    if (need > (size_t)dyn_max)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(kern, dim3(a.ntiles + a.nzero), dim3(NT), need, s, a);
    return hipGetLastError();
}

/* ------------------------------------------------------------------ */
/* Kernel shape and tuning                                             */
/* ------------------------------------------------------------------ */

/* Render kernel shape: columns per workgroup tile (8: 32-byte row
 * segments; 4: 16-byte segments), one 1024-thread or two 512-thread
 * workgroups per CU (two let one tile's barrier and gather stalls overlap
 * the other's work -- worth it when there are many more tiles than CUs, as
 * in batches; they need the LDS of two tiles, so their depth chunk is
 * shorter: at most 512 slices for 8 columns, 1024 for 4), and the
 * depth-chunk length.  Each context carries its own tuning (fsx_set_tuning,
 * the autotuner), which starts from the FSX_RENDER_TC /
 * FSX_RENDER_WG_PER_CU / FSX_RENDER_CHUNK environment variables; 0 means
 * automatic. */

struct Shape {
    int tc, nt, wg, chunk, cmax;
};

constexpr size_t kStaticLds = 12288;  /* headroom for the kernels' static shared memory */

inline Shape resolve_shape(int32_t rows, const fsk_tuning &t)
{
    Shape s;
    s.cmax = kChunk;
    int chunk = (t.chunk >= 256 && t.chunk <= kChunk) ? t.chunk : 0;
    int c8 = chunk && chunk <= 512 ? chunk : 512;
    auto fits8 = [&](int c) {
        return render_lds_bytes(8, 1024, rows, c) + kStaticLds <= lds_per_cu();
    };
    if (t.tc != 4 && t.wg_per_cu == 2 &&
        2 * (render_lds_bytes(8, 512, rows, c8) + kStaticLds) <= lds_per_cu()) {
        s.tc = 8;
        s.nt = 512;
        s.wg = 2;
        s.chunk = c8;
        s.cmax = 512;
    } else if (t.tc != 4 && (chunk ? fits8(chunk) : fits8(kChunk) || fits8(kChunk / 2))) {
        /* left automatic, a frame too tall for 8-column tiles of the whole
         * chunk keeps them with half the chunk: measured faster than
         * 4-column tiles (2160 rows, 4096 slices: 49 against 55 us) */
        s.tc = 8;
        s.nt = 1024;
        s.wg = 1;
        s.chunk = chunk ? chunk : fits8(kChunk) ? kChunk : kChunk / 2;
    } else {
        s.tc = 4;
        int c2 = chunk ? chunk : 1024;
        c2 -= c2 % 128;
        if (t.wg_per_cu == 2 && 2 * (render_lds_bytes(4, 512, rows, c2) + kStaticLds) <=
                                    lds_per_cu()) {
            s.nt = 512;
            s.wg = 2;
            s.chunk = c2;
        } else {
            s.nt = 1024;
            s.wg = 1;
            s.chunk = chunk ? chunk : kChunk;
        }
    }
    /* a chunk holds whole groups of kChunkAlign slices per slice group;
     * tall frames take shorter chunks until the tile fits in LDS */
    auto need = [&](int c) {
        return render_lds_bytes(s.tc, s.nt, rows, c) + kStaticLds;
    };
    auto fit = [&]() {
        const int ga = (s.nt / s.tc) * kChunkAlign;
        s.chunk = s.chunk < ga ? ga : s.chunk - s.chunk % ga;
        while (s.chunk > ga && (size_t)s.wg * need(s.chunk) > lds_per_cu())
            s.chunk = (s.chunk / 2) - (s.chunk / 2) % ga;
        return (size_t)s.wg * need(s.chunk) <= lds_per_cu();
    };
    if (!fit() && s.tc == 4 && s.nt == 1024) {
        /* the tallest frames: 512-thread tiles, whose chunks go down to
         * 512 slices */
        s.nt = 512;
        s.wg = 1;
        fit();
    }
    return s;
}

template <int HBITS, bool HP2, bool CP2, int MODE = kOutColour>
hipError_t launch_render(RenderArgs a, const fsk_tuning &tune, hipStream_t s,
                         int32_t *nzero = nullptr)
{
    if (a.f.rows > kMaxRows) {
        /* taller than a tile: the column kernel (no gradient clear) */
        if (nzero)
            *nzero = 0;
        if (a.f.rows > kMaxFrameRows)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_render_column<HBITS, HP2, CP2, MODE>), dim3((a.f.cols + 63) / 64),
                           dim3(64), 0, s, a);
        return hipGetLastError();
    }
    Shape sh = resolve_shape(a.f.rows, tune);
    a.chunk_max = sh.chunk;
    a.ntiles = (a.f.cols + sh.tc - 1) / sh.tc;
    /* slice mode: the workgroup slots the tiles leave free clear the
     * gradient, if there are enough of them */
    const int32_t spare = cu_count() * sh.wg - a.ntiles;
    a.nzero = (MODE == kOutSlice && a.zero && a.zero_n4 && spare >= kMinZeroBlocks) ? spare : 0;
    if (nzero)
        *nzero = a.nzero;
    if (sh.tc == 8) {
        return sh.nt == 512 ? launch_render_cfg<8, 512, 512, HBITS, HP2, CP2, 4, MODE>(a, s)
                            : launch_render_cfg<8, 1024, kChunk, HBITS, HP2, CP2, 4, MODE>(a, s);
    }
    return sh.nt == 512 ? launch_render_cfg<4, 512, kChunk, HBITS, HP2, CP2, 4, MODE>(a, s)
                        : launch_render_cfg<4, 1024, kChunk, HBITS, HP2, CP2, 4, MODE>(a, s);
}

/* ------------------------------------------------------------------ */
/* Batched persistent render                                           */
/* ------------------------------------------------------------------ */

struct BatchArgs {
    const fsk_batch_view *views;
    const uint32_t *items;
    const uint32_t *qmeta;
    uint32_t *counters;
    uint32_t *out;
    int32_t nqueues, chunk_max;
};

/* XCD this wave runs on (hardware register XCC_ID). */
__device__ __forceinline__ uint32_t xcc_id()
{
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 0xFu;
}

__global__ void k_batch_reset(uint32_t *counters, int32_t n)
{
    int32_t i = threadIdx.x;
    if (i < n)
        counters[i] = 0u;
}

/* Persistent workgroups claim (view, tile) items from per-XCD queues: the
 * host deals each view's tiles out to the queues in contiguous blocks, so
 * neighbouring tiles of a view stay on one L2; a workgroup whose own queue
 * is drained takes from the next queues.  Every workgroup exits once all
 * queues report exhausted, after at most nqueues claims. */
template <int TC, int NT, int CMAX, int HBITS, bool HP2, bool CP2, int WPE>
__global__ __launch_bounds__(NT, WPE) void k_render_batch(BatchArgs b)
{
    extern __shared__ __align__(16) uint32_t lds[];  /* float4 slice table at its start */
    __shared__ TileShared<TC> sh;
    __shared__ uint32_t cur;
    {
        TileTables<TC, NT> tables;
        tables.load(threadIdx.x);
        tables.store(sh, threadIdx.x);    /* the claim loop's barrier publishes them */
    }
    uint32_t q = xcc_id() % (uint32_t)b.nqueues;
    int32_t tried = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t v = 0xFFFFFFFFu;
            while (tried < b.nqueues) {
                uint32_t k = atomicAdd(&b.counters[q], 1u);
                if (k < b.qmeta[b.nqueues + q]) {
                    v = b.items[b.qmeta[q] + k];
                    break;
                }
                tried++;
                q = (q + 1u) % (uint32_t)b.nqueues;
            }
            cur = v;
        }
        __syncthreads();
        uint32_t it = cur;
        __syncthreads();
        if (it == 0xFFFFFFFFu)
            break;
        const fsk_batch_view &bv = b.views[it >> 16];
        render_tile<TC, NT, CMAX, HBITS, HP2, CP2, kOutColour>(
            bv.m, bv.f, b.out + bv.offset, bv.f.cols, it & 0xFFFFu, b.chunk_max, lds, sh);
    }
}

template <int TC, int NT, int CMAX, int HBITS, bool HP2, bool CP2, int WPE>
hipError_t launch_batch_cfg(const fsk_batch *fb, int chunk, hipStream_t s)
{
    auto kern = k_render_batch<TC, NT, CMAX, HBITS, HP2, CP2, WPE>;
    static std::atomic<int> dyn_cache[kMaxDevices];
    hipError_t e = hipSuccess;
    const int dyn_max = dyn_lds_max((const void *)kern, dyn_cache, &e);
    if (!dyn_max)
        return e;
    size_t need = render_lds_bytes(TC, NT, fb->max_rows, chunk);
    if (need > (size_t)dyn_max)
        return hipErrorInvalidValue;
    BatchArgs b;
    b.views = fb->views;
    b.items = fb->items;
    b.qmeta = fb->qmeta;
    b.counters = fb->counters;
    b.out = fb->out;
    b.nqueues = fb->nqueues;
    b.chunk_max = chunk;
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, fb->counters, fb->nqueues);
    hipLaunchKernelGGL(kern, dim3(fb->grid), dim3(NT), need, s, b);
    return hipGetLastError();
}

template <int HBITS, bool HP2, bool CP2>
hipError_t launch_batch(const fsk_batch *fb, hipStream_t s)
{
    Shape sh = resolve_shape(fb->max_rows, fb->tune);
    if (sh.tc == 8)
        return sh.nt == 512 ? launch_batch_cfg<8, 512, 512, HBITS, HP2, CP2, 4>(fb, sh.chunk, s)
                            : launch_batch_cfg<8, 1024, kChunk, HBITS, HP2, CP2, 4>(fb, sh.chunk, s);
    return sh.nt == 512 ? launch_batch_cfg<4, 512, kChunk, HBITS, HP2, CP2, 4>(fb, sh.chunk, s)
                        : launch_batch_cfg<4, 1024, kChunk, HBITS, HP2, CP2, 4>(fb, sh.chunk, s);
}

/* ------------------------------------------------------------------ */
/* Differentiable renderer: f32 colour planes                           */
/* ------------------------------------------------------------------ */

/* The f32 relaxation of the colour path: planes cp[3][hc][wc] in [0,1]
 * replace the u32 ARGB map, each texel scaled by its shade factor
 * sf = sqrt(1/(4*count + 1)) (argb.mix(4*count, black, 1, c) without
 * quantisation, fut/effects.fut:123), and argb.mix's channel rule
 * sqrt(m1'*a^2 + m2'*b^2) without u8 rounding (matte colour.fut, as
 * png_color_filtered calls it at fut/render_functions.fut:101-103).
 * Visibility is the integer renderer's, held fixed: src[r][c] is the
 * depth slice whose sample shows at the pixel, -1 for sky.  Sky pixels
 * are sky_base * sky_scale with sky_base = (0x90, 0x90, 0xE0) / 255
 * (argb.scale of 0xFF9090e0, fut/interactive.fut:163). */


/* Sum of n per-workgroup partials in a fixed order by one workgroup of NT
 * threads, stored into *dst or added to it: thread t adds part[t],
 * part[t + NT], ... in that order (eight loads issued before the first
 * add), then the wave totals are added in wave order.  Every thread of the
 * workgroup must call it. */
template <int NT>
__device__ __forceinline__ void block_sum_partials(const float *__restrict__ part, int32_t n,
                                                   float *__restrict__ dst, int32_t accumulate)
{
    __shared__ float wsum[NT / 64];
    float acc = 0.0f;
    for (int32_t base = 0; base < n; base += 8 * NT) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int32_t i = base + u * NT + (int32_t)threadIdx.x;
            v[u] = i < n ? part[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            acc += v[u];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0)
        wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int w = 0; w < NT / 64; w++)
            t += wsum[w];
        *dst = accumulate ? *dst + t : t;
    }
}

struct DiffArgs {
    fsk_frame f;
    const float *cp;
    const float *sf;
    int32_t hc, wc;
    float sky_scale;
    const int32_t *src;
    const float *gimg;     /* backward: dL/dimg [3][rows][cols] */
    float *img;            /* forward output [3][rows][cols] */
    float *gcp;            /* backward: dL/dcp accumulated [3][hc][wc] */
    float *gsky;           /* backward: dL/dsky_scale accumulated */
    float *partials;       /* per-workgroup partial sums (loss or sky) */
    const float *tgt;      /* fused step: the target frame [3][rows][cols] */
    float *gout;           /* fused step: dL/dimg written out (may be null) */
    float *sky_part;       /* fused step: per-workgroup sky terms */
};

/* The relaxed argb.mix channel: sqrt(m1' a^2 + m2' b^2) with weights
 * m1', m2' already normalised by norm_weights. */
__device__ __forceinline__ float mixn(float m1n, float a, float m2n, float b)
{
    return sqrt_rn((m1n * (a * a)) + (m2n * (b * b)));
}

/* One visible sample: texel offsets (fy,fx) (fy,cx) (cy,fx) (cy,cx) and
 * the bilinear weights, as png_color_filtered forms them
 * (fut/render_functions.fut:95-105) with floor-mod indices. */
struct Bilin {
    uint32_t t[4];
    float wx0, wx1, wy0, wy1;
    bool pair;               /* fx is not the map's last column */
};

/* The four texels of a bilinear quad from one plane.  When no lane of
 * the wave has its quad wrap round the map's last column, texels t1 and
 * t3 follow t0 and t2 in memory, and each pair is one 8-byte load (4-byte
 * aligned, which global loads allow).  At an integer x the quad's second
 * texel is its first (cx = fx) and the pair loads the next one instead;
 * the x weights are then 0 and their sum 0, so the colour is NaN whatever
 * the texel holds. */
typedef float float2_a4 __attribute__((ext_vector_type(2), aligned(4)));
__device__ __forceinline__ void quad_load(const float *P, const Bilin &b, bool pairs,
                                          float v[4])
{
    if (pairs) {
        const float2_a4 lo = *(const float2_a4 *)(P + b.t[0]);
        const float2_a4 hi = *(const float2_a4 *)(P + b.t[2]);
        v[0] = lo.x;
        v[1] = lo.y;
        v[2] = hi.x;
        v[3] = hi.y;
    } else {
#pragma unroll
        for (int t = 0; t < 4; t++)
            v[t] = P[b.t[t]];
    }
}

/* argb.mix's weights divided by their sum, for the x pair and the y pair
 * of a bilinear sample (w = wx0', wx1', wy0', wy1').  A pair sums to
 * exactly 1 unless its coordinate is an integer (sum 0) or within 1 of the
 * map origin, where the division is an exact no-op when skipped, so the
 * divisions run on a wave-uniform branch only when some lane of the wave
 * needs them (the empty asm keeps the compiler from evaluating them on
 * every wave). */
__device__ __forceinline__ void norm_weights(const Bilin &b, float w[4])
{
    w[0] = b.wx0;
    w[1] = b.wx1;
    w[2] = b.wy0;
    w[3] = b.wy1;
    const float sx = b.wx0 + b.wx1, sy = b.wy0 + b.wy1;
    const bool dx = sx != 1.0f, dy = sy != 1.0f;
    if (__ballot(dx || dy)) {
        float x0 = b.wx0, x1 = b.wx1, y0 = b.wy0, y1 = b.wy1;
        asm volatile("" : "+v"(x0), "+v"(x1), "+v"(y0), "+v"(y1));
        if (dx) {
            w[0] = x0 / sx;
            w[1] = x1 / sx;
        }
        if (dy) {
            w[2] = y0 / sy;
            w[3] = y1 / sy;
        }
    }
}

/* Texel offsets are 32-bit: a plane holds at most 2^31 texels (checked by
 * the host).  P2: both plane sides are powers of two, so the floor-mod is
 * a mask (also for the INT_MIN of an out-of-range cast). */
template <bool P2, bool TAME = false>
__device__ __forceinline__ Bilin bilin_of(const fsk_frame &f, int32_t k, int32_t col, int32_t hc,
                                          int32_t wc)
{
    Slice sl = TAME || f.tame ? slice_params_tame(f, k) : slice_params(f, k);
    float fcol = (float)col;
    float x = sl.sx + (fcol * sl.dx);
    float y = sl.sy + (fcol * sl.dy);
    float fx = floorf(x), cx = ceilf(x), fy = floorf(y), cy = ceilf(y);
    /* TAME: a tame frame keeps |x|, |y| < 2^22, where the x86 conversion
     * is the plain one */
    auto cvt = [](float v) { return TAME ? (int32_t)v : i32_f32_x86(v); };
    int32_t ify = wrap<P2>(cvt(fy), hc), icy = wrap<P2>(cvt(cy), hc);
    int32_t ifx = wrap<P2>(cvt(fx), wc), icx = wrap<P2>(cvt(cx), wc);
    const uint32_t ry = P2 ? (uint32_t)ify << __builtin_ctz((uint32_t)wc)
                           : (uint32_t)ify * (uint32_t)wc;
    const uint32_t rc = P2 ? (uint32_t)icy << __builtin_ctz((uint32_t)wc)
                           : (uint32_t)icy * (uint32_t)wc;
    Bilin b;
    b.t[0] = ry + (uint32_t)ifx;
    b.t[1] = ry + (uint32_t)icx;
    b.t[2] = rc + (uint32_t)ifx;
    b.t[3] = rc + (uint32_t)icx;
    b.wx0 = cx - x;
    b.wx1 = x - fx;
    b.wy0 = cy - y;
    b.wy1 = y - fy;
    b.pair = ifx < wc - 1;
    return b;
}

__device__ __forceinline__ float sky_base(int ch)
{
    return ch == 2 ? 224.0f / 255.0f : 144.0f / 255.0f;
}

/* The shaded texel values of a sample's quad: sf[t] and the three planes'
 * cv[ch][t].  All sixteen gathers are issued here, before any mix: the
 * mixes branch, and loads are not moved above a branch, so gathers
 * interleaved with them would wait for each other. */
__device__ __forceinline__ void sample_gather(const DiffArgs &a, const Bilin &b, size_t plane,
                                              float cv[3][4], float sf[4])
{
    /* one load per texel: the paired loads of quad_load measured slower
     * in the adjoint (its gathers sit under a per-lane branch) */
#pragma unroll
    for (int t = 0; t < 4; t++)
        sf[t] = a.sf[b.t[t]];
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        const float *P = a.cp + ch * plane;
#pragma unroll
        for (int t = 0; t < 4; t++)
            cv[ch][t] = P[b.t[t]];
    }
}

/* The colour of a visible sample from its gathered quad, per channel (the
 * relaxed png_color_filtered); sv[ch][t] = cv[ch][t] * sf[t]. */
__device__ __forceinline__ void sample_mix(const Bilin &b, const float cv[3][4],
                                           const float sf[4], float sv[3][4], float px[3])
{
    float w[4];
    norm_weights(b, w);
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
#pragma unroll
        for (int t = 0; t < 4; t++)
            sv[ch][t] = cv[ch][t] * sf[t];
        float X1 = mixn(w[0], sv[ch][0], w[1], sv[ch][1]);
        float X2 = mixn(w[0], sv[ch][2], w[1], sv[ch][3]);
        px[ch] = mixn(w[2], X1, w[3], X2);
    }
}

/* Forward colour, one thread per pixel (2-D blocks of 64 columns x 4
 * rows: the gathers of neighbouring pixels share texels).  This kernel
 * gathers each channel's quad just before its mixes: with one pixel per
 * thread and many waves in flight that measured faster than issuing all
 * sixteen gathers first (which needs more registers per wave), and so did
 * one pixel per thread against two or four, and against listing each
 * column's runs to colour each visible sample once. */
template <bool P2, bool TAME>
__global__ __launch_bounds__(256) void k_diff_color(DiffArgs a)
{
    const int32_t col = blockIdx.x * 64 + (threadIdx.x & 63);
    const int32_t row = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int32_t H = a.f.rows, W = a.f.cols;
    if (col >= W || row >= H)
        return;
    const size_t plane = (size_t)a.hc * a.wc, fplane = (size_t)H * W;
    const size_t p = (size_t)row * W + col;
    const int32_t k = a.src[p];
    float px[3];
    if (k < 0) {
#pragma unroll
        for (int ch = 0; ch < 3; ch++)
            px[ch] = sky_base(ch) * a.sky_scale;
    } else {
        const Bilin b = bilin_of<P2, TAME>(a.f, k, col, a.hc, a.wc);
        float sf[4], w[4];
        const bool pairs = __ballot(!b.pair) == 0;
        quad_load(a.sf, b, pairs, sf);
        norm_weights(b, w);
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            float sv[4];
            quad_load(a.cp + ch * plane, b, pairs, sv);
#pragma unroll
            for (int t = 0; t < 4; t++)
                sv[t] *= sf[t];
            float X1 = mixn(w[0], sv[0], w[1], sv[1]);
            float X2 = mixn(w[0], sv[2], w[1], sv[3]);
            px[ch] = mixn(w[2], X1, w[3], X2);
        }
    }
#pragma unroll
    for (int ch = 0; ch < 3; ch++)
        a.img[ch * fplane + p] = px[ch];
}

constexpr uint32_t kNoKey = 0xFFFFFFFFu;     /* an empty texel slot, a lane with no sample */

/* One step of segment_sum_quad: each lane takes (v, f) of the lane CTRL
 * names (DPP) unless its accumulated range already holds its segment's
 * first lane (f).  Lanes with no source (shifted out of their row, or in
 * rows that ROWS leaves out) take (0, 0), which changes nothing. */
template <int CTRL, int ROWS, int N>
__device__ __forceinline__ void segment_step(float (&v)[N], int32_t &f)
{
    constexpr bool kShift = CTRL < 0x140;      /* row shifts read 0 outside the row */
    float d[N];
#pragma unroll
    for (int c = 0; c < N; c++)
        d[c] = __int_as_float(
            __builtin_amdgcn_update_dpp(0, __float_as_int(v[c]), CTRL, ROWS, 0xf, kShift));
    const int32_t fd = __builtin_amdgcn_update_dpp(0, f, CTRL, ROWS, 0xf, kShift);
    if (!f) {
#pragma unroll
        for (int c = 0; c < N; c++)
            v[c] += d[c];
        f = fd;
    }
}

/* Segmented inclusive prefix sum of a sample's four texel values over the
 * lanes of a wave, in registers: lanes holding the same texel quad (all
 * four keys equal, none of them kNoKey) in a contiguous run form a
 * segment, and each lane ends up with the sums of val from its segment's
 * first lane to itself, so the segment's last lane holds the segment's
 * totals.  Steps of 1, 2, 4 and 8 lanes scan each row of 16 lanes (DPP
 * row shifts); then the last lane of rows 0 and 2 is carried into rows 1
 * and 3 (row_bcast:15) and lane 31 into rows 2 and 3 (row_bcast:31).  A
 * wave in which no lane shares its quad with the lane before skips the
 * scan (every lane is a segment of its own).  Returns whether this lane
 * ends its segment. */
__device__ __forceinline__ bool segment_sum_quad(const uint32_t key[4], float (&val)[4][3])
{
    /* the keys of lane - 1 and lane + 1 (wave shifts; the first and the
     * last lane see a key that differs from their own).  Every lane reads
     * both neighbours: a DPP read from a lane that a branch has switched
     * off does not see that lane's key, so the comparisons are combined
     * without short-circuiting */
    uint32_t diff_prev = key[0] == kNoKey ? 1u : 0u, diff_next = diff_prev;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int32_t k = (int32_t)key[t];
        diff_prev |= (uint32_t)(__builtin_amdgcn_update_dpp(~k, k, 0x138, 0xf, 0xf, false) ^ k);
        diff_next |= (uint32_t)(__builtin_amdgcn_update_dpp(~k, k, 0x130, 0xf, 0xf, false) ^ k);
    }
    const bool same_prev = diff_prev == 0u, same_next = diff_next == 0u;
    if (!__ballot(same_prev))
        return true;
    float v[12];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            v[t * 3 + c] = val[t][c];
    int32_t f = same_prev ? 0 : 1;
    segment_step<0x111, 0xf>(v, f);            /* row_shr:1 */
    segment_step<0x112, 0xf>(v, f);            /* row_shr:2 */
    segment_step<0x114, 0xf>(v, f);            /* row_shr:4 */
    segment_step<0x118, 0xf>(v, f);            /* row_shr:8 */
    segment_step<0x142, 0xa>(v, f);            /* row_bcast:15 */
    segment_step<0x143, 0xc>(v, f);            /* row_bcast:31 */
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            val[t][c] = v[t * 3 + c];
    return !same_next;
}

constexpr int kGradThreads = 512;
/* Rows per thread in the adjoint: 2, or 4 in frames of kGradTallPixels
 * pixels or more.  A run of equal src down a thread's rows is merged in
 * registers, so more rows mean fewer texel inserts per pixel, which pays
 * where the adjoint is bound by its VALU work (3840 x 2160: 16.3 against
 * 18.2 ms per fitting step), while smaller frames need the waves of the
 * shorter blocks to hide their gathers' latency (1280 x 720: 78 against
 * 91 us per fused step). */
constexpr int64_t kGradTallPixels = (int64_t)1 << 22;
constexpr int kGradCapLog = 12;
constexpr int kGradCap = 1 << kGradCapLog;   /* texel slots of a workgroup's LDS table */
constexpr int kGradProbes = 2;

/* Add v[0..2] to texel `key` of the workgroup's LDS table (open addressing,
 * linear probing); returns whether this call took a new slot.  A texel that
 * finds no slot within kGradProbes, or any texel once half of the table
 * is taken (far-range blocks, where every sample has texels of its own),
 * goes straight to the global gradient.  The global f32 atomics are what
 * bounds the adjoint at 1280 x 720: stopping at a quarter of the table
 * had the fused step at 74 against 68 us. */
template <int CAPLOG = kGradCapLog, int FILL_LIMIT = kGradCap / 2, int PROBES = kGradProbes>
__device__ __forceinline__ bool table_add(uint32_t *keys, float *vals, const uint32_t *fill,
                                          float *gcp, size_t plane, uint32_t key, const float v[3])
{
    constexpr int kCap = 1 << CAPLOG;
    uint32_t h = (key * 2654435761u) >> (32 - CAPLOG);
    /* a relaxed atomic load, which becomes an LDS read: a volatile read
     * through this generic pointer was a flat load that waited for every
     * global atomic of the wave still in flight */
    const bool open = __hip_atomic_load(fill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) <
                      (uint32_t)FILL_LIMIT;
#pragma unroll 1
    for (int p = 0; open && p < PROBES; p++) {
        uint32_t k = keys[h];
        bool mine = false;
        if (k == kNoKey) {
            k = atomicCAS(&keys[h], kNoKey, key);
            mine = k == kNoKey;
        }
        if (mine || k == key) {
#pragma unroll
            for (int ch = 0; ch < 3; ch++)
                if (v[ch] != 0.0f)
                    atomicAdd(&vals[ch * kCap + h], v[ch]);
            return mine;
        }
        h = (h + 1) & (kCap - 1);
    }
#pragma unroll
    for (int ch = 0; ch < 3; ch++)
        if (v[ch] != 0.0f)
            unsafeAtomicAdd(&gcp[ch * plane + key], v[ch]);
    return false;
}

/* The texel terms of one visible sample k at column col whose pixels'
 * incoming gradients sum to run[3]: key[t] = texel t of the bilinear quad,
 * val[t][ch] = dL/d plane_ch at that texel.  For px = mix(wy0, X1, wy1, X2),
 * X1 = mix(wx0, s0, wx1, s1): d px / d s0 = wy0' * wx0' * s0 / px (X1
 * cancels), likewise for the other texels, 0 where px == 0. */
struct SampleEval {
    Bilin b;
    float sf[4];             /* shade factors of the quad */
    float sv[3][4], px[3];   /* shaded texel values, pixel colour */
};

template <bool P2>
__device__ __forceinline__ void sample_eval(const DiffArgs &a, int32_t k, int32_t col,
                                            size_t plane, SampleEval &e)
{
    e.b = bilin_of<P2>(a.f, k, col, a.hc, a.wc);
    float cv[3][4];
    sample_gather(a, e.b, plane, cv, e.sf);
    sample_mix(e.b, cv, e.sf, e.sv, e.px);
}

__device__ __forceinline__ void sample_grad(const DiffArgs &a, const SampleEval &e,
                                            const float run[3], uint32_t key[4], float val[4][3])
{
    const Bilin &b = e.b;
    const float(&sv)[3][4] = e.sv;
    const float(&px)[3] = e.px;
    float w[4];
    norm_weights(b, w);
    const float wt[4] = {w[2] * w[0], w[2] * w[1], w[3] * w[0], w[3] * w[1]};
    /* gradients are not bit-exact quantities: the hardware reciprocal
     * (within an ulp) stands in for the division, and the texel's weight
     * and shade factor are multiplied once */
    float q[3];
#pragma unroll
    for (int ch = 0; ch < 3; ch++)
        q[ch] = px[ch] > 0.0f ? run[ch] * __builtin_amdgcn_rcpf(px[ch]) : 0.0f;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const float ws = wt[t] * e.sf[t];
        key[t] = b.t[t];
#pragma unroll
        for (int ch = 0; ch < 3; ch++)
            val[t][ch] = (q[ch] * ws) * sv[ch][t];
    }
}

template <bool P2>
__device__ __forceinline__ void sample_terms(const DiffArgs &a, int32_t k, int32_t col,
                                             size_t plane, const float run[3], uint32_t key[4],
                                             float val[4][3])
{
    SampleEval e;
    sample_eval<P2>(a, k, col, plane, e);
    sample_grad(a, e, run, key, val);
}

/* Adjoint.  A workgroup covers 64 adjacent columns x 16 rows; each thread
 * holds kGradRows rows of one column.  A run of equal src down the column
 * is one visible sample: its incoming gradient is summed in registers and
 * pushed through the three mixes once.  For px = mix(wy0, X1, wy1, X2),
 * X1 = mix(wx0, s0, wx1, s1): d px / d s0 = wy0' * wx0' * s0 / px (X1
 * cancels), likewise for the other texels, 0 where px == 0.  The texel
 * terms are summed in two stages before any global atomic: lanes of a
 * wave that end a run on the same row and hit the same texel (adjacent
 * columns magnifying one near texel) are merged by a segmented reduction
 * across the wave, so that the LDS table below sees distinct texels per
 * instruction; then an LDS hash table sums the terms per texel over the
 * whole 64 x 16 block (neighbouring samples share texels: a block touches
 * several times fewer texels than it has sample-texel terms).  Short
 * per-thread row runs keep many waves in flight for the gathers.  Each
 * table entry is then added into the global f32 gradient with one atomic
 * per channel.
 *
 * FWD (the fused L2 step): the incoming gradient is not read but formed
 * here from the target, g = 2 (px - target), with the colour px of each
 * run's sample evaluated where the run starts (the same evaluation as
 * k_diff_color's, so img and g are bit-identical to the separate calls);
 * img and g are written only where the caller asks for them, and each
 * workgroup leaves one partial of the loss, sum (px - target)^2, and one of
 * the sky term.  Otherwise each workgroup leaves one partial of the sky
 * term. */
template <bool P2, bool FWD, int kGradRows>
__global__ __launch_bounds__(kGradThreads) void k_diff_grad(DiffArgs a)
{
    constexpr int kGradBlockRows = (kGradThreads / 64) * kGradRows;
    __shared__ uint32_t keys[kGradCap];
    __shared__ float vals[3 * kGradCap];
    __shared__ uint32_t fill;
    for (int32_t i = threadIdx.x; i < kGradCap; i += kGradThreads) {
        keys[i] = kNoKey;
        vals[i] = 0.0f;
        vals[kGradCap + i] = 0.0f;
        vals[2 * kGradCap + i] = 0.0f;
    }
    if (threadIdx.x == 0)
        fill = 0u;
    const int32_t col = blockIdx.x * 64 + (threadIdx.x & 63);
    const int32_t r0 = blockIdx.y * kGradBlockRows + (threadIdx.x >> 6) * kGradRows;
    const int32_t H = a.f.rows, W = a.f.cols;
    const bool ok = col < W && r0 < H;
    const size_t plane = (size_t)a.hc * a.wc, fplane = (size_t)H * W;
    /* global-space views of the per-pixel inputs (loads through the
     * generic pointers came out as flat loads) */
    const FS_GLOBAL float *gin = (const FS_GLOBAL float *)(FWD ? a.tgt : a.gimg);
    const FS_GLOBAL int32_t *srcp = (const FS_GLOBAL int32_t *)a.src;
    int32_t src[kGradRows];
    float g[3][kGradRows];     /* FWD: the target until the pixel's g replaces it */
#pragma unroll
    for (int j = 0; j < kGradRows; j++) {
        const int32_t r = r0 + j;
        const bool in = ok && r < H;
        const size_t p = (size_t)r * W + col;
        src[j] = in ? srcp[p] : INT32_MIN;
#pragma unroll
        for (int ch = 0; ch < 3; ch++)
            g[ch][j] = in ? gin[ch * fplane + p] : 0.0f;
    }
    __syncthreads();
    float sky_acc = 0.0f, loss_acc = 0.0f;
    float run[3] = {0.0f, 0.0f, 0.0f};
    SampleEval e;              /* FWD: the current run's sample */
#pragma unroll
    for (int j = 0; j < kGradRows; j++) {
        const int32_t k = src[j];
        if constexpr (FWD) {
            const bool starts = k != INT32_MIN && (j == 0 || src[j - 1] != k);
            if (starts && k < 0) {
#pragma unroll
                for (int ch = 0; ch < 3; ch++)
                    e.px[ch] = sky_base(ch) * a.sky_scale;
            }
            const bool eval = starts && k >= 0;
            if (__ballot(eval)) {
                if (eval)
                    sample_eval<P2>(a, k, col, plane, e);
            }
            if (k != INT32_MIN) {
                const size_t p = (size_t)(r0 + j) * W + col;
#pragma unroll
                for (int ch = 0; ch < 3; ch++) {
                    const float d = e.px[ch] - g[ch][j];
                    const float gd = 2.0f * d;
                    if (a.img)
                        a.img[ch * fplane + p] = e.px[ch];
                    if (a.gout)
                        a.gout[ch * fplane + p] = gd;
                    loss_acc += d * d;
                    g[ch][j] = gd;
                }
            }
        }
#pragma unroll
        for (int ch = 0; ch < 3; ch++)
            run[ch] += g[ch][j];
        const bool ends = (j == kGradRows - 1 || src[j + 1] != k) && k != INT32_MIN;
        if (ends && k < 0) {
#pragma unroll
            for (int ch = 0; ch < 3; ch++)
                sky_acc += run[ch] * sky_base(ch);
        }
        const bool flush = ends && k >= 0;
        if (__ballot(flush)) {
            float val[4][3];
            uint32_t key[4] = {kNoKey, kNoKey, kNoKey, kNoKey};
#pragma unroll
            for (int t = 0; t < 4; t++)
                val[t][0] = val[t][1] = val[t][2] = 0.0f;
            if (flush) {
                if constexpr (FWD)
                    sample_grad(a, e, run, key, val);
                else
                    sample_terms<P2>(a, k, col, plane, run, key, val);
            }
            const bool last = segment_sum_quad(key, val);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                bool inserted = false;
                if (last && key[t] != kNoKey)
                    inserted = table_add(keys, vals, &fill, a.gcp, plane, key[t], val[t]);
                /* the slot count is kept per wave, not per insertion: one
                 * LDS atomic instead of one per inserting lane */
                const uint64_t ins = __ballot(inserted);
                if (ins && (threadIdx.x & 63) == 0)
                    atomicAdd(&fill, (uint32_t)__popcll(ins));
            }
        }
        if (ends) {
            run[0] = run[1] = run[2] = 0.0f;
        }
    }
    /* per-workgroup partials, summed in a fixed order by the next launch */
    __shared__ float wsky[kGradThreads / 64], wloss[kGradThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sky_acc += __shfl_xor(sky_acc, off);
        if (FWD)
            loss_acc += __shfl_xor(loss_acc, off);
    }
    if ((threadIdx.x & 63) == 0) {
        wsky[threadIdx.x >> 6] = sky_acc;
        if (FWD)
            wloss[threadIdx.x >> 6] = loss_acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int32_t blk = blockIdx.y * gridDim.x + blockIdx.x;
        float t = 0.0f, tl = 0.0f;
        for (int w = 0; w < kGradThreads / 64; w++) {
            t += wsky[w];
            if (FWD)
                tl += wloss[w];
        }
        if (FWD) {
            a.sky_part[blk] = t;
            a.partials[blk] = tl;
        } else {
            a.partials[blk] = t;
        }
    }
    for (int32_t i = threadIdx.x; i < kGradCap; i += kGradThreads) {
        const uint32_t key = keys[i];
        if (key == kNoKey)
            continue;
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float v = vals[ch * kGradCap + i];
            if (v != 0.0f)
                unsafeAtomicAdd(&a.gcp[ch * plane + key], v);
        }
    }
}

/* The fused step's two sums: its loss partials into *loss and its sky
 * partials into *gsky (= or +=), each in block_sum_partials' order (so
 * bit-identical to two calls of it), the two arrays' loads, shuffles and
 * final adds interleaved so that their latencies overlap. */
__global__ __launch_bounds__(1024) void k_sum_step(const float *__restrict__ loss_part,
                                                   const float *__restrict__ sky_part, int32_t n,
                                                   float *__restrict__ loss, int32_t loss_acc,
                                                   float *__restrict__ gsky, int32_t sky_acc)
{
    constexpr int NT = 1024;
    __shared__ float wsum[2][NT / 64];
    float a = 0.0f, b = 0.0f;
    for (int32_t base = 0; base < n; base += 8 * NT) {
        float va[8], vb[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int32_t i = base + u * NT + (int32_t)threadIdx.x;
            va[u] = i < n ? loss_part[i] : 0.0f;
            vb[u] = i < n ? sky_part[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            a += va[u];
            b += vb[u];
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if ((threadIdx.x & 63) == 0) {
        wsum[0][threadIdx.x >> 6] = a;
        wsum[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0 || threadIdx.x == 64) {
        const int k = threadIdx.x >> 6;
        float t = 0.0f;
        for (int w = 0; w < NT / 64; w++)
            t += wsum[k][w];
        float *dst = k ? gsky : loss;
        *dst = (k ? sky_acc : loss_acc) ? *dst + t : t;
    }
}

/* L2 loss of the fitting harness: g = 2 * (img - target) and one partial
 * sum((img - target)^2) per workgroup. */
__global__ __launch_bounds__(256) void k_l2_loss(const float *__restrict__ img,
                                                 const float *__restrict__ tgt, size_t n,
                                                 float *__restrict__ g, float *__restrict__ part)
{
    __shared__ float wpart[4];
    float acc = 0.0f;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t n4 = n / 4;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        float4 a = ((const float4 *)img)[i], b = ((const float4 *)tgt)[i];
        float4 d = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
        ((float4 *)g)[i] = make_float4(2.0f * d.x, 2.0f * d.y, 2.0f * d.z, 2.0f * d.w);
        acc += ((d.x * d.x) + (d.y * d.y)) + ((d.z * d.z) + (d.w * d.w));
    }
    for (size_t i = n4 * 4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float d = img[i] - tgt[i];
        g[i] = 2.0f * d;
        acc += d * d;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0)
        wpart[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.x] = (wpart[0] + wpart[1]) + (wpart[2] + wpart[3]);
}

/* One Adam step (no weight decay, bias-corrected moments) on n f32
 * parameters, then the parameters clamped to [lo, hi]: the fitting loop's
 * optimizer step and its projection back into the relaxation's domain in
 * one pass over parameter, gradient and moments (float4 where aligned).
 *   m = b1 m + (1 - b1) g,  v = b2 v + (1 - b2) g^2,
 *   p = clamp(p - (lr / (1 - b1^t)) m / (sqrt(v) / sqrt(1 - b2^t) + eps)) */
struct AdamArgs {
    float lr_t;          /* lr / (1 - b1^t) */
    float rbc2;          /* 1 / sqrt(1 - b2^t) */
    float b1, b2, omb1, omb2, eps, lo, hi;   /* omb = 1 - b, rounded once from double */
};

__device__ __forceinline__ float adam_one(float &p, float g, float &m, float &v,
                                          const AdamArgs &c)
{
    m = (c.b1 * m) + (c.omb1 * g);
    v = (c.b2 * v) + (c.omb2 * (g * g));
    /* the hardware sqrt and reciprocal (within an ulp or two): an
     * optimizer step has no bit-exact counterpart, and the correctly
     * rounded sequences would make this memory-bound pass VALU-bound */
    const float denom = (__builtin_amdgcn_sqrtf(v) * c.rbc2) + c.eps;
    p = fminf(fmaxf(p - (c.lr_t * (m * __builtin_amdgcn_rcpf(denom))), c.lo), c.hi);
    return p;
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v,
                                              size_t n, AdamArgs c)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t n4 = n / 4;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        float4 pp = ((float4 *)p)[i], mm = ((float4 *)m)[i], vv = ((float4 *)v)[i];
        const float4 gg = ((const float4 *)g)[i];
        adam_one(pp.x, gg.x, mm.x, vv.x, c);
        adam_one(pp.y, gg.y, mm.y, vv.y, c);
        adam_one(pp.z, gg.z, mm.z, vv.z, c);
        adam_one(pp.w, gg.w, mm.w, vv.w, c);
        ((float4 *)p)[i] = pp;
        ((float4 *)m)[i] = mm;
        ((float4 *)v)[i] = vv;
    }
    for (size_t i = n4 * 4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float pp = p[i], mm = m[i], vv = v[i];
        adam_one(pp, g[i], mm, vv, c);
        p[i] = pp;
        m[i] = mm;
        v[i] = vv;
    }
}

__global__ __launch_bounds__(1024) void k_sum_partials(const float *__restrict__ part, int32_t n,
                                                       float *__restrict__ dst, int32_t accumulate)
{
    block_sum_partials<1024>(part, n, dst, accumulate);
}

/* A frame of one colour (the degenerate frames of the state's #math mode). */
__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

/* Zero fill (a kernel, not a memset node, so that callers may capture it). */
__global__ void k_zero_f32(float *__restrict__ p, size_t n)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = 0.0f;
}

/* Shade factor of the relaxation: the same sun-ray count as k_shade,
 * then sf = sqrt(1 / (4*count + 1)). */
__global__ void k_shade_factor(const uint16_t *__restrict__ hp, int32_t hrows, int32_t hcols,
                               float s0, float s1, float s2, int32_t size, float *__restrict__ sf)
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
    sf[(size_t)y * size + x] = __builtin_sqrtf(1.0f / ((step_size * (float)count) + 1.0f));
}

inline bool pow2(int32_t v)
{
    return v > 0 && (v & (v - 1)) == 0;
}

/* Host: u32.f32(sqrtf(t) * 255) for the f32 with bits tb (sqrtf is
 * correctly rounded, no contraction: this file is built with
 * -ffp-contract=off). */
uint32_t q_exact(uint32_t tb)
{
    float t;
    memcpy(&t, &tb, sizeof t);
    return (uint32_t)(sqrtf(t) * 255.0f);
}

/* The bucket table of mix_q (see kQLo): base byte per bucket and the offset
 * of its single step, found by bisection. */
void qtab_build(uint32_t *qt)
{
    for (int i = 0; i < kQBuckets; i++) {
        const uint32_t lo = kQLo + ((uint32_t)i << 16);
        const uint32_t hi = std::min(lo + 0xFFFFu, 0x3F800000u);
        const uint32_t base = q_exact(lo);
        uint32_t step = 0x10000u;
        if (q_exact(hi) != base) {
            uint32_t a = lo, b = hi;          /* q(a) == base < q(b) */
            while (b - a > 1) {
                const uint32_t m = a + (b - a) / 2;
                if (q_exact(m) > base)
                    b = m;
                else
                    a = m;
            }
            step = b - lo;
        }
        qt[i] = (base << 17) + 0x20000u - step - ((uint32_t)i << 16);
    }
}

std::mutex g_init_mu;
bool g_init_done[64];
float *g_scratch_ptr[64];    /* device address of `scratch` on each device */

}  // namespace

extern "C" {

int fsk_device_init(void)
{
    static uint32_t host_qt[kQBuckets];
    static bool built = false;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= 64)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_init_mu);
    if (!built) {
        qtab_build(host_qt);
        built = true;
    }
    if (!g_init_done[dev]) {
        e = hipMemcpyToSymbol(HIP_SYMBOL(g_qtab), host_qt, sizeof host_qt, 0,
                              hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return e;
        /* this device's copy of the scratch array: its address, not the
         * array's host-side name, is what kernels and copies may use */
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[dev] = (float *)p;
        g_init_done[dev] = true;
    }
    return hipSuccess;
}

float *fsk_scratch(int32_t device, int64_t *n_floats)
{
    if (n_floats)
        *n_floats = kScratchFloats;
    if (device < 0 || device >= 64)
        return nullptr;
    std::lock_guard<std::mutex> lock(g_init_mu);
    return g_init_done[device] ? g_scratch_ptr[device] : nullptr;
}

int32_t fsk_tile_columns(int32_t max_rows, const fsk_tuning *t)
{
    return resolve_shape(max_rows, *t).tc;
}

int32_t fsk_batch_grid(int32_t max_rows, const fsk_tuning *t)
{
    Shape sh = resolve_shape(max_rows, *t);
    return cu_count() * sh.wg;
}

void fsk_default_tuning(fsk_tuning *t)
{
    const char *e = getenv("FSX_RENDER_TC");
    t->tc = e ? atoi(e) : 0;
    e = getenv("FSX_RENDER_WG_PER_CU");
    t->wg_per_cu = e ? atoi(e) : 0;
    e = getenv("FSX_RENDER_CHUNK");
    t->chunk = e ? atoi(e) : 0;
}

int fsk_batch_run(const fsk_batch *b, hipStream_t s)
{
    if (b->max_rows > kMaxRows || b->nqueues < 1 || b->nqueues > 64)
        return hipErrorInvalidValue;
    if (b->hbits == 8)
        return b->pow2 ? launch_batch<8, true, true>(b, s) : launch_batch<8, false, false>(b, s);
    return b->pow2 ? launch_batch<16, true, true>(b, s) : launch_batch<16, false, false>(b, s);
}

int fsk_render_slices_clear(const fsk_map *m, const fsk_frame *f, const fsk_tuning *t,
                            int32_t *d_src, int64_t pitch, float *d_zero, size_t zero_n,
                            int32_t *cleared, hipStream_t s)
{
    if (cleared)
        *cleared = 0;
    if (f->rows <= 0 || f->cols <= 0)
        return hipSuccess;
    RenderArgs a;
    a.m = *m;
    a.f = *f;
    a.f.tame = fsk_frame_tame(f);
    if (a.f.n < 0)
        a.f.n = 0;
    a.out = (uint32_t *)d_src;
    a.pitch = pitch;
    a.ntiles = 0;
    a.chunk_max = kChunk;
    a.zero = nullptr;
    a.zero_n4 = 0;
    a.nzero = 0;
    /* float4 stores only */
    const bool z = d_zero && cleared && (((uintptr_t)d_zero) & 15u) == 0 && zero_n % 4 == 0;
    a.zero = z ? (float4 *)d_zero : nullptr;
    a.zero_n4 = z ? zero_n / 4 : 0;
    a.nzero = 0;
    int32_t nz = 0;
    bool hp2 = pow2(m->hrows) && pow2(m->hcols);
    hipError_t e;
    if (m->hbits == 8)
        e = hp2 ? launch_render<8, true, false, kOutSlice>(a, *t, s, &nz)
                : launch_render<8, false, false, kOutSlice>(a, *t, s, &nz);
    else
        e = hp2 ? launch_render<16, true, false, kOutSlice>(a, *t, s, &nz)
                : launch_render<16, false, false, kOutSlice>(a, *t, s, &nz);
    if (e == hipSuccess && cleared)
        *cleared = nz > 0;
    return e;
}

int fsk_render_slices(const fsk_map *m, const fsk_frame *f, const fsk_tuning *t, int32_t *d_src,
                      int64_t pitch, hipStream_t s)
{
    return fsk_render_slices_clear(m, f, t, d_src, pitch, nullptr, 0, nullptr, s);
}

static DiffArgs diff_args(const fsk_frame *f, const float *cp, const float *sf, int32_t hc,
                          int32_t wc, const int32_t *src)
{
    DiffArgs a;
    a.f = *f;
    a.f.tame = fsk_frame_tame(f);
    a.cp = cp;
    a.sf = sf;
    a.hc = hc;
    a.wc = wc;
    a.sky_scale = 0.0f;
    a.src = src;
    a.gimg = nullptr;
    a.img = nullptr;
    a.gcp = nullptr;
    a.gsky = nullptr;
    a.partials = nullptr;
    a.tgt = nullptr;
    a.gout = nullptr;
    a.sky_part = nullptr;
    return a;
}

int fsk_render_nearest(const fsk_map *m, const fsk_frame *f, const fsk_tuning *t,
                       uint32_t *d_out, int64_t pitch, hipStream_t s)
{
    if (f->rows <= 0 || f->cols <= 0)
        return hipSuccess;
    RenderArgs a;
    a.m = *m;
    a.f = *f;
    a.f.tame = fsk_frame_tame(f);
    if (a.f.n < 0)
        a.f.n = 0;
    a.out = d_out;
    a.pitch = pitch;
    a.ntiles = 0;
    a.chunk_max = kChunk;
    a.zero = nullptr;
    a.zero_n4 = 0;
    a.nzero = 0;
    return pow2(m->hrows) && pow2(m->hcols) ? launch_render<8, true, false, kOutNearest>(a, *t, s)
                                            : launch_render<8, false, false, kOutNearest>(a, *t, s);
}

int fsk_diff_color(const fsk_frame *f, const float *d_cp, const float *d_sf, int32_t hc,
                   int32_t wc, float sky_scale, const int32_t *d_src, float *d_img, hipStream_t s)
{
    if (f->rows <= 0 || f->cols <= 0)
        return hipSuccess;
    DiffArgs a = diff_args(f, d_cp, d_sf, hc, wc, d_src);
    a.sky_scale = sky_scale;
    a.img = d_img;
    dim3 grid((f->cols + 63) / 64, (f->rows + 3) / 4);
    const bool p2 = pow2(hc) && pow2(wc), tame = a.f.tame != 0;
    void (*kern)(DiffArgs) = p2 ? (tame ? k_diff_color<true, true> : k_diff_color<true, false>)
                                : (tame ? k_diff_color<false, true> : k_diff_color<false, false>);
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

int fsk_l2_loss(const float *d_img, const float *d_tgt, size_t n, float *d_g, float *d_loss,
                int32_t accumulate, float *d_partials, int32_t n_partials, hipStream_t s)
{
    size_t blocks = (n / 4 + 255) / 256;
    if (blocks > 4096)
        blocks = 4096;
    if (blocks == 0)
        blocks = 1;
    if ((int64_t)blocks > n_partials)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_l2_loss, dim3((unsigned)blocks), dim3(256), 0, s, d_img, d_tgt, n, d_g,
                       d_partials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, d_partials, (int32_t)blocks,
                       d_loss, accumulate);
    return hipGetLastError();
}

int fsk_adam(float *d_p, const float *d_g, float *d_m, float *d_v, size_t n, double lr,
             double beta1, double beta2, double eps, int32_t step, float lo, float hi,
             hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    /* the hyper-parameters' derived constants in double, each rounded to
     * f32 once */
    AdamArgs c;
    c.lr_t = (float)(lr / (1.0 - pow(beta1, (double)step)));
    c.rbc2 = (float)(1.0 / sqrt(1.0 - pow(beta2, (double)step)));
    c.b1 = (float)beta1;
    c.b2 = (float)beta2;
    c.omb1 = (float)(1.0 - beta1);
    c.omb2 = (float)(1.0 - beta2);
    c.eps = (float)eps;
    c.lo = lo;
    c.hi = hi;
    size_t blocks = (n / 4 + 255) / 256;
    if (blocks > 16384)
        blocks = 16384;
    if (blocks == 0)
        blocks = 1;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks), dim3(256), 0, s, d_p, d_g, d_m, d_v, n, c);
    return hipGetLastError();
}

int fsk_fill_u32(uint32_t *d_p, size_t n, uint32_t v, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, d_p, n, v);
    return hipGetLastError();
}

int fsk_zero_f32(float *d_p, size_t n, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)blocks), dim3(256), 0, s, d_p, n);
    return hipGetLastError();
}

int fsk_diff_grad(const fsk_frame *f, const float *d_cp, const float *d_sf, int32_t hc,
                  int32_t wc, const int32_t *d_src, const float *d_gimg, float *d_gcp,
                  float *d_gsky, int32_t sky_accumulate, float *d_partials, int32_t n_partials,
                  hipStream_t s)
{
    if (f->rows <= 0 || f->cols <= 0)
        return hipSuccess;
    DiffArgs a = diff_args(f, d_cp, d_sf, hc, wc, d_src);
    a.gimg = d_gimg;
    a.gcp = d_gcp;
    a.gsky = d_gsky;
    a.partials = d_partials;
    const int32_t br = fsk_grad_block_rows(f->cols, f->rows);
    dim3 grid((f->cols + 63) / 64, (f->rows + br - 1) / br);
    if ((int64_t)grid.x * grid.y > n_partials)
        return hipErrorInvalidValue;
    const bool p2 = pow2(hc) && pow2(wc), tall = br > 16;
    void (*kern)(DiffArgs) = p2 ? (tall ? k_diff_grad<true, false, 4> : k_diff_grad<true, false, 2>)
                                : (tall ? k_diff_grad<false, false, 4> : k_diff_grad<false, false, 2>);
    hipLaunchKernelGGL(kern, grid, dim3(kGradThreads), 0, s, a);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, d_partials,
                       (int32_t)(grid.x * grid.y), d_gsky, sky_accumulate);
    return hipGetLastError();
}

int fsk_diff_step_l2(const fsk_frame *f, const float *d_cp, const float *d_sf, int32_t hc,
                     int32_t wc, float sky_scale, const int32_t *d_src, const float *d_tgt,
                     float *d_img, float *d_g, float *d_gcp, float *d_gsky,
                     int32_t sky_accumulate, float *d_loss, int32_t loss_accumulate,
                     float *d_loss_part, float *d_sky_part, int32_t n_partials, hipStream_t s)
{
    if (f->rows <= 0 || f->cols <= 0)
        return hipSuccess;
    DiffArgs a = diff_args(f, d_cp, d_sf, hc, wc, d_src);
    a.sky_scale = sky_scale;
    a.tgt = d_tgt;
    a.img = d_img;
    a.gout = d_g;
    a.gcp = d_gcp;
    a.gsky = d_gsky;
    a.partials = d_loss_part;
    a.sky_part = d_sky_part;
    const int32_t br = fsk_grad_block_rows(f->cols, f->rows);
    dim3 grid((f->cols + 63) / 64, (f->rows + br - 1) / br);
    const int64_t nb = (int64_t)grid.x * grid.y;
    if (nb > n_partials)
        return hipErrorInvalidValue;
    const bool p2 = pow2(hc) && pow2(wc), tall = br > 16;
    void (*kern)(DiffArgs) = p2 ? (tall ? k_diff_grad<true, true, 4> : k_diff_grad<true, true, 2>)
                                : (tall ? k_diff_grad<false, true, 4> : k_diff_grad<false, true, 2>);
    hipLaunchKernelGGL(kern, grid, dim3(kGradThreads), 0, s, a);
    hipLaunchKernelGGL(k_sum_step, dim3(1), dim3(1024), 0, s, d_loss_part, d_sky_part,
                       (int32_t)nb, d_loss, loss_accumulate, d_gsky, sky_accumulate);
    return hipGetLastError();
}

int fsk_shade_factor(const uint16_t *d_hplain, int32_t hrows, int32_t hcols, float sun0,
                     float sun1, float sun2, int32_t size, float *d_sf, hipStream_t s)
{
    dim3 grid((size + 255) / 256, size);
    hipLaunchKernelGGL(k_shade_factor, grid, dim3(256), 0, s, d_hplain, hrows, hcols, sun0, sun1,
                       sun2, size, d_sf);
    return hipGetLastError();
}

int fsk_render2(const fsk_map *m, const fsk_frame *f, const float *d_cs, const float *d_sn,
                const uint32_t *d_color, int32_t n, uint32_t *d_out, int64_t pitch, hipStream_t s)
{
    if (f->rows <= 0 || f->cols <= 0)
        return hipSuccess;
    Render2Args a;
    a.m = *m;
    a.color = d_color;
    a.cs = d_cs;
    a.sn = d_sn;
    a.out = d_out;
    a.pitch = pitch;
    a.cam_x = f->cam_x;
    a.cam_y = f->cam_y;
    a.cam_h = f->cam_h;
    a.horizon = f->horizon;
    a.n = n;
    a.rows = f->rows;
    a.cols = f->cols;
    dim3 grid((f->cols + 63) / 64, (f->rows + 3) / 4);
    const bool p2 = pow2(m->hrows) && pow2(m->hcols);
    void (*kern)(Render2Args) = m->hbits == 8 ? (p2 ? k_render2<8, true> : k_render2<8, false>)
                                              : (p2 ? k_render2<16, true> : k_render2<16, false>);
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

int32_t fsk_frame_tame(const fsk_frame *f)
{
    const float v[] = {f->cam_x, f->cam_y, f->cam_h, f->horizon, f->Lx, f->Ly,
                       f->Rx,    f->Ry,    f->d,     f->z0,      f->scale, f->fw};
    for (float x : v)
        if (!std::isfinite(x))
            return 0;
    if (!(f->z0 > 0.0f) || !(f->d >= 0.0f) || !(f->scale > 0.0f) || !(f->fw >= 1.0f) ||
        f->n <= 0)
        return 0;
    /* z_k = (k/2)(2 z0 + (k-1) d) is non-decreasing from z_1 = z0 (every
     * f32 step of get_zs' term is monotone for z0 > 0, d >= 0), so
     * inv_z <= scale / z0; a 1e-3 margin covers the roundings */
    const double n = (double)f->n, margin = 1.001;
    const double zmax = (n / 2.0) * (2.0 * (double)f->z0 + (n - 1.0) * (double)f->d) * margin;
    const double izmax = (double)f->scale / (double)f->z0 * margin;
    /* x = (lx + cam_x) + col * (rx - lx) / w over columns up to w + 7 (the
     * last tile's lanes past the frame) */
    const double span = ((double)f->fw + 8.0) / (double)f->fw;
    const double ax = std::fabs((double)f->Lx), bx = std::fabs((double)f->Rx);
    const double ay = std::fabs((double)f->Ly), by = std::fabs((double)f->Ry);
    const double xmax = (std::fabs((double)f->cam_x) + (ax + (ax + bx) * span) * zmax) * margin;
    const double ymax = (std::fabs((double)f->cam_y) + (ay + (ay + by) * span) * zmax) * margin;
    const double relmax =
        ((std::fabs((double)f->cam_h) + 65536.0) * izmax + std::fabs((double)f->horizon)) * margin;
    /* slice depths inside [2^-100, 2^100], where the table's reciprocal
     * (rcp_rn) is exact */
    const bool zok = (double)f->z0 >= 0x1p-100 && zmax <= 0x1p100;
    return zok && xmax < 4194304.0 && ymax < 4194304.0 && relmax < 1073741824.0;
}

int32_t fsk_render_max_rows(void)
{
    return kMaxFrameRows;
}

int32_t fsk_batch_max_rows(void)
{
    return kMaxRows;
}

int32_t fsk_grad_block_rows(int32_t w, int32_t h)
{
    return (kGradThreads / 64) * ((int64_t)w * (int64_t)h >= kGradTallPixels ? 4 : 2);
}

int fsk_selftest(int which, unsigned long long *d_bad, hipStream_t s)
{
    if (which < 0 || which > 7)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_selftest, dim3(4096), dim3(256), 0, s, which, d_bad);
    return hipGetLastError();
}

int fsk_render(const fsk_map *m, const fsk_frame *f, const fsk_tuning *t, uint32_t *d_out,
               int64_t pitch, hipStream_t s)
{
    if (f->rows <= 0 || f->cols <= 0)
        return hipSuccess;
    RenderArgs a;
    a.m = *m;
    a.f = *f;
    a.f.tame = fsk_frame_tame(f);
    if (a.f.n < 0)
        a.f.n = 0;
    a.out = d_out;
    a.pitch = pitch;
    a.ntiles = 0;
    a.chunk_max = kChunk;
    a.zero = nullptr;
    a.zero_n4 = 0;
    a.nzero = 0;
    bool hp2 = pow2(m->hrows) && pow2(m->hcols);
    bool cp2 = pow2(m->crows) && pow2(m->ccols);
    hipError_t e;
    if (m->hbits == 8) {
        if (hp2 && cp2) e = launch_render<8, true, true>(a, *t, s);
        else if (hp2) e = launch_render<8, true, false>(a, *t, s);
        else if (cp2) e = launch_render<8, false, true>(a, *t, s);
        else e = launch_render<8, false, false>(a, *t, s);
    } else {
        if (hp2 && cp2) e = launch_render<16, true, true>(a, *t, s);
        else if (hp2) e = launch_render<16, true, false>(a, *t, s);
        else if (cp2) e = launch_render<16, false, true>(a, *t, s);
        else e = launch_render<16, false, false>(a, *t, s);
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
