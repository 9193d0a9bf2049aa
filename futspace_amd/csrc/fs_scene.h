/*
 * fs_scene.h -- seeded synthetic Voxel Space scenes.
 *
 * Inputs for tests and benchmarks: toroidal diamond-square heightmaps,
 * height-banded ARGB colormaps and camera paths.  Random numbers come
 * from SplitMix64 (seeding) feeding per-stream PCG32 generators, so every
 * scene is a pure function of (seed, size, parameters) and identical on
 * every x86-64 host.  No reference counterpart: the reference loads PNG
 * maps (c/interactive.c:25-57).
 */
#ifndef FS_SCENE_H
#define FS_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One camera of a path: x, y, height, angle, horizon. */
typedef struct {
    float x, y, height, angle, horizon;
} fss_pose;

/* Diamond-square heightmap on a size x size torus (size a power of two),
 * displacement amplitude multiplied by `roughness` per level, rescaled to
 * integers 0..max_value.  Returns 0 on success. */
int fss_heightmap(uint64_t seed, int32_t size, double roughness, int32_t max_value,
                  int32_t *out);

/* Height-banded palette (water, sand, grass, rock, snow) plus uniform
 * +-noise per channel, alpha 0xFF.  `heights` holds 0..max_value. */
int fss_colormap(uint64_t seed, const int32_t *heights, int32_t h, int32_t w,
                 int32_t max_value, int32_t noise, uint32_t *out);

/* Closed Catmull-Rom fly-through over a map_size torus: n_ctrl control
 * points with uniform positions and heights in [alt_lo, alt_hi], yaw
 * sweeping 2*pi over the n frames, horizon = horizon_mid + horizon_amp *
 * sin(4*pi*t + phase). */
int fss_flythrough(uint64_t seed, int32_t n, int32_t n_ctrl, int32_t map_size,
                   float alt_lo, float alt_hi, float horizon_mid, float horizon_amp,
                   fss_pose *out);

/* Uniform integer in [lo, hi] and float in [lo, hi) from a stream; used by
 * callers that draw batched view parameters. */
typedef struct {
    uint64_t state, inc;
} fss_rng;
void     fss_rng_seed(fss_rng *r, uint64_t seed, uint64_t stream);
uint32_t fss_rng_u32(fss_rng *r);
int32_t  fss_rng_int(fss_rng *r, int32_t lo, int32_t hi);
double   fss_rng_unit(fss_rng *r);

#ifdef __cplusplus
}
#endif
#endif
