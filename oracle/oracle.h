/*
 * oracle.h — CPU oracle of the path-tracing hot path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so, and only as the checker or the timed CPU baseline.  The
 * product library (libkuafu_amd.so) never links, loads or calls it.
 *
 * It restates the reference's GLSL shaders (resources/shaders/PathTrace.*,
 * base/ shader sources) on the CPU:
 *   - FP32 flavour: same operations in the same order as the HIP kernels, with
 *     the oracle's own deterministic sin/cos/exp2 -> per-sample parity;
 *   - FP64 flavour: libm math -> converged (high-spp) reference images.
 * Layouts come from include/kuafu_amd.h (the reference's std430 layouts).
 */
#ifndef KUAFU_ORACLE_H
#define KUAFU_ORACLE_H
#include "../include/kuafu_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RNG and math, batched for known-answer tests. */
void ko_tea_batch(const uint32_t* a, const uint32_t* b, uint32_t n, uint32_t* out);
void ko_rnd_sequence(uint32_t seed, uint32_t n, float* out, uint32_t* final_state);
void ko_sincos_batch(const float* x, uint32_t n, float* s, float* c);
void ko_exp2_batch(const float* x, uint32_t n, float* out);

/* Single watertight ray/triangle test (Woo, Benthin, Wald, JCGT 2013).
 * Returns 1 on hit and fills t, b1, b2, back. */
int ko_intersect_triangle(const float org[3], const float dir[3], float tmin, float tmax,
                          const float v0[3], const float v1[3], const float v2[3],
                          float* t, float* b1, float* b2, uint32_t* back);

/* world -> object 3x4 rows of an instance transform (double cofactors). */
void ko_world_to_object(const float transform[16], float out_rows[12]);

typedef struct ko_scene ko_scene;
ko_scene* ko_scene_create(const kf_scene_desc* d);
void ko_scene_destroy(ko_scene* s);

/* Closest / any hit for a ray batch.  brute = 1 loops over every triangle of
 * every instance; brute = 0 uses the oracle's own CPU BVH. */
void ko_trace_closest(const ko_scene* s, const kf_ray* rays, uint32_t n, kf_hit* hits, int brute,
                      int threads);
void ko_trace_any(const ko_scene* s, const kf_ray* rays, uint32_t n, uint8_t* occluded, int brute,
                  int threads);

/* One frame of PathTrace.rgen over rows [y0, y1).  precision 32 or 64.
 * rgba/albedo/normal: W*H*4 floats (rgba is the running mean, read and
 * written).  samples: NULL or spp*W*H*3 floats (sample-major).
 * counts: NULL or 3 x uint64 (primary, extension, shadow rays).
 * Returns 0, or -1 on bad arguments. */
int ko_render(const ko_scene* s, const kf_camera_ubo* cam, const kf_push_constants* pc, uint32_t seed,
              uint32_t width, uint32_t height, uint32_t y0, uint32_t y1, int precision, int threads,
              float* rgba, float* albedo, float* normal, float* samples, uint64_t* counts);

/* relMSE = mean over pixels and RGB of (x - y)^2 / (y^2 + 1e-2) (FP64 sum). */
double ko_relmse(const float* x, const float* y, uint32_t n_pixels, uint32_t stride_x, uint32_t stride_y);

/* RGBA32F -> B8G8R8A8_SRGB bytes (clamp, IEC 61966-2-1 encode, round). */
void ko_to_bgra8_srgb(const float* rgba, uint32_t n_pixels, uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif
