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

/* The same over the pixel rectangle [x0, x1) x [y0, y1) of the W x H frame
 * (buffers are still W x H). */
int ko_render_region(const ko_scene* s, const kf_camera_ubo* cam, const kf_push_constants* pc, uint32_t seed,
                     uint32_t width, uint32_t height, uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1,
                     int precision, int threads, float* rgba, float* albedo, float* normal, float* samples,
                     uint64_t* counts);

/* relMSE = mean over pixels and RGB of (x - y)^2 / (y^2 + 1e-2) (FP64 sum). */
double ko_relmse(const float* x, const float* y, uint32_t n_pixels, uint32_t stride_x, uint32_t stride_y);

/* RGBA32F -> B8G8R8A8_SRGB bytes (clamp, IEC 61966-2-1 encode, round). */
void ko_to_bgra8_srgb(const float* rgba, uint32_t n_pixels, uint8_t* out);

/* Unit access to the shading terms, so that tests can pin each against the
 * closed forms the mathematics fixes.  precision 32 or 64; n rows of inputs
 * (row-major, doubles, converted to the precision) -> n rows of outputs.
 *   KO_SCHLICK          (cos, ior)              -> 1   Random.glsl:68-73
 *   KO_GGX_D            (NdotH, a2)             -> 1   Random.glsl:75-79
 *   KO_G1               (x, a2)                 -> 1   Random.glsl:117-121
 *   KO_GEOM_SHADOW      (NdotV, NdotL, a2)      -> 1   Random.glsl:123-127
 *   KO_SCHLICK_FRESNEL  (f0, lDotH)             -> 1   Random.glsl:133-136 (scalar f0)
 *   KO_REFRACT          (I xyz, N xyz, eta)     -> 3   GLSL refract
 *   KO_REFLECT          (I xyz, N xyz)          -> 3   GLSL reflect
 *   KO_LOCAL_TO_WORLD   (d xyz, n xyz)          -> 3   Sampling.glsl:6-23
 *   KO_PERPENDICULAR    (u xyz)                 -> 3   Random.glsl:58-65
 * Returns 0, or -1 for an unknown term or precision. */
enum {
    KO_SCHLICK = 0, KO_GGX_D = 1, KO_G1 = 2, KO_GEOM_SHADOW = 3, KO_SCHLICK_FRESNEL = 4,
    KO_REFRACT = 5, KO_REFLECT = 6, KO_LOCAL_TO_WORLD = 7, KO_PERPENDICULAR = 8
};
int ko_shading_term(int term, int precision, const double* in, uint32_t n, double* out);

/* n successive draws of a sampler from one rnd stream starting at seed:
 *   KO_SAMPLE_COSINE  param N xyz         -> 3   Sampling.glsl:25-35
 *   KO_SAMPLE_GGX     param a2, N xyz     -> 3   Random.glsl:83-101 (half vector)
 *   KO_SAMPLE_SPHERE  no param            -> 3   Sampling.glsl:72-83 (point in the unit ball)
 *   KO_SAMPLE_DISK    no param            -> 2   Sampling.glsl:124-134
 * *draws receives the number of rnd() calls the n samples consumed. */
enum { KO_SAMPLE_COSINE = 0, KO_SAMPLE_GGX = 1, KO_SAMPLE_SPHERE = 2, KO_SAMPLE_DISK = 3 };
int ko_sample(int sampler, int precision, uint32_t seed, const double* param, uint32_t n, double* out,
              uint64_t* draws);

/* calcDirectContribution (PathTrace.rchit:103-171) for one shading point:
 * shade = N xyz, V xyz, L xyz, f, a2, diffuseColor rgb, specularColor rgb,
 * transmissionColor rgb, inside (0/1), lightEmission rgb (24 doubles).
 * out = the returned rgb; *seed is the ray seed, advanced by the draws. */
int ko_direct(int precision, const double* shade, uint32_t* seed, double* out);

#ifdef __cplusplus
}
#endif
#endif
