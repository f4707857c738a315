/*
 * kuafu_amd.h — C ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * The reference exposes a C++ API only (kuafu.hpp:31-93); the C++ facade in
 * include/kuafu.hpp keeps that API.  This C ABI is the seam underneath it:
 * plain pointers and sizes, no C++ or torch types, so that ctypes / cgo / JNI
 * harnesses (and the parity tests) can drive the same code the facade drives.
 *
 * Every POD struct below reproduces a std430 / host layout of the reference
 * byte for byte, so that a scene uploaded by the facade and a scene handed in
 * through this ABI are the same bytes in HBM.
 *
 * Error convention: functions return 0 on success and a negative code on
 * failure; kf_last_error() returns the message.  The C++ facade turns a
 * negative code into std::runtime_error, the reference's convention
 * (stdafx.hpp:61-67).
 */
#ifndef KUAFU_AMD_H
#define KUAFU_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Layouts shared with the reference                                         */
/* ------------------------------------------------------------------------ */

/* Vertex, 48 B: vertex.hpp:28-34; unpacked as 3 x vec4 in PathTrace.rchit:52-64. */
typedef struct kf_vertex {
    float pos[3];
    float normal[3];
    float color[3];
    float uv[2];
    float pad;
} kf_vertex;

/* NiceMaterialSSBO, 80 B: geometry.hpp:70-87, Geometry.glsl:1-19. */
typedef struct kf_material {
    float diffuse[4];   /* rgb + pad */
    float emission[4];  /* rgb + strength */
    float alpha, metallic, specular, roughness, ior, transmission;
    int32_t diffuse_tex, metallic_tex, roughness_tex, transmission_tex;
    int32_t pad0, pad1;
} kf_material;

/* GeometryInstanceSSBO, 80 B: geometry.hpp:91-99.  transform is a
 * column-major glm::mat4 (object -> world). */
typedef struct kf_instance {
    float transform[16];
    uint32_t geometry_index;
    uint32_t pad[3];
} kf_instance;

/* DirectionalLightUBO, 32 B: light.hpp:41-44 (direction is normalized by the
 * host, scene.cpp:254). */
typedef struct kf_dir_light {
    float direction[4]; /* xyz + softness */
    float rgbs[4];      /* rgb + strength */
} kf_dir_light;

#define KF_MAX_POINT_LIGHTS 32
#define KF_MAX_ACTIVE_LIGHTS 8

/* PointLightsUBO, 1024 B: light.hpp:46-49. */
typedef struct kf_point_lights {
    float posr[KF_MAX_POINT_LIGHTS][4]; /* xyz + radius */
    float rgbs[KF_MAX_POINT_LIGHTS][4]; /* rgb + strength (w > 0 => active) */
} kf_point_lights;

/* ActiveLightsUBO, 1536 B: light.hpp:51-58, filled as in scene.cpp:266-287. */
typedef struct kf_active_lights {
    float view[KF_MAX_ACTIVE_LIGHTS][16];
    float proj[KF_MAX_ACTIVE_LIGHTS][16];
    float front[KF_MAX_ACTIVE_LIGHTS][4];    /* xyz + in-use flag (w > 0) */
    float rgbs[KF_MAX_ACTIVE_LIGHTS][4];     /* rgb + strength */
    float position[KF_MAX_ACTIVE_LIGHTS][4];
    float sftp[KF_MAX_ACTIVE_LIGHTS][4];     /* softness, fov, texID, pad */
} kf_active_lights;

/* CameraUBO, 320 B: camera.hpp:181-192, Camera.glsl:1-13. */
typedef struct kf_camera_ubo {
    float view[16], proj[16], view_inverse[16], proj_inverse[16];
    float position[4]; /* xyz + aperture */
    float front[4];    /* xyz + focal distance */
    float pad1[4], pad2[4];
} kf_camera_ubo;

/* RtPushConstants, 48 B: rt.hpp:31-42, PushConstants.glsl:1-16. */
typedef struct kf_push_constants {
    float clear_color[4];
    int32_t frame_count;
    uint32_t spp;
    uint32_t path_depth;
    uint32_t use_env_map;
    uint32_t russian_roulette;
    uint32_t rr_min_bounces;
    uint32_t nee;            /* pushed but unused by the shaders */
    uint32_t nee_min_bounces;
} kf_push_constants;

/* The layouts above are the contract with the reference's std430 blocks and
 * host structs: pinned at compile time for every client of this header. */
#ifdef __cplusplus
#define KF_LAYOUT_ASSERT(cond, msg) static_assert(cond, msg)
#else
#define KF_LAYOUT_ASSERT(cond, msg) _Static_assert(cond, msg)
#endif
KF_LAYOUT_ASSERT(sizeof(kf_vertex) == 48, "Vertex is 48 B (vertex.hpp:28-34)");
KF_LAYOUT_ASSERT(offsetof(kf_vertex, normal) == 12 && offsetof(kf_vertex, color) == 24 &&
                     offsetof(kf_vertex, uv) == 36, "Vertex field offsets (Geometry.glsl unpack)");
KF_LAYOUT_ASSERT(sizeof(kf_material) == 80, "NiceMaterialSSBO is 80 B (geometry.hpp:70-87)");
KF_LAYOUT_ASSERT(offsetof(kf_material, emission) == 16 && offsetof(kf_material, alpha) == 32 &&
                     offsetof(kf_material, ior) == 48 && offsetof(kf_material, diffuse_tex) == 56 &&
                     offsetof(kf_material, transmission_tex) == 68, "NiceMaterialSSBO field offsets");
KF_LAYOUT_ASSERT(sizeof(kf_instance) == 80 && offsetof(kf_instance, geometry_index) == 64,
                 "GeometryInstanceSSBO is 80 B (geometry.hpp:91-99)");
KF_LAYOUT_ASSERT(sizeof(kf_dir_light) == 32, "DirectionalLightUBO is 32 B (light.hpp:41-44)");
KF_LAYOUT_ASSERT(sizeof(kf_point_lights) == 1024 && offsetof(kf_point_lights, rgbs) == 512,
                 "PointLightsUBO is 1024 B (light.hpp:46-49)");
KF_LAYOUT_ASSERT(sizeof(kf_active_lights) == 1536 && offsetof(kf_active_lights, proj) == 512 &&
                     offsetof(kf_active_lights, front) == 1024 && offsetof(kf_active_lights, sftp) == 1408,
                 "ActiveLightsUBO is 1536 B (light.hpp:51-58)");
KF_LAYOUT_ASSERT(sizeof(kf_camera_ubo) == 320 && offsetof(kf_camera_ubo, proj_inverse) == 192 &&
                     offsetof(kf_camera_ubo, position) == 256 && offsetof(kf_camera_ubo, front) == 272,
                 "CameraUBO is 320 B (camera.hpp:181-192)");
KF_LAYOUT_ASSERT(sizeof(kf_push_constants) == 48 && offsetof(kf_push_constants, frame_count) == 16 &&
                     offsetof(kf_push_constants, nee_min_bounces) == 44,
                 "RtPushConstants is 48 B (rt.hpp:31-42)");

/* RGBA8 image, row-major, 4 bytes per texel, sRGB-encoded (vkCore.hpp:585). */
typedef struct kf_texture_desc {
    uint32_t width, height;
    const uint8_t* rgba8;
} kf_texture_desc;

enum {
    KF_GEOM_OPAQUE = 1u,   /* Geometry::isOpaque (rt.cpp:101-102) */
    KF_GEOM_HIDDEN = 2u    /* Geometry::hideRender -> empty BLAS (rt.cpp:156) */
};

enum {
    KF_ENV_CLEAR_COLOR = 0, /* rmiss:21: clearColor.rgb * clearColor.a */
    KF_ENV_CUBEMAP = 1,     /* rmiss:19: samplerCube lookup */
    KF_ENV_GRADIENT = 2     /* procedural gradient (extension, not in the reference) */
};

/* Whole-scene description.  Geometries are concatenated: geometry g owns
 * vertices [geom_vertex_offset[g], geom_vertex_offset[g+1]) and triangles
 * [geom_tri_offset[g], geom_tri_offset[g+1]); indices are local to the
 * geometry (Scene::uploadGeometries, scene.cpp:423-440). */
typedef struct kf_scene_desc {
    uint32_t num_geometries;
    const uint32_t* geom_vertex_offset; /* num_geometries + 1 */
    const uint32_t* geom_tri_offset;    /* num_geometries + 1 */
    const uint32_t* geom_flags;         /* num_geometries, KF_GEOM_* */
    const kf_vertex* vertices;
    const uint32_t* indices;            /* 3 per triangle */
    const uint32_t* mat_index;          /* 1 per triangle (global material index) */
    uint32_t num_instances;
    const kf_instance* instances;
    uint32_t num_materials;
    const kf_material* materials;
    uint32_t num_textures;
    const kf_texture_desc* textures;
    const kf_dir_light* dir_light;
    const kf_point_lights* point_lights;
    const kf_active_lights* active_lights;
    uint32_t env_kind;                  /* KF_ENV_* */
    const kf_texture_desc* env_faces;   /* 6 faces +X,-X,+Y,-Y,+Z,-Z for KF_ENV_CUBEMAP */
    float env_gradient[8];              /* bottom rgb, pad, top rgb, pad for KF_ENV_GRADIENT */
    /* Batched environments (no counterpart in the reference, which traces one
     * Scene per RayTracer::trace): instances [env_instance_offset[e],
     * env_instance_offset[e+1]) form environment e, with a TLAS of its own;
     * rays of environment e see only those instances.  Geometries, materials,
     * textures, lights and the sky are shared.  num_envs 0 or 1: one
     * environment holding every instance (env_instance_offset unused). */
    uint32_t num_envs;
    const uint32_t* env_instance_offset; /* num_envs + 1, from 0 to num_instances */
    /* 0: geometries whose triangles equal an earlier geometry's (the same
     * vertex positions at every corner, bit for bit) traverse that
     * geometry's BVH (no counterpart in the reference, which builds one BLAS
     * per geometry; results are the same, the traversal's working set is
     * smaller); 1: every geometry traverses its own BVH. */
    uint32_t separate_geometry;
} kf_scene_desc;

/* A ray of a fixed batch: origin, tmin, direction, tmax (32 B). */
typedef struct kf_ray {
    float org[3], tmin;
    float dir[3], tmax;
} kf_ray;

/* Closest-hit record.  inst = ~0u means miss.  b1/b2 are gl_HitAttributeEXT
 * (weights of vertex 1 and 2); back = 1 for gl_HitKindBackFacingTriangleEXT. */
typedef struct kf_hit {
    float t;
    uint32_t inst;
    uint32_t prim;
    float b1, b2;
    uint32_t back;
} kf_hit;

/* Per-frame render statistics, counted on the device. */
typedef struct kf_render_stats {
    uint64_t primary_rays;
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t node_visits;     /* BVH4 node records (128 B); only with the counting traversal */
    uint64_t tri_tests;       /* triangle records (48 B) */
    float trace_ms;           /* kernel wall time of the frame (HIP events) */
    float extend_ms;          /* closest-hit traversal kernels */
    float shadow_ms;          /* any-hit traversal kernels */
    float other_ms;
    uint32_t extend_launches; /* closest-hit traversal kernel launches */
    uint32_t shadow_launches; /* any-hit traversal kernel launches */
    uint64_t instance_visits; /* instance records (64 B) */
    uint64_t speculative_rays; /* next-bounce rays traced for paths Russian roulette
                                  then ended (not in extension_rays) */
    uint64_t tlas_node_visits; /* of node_visits, the TLAS ones */
    uint64_t tail_rays;        /* closest-hit rays traced by the tail kernel (in primary/extension_rays) */
    uint64_t graph_captures;   /* frame graphs this renderer has captured so far (kf_render_opts.use_graph) */
} kf_render_stats;

/* Render options that the reference takes from Config / push constants. */
typedef struct kf_render_opts {
    uint32_t seed;            /* replaces clockARB() seeding (rgen:23, 32) */
    uint32_t samples_per_batch; /* wavefront batch size in samples per pixel; 0 = auto */
    uint32_t keep_samples;    /* keep per-sample radiance for kf_renderer_read_samples */
    uint32_t count_traversal; /* count node visits / triangle tests */
    uint32_t use_graph;       /* capture the frame's launches in a HIP graph once and replay it while
                                 the scene, target, options and schedule stay the same (seed, frame
                                 count, cameras and lights may change); stats then carry the frame's
                                 kernel time only (extend_ms / shadow_ms are 0) */
    uint32_t full_tlas_camera; /* camera rays start from their environment's TLAS root instead of
                                 a root over the instances their 8x8 pixel tile can see (same
                                 images, more TLAS steps; for comparison) */
} kf_render_opts;

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */

/* Replaces Context::init's device selection (context.cpp:88-296): picks the
 * HIP device and creates the library's stream.  Idempotent. */
int kf_init(int device);
const char* kf_last_error(void);
const char* kf_version(void);

/* ------------------------------------------------------------------------ */
/* Host math used by the facade (camera.cpp, scene.cpp)                      */
/* ------------------------------------------------------------------------ */

/* CameraUBO contents for a pose + intrinsics, as Scene::uploadCameraBuffer
 * builds them (scene.cpp:236-250, camera.cpp:92-128). */
int kf_camera_ubo_compute(const float position[3], const float front[3], const float up[3],
                          int width, int height, float fx, float fy, float cx, float cy, float skew,
                          float aperture, float focal_length, kf_camera_ubo* out);

/* World -> object 3x4 (row-major) of an instance transform; the matrix the
 * traversal applies to rays (gl_WorldToObjectEXT). */
int kf_world_to_object(const float transform[16], float out_rows[12]);

/* Environment cube map from a KTX file (KTX 1 or 2; uncompressed 8-bit
 * RGBA, 6 square faces, base level), as vkCore::Cubemap::init loads it for
 * Scene::setEnvironmentMap (vkCore.hpp:1852-1970, scene.cpp:301-303): *size =
 * face width; faces (may be NULL to query the size) receives 6 * size * size
 * * 4 bytes in the order +X, -X, +Y, -Y, +Z, -Z, to be used as sRGB (the
 * reference uploads them as R8G8B8A8_SRGB whatever the file's format).
 * Host only. */
int kf_ktx_cubemap_read(const char* path, uint32_t* size, uint8_t* faces, uint64_t max_bytes);

/* Active-light UBO slot for a light, as scene.cpp:266-281 fills it. */
int kf_active_light_fill(kf_active_lights* ubo, int slot, const float view[16], const float color[3],
                         float fov, float strength, float softness, int tex_id);

/* ------------------------------------------------------------------------ */
/* Acceleration structures + traversal                                      */
/* ------------------------------------------------------------------------ */

typedef struct kf_scene kf_scene;

/* Upload the scene and build every BLAS plus the TLAS on the GPU
 * (Context::update -> uploadGeometries, createBottomLevelAS, buildTlas:
 * context.cpp:323-336, rt.cpp:142-495). */
int kf_scene_create(const kf_scene_desc* desc, kf_scene** out);
void kf_scene_destroy(kf_scene* s);

/* Replace instance transforms and rebuild the TLAS on the GPU (the per-frame
 * RayTracer::updateTlas, rt.cpp:372-404, context.cpp:337-341). */
int kf_scene_update_instances(kf_scene* s, const kf_instance* inst, uint32_t n);

/* Move the vertices of geometry g (same vertex count and triangles) and
 * refit its BLAS on the GPU: triangle slots, BVH2 and BVH4 boxes of that
 * geometry only, then the TLAS.  The reference rebuilds every BLAS when
 * anything changes (context.cpp:328-336, rt.cpp:142-161); this is the
 * incremental path for dirty geometry.  The trees keep their shape, so
 * traversal cost can grow with large deformations (rebuild the scene then);
 * hits do not depend on the tree. */
int kf_scene_update_geometry(kf_scene* s, uint32_t g, const kf_vertex* vertices, uint32_t n);

/* Replace the material parameters (same count): the material upload of
 * Scene::uploadGeometries (scene.cpp:414-415) without a rebuild. */
int kf_scene_update_materials(kf_scene* s, const kf_material* materials, uint32_t n);

/* Replace the lights (Scene::uploadLightBuffers, scene.cpp:252-292). */
int kf_scene_set_lights(kf_scene* s, const kf_dir_light* d, const kf_point_lights* p,
                        const kf_active_lights* a);

/* Statistics of the built BVHs. */
typedef struct kf_bvh_info {
    uint32_t blas_nodes, tlas_nodes, triangles, instances;  /* BVH2 nodes */
    uint32_t max_blas_depth, max_tlas_depth;
    float build_ms;
    uint32_t blas4_nodes, tlas4_nodes;  /* BVH4 nodes read by the traversal */
} kf_bvh_info;
int kf_scene_bvh_info(kf_scene* s, kf_bvh_info* out);

/* Copy the BLAS of one geometry back: nodes are 16 floats each (left box
 * lo/hi xyz, right box lo/hi xyz as in the device layout, then child refs as
 * int bits) and the triangle order.  For BVH invariant tests. */
int kf_scene_read_blas(kf_scene* s, uint32_t geometry, float* nodes, uint32_t max_nodes,
                       uint32_t* n_nodes, int32_t* root_ref, uint32_t* tri_order, uint32_t max_tris);

/* Copy a BVH4 back (which = 0: the BLAS of all geometries, roots[g] per
 * geometry; 1: the TLAS, roots[0], or roots[e] per environment of a batched
 * scene, whose environments' nodes sit in ranges of the array with unused
 * slots between them).  Nodes are 32 floats each: lo.x, hi.x,
 * lo.y, hi.y, lo.z, hi.z of the four children (4 floats each), then the four
 * children and 4 pad words as uint32 bits.  A child is kind << 30 | index:
 * kind 0 / 1 an inner TLAS / BLAS node of this array, 2 an instance, 3 a
 * triangle slot; 0xffffffff an unused slot.  Roots are refs: >= 0 a node of
 * this array, < 0 ~leaf, 0x7fffffff empty.
 * For BVH invariant tests; no counterpart in the reference, whose
 * acceleration structures are opaque driver objects. */
int kf_scene_read_bvh4(kf_scene* s, int which, float* nodes, uint32_t max_nodes, uint32_t* n_nodes, int32_t* roots,
                       uint32_t max_roots);

/* Fixed ray batches (host pointers, synchronous): the traceRayEXT calls of
 * PathTrace.rgen:97-107 (closest hit) and PathTrace.rchit:186-197
 * (terminate-on-first-hit shadow rays).  A batched scene's rays see
 * environment 0. */
int kf_trace_closest(kf_scene* s, const kf_ray* rays, uint32_t n, kf_hit* hits);
int kf_trace_any(kf_scene* s, const kf_ray* rays, uint32_t n, uint8_t* occluded);

/* kf_trace_closest without any allocation, for callers that cannot allocate
 * (tight query loops, harnesses that forbid hipMalloc); no counterpart in the
 * reference.  Rays and hits are staged through a fixed device scratch of the
 * library (KF_TRACE_NOALLOC_MAX_RAYS rays: 14.7 MB per GPU, held from load
 * until the process ends); n above that is rejected with an error.  Same
 * results as kf_trace_closest.  Call kf_init first.  Every call on a GPU
 * shares that GPU's scratch: calls that overlap on one GPU (several host
 * threads) race. */
#define KF_TRACE_NOALLOC_MAX_RAYS (1u << 18)
int kf_trace_closest_noalloc(kf_scene* s, const kf_ray* rays, uint32_t n, kf_hit* hits);

/* Device-resident variants for measurement: rays/hits live in HBM. */
int kf_trace_closest_device(kf_scene* s, const kf_ray* d_rays, uint32_t n, kf_hit* d_hits,
                            float* kernel_ms);

/* ------------------------------------------------------------------------ */
/* Renderer: raygen + closest hit + miss + shadow (PathTrace.*)              */
/* ------------------------------------------------------------------------ */

typedef struct kf_renderer kf_renderer;

/* Render targets rgba/albedo/normal RGBA32F (rt.cpp:497-512). */
int kf_renderer_create(uint32_t width, uint32_t height, kf_renderer** out);
void kf_renderer_destroy(kf_renderer* r);

/* One frame: RayTracer::trace -> vkCmdTraceRaysKHR(W,H,1) (rt.cpp:637-666)
 * with the push constants of context.cpp:605-614. */
int kf_render(kf_renderer* r, kf_scene* s, const kf_camera_ubo* cam, const kf_push_constants* pc,
              const kf_render_opts* opts, kf_render_stats* stats);

/* Batched environments in one frame (no counterpart in the reference, where
 * each environment is its own Scene rendered by its own setCamera + run).
 * The scene holds num_envs environments (kf_scene_desc.env_instance_offset)
 * and the renderer's H rows are num_envs images of W x He, He = H / num_envs,
 * stacked: rows [e*He, (e+1)*He) show environment e through cams[e].  The
 * per-pixel seeds, jitter and raster coordinates are those of a W x He image,
 * so for opaque geometry environment e's rows equal kf_render of a scene
 * holding only e's instances on a W x He renderer (the alpha test's coin is
 * keyed by the global instance index).  All paths of all environments share
 * the wavefront queues and persistent traversal launches. */
int kf_render_envs(kf_renderer* r, kf_scene* s, uint32_t num_envs, const kf_camera_ubo* cams,
                   const kf_push_constants* pc, const kf_render_opts* opts, kf_render_stats* stats);

/* Autotuner of the wavefront schedule (no counterpart in the reference,
 * whose schedule is the driver's): times frames of this scene and target
 * over the batch size (samples per batch, i.e. the wavefront queue length),
 * the traversal's refill threshold and its work reservation per wave (the
 * tile a persistent wave takes from its XCD's queue) and the bounce from
 * which one thread carries each remaining path to its end instead of the
 * wavefront: batch size, then reservation, refill threshold, reservation
 * again and that bounce, each candidate timed
 * in a forward and a backward sweep; keeps the fastest for later frames
 * rendered with samples_per_batch = 0.  Images do not depend on the
 * schedule.  Renders into r's targets (frame_count 0). */
typedef struct kf_tune_result {
    uint32_t samples_per_batch, refill, grab;
    float frame_ms;    /* kernel time of a frame with the picked schedule */
    float default_ms;  /* ... with the built-in default */
    uint32_t trials;   /* frames timed */
    uint32_t tail_bounce; /* bounce from which one thread carries each path to its end (0: never) */
    uint32_t grab0;       /* queries reserved per wave for the camera rays (0: grab) */
    uint32_t grab_shadow; /* ... for the shadow rays (0: grab) */
    uint32_t tail_pairs;  /* 1: two lanes per path carry the tail bounces while few paths remain (at most a
                             quarter of the launch's threads; no soft light), 0: one lane per path */
} kf_tune_result;
int kf_renderer_autotune(kf_renderer* r, kf_scene* s, uint32_t num_envs, const kf_camera_ubo* cams,
                         const kf_push_constants* pc, kf_tune_result* out);
/* Set the schedule directly (e.g. one kf_renderer_autotune picked earlier);
 * 0 keeps the built-in default of that parameter (for tail_bounce: the
 * wavefront runs every bounce; for grab0 and grab_shadow, the camera rays'
 * and the shadow rays' reservations: grab).  refill is clamped to [1, 64],
 * the reservations to [16, 4096]. */
int kf_renderer_set_schedule(kf_renderer* r, uint32_t samples_per_batch, uint32_t refill, uint32_t grab,
                             uint32_t tail_bounce, uint32_t grab0, uint32_t grab_shadow);
/* Choose the kernel that carries the bounces from tail_bounce on (cleared to
 * 1 by kf_renderer_set_schedule): 1 lets two lanes per path take them while
 * at most a quarter of the launch's threads' worth of paths is left and no
 * light has a random perturbation (kf_tune_result.tail_pairs), 0 keeps one
 * lane per path.  Images do not depend on it. */
int kf_renderer_set_tail_pairs(kf_renderer* r, uint32_t tail_pairs);

/* Read a render target: 0 = rgba, 1 = albedo, 2 = normal; W*H*4 floats. */
int kf_renderer_read(kf_renderer* r, int target, float* out);
/* Overwrite the rgba target (running-mean state), W*H*4 floats. */
int kf_renderer_write(kf_renderer* r, int target, const float* in);
/* Per-sample radiance of the last frame: spp * W * H * 3 floats, sample-major. */
int kf_renderer_read_samples(kf_renderer* r, float* out, uint32_t max_floats);
/* Device pointer of a target (for the denoiser and readback kernels). */
float* kf_renderer_device_target(kf_renderer* r, int target);

/* downloadLatestFrame: RGBA32F -> B8G8R8A8_SRGB bytes (PostProcessing.frag:17,
 * config.hpp:144, vkCore.hpp:3008-3111); W*H*4 bytes, order B,G,R,A. */
int kf_renderer_download_bgra8(kf_renderer* r, uint8_t* out);

/* Single-channel FP32 image of the rgba target, (r + g + b) / 3 per pixel in
 * FP32 (W*H floats): the IR-camera output of active-light sensing.  No
 * counterpart in the reference, which exposes only the BGRA8 download. */
int kf_renderer_download_gray32f(kf_renderer* r, float* out);

/* Measurement utility (no reference counterpart): streaming copy of `bytes`
 * between two HBM buffers, `iters` times; reports the achieved GB/s (read +
 * write bytes / kernel time).  The achievable-bandwidth denominator of the
 * traversal roofline. */
int kf_bench_copy(uint64_t bytes, int iters, float* gbps);

/* ------------------------------------------------------------------------ */
/* Denoise stage (replaces DenoiserOptix, denoiser.cpp:24-344)               */
/* ------------------------------------------------------------------------ */

/* U-Net, 5 levels with 32/64/96/128/160 channels, 3x3 convolutions, BF16
 * weights and activations with FP32 accumulation (MFMA).  Input: noisy RGB,
 * albedo and normal targets; output: denoised RGB. */
typedef struct kf_denoiser kf_denoiser;

/* weights_path: weight file written by kf_denoiser_save; NULL loads the
 * default file (kuafu_amd/weights/unet.kfw next to the library) or, if it is
 * missing, starts from seeded random weights. */
int kf_denoiser_create(const char* weights_path, kf_denoiser** out);
void kf_denoiser_destroy(kf_denoiser* d);
int kf_denoiser_init_random(kf_denoiser* d, uint32_t seed);
int kf_denoiser_save(kf_denoiser* d, const char* path);
int kf_denoiser_load(kf_denoiser* d, const char* path);
/* Number of parameters (FP32 master copy) and host access to them. */
int kf_denoiser_num_params(kf_denoiser* d, uint64_t* n);
int kf_denoiser_get_params(kf_denoiser* d, float* out);
int kf_denoiser_set_params(kf_denoiser* d, const float* in);

/* The denoise stage of a frame: the renderer's rgba target is replaced by the
 * denoised image (context.cpp:650-670 writes the OptiX output back to rgba). */
int kf_denoiser_apply(kf_denoiser* d, kf_renderer* r);  /* after kf_render_envs: per environment */
/* Device pointers: rgba/albedo/normal W*H float4 in HBM -> out W*H float4.
 * kernel_ms (may be NULL) receives the kernels' time; both calls return when
 * the output is written. */
int kf_denoiser_infer_device(kf_denoiser* d, const float* d_rgba, const float* d_albedo, const float* d_normal,
                             uint32_t width, uint32_t height, float* d_out, float* kernel_ms);
/* Batched denoise (no counterpart in the reference): n_images images of
 * width x height stacked in each buffer (the layout of a kf_render_envs
 * frame), each denoised on its own, as one launch per layer. */
int kf_denoiser_infer_device_batch(kf_denoiser* d, const float* d_rgba, const float* d_albedo,
                                   const float* d_normal, uint32_t width, uint32_t height, uint32_t n_images,
                                   float* d_out, float* kernel_ms);

/* Training on host batches: inputs n*h*w*9 (RGB, albedo, normal), targets
 * n*h*w*3; one forward + backward + Adam step; h and w multiples of 16. */
int kf_denoiser_train_step(kf_denoiser* d, const float* inputs, const float* targets, uint32_t n, uint32_t h,
                           uint32_t w, float lr, float* loss);
/* Forward + backward without the update: loss and the gradient of every
 * parameter (same order as kf_denoiser_get_params), for gradient checks. */
int kf_denoiser_gradients(kf_denoiser* d, const float* inputs, const float* targets, uint32_t n, uint32_t h,
                          uint32_t w, float* loss, float* grads);
/* Forward only: denoised linear RGB, n*h*w*3. */
int kf_denoiser_forward(kf_denoiser* d, const float* inputs, uint32_t n, uint32_t h, uint32_t w, float* out);

/* Single 3x3 convolution layer on host arrays (BF16 bit patterns), NHWC:
 * forward y = act(conv(x, w) + b); backward dx (masked by x > 0 when
 * mask_by_x), dw and db.  For layer-level gradient checks. */
int kf_conv3x3_forward(const uint16_t* x, uint32_t cin, const uint16_t* w, const float* bias, uint32_t cout,
                       uint32_t n, uint32_t h, uint32_t wd, int relu, uint16_t* y);
int kf_conv3x3_backward(const uint16_t* x, uint32_t cin, const uint16_t* w, uint32_t cout, const uint16_t* dy,
                        uint32_t n, uint32_t h, uint32_t wd, int mask_by_x, uint16_t* dx, float* dw, float* db);

#ifdef __cplusplus
}
#endif

#endif /* KUAFU_AMD_H */
