// pathtrace.hip — wavefront path tracer for gfx950.
//
// Replaces RayTracer::trace -> vkCmdTraceRaysKHR(W,H,1) (rt.cpp:637-666) and
// the ray-tracing programs it runs (PathTrace.rgen/.rchit/.rahit/.rmiss,
// PathTraceShadow.rmiss).  The per-pixel megakernel of the reference becomes
// a wavefront over (pixel, sample) paths of one batch:
//   k_raygen   camera rays, seeds                          (rgen:22-66)
//   k_extend   closest-hit traversal of the active queue   (rgen:97-107)
//   k_shade    closest hit / miss: BSDF sample, emission,
//              NEE shadow-ray generation                   (rchit:322-469, rmiss)
//   k_shadow   any-hit traversal of the shadow queue       (rchit:175-200)
//   k_connect  light contributions (in slot order), path
//              accumulation, Russian roulette, compaction  (rgen:109-138)
//   k_resolve  per-pixel sample sum, running mean, AOVs    (rgen:140-166)
// Every path consumes its random numbers in the reference's order, so the
// per-sample radiance equals the FP32 oracle's.  When a light needs random
// perturbation (softness / radius != 0) the draws of one light depend on the
// visibility of the previous one, so k_shade then traces its shadow rays
// inline (NEE_INLINE) instead of deferring them to k_shadow.
#include "kf_internal.h"
#include "shading.h"
#include "traverse.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace kf {

namespace {

constexpr int TB = 256;
static_assert(TB == (int)kRingThreads, "traversal kernels index their LDS stack ring by a 256-thread block");

// Device counters of one batch: per bounce b: [0] active paths, [1] shadow
// rays, [2] next queue, [3..10] work chunks of k_extend, [11..18] work chunks
// of k_shadow, [19] rays k_extend traces at bounce b.  Global, at the end:
// primary / extension / shadow rays, BVH4 node records / triangle tests
// (closest-hit traversal), stack overflow, instance records, speculative
// rays, TLAS node records (part of the node records), closest-hit rays
// traced by k_tail (part of the extension rays).
constexpr int kMaxBounces = 64;
constexpr int kCntPerBounce = 20;
// throughput and depth of a path before its first bounce
__device__ constexpr float4 kPathStart = {1.0f, 1.0f, 1.0f, 0.0f};  // .w: depth 0 as uint bits
constexpr int kCntExtendChunks = 3;
constexpr int kCntShadowChunks = 11;
constexpr int kCntExtQueue = 19;
constexpr int kCntGlobal = kMaxBounces * kCntPerBounce;
constexpr int kCntSize = kCntGlobal + 10;

// Per-frame values the kernels read from device memory, so that a captured
// frame graph replays with new seeds and running-mean weights.
struct FrameVars {
    uint32_t S0;          // frame seed
    int32_t frame_count;  // running-mean position (rgen:153-163)
};

// The render target holds n_env images of W x He stacked vertically (one
// image, He = H, without environments): pixel = env * WHe + (y * W + x), and
// seeds, jitter and raster coordinates are those of the W x He image.
struct Frame {
    const kf_camera_ubo* cams;  // n_env cameras (device)
    kf_push_constants pc;
    uint32_t W, H, WH;
    uint32_t He, WHe, n_env;
    uint32_t s0, nsamp;   // samples [s0, s0 + nsamp) in this batch
    const FrameVars* fv;  // frame seed and count (device)
    uint32_t slots;       // NEE slots per path
    int inline_nee;
    int count;            // count traversal steps
};

KF_DEV void wave_add(unsigned long long* p, uint32_t v) {
    // wave64 reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(p, (unsigned long long)v);
}

// Wave-aggregated queue append; returns the slot index for lanes with pred.
KF_DEV uint32_t wave_append(unsigned long long* counter, bool pred) {
    unsigned long long m = __ballot(pred);
    uint32_t lane = threadIdx.x & 63;
    uint32_t leader = m ? (uint32_t)__ffsll((long long)m) - 1 : 0;
    uint32_t base = 0;
    if (m && lane == leader) base = (uint32_t)atomicAdd(counter, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    return base + (uint32_t)__popcll(m & lt);
}

// Block-aggregated queue append: one atomic per TB-thread block instead of
// one per wave (appends to a single counter serialise in the memory system).
// Every thread of the block must call it at the same point.
KF_DEV uint32_t block_append(unsigned long long* counter, bool pred) {
    __shared__ uint32_t s_wave[TB / 64];
    __shared__ uint32_t s_base;
    const unsigned long long m = __ballot(pred);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    if (lane == 0) s_wave[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int i = 0; i < TB / 64; ++i) tot += s_wave[i];
        s_base = tot ? (uint32_t)atomicAdd(counter, (unsigned long long)tot) : 0u;
    }
    __syncthreads();
    uint32_t off = s_base;
    for (uint32_t i = 0; i < w; ++i) off += s_wave[i];
    const uint32_t r = off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();  // s_wave and s_base are reused by the next call
    return r;
}

__global__ void k_zero(unsigned long long* c, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
}

// ------------------------------------------------------------ camera-ray TLAS
// The camera rays of an 8x8 pixel tile leave one point (the camera position,
// pinhole cameras) in directions inside the tile's pyramid, so an instance
// whose world box lies wholly outside that pyramid is hit by none of them.
// k_tile_roots lists, per tile, the instances of its environment whose boxes
// meet the pyramid and gives the tile a root item over just those: none ->
// the empty item, one -> the instance, 2-4 -> one node, 5-16 -> a node over
// up to four groups in depth order; more, or a camera with an aperture (rays
// from a disc), -> the environment's TLAS root.  The closest hit does not
// depend on the order in which records are visited (smallest t, ties by
// instance and primitive), each instance keeps its TLAS leaf box, and a node
// box here is the exact union of its children's, so camera rays find the hits
// they find through the full TLAS.  Later bounces and shadow rays start from
// the environment's root.
constexpr uint32_t kTile = 8, kTileNodes = 5, kTileMax = 16;
// The pyramid's planes pass half a pixel outside the tile, and each of the
// tile's rays must lie at least this angle (radians) inside every plane, far
// more than the float rounding of a ray's direction (about 1e-6); tiles of
// narrower pixels use the environment's root.
constexpr double kTileMinAngle = 2e-5;
// A box is outside a plane when its nearest corner lies beyond it by this
// fraction of the corner's distance from the camera.
constexpr double kTileBoxEps = 1e-5;

struct DV {
    double x, y, z;
};
KF_DEV DV dv_sub(DV a, DV b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
KF_DEV double dv_dot(DV a, DV b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KF_DEV DV dv_cross(DV a, DV b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
KF_DEV double dv_len(DV a) { return sqrt(dv_dot(a, a)); }
KF_DEV DV dv_scale(DV a, double s) { return {a.x * s, a.y * s, a.z * s}; }

// k_raygen's raster -> world direction map (not normalised) is affine in the
// raster point: dir(px, py) = a + px bx + py by for a W x He image, where
// target = proj_inverse (2 px / W - 1, 2 py / He - 1, 1, 1) and dir =
// view_inverse (target.xyz, 0).  In double.
struct DirMap {
    DV a, bx, by;
};
KF_DEV DirMap dir_map(const kf_camera_ubo& cam, double W, double He) {
    const float* P = cam.proj_inverse;
    const float* V = cam.view_inverse;
    auto rot = [&](double t0, double t1, double t2) -> DV {
        return {(double)V[0] * t0 + (double)V[4] * t1 + (double)V[8] * t2,
                (double)V[1] * t0 + (double)V[5] * t1 + (double)V[9] * t2,
                (double)V[2] * t0 + (double)V[6] * t1 + (double)V[10] * t2};
    };
    auto p = [&](int i) { return (double)P[i]; };
    const double sx = 2.0 / W, sy = 2.0 / He;
    DirMap m;
    m.a = rot(p(8) + p(12) - p(0) - p(4), p(9) + p(13) - p(1) - p(5), p(10) + p(14) - p(2) - p(6));
    m.bx = rot(p(0) * sx, p(1) * sx, p(2) * sx);
    m.by = rot(p(4) * sy, p(5) * sy, p(6) * sy);
    return m;
}
KF_DEV DV dir_at(const DirMap& m, double px, double py) {
    return {m.a.x + px * m.bx.x + py * m.by.x, m.a.y + px * m.bx.y + py * m.by.y, m.a.z + px * m.bx.z + py * m.by.z};
}

// Child slot c of a node: box (lo, hi) and item; an unused slot is the
// empty item with a zero box.
KF_DEV void tile_child(Node4* __restrict__ nd, uint32_t c, float4 lo, float4 hi, uint32_t item) {
    float* f = reinterpret_cast<float*>(nd);
    f[c] = lo.x;
    f[4 + c] = hi.x;
    f[8 + c] = lo.y;
    f[12 + c] = hi.y;
    f[16 + c] = lo.z;
    f[20 + c] = hi.z;
    uint32_t* u = reinterpret_cast<uint32_t*>(nd);
    u[24 + c] = item;
    u[28 + c] = (uint32_t)kEmptyRef;
}

// One wave per tile (n_env * tx * ty tiles, tx x ty per environment): the
// lanes test the environment's instances 64 at a time, then write the tile's
// nodes at out[n_tlas + kTileNodes * tile ...] one child slot per lane (the
// root at the first, a group of 2-4 instances at 1 + group).  The scene's n_tlas
// TLAS nodes are copied to out[0, n_tlas) first, so that the environments'
// root items index the same nodes there.
__global__ __launch_bounds__(TB) void k_tile_roots(SceneView S, Frame F, const Node4* __restrict__ tlas,
                                                   uint32_t n_tlas, const float4* __restrict__ ilo,
                                                   const float4* __restrict__ ihi,
                                                   const uint32_t* __restrict__ env_off, uint32_t tx, uint32_t ty,
                                                   Node4* __restrict__ out, uint32_t* __restrict__ tile_root) {
    __shared__ uint32_t s_cand[TB / 64][kTileMax];
    __shared__ uint32_t s_sorted[TB / 64][kTileMax];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nth = gridDim.x * blockDim.x;
    {
        const float4* src = (const float4*)tlas;
        float4* dst = (float4*)out;
        for (size_t i = tid; i < (size_t)n_tlas * 8u; i += nth) dst[i] = src[i];
    }
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t per_env = tx * ty, n_tiles = per_env * F.n_env;
    for (uint32_t t = tid >> 6; t < n_tiles; t += nth >> 6) {  // wave-uniform
        const uint32_t e = t / per_env, r = t - e * per_env, by = r / tx, bx = r - by * tx;
        const uint32_t env_item = env_root_item(S, e);
        uint32_t root = env_item;
        const kf_camera_ubo& cam = F.cams[e];
        // camera position as k_raygen computes it (the rays' origin, bit for bit)
        const float o4[4] = {0.0f, 0.0f, 0.0f, 1.0f};
        float org[4];
        mat4_mul(cam.view_inverse, o4, org);
        const DV o = {org[0], org[1], org[2]};
        const double W = F.W, He = F.He;
        const double xa = (double)(bx * kTile), xb = (double)min((bx + 1u) * kTile, F.W);
        const double ya = (double)(by * kTile), yb = (double)min((by + 1u) * kTile, F.He);
        bool ok = env_item != kEmptyItem && !(cam.position[3] > 0.0f) && isfinite(org[0]) && isfinite(org[1]) &&
                  isfinite(org[2]);
        const DirMap dm = dir_map(cam, W, He);
        const DV ctr = dir_at(dm, 0.5 * (xa + xb), 0.5 * (ya + yb));  // the tile's central ray
        DV n[4];
        if (ok) {
            const DV c[4] = {dir_at(dm, xa - 0.5, ya - 0.5), dir_at(dm, xb + 0.5, ya - 0.5),
                             dir_at(dm, xb + 0.5, yb + 0.5), dir_at(dm, xa - 0.5, yb + 0.5)};
            const DV q[4] = {dir_at(dm, xa, ya), dir_at(dm, xb, ya), dir_at(dm, xb, yb), dir_at(dm, xa, yb)};
            double sgn = 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                n[i] = dv_cross(c[i], c[(i + 1) & 3]);
                const double s = dv_dot(n[i], ctr);
                if (i == 0) sgn = s > 0.0 ? 1.0 : -1.0;
                const double len = dv_len(n[i]);
                ok = ok && s * sgn > 0.0 && len > 0.0 && isfinite(len);
                n[i] = dv_scale(n[i], ok ? sgn / len : 0.0);  // unit normal pointing into the pyramid
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double ql = dv_len(q[j]);
#pragma unroll
                for (int i = 0; i < 4; ++i) ok = ok && ql > 0.0 && dv_dot(n[i], q[j]) >= kTileMinAngle * ql;
            }
        }
        if (ok) {
            const uint32_t i0 = env_off ? env_off[e] : 0u, i1 = env_off ? env_off[e + 1] : S.n_inst;
            uint32_t k = 0;
            for (uint32_t base = i0; base < i1 && k <= kTileMax; base += 64u) {
                const uint32_t i = base + lane;
                bool in = i < i1;
                if (in) {
                    const float4 lo = ilo[i], hi = ihi[i];
                    // lo > hi: an instance without triangles to hit
                    in = lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z;
#pragma unroll
                    for (int pl = 0; pl < 4; ++pl) {
                        const DV pv = {n[pl].x > 0.0 ? hi.x : lo.x, n[pl].y > 0.0 ? hi.y : lo.y,
                                       n[pl].z > 0.0 ? hi.z : lo.z};
                        const DV v = dv_sub(pv, o);
                        const double dd = dv_dot(n[pl], v);  // outside: dd < -kTileBoxEps |v|
                        in = in && !(dd < 0.0 && dd * dd > kTileBoxEps * kTileBoxEps * dv_dot(v, v));
                    }
                }
                const unsigned long long m = __ballot(in);
                const uint32_t at = k + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                if (in && at < kTileMax) s_cand[w][at] = i;
                k += (uint32_t)__popcll(m);
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            if (k == 0) {
                root = kEmptyItem;
            } else if (k == 1) {
                root = make_item(kInstance, s_cand[w][0]);
            } else if (k <= kTileMax) {
                // lane j < k holds candidate j; with more than four, lanes
                // rank them by their box centres' depth along the tile's
                // central ray (ties by list position) and consecutive ranks
                // form up to four groups of 2-4 (or single instances)
                const uint32_t node0 = n_tlas + kTileNodes * t;
                const bool mine = lane < k;
                const uint32_t ci = s_cand[w][mine ? lane : 0u];
                uint32_t rank = lane;
                if (k > 4) {
                    const float4 lo = ilo[ci], hi = ihi[ci];
                    const DV m = {0.5 * ((double)lo.x + hi.x), 0.5 * ((double)lo.y + hi.y),
                                  0.5 * ((double)lo.z + hi.z)};
                    const double key = dv_dot(dv_sub(m, o), ctr);
                    rank = 0;
                    for (uint32_t j = 0; j < k; ++j) {
                        const double kj = __shfl(key, (int)j);
                        rank += (kj < key || (kj == key && j < lane)) ? 1u : 0u;
                    }
                }
                if (mine) s_sorted[w][rank] = ci;
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (k <= 4) {  // one node over the candidates
                    if (lane < 4) {
                        const uint32_t c = s_sorted[w][lane < k ? lane : 0u];
                        if (lane < k) tile_child(out + node0, lane, ilo[c], ihi[c], make_item(kInstance, c));
                        else tile_child(out + node0, lane, z, z, kEmptyItem);
                    }
                } else {
                    const uint32_t g = (k + 3u) / 4u, base = k / g, rem = k % g;
                    auto gsize = [&](uint32_t gi) { return base + (gi < rem ? 1u : 0u); };
                    auto gstart = [&](uint32_t gi) { return gi * base + min(gi, rem); };
                    if (lane < 16) {  // (group, slot): the groups' nodes
                        const uint32_t gi = lane >> 2, slot = lane & 3u;
                        if (gi < g && gsize(gi) >= 2) {
                            Node4* nd = out + node0 + 1u + gi;
                            if (slot < gsize(gi)) {
                                const uint32_t c = s_sorted[w][gstart(gi) + slot];
                                tile_child(nd, slot, ilo[c], ihi[c], make_item(kInstance, c));
                            } else {
                                tile_child(nd, slot, z, z, kEmptyItem);
                            }
                        }
                    } else if (lane < 20) {  // the root's children: groups, or single instances
                        const uint32_t gi = lane - 16u;
                        if (gi < g) {
                            const uint32_t s0 = gstart(gi), n = gsize(gi);
                            float4 l = ilo[s_sorted[w][s0]], h = ihi[s_sorted[w][s0]];
                            for (uint32_t a = 1; a < n; ++a) {
                                const uint32_t c = s_sorted[w][s0 + a];
                                const float4 la = ilo[c], ha = ihi[c];
                                l = make_float4(fminf(l.x, la.x), fminf(l.y, la.y), fminf(l.z, la.z), 0.0f);
                                h = make_float4(fmaxf(h.x, ha.x), fmaxf(h.y, ha.y), fmaxf(h.z, ha.z), 0.0f);
                            }
                            const uint32_t item = n == 1 ? make_item(kInstance, s_sorted[w][s0])
                                                         : make_item(kTlasNode, node0 + 1u + gi);
                            tile_child(out + node0, gi, l, h, item);
                        } else {
                            tile_child(out + node0, gi, z, z, kEmptyItem);
                        }
                    }
                }
                root = make_item(kTlasNode, node0);
            }
        }
        if (lane == 0) tile_root[t] = root;
        __builtin_amdgcn_wave_barrier();  // s_cand is reused by the wave's next tile
    }
}

// tile_root: each 8x8 tile's root item (k_tile_roots, tx x ty tiles per
// environment), or null: the environment's TLAS root.
__global__ __launch_bounds__(TB) void k_raygen(SceneView S, Frame F, const uint2* __restrict__ jump,
                                               const uint32_t* __restrict__ tile_root, uint32_t tx, uint32_t ty,
                                               float4* __restrict__ ray_o, float4* __restrict__ ray_d,
                                               uint32_t* __restrict__ seed, uint32_t* __restrict__ queue,
                                               unsigned long long* __restrict__ cnt) {
    const uint32_t n = F.WH * F.nsamp;
    const uint32_t S0 = F.fv->S0;
    // one thread per pixel and all of the batch's samples of it: the jitter
    // stream's seed tea(pixel, S0) is the same for every sample
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < F.WH; r += gridDim.x * blockDim.x) {
        const uint32_t env = r / F.WHe;
        const uint32_t pixel = r - env * F.WHe;  // within the environment's image
        const uint32_t x = pixel % F.W, y = pixel / F.W;
        const kf_camera_ubo& cam = F.cams[env];
        const uint32_t js0 = tea(pixel, S0);
        for (uint32_t p = r; p < n; p += F.WH) {
            const uint32_t i = F.s0 + p / F.WH;
            // jitter stream: tea(pixel, S0) advanced by 2i draws (rgen:23, 35)
            uint2 jm = jump[i];
            uint32_t js = jm.x * js0 + jm.y;
            float r0 = rnd(js);
            float r1 = rnd(js);
            uint32_t rs = tea(pixel, S0 + 1u + i);
            float posx = float(x) + r0, posy = float(y) + r1;
            float npx = posx / float(F.W), npy = posy / float(F.He);
            float dx = npx * 2.0f - 1.0f, dy = npy * 2.0f - 1.0f;
            float aperture = cam.position[3];
            float focus = cam.front[3];
            float ox, oy;
            do {  // diskSampling (Sampling.glsl:124-134), always drawn (rgen:42)
                float a = rnd(rs), b = rnd(rs);
                ox = 2.0f * a - 1.0f;
                oy = 2.0f * b - 1.0f;
            } while (ox * ox + oy * oy >= 1.0f);
            float half = aperture / 2.0f;
            ox = half * ox;
            oy = half * oy;
            float tv[4] = {dx, dy, 1.0f, 1.0f};
            float target[4];
            mat4_mul(cam.proj_inverse, tv, target);
            float org[4], dir4[4];
            if (aperture > 0.0f) {
                float o4[4] = {ox, oy, 0.0f, 1.0f};
                mat4_mul(cam.view_inverse, o4, org);
                v3 nd = normalize(mk(target[0], target[1], target[2]) * focus - mk(ox, oy, 0.0f));
                float d4[4] = {nd.x, nd.y, nd.z, 0.0f};
                mat4_mul(cam.view_inverse, d4, dir4);
            } else {
                float o4[4] = {0.0f, 0.0f, 0.0f, 1.0f};
                mat4_mul(cam.view_inverse, o4, org);
                v3 nd = normalize(mk(target[0], target[1], target[2]));
                float d4[4] = {nd.x, nd.y, nd.z, 0.0f};
                mat4_mul(cam.view_inverse, d4, dir4);
            }
            // ray_d.w carries the ray's TLAS root item: the pixel tile's for the
            // camera ray, the environment's from the next bounce on
            const uint32_t root =
                tile_root ? tile_root[(env * ty + y / kTile) * tx + x / kTile] : env_root_item(S, env);
            ray_o[p] = make_float4(org[0], org[1], org[2], 0.0f);
            ray_d[p] = make_float4(dir4[0], dir4[1], dir4[2], __uint_as_float(root));
            // throughput (1, depth 0), colour (0) and BSDF weight (0) of a new
            // path are implicit at bounce 0 (kPathStart): k_shade and k_connect
            // use them there instead of reading them back, which saves 48 B of
            // writes and as many of reads per path
            seed[p] = rs;
            // Camera rays are traced in 8-row bands, column by column, so a
            // wave's 64 rays cover an 8x8 pixel tile (paths stay stored by
            // sample and pixel; only the order of the queue changes).
            {
                const uint32_t band = r / (8u * F.W);
                const uint32_t rows = min(8u, F.H - band * 8u), off = r - band * 8u * F.W;
                const uint32_t x = off / rows, yy = off - x * rows;
                queue[p] = (p - r) + (band * 8u + yy) * F.W + x;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cnt[0 * kCntPerBounce + 0] = n;  // active paths at bounce 0
        cnt[0 * kCntPerBounce + kCntExtQueue] = n;
        atomicAdd(&cnt[kCntGlobal + 0], (unsigned long long)n);
    }
}

// Register budgets of the two traversal kernels that run side by side
// (k_extend of bounce b+1 beside k_shadow of bounce b): 4 waves of k_extend
// at 104 VGPRs and one wave of k_shadow at 96 fill a SIMD's 512, so the
// shadow rays are traced while the closest-hit launch is still resident.
// amdgpu_num_vgpr counts in units of two registers on gfx950 (the unified
// VGPR/AGPR file), hence 52 and 48.
template <bool COUNT>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_num_vgpr(52))) void k_extend(SceneView S, int bounce, uint32_t cap, const uint32_t* __restrict__ queue,
                                               const float4* __restrict__ ray_o, const float4* __restrict__ ray_d,
                                               const uint32_t* __restrict__ seed, float4* __restrict__ hit,
                                               uint32_t* __restrict__ hit_prim, unsigned long long* __restrict__ cnt,
                                               uint32_t refill, uint32_t grab, uint32_t WH, uint32_t WHe,
                                               uint32_t s0, const FrameVars* __restrict__ fv) {
    const uint32_t n = (uint32_t)min(cnt[bounce * kCntPerBounce + kCntExtQueue], (unsigned long long)cap);
    const uint32_t S0 = S.any_alpha ? fv->S0 : 0u;
    uint32_t ovf = 0;
    TCount tot{0, 0, 0, 0};
    auto load = [&](uint32_t q, TRay& r) __attribute__((always_inline)) {
        const uint32_t p = queue[q];
        const float4 o = ray_o[p], d = ray_d[p];
        r.o = xyz(o);
        r.d = xyz(d);
        r.tmin = 0.001f;
        r.tmax = 10000.0f;
        r.root = __float_as_uint(d.w);
        // alpha coin seed: tea(initial seed of the sample's stream, bounce)
        r.alpha_seed = S.any_alpha ? tea(tea((p % WH) % WHe, S0 + 1u + s0 + p / WH), (uint32_t)bounce) : 0u;
        r.payload = p;
    };
    auto store = [&](uint32_t p, THit b, bool) __attribute__((always_inline)) {
        hit[p] = make_float4(b.t, b.b1, b.b2, __uint_as_float(b.inst));
        hit_prim[p] = b.prim | (b.back << 31);
    };
    trace_persistent<false, COUNT>(S, n, cnt + bounce * kCntPerBounce + kCntExtendChunks, refill, grab, S.any_alpha != 0u, load,
                                   store, ovf, tot);
    if (COUNT) {
        wave_add(&cnt[kCntGlobal + 3], tot.nodes);
        wave_add(&cnt[kCntGlobal + 4], tot.tris);
        wave_add(&cnt[kCntGlobal + 6], tot.insts);
        wave_add(&cnt[kCntGlobal + 8], tot.tlas);
    }
    if (ovf) atomicAdd(&cnt[kCntGlobal + 5], 1ull);
}

// Closest-hit / miss for one path; writes next ray, emission, BSDF weight and
// NEE slots (deferred) or the NEE sum (inline).
// Light j of the frame's light list seen from worldPos: direction Ld,
// emitted radiance lem, shadow-ray length maxd; false when the point is
// outside an active light's cone.  PERTURB draws the light's random
// perturbation from rs first (soft directional light, point-light radius,
// active-light softness), as the reference does before tracing.
template <bool PERTURB>
KF_DEV bool light_eval(const SceneView& S, const DevLights* __restrict__ LT, int kind, int li, v3 worldPos, uint32_t& rs,
                       v3& Ld_out, v3& lem_out, float& maxd_out) {
    bool ok = true;
    float ldx, ldy, ldz, lex, ley, lez, maxd;  // scalars: per-branch vector writes would go through memory
    if (kind == 0) {  // traceDirectionalLight (rchit:204-223)
        const v3 lem = mk(LT->dl.rgbs[0], LT->dl.rgbs[1], LT->dl.rgbs[2]) * LT->dl.rgbs[3];
        v3 Ld = -mk(LT->dl.direction[0], LT->dl.direction[1], LT->dl.direction[2]);
        if (PERTURB && LT->dl.direction[3] != 0.0f) {
            float a = rnd(rs), b = rnd(rs), c = rnd(rs);
            Ld = normalize(Ld + LT->dl.direction[3] * mk(a, b, c));
        }
        ldx = Ld.x; ldy = Ld.y; ldz = Ld.z;
        lex = lem.x; ley = lem.y; lez = lem.z;
        maxd = 1e6f;
    } else if (kind == 1) {  // tracePointLights (rchit:227-258)
        v3 rgb = mk(LT->pl.rgbs[li][0], LT->pl.rgbs[li][1], LT->pl.rgbs[li][2]);
        v3 lpos = mk(LT->pl.posr[li][0], LT->pl.posr[li][1], LT->pl.posr[li][2]);
        if (PERTURB && LT->pl.posr[li][3] != 0.0f) {
            v3 pert = uniform_sphere(rs);
            lpos = lpos + LT->pl.posr[li][3] * normalize(pert);
        }
        v3 Lv = lpos - worldPos;
        float dd2 = length(Lv);
        const v3 Ld = normalize(Lv);
        const v3 lem = rgb * LT->pl.rgbs[li][3] / dd2 / dd2;
        ldx = Ld.x; ldy = Ld.y; ldz = Ld.z;
        lex = lem.x; ley = lem.y; lez = lem.z;
        maxd = dd2;
    } else {  // traceActiveLights (rchit:262-316)
        v3 rgb = mk(LT->al.rgbs[li][0], LT->al.rgbs[li][1], LT->al.rgbs[li][2]);
        float softness = LT->al.sftp[li][0];
        v3 lpos = mk(LT->al.position[li][0], LT->al.position[li][1], LT->al.position[li][2]);
        if (PERTURB && softness != 0.0f) {
            v3 pert = uniform_sphere(rs);
            lpos = lpos + softness * normalize(pert);
        }
        v3 Lv = lpos - worldPos;
        float dd2 = length(Lv);
        const v3 Ld = normalize(Lv);
        float fov = LT->al.sftp[li][1];
        v3 ldir = normalize(mk(LT->al.front[li][0], LT->al.front[li][1], LT->al.front[li][2]));
        float halfA = gclamp(fov, 0.0f, kPi) / 2.0f;
        float cs = dot(ldir, -Ld);
        ok = cs > cos_k(halfA);
        v3 color = rgb;
        int texID = (int)LT->al.sftp[li][2];
        if (ok && texID >= 0 && (uint32_t)texID < S.n_tex) {
            float v4[4] = {worldPos.x, worldPos.y, worldPos.z, 1.0f};
            float tc[4];
            mat4_mul(LT->pv[li], v4, tc);
            float u = (tc[0] / tc[3]) * 0.5f + 0.5f;
            float v = (tc[1] / tc[3]) * 0.5f + 0.5f;
            float4 tx = tex_sample(S, S.tex_data, S.tex_info[texID], u, v, true);
            color = color * xyz(tx);
        }
        const v3 lem = color * LT->al.rgbs[li][3] / dd2 / dd2;
        ldx = Ld.x; ldy = Ld.y; ldz = Ld.z;
        lex = lem.x; ley = lem.y; lez = lem.z;
        maxd = dd2;
    }
    Ld_out = mk(ldx, ldy, ldz);
    lem_out = mk(lex, ley, lez);
    maxd_out = maxd;
    return ok;
}

// Shading point of a path whose NEE is deferred to the shadow-ray kernel.
struct NeeCtx {
    ShadeCtx sh;
    uint32_t root;  // the path's TLAS root item
    bool on;
};

// The deferred light loop for the paths of a block (rchit:457-460 order):
// per light, the (ca, theta, cb, mode) record of the contribution and a
// shadow ray, appended with one atomic per block.  Every thread of the
// block calls it; threads without a shading point (nc.on false) append
// nothing.
KF_DEV void nee_deferred(const SceneView& S, const DevLights* __restrict__ LT, const NeeCtx& nc, uint32_t p,
                         uint32_t K, float4* __restrict__ nee_a, float4* __restrict__ nee_b,
                         float4* __restrict__ sh_o, float4* __restrict__ sh_d, uint32_t* __restrict__ sh_root,
                         unsigned long long* __restrict__ shadow_ctr, uint32_t& n_shadow) {
    const int nl = LT->n_slots;
    uint32_t unused_rs = 0u;
    for (int j = 0; j < nl; ++j) {
        bool emit = false;
        v3 Ld = splat(0.0f);
        float maxd = 0.0f;
        if (nc.on) {
            const int kind = LT->kind[j], li = LT->index[j];
            v3 lem;
            const bool ok = light_eval<false>(S, LT, kind, li, nc.sh.worldPos, unused_rs, Ld, lem, maxd);
            if (ok && dot(nc.sh.N, Ld) > 0.0f) {
                v3 ca, cb;
                float th;
                int mode;
                direct_split(nc.sh, Ld, lem, ca, cb, th, mode);
                const size_t slot = (size_t)p * K + j;
                nee_a[slot] = f4(ca, th);
                nee_b[slot] = f4(cb, __int_as_float(mode | (kind << 4)));
                emit = true;
            }
        }
        const uint32_t at = block_append(shadow_ctr, emit);
        if (emit) {
            sh_o[at] = f4(nc.sh.worldPos, __uint_as_float((uint32_t)((size_t)p * K + j)));
            sh_d[at] = f4(Ld, maxd);
            if (S.n_env > 1) sh_root[at] = nc.root;
            n_shadow++;
        }
    }
}

// The per-path inputs of shading (throughput, hit, ray, random state),
// loaded apart from the shading itself so that k_shade can fetch the next
// path's while it shades this one.
struct ShadeIn {
    float4 t4, h, ro, rd;
    uint32_t rs, hp;
};
KF_DEV ShadeIn shade_in(uint32_t p, float4 t4, const float4* __restrict__ hit, const uint32_t* __restrict__ hit_prim,
                        const float4* __restrict__ ray_o, const float4* __restrict__ ray_d,
                        const uint32_t* __restrict__ seed) {
    ShadeIn in;
    in.t4 = t4;
    in.h = hit[p];
    in.ro = ray_o[p];
    in.rd = ray_d[p];
    in.rs = seed[p];
    in.hp = hit_prim[p];
    return in;
}

// Closest hit / miss of path p at this bounce (rchit:322-469, rmiss:12-31):
// emission, BSDF sample and next ray, NEE (shadow-ray records, or the sum
// with the shadow rays traced here when NEE_INLINE), AOVs.  Returns whether
// the path is traced at the next bounce unless Russian roulette ends it.
template <bool NEE_INLINE>
KF_DEV bool shade_path(const SceneView& S, const Frame& F, const DevLights* __restrict__ LT, int bounce, uint32_t p,
                       const ShadeIn& in, bool aov, float4* __restrict__ ray_o, float4* __restrict__ ray_d,
                       float4* __restrict__ rw, float4* __restrict__ em, uint32_t* __restrict__ seed, float4* __restrict__ nee_a,
                       float4* __restrict__ nee_b, uint8_t* __restrict__ vis, float4* __restrict__ sh_o,
                       float4* __restrict__ sh_d, uint32_t* __restrict__ sh_root, float4* __restrict__ albedo_out,
                       float4* __restrict__ normal_out, unsigned long long* __restrict__ cnt, uint32_t& n_shadow,
                       uint32_t& ovf, NeeCtx& nc) {
    const uint32_t K = F.slots;
    const float4 t4 = in.t4;
    const uint32_t depth = __float_as_uint(t4.w);
    bool ext = false;
    nc.on = false;
    const float4 h = in.h;
    const float4 ro = in.ro, rd = in.rd;
    v3 rdir = xyz(rd);
    // root item of the rays from here on (a camera ray carries its tile's)
    const uint32_t root = bounce == 0 ? env_root_item(S, (p % F.WH) / F.WHe) : __float_as_uint(rd.w);
    uint32_t rs = in.rs;
    uint32_t inst = __float_as_uint(h.w);
    for (uint32_t j = 0; j < K; ++j) vis[(size_t)p * K + j] = 0;
    if (inst == ~0u) {
        // miss (rmiss:12-31)
        v3 e = miss_emission(S, F.pc, rdir);
        em[p] = f4(e, __uint_as_float(1u));
        if (aov) {
            albedo_out[p] = make_float4(0.f, 0.f, 0.f, 1.f);
            normal_out[p] = make_float4(0.f, 0.f, 0.f, 1.f);
        }
        if (NEE_INLINE) nee_a[(size_t)p * K] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        uint32_t hp = in.hp;
        uint32_t prim = hp & 0x7fffffffu;
        bool inside = (hp >> 31) != 0;
        Inst I = S.inst[inst];
        // the instance's geometry: where its vertices and indices are read,
        // and its own triangle offset (material indices), on the record's
        // last line, fetched with its world->object rows
        const size_t tri = (size_t)(uint32_t)I.shade.z + prim;
        const size_t mtri = (size_t)(uint32_t)I.shade.y + prim;
        const uint32_t vb = (uint32_t)I.shade.x;
        uint32_t i0 = S.idx[3 * mtri + 0], i1 = S.idx[3 * mtri + 1], i2 = S.idx[3 * mtri + 2];
        const float4* V0 = S.verts + 3 * (size_t)(vb + i0);
        const float4* V1 = S.verts + 3 * (size_t)(vb + i1);
        const float4* V2 = S.verts + 3 * (size_t)(vb + i2);
        float4 a0 = V0[0], a1 = V0[1], a2v = V0[2];
        float4 b0 = V1[0], b1 = V1[1], b2v = V1[2];
        float4 c0 = V2[0], c1 = V2[1], c2v = V2[2];
        float bx = 1.0f - h.y - h.z, by = h.y, bz = h.z;
        v3 n0 = mk(a0.w, a1.x, a1.y), n1 = mk(b0.w, b1.x, b1.y), n2 = mk(c0.w, c1.x, c1.y);
        v3 ln = n0 * bx + n1 * by + n2 * bz;
        v3 wn = mk(ln.x * I.w2o[0].x + ln.y * I.w2o[1].x + ln.z * I.w2o[2].x,
                   ln.x * I.w2o[0].y + ln.y * I.w2o[1].y + ln.z * I.w2o[2].y,
                   ln.x * I.w2o[0].z + ln.y * I.w2o[1].z + ln.z * I.w2o[2].z);
        v3 N = normalize(wn);
        v3 worldPos = xyz(ro) + rdir * h.x;
        float uvx = a2v.y * bx + b2v.y * by + c2v.y * bz;
        float uvy = a2v.z * bx + b2v.z * by + c2v.z * bz;
        // the whole 80-byte material in one round trip (five 16-byte loads
        // issued together), not field by field behind the emission test
        kf_material m;
        {
            const gf4* mp = (const gf4*)(S.mats + S.matidx[tri]);
            const vf4 m0 = mp[0], m1 = mp[1], m2 = mp[2], m3 = mp[3], m4 = mp[4];
            // the empty asm needs all five: the loads are not sunk into the
            // branches below (each sunk load is another round trip)
            asm volatile("" ::"v"(m0), "v"(m1), "v"(m2), "v"(m3), "v"(m4));
            const vf4 mv[5] = {m0, m1, m2, m3, m4};
            static_assert(sizeof(kf_material) == sizeof(mv), "material record is 80 bytes");
            __builtin_memcpy(&m, mv, sizeof(m));
        }
        v3 emission = mk(m.emission[0], m.emission[1], m.emission[2]) * m.emission[3];
        if (!all_zero(emission)) {
            em[p] = f4(emission, __uint_as_float(1u));
            if (aov) {
                albedo_out[p] = make_float4(0.f, 0.f, 0.f, 1.f);
                normal_out[p] = make_float4(0.f, 0.f, 0.f, 1.f);
            }
            if (NEE_INLINE) nee_a[(size_t)p * K] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            auto tex = [&](int id) -> float4 {
                return tex_sample(S, S.tex_data, S.tex_info[id], uvx, uvy, true);
            };
            auto valid_tex = [&](int id) { return id >= 0 && (uint32_t)id < S.n_tex; };
            v3 base = mk(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
            if (valid_tex(m.diffuse_tex)) base = xyz(tex(m.diffuse_tex));
            base = base / kPi;
            float metallic = m.metallic;
            if (valid_tex(m.metallic_tex)) metallic = tex(m.metallic_tex).x;
            float roughness = m.roughness;
            if (valid_tex(m.roughness_tex)) roughness = tex(m.roughness_tex).x;
            float a2 = roughness * roughness;
            float transmission = m.transmission;
            if (valid_tex(m.transmission_tex)) transmission = tex(m.transmission_tex).x;
            float f = gmax(m.ior, 1e-5f);
            float diffuse_weight = (1.0f - gclamp(metallic, 0.0f, 1.0f)) * (1.0f - gclamp(transmission, 0.0f, 1.0f));
            float final_transmission = gclamp(transmission, 0.0f, 1.0f) * (1.0f - gclamp(metallic, 0.0f, 1.0f));
            float specular_weight = 1.0f - final_transmission;
            v3 weight = splat(0.0f);
            v3 V = normalize(-rdir);
            N = inside ? -N : N;
            v3 L = splat(0.0f);
            float NdotV = dot(N, V);
            v3 diffuseColor = diffuse_weight * base;
            v3 specularColor = specular_weight * (base * metallic + splat(m.specular * 0.08f) * (1.0f - metallic));
            v3 transmissionColor = transmission * base;
            float dl = length(diffuseColor), sl = length(specularColor);
            float pD = dl / (dl + sl);
            if (dl == 0.0f && sl == 0.0f) {
                if (all_zero(base)) {
                    if (diffuse_weight == 1.0f) pD = 1.0f;
                    else if (diffuse_weight == 0.0f) pD = 0.0f;
                    else pD = 0.5f;
                } else {
                    pD = 0.0f;
                }
            } else {
                pD *= specular_weight;
            }
            bool chooseDiffuse = rnd(rs) < pD;
            if (chooseDiffuse) {
                L = cosine_hemisphere(rs, N);
                float NdotL = gclamp(dot(N, L), 0.0f, 1.0f);
                weight = kPi * diffuseColor * NdotL / pD;
            } else if (final_transmission == 0.0f) {
                v3 H = sample_ggx(rs, a2, N);
                float HdotV = dot(H, V);
                L = 2.0f * HdotV * H - V;
                float NoV = gmax(dot(N, V), 1e-7f);
                float NoL = dot(N, L);
                float NoH = gmax(dot(N, H), 1e-7f);
                float VoH = gmax(dot(V, H), 1e-7f);
                if (NoL >= 0.0f) {
                    float G = geom_shadow(NoV, NoL, a2);
                    v3 F = schlick_fresnel(specularColor, VoH);
                    weight = kPi * F * G * VoH / (NoH * NoV * (1.0f - pD));
                } else {
                    weight = splat(0.0f);
                }
            } else {
                float ior = inside ? 1.0f / f : f;
                float dd = inside ? NdotV * ior : NdotV;
                v3 refr = refract(-V, N, 1.0f / f);
                float reflectProb = !all_zero(refr) ? schlick(dd, f) : 1.0f;
                if (rnd(rs) >= reflectProb) {
                    L = refr;
                    weight = kPi * transmissionColor / (1.0f - pD);
                } else {
                    L = reflect(-V, N);
                    weight = kPi * transmissionColor / (1.0f - pD);
                }
            }
            ShadeCtx sh;
            sh.N = N; sh.V = V; sh.worldPos = worldPos; sh.f = f; sh.a2 = a2;
            sh.diffuseColor = diffuseColor; sh.specularColor = specularColor;
            sh.transmissionColor = transmissionColor; sh.inside = inside;
            // NEE slots in evaluation order (rchit:457-460).  Deferred NEE
            // hands the shading point to the caller, which runs the light loop
            // with the whole block (nee_deferred).
            v3 sum_d = splat(0.0f), sum_p = splat(0.0f), sum_a = splat(0.0f);
            if (!NEE_INLINE) {
                nc.sh = sh;
                nc.root = root;
                nc.on = true;
            }
            const int nl = NEE_INLINE ? LT->n_slots : 0;
            for (int j = 0; j < nl; ++j) {
                const int kind = LT->kind[j], li = LT->index[j];
                v3 Ld, lem;
                float maxd;
                const bool ok = light_eval<true>(S, LT, kind, li, worldPos, rs, Ld, lem, maxd);
                if (!ok) continue;
                float NdotL = dot(N, Ld);
                if (!(NdotL > 0.0f)) continue;
                THit hh;
                TCount cc{0, 0, 0, 0};
                n_shadow++;
                bool occ = traverse<true, false>(S, root, worldPos, Ld, 0.001f, maxd, false, 0u, hh,
                                                 cc, ovf);
                if (!occ) {
                    v3 ca, cb;
                    float th;
                    int mode;
                    direct_split(sh, Ld, lem, ca, cb, th, mode);
                    v3 cj = direct_resolve(ca, cb, th, mode, rs);
                    if (kind == 0) sum_d = sum_d + cj;
                    else if (kind == 1) sum_p = sum_p + cj;
                    else sum_a = sum_a + cj;
                }
            }
            if (NEE_INLINE) nee_a[(size_t)p * K] = f4((sum_d + sum_p) + sum_a, 0.0f);
            ray_o[p] = f4(worldPos, 0.0f);
            ray_d[p] = f4(L, __uint_as_float(root));
            em[p] = make_float4(0.f, 0.f, 0.f, __uint_as_float(0u));
            rw[p] = f4(weight, 0.0f);
            ext = !all_zero(xyz(t4) * weight) && depth + 1u <= F.pc.path_depth;  // as in k_connect
            if (aov) {
                albedo_out[p] = f4(base, 1.0f);
                normal_out[p] = f4(N, 1.0f);
            }
        }
    }
    seed[p] = rs;
    return ext;
}

// k_shade waits on memory most of its time (the hit, the instance record,
// the vertices and the material, one after another), so it is held to 128
// registers for 4 waves per SIMD: 12 spilled registers cost less than the
// fourth wave gains (489 -> 464 us per launch on the 1080p frame).  With
// inline shadow rays it keeps its registers (2 waves).
template <bool NEE_INLINE>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(NEE_INLINE ? 1 : 4))) void k_shade(SceneView S, Frame F, const DevLights* __restrict__ LT, int bounce,
                                              const uint32_t* __restrict__ queue, float4* __restrict__ ray_o,
                                              float4* __restrict__ ray_d, const float4* __restrict__ thr,
                                              const float4* __restrict__ hit, const uint32_t* __restrict__ hit_prim,
                                              float4* __restrict__ rw, float4* __restrict__ em,
                                              uint32_t* __restrict__ seed, float4* __restrict__ nee_a,
                                              float4* __restrict__ nee_b, uint8_t* __restrict__ vis,
                                              float4* __restrict__ sh_o, float4* __restrict__ sh_d,
                                              uint32_t* __restrict__ sh_root, float4* __restrict__ albedo_out,
                                              float4* __restrict__ normal_out,
                                              uint32_t* __restrict__ ext_queue, unsigned long long* __restrict__ cnt) {
    const uint32_t n = (uint32_t)min(cnt[bounce * kCntPerBounce + 0], (unsigned long long)(F.WH * F.nsamp));
    const uint32_t stride = gridDim.x * blockDim.x, tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t iters = (n + stride - 1) / stride;
    uint32_t n_shadow = 0, ovf = 0;  // shadow rays of this thread
    // A software pipeline over the thread's paths: the queue entry two paths
    // ahead and the inputs of the next path are fetched before this path is
    // shaded, so that their memory round trips overlap its shading.
    auto fetch = [&](uint32_t p) __attribute__((always_inline)) {
        return shade_in(p, bounce == 0 ? kPathStart : thr[p], hit, hit_prim, ray_o, ray_d, seed);
    };
    bool active = tid < n, active1 = iters > 1 && stride + tid < n;
    uint32_t p = active ? queue[tid] : 0, p1 = active1 ? queue[stride + tid] : 0;
    ShadeIn in{};
    if (active) in = fetch(p);
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t q2 = (it + 2) * stride + tid;
        const bool active2 = it + 2 < iters && q2 < n;
        const uint32_t p2 = active2 ? queue[q2] : 0;
        ShadeIn in1{};
        if (active1) in1 = fetch(p1);
        const uint32_t depth = __float_as_uint(in.t4.w);
        // The path is traced at the next bounce unless k_connect ends it for
        // a reason already known here (miss, emitter, zero throughput, depth):
        // only Russian roulette, decided in k_connect, can still end it.
        bool ext = false;
        bool aov = active && depth == 0 && F.s0 == 0 && p < F.WH;  // sample 0, first segment
        NeeCtx nc;
        nc.on = false;
        if (active)
            ext = shade_path<NEE_INLINE>(S, F, LT, bounce, p, in, aov, ray_o, ray_d, rw, em, seed, nee_a, nee_b, vis,
                                         sh_o, sh_d, sh_root, albedo_out, normal_out, cnt, n_shadow, ovf, nc);
        if (!NEE_INLINE)
            nee_deferred(S, LT, nc, p, F.slots, nee_a, nee_b, sh_o, sh_d, sh_root, &cnt[bounce * kCntPerBounce + 1],
                         n_shadow);
        const uint32_t at = block_append(&cnt[(bounce + 1) * kCntPerBounce + kCntExtQueue], ext);
        if (ext) ext_queue[at] = p;
        active = active1;
        p = p1;
        in = in1;
        active1 = active2;
        p1 = p2;
    }
    wave_add(&cnt[kCntGlobal + 2], n_shadow);
    if (NEE_INLINE && ovf) atomicAdd(&cnt[kCntGlobal + 5], 1ull);
}

__global__ __launch_bounds__(TB) __attribute__((amdgpu_num_vgpr(48))) void k_shadow(SceneView S, int bounce, uint32_t cap, const float4* __restrict__ sh_o,
                                               const float4* __restrict__ sh_d, const uint32_t* __restrict__ sh_root,
                                               uint8_t* __restrict__ vis,
                                               unsigned long long* __restrict__ cnt, uint32_t refill, uint32_t grab) {
    const uint32_t n = (uint32_t)min(cnt[bounce * kCntPerBounce + 1], (unsigned long long)cap);
    const uint32_t root0 = env_root_item(S, 0u);
    uint32_t ovf = 0;
    TCount tot{0, 0, 0, 0};
    auto load = [&](uint32_t q, TRay& r) __attribute__((always_inline)) {
        const float4 o = sh_o[q], d = sh_d[q];
        r.o = xyz(o);
        r.d = xyz(d);
        r.tmin = 0.001f;
        r.tmax = d.w;
        r.root = S.n_env > 1 ? sh_root[q] : root0;
        r.alpha_seed = 0u;
        r.payload = __float_as_uint(o.w);
    };
    auto store = [&](uint32_t slot, THit, bool found) __attribute__((always_inline)) {
        if (!found) vis[slot] = 1;
    };
    trace_persistent<true, false>(S, n, cnt + bounce * kCntPerBounce + kCntShadowChunks, refill, grab, false, load,
                                  store, ovf, tot);
    if (ovf) atomicAdd(&cnt[kCntGlobal + 5], 1ull);
}

// Path update of path p after this bounce (rgen:109-138): emission, light
// contributions in light order, throughput, Russian roulette.  Returns
// whether the path continues; a finished path writes its radiance to out.
template <bool NEE_INLINE>
KF_DEV bool connect_path(const Frame& F, int bounce, uint32_t p, float4* __restrict__ thr, float4* __restrict__ col,
                         const float4* __restrict__ rw, const float4* __restrict__ em, uint32_t* __restrict__ seed,
                         const float4* __restrict__ nee_a, const float4* __restrict__ nee_b,
                         const uint8_t* __restrict__ vis, float4* __restrict__ out) {
    const uint32_t K = F.slots;
    const uint32_t maxDepth = F.pc.path_depth;
    bool cont = false;
    const bool first = bounce == 0;  // the path's state is kPathStart, colour 0, weight 0
    float4 t4 = first ? kPathStart : thr[p];
    v3 weight = xyz(t4);
    uint32_t depth = __float_as_uint(t4.w);
    v3 color = first ? splat(0.0f) : xyz(col[p]);
    float4 e4 = em[p];
    bool term = __float_as_uint(e4.w) != 0u;
    uint32_t rs = seed[p];
    v3 shadow = splat(0.0f);
    if (!term) {
        if (NEE_INLINE) {
            shadow = xyz(nee_a[(size_t)p * K]);
        } else {
            v3 sd = splat(0.0f), sp = splat(0.0f), sa = splat(0.0f);
            for (uint32_t j = 0; j < K; ++j) {
                size_t slot = (size_t)p * K + j;
                if (!vis[slot]) continue;
                float4 a = nee_a[slot], b = nee_b[slot];
                int mb = __float_as_int(b.w);
                v3 cj = direct_resolve(xyz(a), xyz(b), a.w, mb & 15, rs);
                int kind = mb >> 4;
                if (kind == 0) sd = sd + cj;
                else if (kind == 1) sp = sp + cj;
                else sa = sa + cj;
            }
            shadow = (sd + sp) + sa;
        }
    }
    // rgen:109-138
    color = color + xyz(e4) * weight;
    // k_shade wrote rw for every path it continues; a path it ended at
    // bounce 0 still has the initial weight 0
    weight = weight * ((first && term) ? splat(0.0f) : xyz(rw[p]));
    color = color + shadow * weight;
    uint32_t d2 = term ? maxDepth + 1 : depth;
    cont = true;
    if (all_zero(weight)) {
        cont = false;
    } else if (F.pc.russian_roulette && d2 >= F.pc.rr_min_bounces) {
        float pr = gmax(weight.x, gmax(weight.y, weight.z));
        float rr = rnd(rs);
        if (rr > pr) cont = false;
        else weight = weight * (1.0f / pr);
    }
    uint32_t nd = d2 + 1;
    if (nd > maxDepth) cont = false;
    if (cont) {
        thr[p] = f4(weight, __uint_as_float(nd));
        col[p] = f4(color, 0.0f);
        seed[p] = rs;
    } else {
        out[p] = f4(color, 0.0f);
    }
    return cont;
}

template <bool NEE_INLINE>
__global__ __launch_bounds__(TB) void k_connect(Frame F, int bounce, const uint32_t* __restrict__ queue,
                                                uint32_t* __restrict__ next, float4* __restrict__ thr,
                                                float4* __restrict__ col, const float4* __restrict__ rw,
                                                const float4* __restrict__ em, uint32_t* __restrict__ seed,
                                                const float4* __restrict__ nee_a, const float4* __restrict__ nee_b,
                                                const uint8_t* __restrict__ vis, float4* __restrict__ out,
                                                unsigned long long* __restrict__ cnt) {
    uint32_t n = (uint32_t)min(cnt[bounce * kCntPerBounce + 0], (unsigned long long)(F.WH * F.nsamp));
    uint32_t stride = gridDim.x * blockDim.x;
    uint32_t iters = (n + stride - 1) / stride;
    for (uint32_t it = 0; it < iters; ++it) {
        uint32_t q = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool cont = false;
        uint32_t p = 0;
        if (q < n) {
            p = queue[q];
            cont = connect_path<NEE_INLINE>(F, bounce, p, thr, col, rw, em, seed, nee_a, nee_b, vis, out);
        }
        uint32_t at = block_append(&cnt[bounce * kCntPerBounce + 2], cont);
        if (cont) next[at] = p;
    }
}

// The last bounces of the paths still active at bounce b0, each path carried
// to its end by one thread: closest hit (as k_extend traces it), shading with
// the shadow rays traced inline, path update, and again.  Few paths survive
// Russian roulette, and the wavefront would then wait, bounce after bounce,
// for the slowest ray of each launch; here a launch lasts as long as the
// longest remaining path.  The random streams and light order are those of
// the wavefront kernels, so the radiance is the same.
__global__ __launch_bounds__(TB) void k_tail(SceneView S, Frame F, const DevLights* __restrict__ LT, int b0,
                                             const uint32_t* __restrict__ queue, float4* __restrict__ ray_o,
                                             float4* __restrict__ ray_d, float4* __restrict__ thr,
                                             float4* __restrict__ col, float4* __restrict__ hit,
                                             uint32_t* __restrict__ hit_prim, float4* __restrict__ rw,
                                             float4* __restrict__ em, uint32_t* __restrict__ seed,
                                             float4* __restrict__ nee_a, float4* __restrict__ nee_b,
                                             uint8_t* __restrict__ vis, float4* __restrict__ out,
                                             unsigned long long* __restrict__ cnt, uint32_t min_n) {
    const uint32_t n = (uint32_t)min(cnt[b0 * kCntPerBounce + 0], (unsigned long long)(F.WH * F.nsamp));
    if (n <= min_n) return;  // k_tail2 takes these paths
    const uint32_t S0 = F.fv->S0;
    uint32_t n_shadow = 0, n_ext = 0, n_traced = 0, ovf = 0;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const uint32_t p = queue[q];
        for (int b = b0; b < kMaxBounces; ++b) {
            // closest hit (rgen:97-107), alpha coin seeded as in k_extend
            const float4 o = ray_o[p], d = ray_d[p];
            const uint32_t aseed =
                S.any_alpha ? tea(tea((p % F.WH) % F.WHe, S0 + 1u + F.s0 + p / F.WH), (uint32_t)b) : 0u;
            THit h;
            TCount c{0, 0, 0, 0};
            traverse<false, false>(S, __float_as_uint(d.w), xyz(o), xyz(d), 0.001f, 10000.0f, S.any_alpha != 0u, aseed, h, c,
                                   ovf);
            ++n_traced;
            hit[p] = make_float4(h.t, h.b1, h.b2, __uint_as_float(h.inst));
            hit_prim[p] = h.prim | (h.back << 31);
            NeeCtx nc;
            shade_path<true>(S, F, LT, b, p, shade_in(p, thr[p], hit, hit_prim, ray_o, ray_d, seed), false, ray_o,
                             ray_d, rw, em, seed, nee_a, nee_b, vis, nullptr, nullptr, nullptr, nullptr, nullptr, cnt,
                             n_shadow, ovf, nc);
            if (!connect_path<true>(F, b, p, thr, col, rw, em, seed, nee_a, nee_b, vis, out)) break;
            ++n_ext;  // traced at the next bounce
        }
    }
    wave_add(&cnt[kCntGlobal + 1], n_ext);
    wave_add(&cnt[kCntGlobal + 2], n_shadow);
    wave_add(&cnt[kCntGlobal + 9], n_traced);
    if (ovf) atomicAdd(&cnt[kCntGlobal + 5], 1ull);
}

// The deferred light loop of one path (rchit:457-460 order) with its shadow
// rays written at the path's slots instead of appended to the shadow queue:
// sh_o[slot] = (shading point, 1 if the slot has a shadow ray), sh_d[slot] =
// (direction, length).  Returns whether any slot has one.
KF_DEV bool nee_slots(const SceneView& S, const DevLights* __restrict__ LT, const NeeCtx& nc, uint32_t p, uint32_t K,
                      float4* __restrict__ nee_a, float4* __restrict__ nee_b, float4* __restrict__ sh_o,
                      float4* __restrict__ sh_d) {
    bool any = false;
    uint32_t unused_rs = 0u;
    for (uint32_t j = 0; j < K; ++j) {
        const size_t slot = (size_t)p * K + j;
        bool emit = false;
        v3 Ld = splat(0.0f);
        float maxd = 0.0f;
        if (nc.on && (int)j < LT->n_slots) {
            const int kind = LT->kind[j], li = LT->index[j];
            v3 lem;
            const bool ok = light_eval<false>(S, LT, kind, li, nc.sh.worldPos, unused_rs, Ld, lem, maxd);
            if (ok && dot(nc.sh.N, Ld) > 0.0f) {
                v3 ca, cb;
                float th;
                int mode;
                direct_split(nc.sh, Ld, lem, ca, cb, th, mode);
                nee_a[slot] = f4(ca, th);
                nee_b[slot] = f4(cb, __int_as_float(mode | (kind << 4)));
                emit = true;
            }
        }
        sh_o[slot] = f4(nc.sh.worldPos, emit ? 1.0f : 0.0f);
        sh_d[slot] = f4(Ld, maxd);
        any = any || emit;
    }
    return any;
}

// The last bounces, for frames whose shadow rays draw no random number before
// they are traced (no soft light): two lanes per path, the first tracing the
// closest-hit ray of bounce b while the second traces the shadow rays of
// bounce b - 1 with the same traversal steps (Trav<MIXED>), as the wavefront
// overlaps its two streams; then the first lane resolves bounce b - 1's
// lights and updates the path, and shades bounce b (its shadow rays written
// per slot).  The closest-hit ray of a path that Russian roulette then ends
// is traced for nothing (counted apart, as in the wavefront).  Random
// streams, light order and radiance are those of the wavefront and k_tail;
// a path's chain of dependent traversals is about half as long, which pays
// while few paths are left: the frame launches k_tail2 for at most max_n
// paths (a quarter of its threads) and k_tail beyond (1080p, 16 spp: 33 k
// paths at bounce 5; 1 spp: hundreds of thousands at bounce 3).
__global__ __launch_bounds__(TB) void k_tail2(SceneView S, Frame F, const DevLights* __restrict__ LT, int b0,
                                              const uint32_t* __restrict__ queue, float4* __restrict__ ray_o,
                                              float4* __restrict__ ray_d, float4* __restrict__ thr,
                                              float4* __restrict__ col, float4* __restrict__ hit,
                                              uint32_t* __restrict__ hit_prim, float4* __restrict__ rw,
                                              float4* __restrict__ em, uint32_t* __restrict__ seed,
                                              float4* __restrict__ nee_a, float4* __restrict__ nee_b,
                                              uint8_t* __restrict__ vis, float4* __restrict__ sh_o,
                                              float4* __restrict__ sh_d, float4* __restrict__ out,
                                              unsigned long long* __restrict__ cnt, uint32_t max_n) {
    const uint32_t n = (uint32_t)min(cnt[b0 * kCntPerBounce + 0], (unsigned long long)(F.WH * F.nsamp));
    if (n > max_n) return;  // k_tail takes these paths
    const uint32_t S0 = F.fv->S0;
    const uint32_t K = F.slots;
    const uint32_t lane = threadIdx.x & 63u;
    const bool second = (lane & 1u) != 0u;  // traces the shadow rays
    const uint32_t npair = (gridDim.x * blockDim.x) >> 1;
    uint32_t n_shadow = 0, n_ext = 0, n_traced = 0, n_spec = 0, ovf = 0, unused = 0;
    const TravScene TS(S);
    PrivStack stk;
    Trav<false, false, true> T;
    T.reset();
    for (uint32_t q = (blockIdx.x * blockDim.x + threadIdx.x) >> 1; q < n; q += npair) {
        const uint32_t p = queue[q];
        const uint32_t root_env = env_root_item(S, (p % F.WH) / F.WHe);
        int b = b0;
        bool closest = true;   // a closest-hit ray of bounce b to trace
        bool pending = false;  // shadow rays of bounce b - 1 to trace
        for (;;) {
            // traversal: the first lane's closest-hit query, the second
            // lane's shadow queries one slot after another
            uint32_t j = 0;
            size_t slot = 0;
            bool alive = false;
            auto start_shadow = [&]() {
                for (; j < K; ++j) {
                    slot = (size_t)p * K + j;
                    const float4 so = sh_o[slot];
                    if (so.w == 0.0f) continue;
                    const float4 sd = sh_d[slot];
                    TRay r;
                    r.o = xyz(so);
                    r.d = xyz(sd);
                    r.tmin = 0.001f;
                    r.tmax = sd.w;
                    r.root = root_env;
                    r.alpha_seed = 0u;
                    r.payload = 0u;
                    T.start(true, r, r.root, false);
                    T.any_q = true;
                    ++n_shadow;
                    if (!(r.tmin <= r.tmax) || r.root == kEmptyItem) {
                        vis[slot] = 1;  // nothing can occlude it
                        continue;
                    }
                    ++j;
                    return true;
                }
                return false;
            };
            if (!second && closest) {
                const float4 o = ray_o[p], d = ray_d[p];
                TRay r;
                r.o = xyz(o);
                r.d = xyz(d);
                r.tmin = 0.001f;
                r.tmax = 10000.0f;
                r.root = __float_as_uint(d.w);
                // alpha coin seed as in k_extend
                r.alpha_seed = S.any_alpha ? tea(tea((p % F.WH) % F.WHe, S0 + 1u + F.s0 + p / F.WH), (uint32_t)b) : 0u;
                r.payload = 0u;
                T.start(true, r, r.root, S.any_alpha != 0u);
                T.any_q = false;
                ++n_traced;
                alive = r.root != kEmptyItem;
                if (!alive) {
                    const THit h = T.result();
                    hit[p] = make_float4(h.t, h.b1, h.b2, __uint_as_float(h.inst));
                    hit_prim[p] = h.prim | (h.back << 31);
                }
            }
            if (second && pending) alive = start_shadow();
            while (__ballot(alive) != 0ull) {
                const bool more = T.template step<true>(TS, stk, ovf, alive);
                if (alive && !more) {
                    alive = false;
                    if (!second) {
                        const THit h = T.result();
                        hit[p] = make_float4(h.t, h.b1, h.b2, __uint_as_float(h.inst));
                        hit_prim[p] = h.prim | (h.back << 31);
                    } else {
                        vis[slot] = T.found ? 0 : 1;
                        alive = start_shadow();
                    }
                }
            }
            // the second lane's visibility bytes before the first lane reads them
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            int flags = 0;  // bit 0: shade bounce b, bit 1: its ray continues, bit 2: it has shadow rays
            if (!second) {
                bool done = false;
                if (b > b0) {  // bounce b - 1: lights in order, path update, Russian roulette
                    if (connect_path<false>(F, b - 1, p, thr, col, rw, em, seed, nee_a, nee_b, vis, out)) {
                        ++n_ext;
                    } else {
                        done = true;
                        if (closest) ++n_spec;
                    }
                }
                if (!done && closest) {
                    NeeCtx nc;
                    nc.on = false;
                    const bool ext = shade_path<false>(S, F, LT, b, p,
                                                       shade_in(p, thr[p], hit, hit_prim, ray_o, ray_d, seed), false,
                                                       ray_o, ray_d, rw, em, seed, nee_a, nee_b, vis, nullptr,
                                                       nullptr, nullptr, nullptr, nullptr, cnt, unused, ovf, nc);
                    const bool sh = nee_slots(S, LT, nc, p, K, nee_a, nee_b, sh_o, sh_d);
                    flags = 1 | (ext ? 2 : 0) | (sh ? 4 : 0);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            flags = __shfl(flags, (int)(lane & ~1u));
            if (!(flags & 1)) break;
            closest = (flags & 2) != 0;
            pending = (flags & 4) != 0;
            ++b;
        }
    }
    wave_add(&cnt[kCntGlobal + 1], n_ext);
    wave_add(&cnt[kCntGlobal + 2], n_shadow);
    wave_add(&cnt[kCntGlobal + 7], n_spec);
    wave_add(&cnt[kCntGlobal + 9], n_traced);
    if (ovf) atomicAdd(&cnt[kCntGlobal + 5], 1ull);
}

__global__ void k_next_count(int bounce, unsigned long long* cnt) {
    unsigned long long nx = cnt[bounce * kCntPerBounce + 2];
    cnt[(bounce + 1) * kCntPerBounce + 0] = nx;
    cnt[kCntGlobal + 1] += nx;
    // next-bounce rays traced for paths Russian roulette then ended
    cnt[kCntGlobal + 7] += cnt[(bounce + 1) * kCntPerBounce + kCntExtQueue] - nx;
}

// Sum the batch's samples into acc in sample order (rgen:140), then on the
// last batch: average, running mean over frames (rgen:153-163).
__global__ __launch_bounds__(TB) void k_resolve(Frame F, const float4* __restrict__ out, float4* __restrict__ acc,
                                                float4* __restrict__ rgba, float* __restrict__ keep, int last) {
    uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= F.WH) return;
    v3 c = F.s0 == 0 ? splat(0.0f) : xyz(acc[px]);
    for (uint32_t i = 0; i < F.nsamp; ++i) {
        v3 s = xyz(out[(size_t)i * F.WH + px]);
        c = c + s;
        if (keep) {
            float* k = keep + 3 * ((size_t)(F.s0 + i) * F.WH + px);
            k[0] = s.x; k[1] = s.y; k[2] = s.z;
        }
    }
    if (!last) {
        acc[px] = f4(c, 0.0f);
        return;
    }
    v3 fc = c / float(F.pc.spp);
    const int32_t frame_count = F.fv->frame_count;
    if (frame_count <= 0) {
        rgba[px] = f4(fc, 1.0f);
    } else {
        float a = 1.0f / float(frame_count + 1);
        v3 old = xyz(rgba[px]);
        v3 m = old * (1.0f - a) + fc * a;
        rgba[px] = f4(m, 1.0f);
    }
}

// B8G8R8A8_SRGB bytes of the rgba target (PostProcessing.frag:17 samples the
// texel; the attachment's sRGB encode clamps, encodes and rounds).
__global__ void k_to_bgra8(uint32_t n, const float4* __restrict__ rgba, const double* __restrict__ thr,
                           uchar4* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c = rgba[i];
    auto enc = [&](float x) -> unsigned char {
        if (!(x > 0.0f)) return 0;
        int lo = 0, hi = 255;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if ((double)x >= thr[mid]) lo = mid + 1; else hi = mid;
        }
        return (unsigned char)lo;
    };
    unsigned char a;
    if (!(c.w > 0.0f)) a = 0;
    else if (c.w >= 1.0f) a = 255;
    else a = (unsigned char)(int)rintf(c.w * 255.0f);
    out[i] = make_uchar4(enc(c.z), enc(c.y), enc(c.x), a);
}

template <class T> T* dmalloc(size_t n) {
    T* p = nullptr;
    if (n) KF_CHECK(hipMalloc((void**)&p, n * sizeof(T)));
    return p;
}

}  // namespace

Renderer::~Renderer() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (st2) {
        (void)hipStreamSynchronize(st2);
        (void)hipStreamDestroy(st2);
    }
    for (auto e : ev) (void)hipEventDestroy(e);
    for (auto e : sync_ev) (void)hipEventDestroy(e);
    if (h_totals) (void)hipHostFree(h_totals);
    void* ptrs[] = {d_rgba, d_albedo, d_normal, d_bgra, ray_o, ray_d, thr, col, rw, em, hit, hit_prim, seed,
                    queue_a, queue_b, queue_e, sh_o, sh_d, sh_root, nee_a, nee_b, vis, samples, counters, d_lights,
                    d_jump, d_acc, d_out, d_cams, d_fv, d_ftlas, d_tile_root};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

void Renderer::ensure(size_t paths, uint32_t nslots, uint32_t spp) {
    if (paths > cap || nslots > slots) {
        void* ptrs[] = {ray_o, ray_d, thr, col, rw, em, hit, hit_prim, seed, queue_a, queue_b, queue_e,
                        sh_o, sh_d, sh_root, nee_a, nee_b, vis};
        for (void* p : ptrs)
            if (p) KF_CHECK(hipFree(p));
        cap = std::max(cap, paths);
        slots = std::max(slots, std::max<uint32_t>(nslots, 1));
        ray_o = dmalloc<float4>(cap);
        ray_d = dmalloc<float4>(cap);
        thr = dmalloc<float4>(cap);
        col = dmalloc<float4>(cap);
        rw = dmalloc<float4>(cap);
        em = dmalloc<float4>(cap);
        hit = dmalloc<float4>(cap);
        hit_prim = dmalloc<uint32_t>(cap);
        seed = dmalloc<uint32_t>(cap);
        queue_a = dmalloc<uint32_t>(cap);
        queue_b = dmalloc<uint32_t>(cap);
        queue_e = dmalloc<uint32_t>(cap);
        sh_o = dmalloc<float4>(cap * slots);
        sh_d = dmalloc<float4>(cap * slots);
        sh_root = dmalloc<uint32_t>(cap * slots);
        nee_a = dmalloc<float4>(cap * slots);
        nee_b = dmalloc<float4>(cap * slots);
        vis = dmalloc<uint8_t>(cap * slots);
    }
    size_t sneed = (size_t)spp * W * H * 3;
    if (spp && sneed > samples_cap) {
        if (samples) KF_CHECK(hipFree(samples));
        samples = dmalloc<float>(sneed);
        samples_cap = sneed;
    }
    if (!counters) counters = dmalloc<unsigned long long>(kCntSize);
}

// ------------------------------------------------------------------ frame
struct FrameOut {
    float trace_ms, extend_ms, shadow_ms;
};

// Blocks of a TB-thread kernel that are resident at once on n_cu CUs.
template <class Kernel>
uint32_t resident_grid(Kernel kernel, int n_cu) {
    int nb = 0;
    KF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, TB, 0));
    return (uint32_t)std::max(nb, 1) * (uint32_t)n_cu;
}

constexpr uint32_t kLastShadowGrab = 32;  // queries per wave reservation, last wavefront bounce's shadow rays

// Default batch: up to 32M paths (16 spp at 1920x1080).  Each bounce's
// traversal launch then holds tens of rays per lane, so the tail of the
// slowest rays is a small part of it; the state costs about 200 B per path
// (6.6 GB at 32M paths).
uint32_t default_batch(uint32_t spp, uint32_t WH) {
    return std::max<uint32_t>(1, std::min<uint32_t>(spp, (32u << 20) / std::max<uint32_t>(WH, 1)));
}

void build_light_list(const Scene& s, DevLights& L, bool& inline_nee) {
    L.dl = s.dl;
    L.pl = s.pl;
    L.al = s.al;
    inline_nee = false;
    int n = 0;
    // traceDirectionalLight returns early when rgb * strength == 0
    float em[3] = {s.dl.rgbs[0] * s.dl.rgbs[3], s.dl.rgbs[1] * s.dl.rgbs[3], s.dl.rgbs[2] * s.dl.rgbs[3]};
    if (em[0] != 0.f || em[1] != 0.f || em[2] != 0.f) {
        L.kind[n] = 0; L.index[n] = 0; ++n;
        if (s.dl.direction[3] != 0.f) inline_nee = true;
    }
    for (int i = 0; i < KF_MAX_POINT_LIGHTS; ++i) {
        if (!(s.pl.rgbs[i][3] > 0.f)) continue;
        float r = s.pl.rgbs[i][0] * s.pl.rgbs[i][3], g = s.pl.rgbs[i][1] * s.pl.rgbs[i][3],
              b = s.pl.rgbs[i][2] * s.pl.rgbs[i][3];
        if (sqrtf(r * r + g * g + b * b) == 0.f) continue;
        L.kind[n] = 1; L.index[n] = i; ++n;
        if (s.pl.posr[i][3] != 0.f) inline_nee = true;
    }
    for (int i = 0; i < KF_MAX_ACTIVE_LIGHTS; ++i) {
        if (!(s.al.front[i][3] > 0.f)) continue;
        float r = s.al.rgbs[i][0] * s.al.rgbs[i][3], g = s.al.rgbs[i][1] * s.al.rgbs[i][3],
              b = s.al.rgbs[i][2] * s.al.rgbs[i][3];
        if (sqrtf(r * r + g * g + b * b) == 0.f) continue;
        L.kind[n] = 2; L.index[n] = i; ++n;
        if (s.al.sftp[i][0] != 0.f) inline_nee = true;
        // proj * view, column order (rchit:303)
        for (int c = 0; c < 4; ++c)
            for (int r2 = 0; r2 < 4; ++r2)
                L.pv[i][c * 4 + r2] = s.al.proj[i][0 * 4 + r2] * s.al.view[i][c * 4 + 0] +
                                      s.al.proj[i][1 * 4 + r2] * s.al.view[i][c * 4 + 1] +
                                      s.al.proj[i][2 * 4 + r2] * s.al.view[i][c * 4 + 2] +
                                      s.al.proj[i][3 * 4 + r2] * s.al.view[i][c * 4 + 3];
    }
    L.n_slots = n;
}

}  // namespace kf

using namespace kf;

namespace kf {
namespace {

// One frame of n_env environments (n_env = 1: the plain kf_render).
void render_frame(Renderer* r, Scene* s, uint32_t n_env, const kf_camera_ubo* cams, const kf_push_constants* pc,
                  const kf_render_opts* opts, kf_render_stats* stats) {
    if (!r || !s || !cams || !pc) throw Error("kf_render: null argument");
    if (pc->spp == 0) throw Error("kf_render: spp must be > 0");
    if (pc->path_depth + 2 >= (uint32_t)kMaxBounces) throw Error("kf_render: path depth too large");
    if (n_env == 0 || n_env != s->n_env)
        throw Error("kf_render_envs: the number of cameras must equal the scene's number of environments");
    if (r->H % n_env) throw Error("kf_render_envs: the target height must be a multiple of the number of environments");
    hipStream_t st = stream();
    kf_render_opts o{};
    if (opts) o = *opts;
    DevLights L;
    memset(&L, 0, sizeof(L));
    bool inline_nee = false;
    build_light_list(*s, L, inline_nee);
    uint32_t WH = r->W * r->H;
    uint32_t spb = o.samples_per_batch ? o.samples_per_batch : r->tune_spb;
    if (spb == 0) spb = default_batch(pc->spp, WH);
    spb = std::min(spb, pc->spp);
    uint32_t K = inline_nee ? 1u : (uint32_t)std::max(L.n_slots, 1);
    r->ensure((size_t)WH * spb, K, o.keep_samples ? pc->spp : 0);
    r->samples_spp = o.keep_samples ? pc->spp : 0;
    // device light list (per frame, like the UBO upload)
    if (!r->d_lights) KF_CHECK(hipMalloc(&r->d_lights, sizeof(DevLights)));
    DevLights* d_lights = (DevLights*)r->d_lights;
    KF_CHECK(hipMemcpyAsync(d_lights, &L, sizeof(DevLights), hipMemcpyHostToDevice, st));
    // LCG jump-ahead for the jitter stream: sample i starts after 2i draws
    std::vector<uint2> jump(pc->spp);
    {
        uint32_t a = 1, c = 0;  // identity
        for (uint32_t i = 0; i < pc->spp; ++i) {
            jump[i] = make_uint2(a, c);
            for (int k = 0; k < 2; ++k) {  // compose with one LCG step
                a = 1664525u * a;
                c = 1664525u * c + 1013904223u;
            }
        }
    }
    if (pc->spp > r->jump_cap) {
        if (r->d_jump) KF_CHECK(hipFree(r->d_jump));
        KF_CHECK(hipMalloc((void**)&r->d_jump, pc->spp * sizeof(uint2)));
        r->jump_cap = pc->spp;
    }
    uint2* d_jump = r->d_jump;
    KF_CHECK(hipMemcpyAsync(d_jump, jump.data(), pc->spp * sizeof(uint2), hipMemcpyHostToDevice, st));
    if (!r->d_acc) KF_CHECK(hipMalloc((void**)&r->d_acc, (size_t)WH * sizeof(float4)));
    float4* d_acc = r->d_acc;
    if ((size_t)WH * spb > r->out_cap) {
        if (r->d_out) KF_CHECK(hipFree(r->d_out));
        KF_CHECK(hipMalloc((void**)&r->d_out, (size_t)WH * spb * sizeof(float4)));
        r->out_cap = (size_t)WH * spb;
    }
    float4* d_out = r->d_out;

    if (n_env > r->cams_cap) {
        if (r->d_cams) KF_CHECK(hipFree(r->d_cams));
        KF_CHECK(hipMalloc((void**)&r->d_cams, n_env * sizeof(kf_camera_ubo)));
        r->cams_cap = n_env;
    }
    KF_CHECK(hipMemcpyAsync(r->d_cams, cams, n_env * sizeof(kf_camera_ubo), hipMemcpyHostToDevice, st));

    if (!r->d_fv) KF_CHECK(hipMalloc(&r->d_fv, sizeof(FrameVars)));
    FrameVars fv{tea_host(o.seed, (uint32_t)pc->frame_count), pc->frame_count};
    KF_CHECK(hipMemcpyAsync(r->d_fv, &fv, sizeof(fv), hipMemcpyHostToDevice, st));

    SceneView S = s->view();
    // Camera-ray roots per 8x8 pixel tile, over a copy of the TLAS followed
    // by the tiles' nodes (pinhole cameras; k_tile_roots falls back to the
    // environment's root per tile where that does not apply)
    const uint32_t He = r->H / n_env;
    const uint32_t tx = (r->W + kTile - 1) / kTile, ty = (He + kTile - 1) / kTile;
    const size_t n_tiles = (size_t)tx * ty * n_env;
    const size_t n_fnodes = (size_t)S.n_tlas_nodes + kTileNodes * n_tiles;
    const bool tiles = !o.full_tlas_camera && s->n_inst > 0 && n_fnodes < (size_t)kIdxMask;
    if (tiles) {
        if (n_fnodes > r->ftlas_cap) {
            if (r->d_ftlas) KF_CHECK(hipFree(r->d_ftlas));
            r->d_ftlas = nullptr;
            KF_CHECK(hipMalloc((void**)&r->d_ftlas, n_fnodes * sizeof(Node4)));
            r->ftlas_cap = n_fnodes;
        }
        if (n_tiles > r->tile_cap) {
            if (r->d_tile_root) KF_CHECK(hipFree(r->d_tile_root));
            r->d_tile_root = nullptr;
            KF_CHECK(hipMalloc((void**)&r->d_tile_root, n_tiles * sizeof(uint32_t)));
            r->tile_cap = n_tiles;
        }
    }
    const Node4* scene_tlas = S.tlas;
    const uint32_t n_scene_tlas = S.n_tlas_nodes;
    if (tiles) {
        S.tlas = r->d_ftlas;
        S.n_tlas_nodes = (uint32_t)n_fnodes;
    }
    Frame F;
    memset(&F, 0, sizeof(F));  // F is compared bytewise as part of the graph key
    F.cams = r->d_cams;
    F.pc = *pc;
    F.pc.frame_count = 0;  // read from fv: the same graph serves every frame
    F.fv = (const FrameVars*)r->d_fv;
    F.W = r->W;
    F.H = r->H;
    F.WH = WH;
    F.n_env = n_env;
    F.He = r->H / n_env;
    F.WHe = r->W * F.He;
    F.slots = K;
    F.inline_nee = inline_nee;
    F.count = o.count_traversal;
    r->n_env_last = n_env;

    if (r->ev.empty()) {
        r->ev.resize(2 + 4 * kMaxBounces);
        for (auto& e : r->ev) KF_CHECK(hipEventCreate(&e));
        r->sync_ev.resize(2 * kMaxBounces);
        for (auto& e : r->sync_ev) KF_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        KF_CHECK(hipStreamCreateWithFlags(&r->st2, hipStreamNonBlocking));  // default priority
        KF_CHECK(hipHostMalloc((void**)&r->h_totals, 10 * sizeof(unsigned long long), hipHostMallocDefault));
    }
    int dev = 0;
    KF_CHECK(hipGetDevice(&dev));
    int n_cu = 256;
    KF_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t grid = (uint32_t)n_cu * 8;
    // k_connect runs beside k_extend with a grid of 8 blocks per CU (all
    // resident at its occupancy): on the 128-environment frame, whose
    // connect passes carry two lights per path, that is 5 % faster than 2
    // blocks per CU; on the 1080p frame the two are within noise
    const uint32_t grid_conn = (uint32_t)n_cu * 8;
    const uint32_t refill = std::min<uint32_t>(64u, std::max<uint32_t>(1u, r->tune_refill ? r->tune_refill
                                                                                         : traversal_refill()));
    const uint32_t grab = std::min<uint32_t>(4096u, std::max<uint32_t>(16u, r->tune_grab ? r->tune_grab : kGrab));
    // camera rays (bounce 0): one long, coherent queue that may want another
    // reservation than the scattered later bounces
    const uint32_t grab0 = r->tune_grab0 ? std::min<uint32_t>(4096u, std::max<uint32_t>(16u, r->tune_grab0)) : grab;
    // shadow rays: short any-hit queries, which favour smaller reservations
    const uint32_t grabs = r->tune_grabs ? std::min<uint32_t>(4096u, std::max<uint32_t>(16u, r->tune_grabs)) : grab;
    // Persistent traversal kernels: exactly the blocks that fit at once.
    static const uint32_t grid_ext0 = resident_grid(k_extend<false>, n_cu);
    static const uint32_t grid_ext1 = resident_grid(k_extend<true>, n_cu);
    static const uint32_t grid_sh = resident_grid(k_shadow, n_cu);
    const uint32_t nb = pc->path_depth + 1;
    // bounces from b_tail on run in k_tail (not in counting frames, whose
    // node counts come from the wavefront kernels)
    const uint32_t b_tail = (o.count_traversal || r->tune_tail == 0 || r->tune_tail >= nb) ? nb : r->tune_tail;

    // One batch of samples [s0, s0 + nsamp).  timed: record the per-kernel
    // events (not while capturing a graph).  The first batch clears every
    // counter; later ones only the per-bounce ones, so the global counters
    // add up over the frame.
    uint32_t n_ext = 0, n_sh = 0;
    auto enqueue_batch = [&](uint32_t s0, bool timed) {
        F.s0 = s0;
        F.nsamp = std::min(spb, pc->spp - s0);
        const uint32_t B = WH * F.nsamp;
        const int nz = s0 == 0 ? kCntSize : kCntGlobal;
        k_zero<<<(nz + TB - 1) / TB, TB, 0, st>>>(r->counters, nz);
        if (timed) KF_CHECK(hipEventRecord(r->ev[0], st));
        if (tiles && s0 == 0)
            k_tile_roots<<<(uint32_t)std::min<size_t>((n_tiles + TB / 64 - 1) / (TB / 64), 65535u), TB, 0, st>>>(
                S, F, scene_tlas, n_scene_tlas, s->d_ilo, s->d_ihi, n_env > 1 ? s->d_env_off : nullptr, tx, ty,
                r->d_ftlas, r->d_tile_root);
        k_raygen<<<std::min<uint32_t>((WH + TB - 1) / TB, 65535u), TB, 0, st>>>(
            S, F, d_jump, tiles ? r->d_tile_root : nullptr, tx, ty, r->ray_o, r->ray_d, r->seed, r->queue_a,
            r->counters);
        uint32_t* qa = r->queue_a;
        uint32_t* qb = r->queue_b;
        // Per bounce b:  A: shade(b) -> B: [shadow(b)], connect(b) ; A:
        // extend(b+1) -> (wait B) shade(b+1).  k_shade lists the paths that
        // continue unless Russian roulette ends them, so the closest-hit
        // rays of bounce b+1 are traced on the main stream while the second
        // stream traces the shadow rays of bounce b and connects the paths
        // (k_connect touches neither the rays nor the hits k_extend uses).
        auto launch_extend = [&](uint32_t b, const uint32_t* queue) {
            if (timed) KF_CHECK(hipEventRecord(r->ev[2 + 4 * b + 0], st));
            if (o.count_traversal)
                k_extend<true><<<grid_ext1, TB, 0, st>>>(S, (int)b, B, queue, r->ray_o, r->ray_d, r->seed, r->hit,
                                                         r->hit_prim, r->counters, refill, b == 0 ? grab0 : grab, WH,
                                                         F.WHe, F.s0, F.fv);
            else
                k_extend<false><<<grid_ext0, TB, 0, st>>>(S, (int)b, B, queue, r->ray_o, r->ray_d, r->seed, r->hit,
                                                          r->hit_prim, r->counters, refill, b == 0 ? grab0 : grab,
                                                          WH, F.WHe, F.s0, F.fv);
            if (timed) KF_CHECK(hipEventRecord(r->ev[2 + 4 * b + 1], st));
            ++n_ext;
        };
        launch_extend(0, qa);
        for (uint32_t b = 0; b < nb; ++b) {
            if (b == b_tail) {
                // the path count is on the device: k_tail2 (two lanes per
                // path) runs for at most a quarter of the grid's threads,
                // k_tail beyond, and the other launch returns at once
                const uint32_t pair_max = (inline_nee || !r->tune_pairs) ? 0u : grid * TB / 4u;
                if (!inline_nee)
                    k_tail2<<<grid, TB, 0, st>>>(S, F, d_lights, (int)b, qa, r->ray_o, r->ray_d, r->thr, r->col,
                                                 r->hit, r->hit_prim, r->rw, r->em, r->seed, r->nee_a, r->nee_b,
                                                 r->vis, r->sh_o, r->sh_d, d_out, r->counters, pair_max);
                k_tail<<<grid, TB, 0, st>>>(S, F, d_lights, (int)b, qa, r->ray_o, r->ray_d, r->thr, r->col, r->hit,
                                            r->hit_prim, r->rw, r->em, r->seed, r->nee_a, r->nee_b, r->vis, d_out,
                                            r->counters, pair_max);
                break;
            }
            if (inline_nee)
                k_shade<true><<<grid, TB, 0, st>>>(S, F, d_lights, (int)b, qa, r->ray_o, r->ray_d, r->thr, r->hit,
                                                   r->hit_prim, r->rw, r->em, r->seed, r->nee_a, r->nee_b, r->vis,
                                                   r->sh_o, r->sh_d, r->sh_root, r->d_albedo, r->d_normal,
                                                   r->queue_e, r->counters);
            else
                k_shade<false><<<grid, TB, 0, st>>>(S, F, d_lights, (int)b, qa, r->ray_o, r->ray_d, r->thr, r->hit,
                                                    r->hit_prim, r->rw, r->em, r->seed, r->nee_a, r->nee_b, r->vis,
                                                    r->sh_o, r->sh_d, r->sh_root, r->d_albedo, r->d_normal,
                                                    r->queue_e, r->counters);
            KF_CHECK(hipEventRecord(r->sync_ev[2 * b], st));
            KF_CHECK(hipStreamWaitEvent(r->st2, r->sync_ev[2 * b], 0));
            if (!inline_nee) {
                if (timed) KF_CHECK(hipEventRecord(r->ev[2 + 4 * b + 2], r->st2));
                // the last wavefront bounce's shadow rays run with nothing
                // beside them and the tail waiting: small reservations, so
                // that few waves are left holding the longest rays
                const uint32_t gsh = b + 1 < b_tail ? grabs : std::min<uint32_t>(grabs, kLastShadowGrab);
                k_shadow<<<grid_sh, TB, 0, r->st2>>>(S, (int)b, B * K, r->sh_o, r->sh_d, r->sh_root, r->vis,
                                                     r->counters, refill, gsh);
                if (timed) KF_CHECK(hipEventRecord(r->ev[2 + 4 * b + 3], r->st2));
                ++n_sh;
            }
            if (inline_nee)
                k_connect<true><<<grid_conn, TB, 0, r->st2>>>(F, (int)b, qa, qb, r->thr, r->col, r->rw, r->em,
                                                              r->seed, r->nee_a, r->nee_b, r->vis, d_out, r->counters);
            else
                k_connect<false><<<grid_conn, TB, 0, r->st2>>>(F, (int)b, qa, qb, r->thr, r->col, r->rw, r->em,
                                                               r->seed, r->nee_a, r->nee_b, r->vis, d_out, r->counters);
            k_next_count<<<1, 1, 0, r->st2>>>((int)b, r->counters);
            KF_CHECK(hipEventRecord(r->sync_ev[2 * b + 1], r->st2));
            if (b + 1 < b_tail) launch_extend(b + 1, r->queue_e);
            KF_CHECK(hipStreamWaitEvent(st, r->sync_ev[2 * b + 1], 0));
            std::swap(qa, qb);
        }
        const bool last = s0 + F.nsamp >= pc->spp;
        k_resolve<<<(WH + TB - 1) / TB, TB, 0, st>>>(F, d_out, d_acc, r->d_rgba,
                                                     o.keep_samples ? r->samples : nullptr, last ? 1 : 0);
        if (timed) KF_CHECK(hipEventRecord(r->ev[1], st));
        KF_CHECK(hipGetLastError());
    };

    // The frame's global counters follow its last kernel on the stream into
    // pinned memory, so that one host wait covers both.
    auto read_totals = [&]() {
        KF_CHECK(hipMemcpyAsync(r->h_totals, r->counters + kCntGlobal, 10 * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, st));
        KF_CHECK(hipStreamSynchronize(st));
    };
    float ms_total = 0.f, ms_ext = 0.f, ms_sh = 0.f;
    if (o.use_graph) {
        // Everything the captured kernel arguments hold: scene view, frame
        // constants (seed and count excluded), schedule and buffers.
        std::vector<uint8_t> key;
        auto put = [&](const void* p, size_t n) {
            const uint8_t* b = (const uint8_t*)p;
            key.insert(key.end(), b, b + n);
        };
        const uint32_t sched[12] = {spb, K, nb, refill, grab, grid, (uint32_t)o.count_traversal,
                                    (uint32_t)o.keep_samples, b_tail, grab0, grabs, r->tune_pairs};
        const void* bufs[] = {r->ray_o, r->ray_d, r->thr, r->col, r->rw, r->em, r->hit, r->hit_prim, r->seed,
                              r->queue_a, r->queue_b, r->queue_e, r->sh_o, r->sh_d, r->sh_root, r->nee_a,
                              r->nee_b, r->vis, r->samples, r->counters, d_lights, d_jump, d_acc, d_out,
                              r->d_rgba, r->d_albedo, r->d_normal, (const void*)r->st2,
                              tiles ? r->d_tile_root : nullptr, scene_tlas, s->d_ilo, s->d_ihi};
        put(&S, sizeof(S));
        put(&F, sizeof(F));
        put(sched, sizeof(sched));
        put(bufs, sizeof(bufs));
        if (!r->graph || key != r->graph_key) {
            if (r->graph) KF_CHECK(hipGraphExecDestroy(r->graph));
            r->graph = nullptr;
            hipGraph_t g = nullptr;
            KF_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                for (uint32_t s0 = 0; s0 < pc->spp; s0 += spb) enqueue_batch(s0, false);
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            KF_CHECK(hipStreamEndCapture(st, &g));
            hipError_t e = hipGraphInstantiate(&r->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                r->graph = nullptr;
                throw Error(std::string("kf_render: graph instantiation failed: ") + hipGetErrorString(e));
            }
            r->graph_key = key;
            ++r->graph_captures;
            r->graph_ext = n_ext;
            r->graph_sh = n_sh;
        }
        n_ext = r->graph_ext;
        n_sh = r->graph_sh;
        KF_CHECK(hipEventRecord(r->ev[0], st));
        KF_CHECK(hipGraphLaunch(r->graph, st));
        KF_CHECK(hipEventRecord(r->ev[1], st));
        read_totals();
        KF_CHECK(hipEventElapsedTime(&ms_total, r->ev[0], r->ev[1]));
    } else {
        for (uint32_t s0 = 0; s0 < pc->spp; s0 += spb) {
            enqueue_batch(s0, true);
            if (s0 + spb >= pc->spp)
                read_totals();
            else
                KF_CHECK(hipEventSynchronize(r->ev[1]));
            float ms = 0.f;
            KF_CHECK(hipEventElapsedTime(&ms, r->ev[0], r->ev[1]));
            ms_total += ms;
            for (uint32_t b = 0; b < b_tail; ++b) {  // the tail kernel counts as other time
                float a = 0.f, c = 0.f;
                KF_CHECK(hipEventElapsedTime(&a, r->ev[2 + 4 * b + 0], r->ev[2 + 4 * b + 1]));
                if (!inline_nee) KF_CHECK(hipEventElapsedTime(&c, r->ev[2 + 4 * b + 2], r->ev[2 + 4 * b + 3]));
                ms_ext += a;
                ms_sh += c;
            }
        }
    }
    const unsigned long long* totals = r->h_totals;
    if (totals[5]) throw Error("kf_render: traversal stack overflow (BVH deeper than the stack)");
    if (stats) {
        stats->primary_rays = totals[0];
        stats->extension_rays = totals[1];
        stats->shadow_rays = totals[2];
        stats->node_visits = totals[3];
        stats->tri_tests = totals[4];
        stats->instance_visits = totals[6];
        stats->speculative_rays = totals[7];
        stats->tlas_node_visits = totals[8];
        stats->tail_rays = totals[9];
        stats->graph_captures = r->graph_captures;
        stats->trace_ms = ms_total;
        stats->extend_ms = ms_ext;
        stats->shadow_ms = ms_sh;
        stats->other_ms = ms_total - ms_ext - ms_sh;
        stats->extend_launches = n_ext;
        stats->shadow_launches = n_sh;
    }
}

}  // namespace
}  // namespace kf

extern "C" int kf_render(kf_renderer* r, kf_scene* s, const kf_camera_ubo* cam, const kf_push_constants* pc,
                         const kf_render_opts* opts, kf_render_stats* stats) {
    try {
        if (s && s->n_env > 1) throw Error("kf_render: the scene has environments; use kf_render_envs");
        render_frame(r, s, 1u, cam, pc, opts, stats);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

extern "C" int kf_render_envs(kf_renderer* r, kf_scene* s, uint32_t num_envs, const kf_camera_ubo* cams,
                              const kf_push_constants* pc, const kf_render_opts* opts, kf_render_stats* stats) {
    try {
        render_frame(r, s, num_envs, cams, pc, opts, stats);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

extern "C" int kf_renderer_autotune(kf_renderer* r, kf_scene* s, uint32_t num_envs, const kf_camera_ubo* cams,
                                    const kf_push_constants* pc, kf_tune_result* out) {
    try {
        if (!r || !s || !cams || !pc || !out) throw Error("kf_renderer_autotune: null argument");
        kf_push_constants p = *pc;
        p.frame_count = 0;
        kf_render_opts o{};
        uint32_t trials = 0;
        // kernel time of one frame under a schedule (0 = default); the best
        // of three (of one for frames over 50 ms)
        uint32_t b_tail = 0, b_grab0 = 0, b_grabs = 0, b_pairs = 1;
        auto run = [&](uint32_t spb, uint32_t refill, uint32_t grab) {
            r->tune_spb = spb;
            r->tune_refill = refill;
            r->tune_grab = grab;
            r->tune_tail = b_tail;
            r->tune_grab0 = b_grab0;
            r->tune_grabs = b_grabs;
            r->tune_pairs = b_pairs;
            float best = INFINITY;
            for (int k = 0; k < 3; ++k) {
                kf_render_stats st{};
                render_frame(r, s, num_envs, cams, &p, &o, &st);
                ++trials;
                best = std::min(best, st.trace_ms);
                if (st.trace_ms > 50.0f) break;
            }
            return best;
        };
        r->tune_tail = 0;
        r->tune_grab0 = 0;
        r->tune_grabs = 0;
        run(0, 0, 0);  // warm-up: buffers, code objects, clocks
        run(0, 0, 0);
        const float t_def = run(0, 0, 0);
        const uint32_t WH = r->W * r->H;
        uint32_t b_spb = 0, b_refill = 0, b_grab = 0;
        float t_best = t_def;
        // One parameter at a time.  Every candidate is timed in a forward
        // and a backward sweep and keeps its faster time, so a drift of the
        // clocks over the sweep does not favour the candidates timed last.
        auto sweep = [&](const std::vector<uint32_t>& cand, uint32_t& best, int which) {
            std::vector<float> t(cand.size(), INFINITY);
            auto one = [&](size_t k) {
                const uint32_t v = cand[k];
                if (which == 3) b_tail = v;
                if (which == 4) b_grab0 = v;
                if (which == 5) b_grabs = v;
                if (which == 6) b_pairs = v;
                const float m = which == 0 ? run(v, b_refill, b_grab)
                                           : (which == 1 ? run(b_spb, v, b_grab) : run(b_spb, b_refill,
                                                                                       which == 2 ? v : b_grab));
                t[k] = std::min(t[k], m);
            };
            for (size_t k = 0; k < cand.size(); ++k) one(k);
            for (size_t k = cand.size(); k-- > 0;) one(k);
            // the candidates include the current value, so the sweep's
            // fastest is chosen among times taken side by side
            size_t kb = 0;
            for (size_t k = 1; k < cand.size(); ++k)
                if (t[k] < t[kb]) kb = k;
            best = cand[kb];
            t_best = t[kb];
        };
        std::vector<uint32_t> spbs;
        for (uint32_t spb = 1; spb <= p.spp; spb *= 2) {
            if ((uint64_t)WH * spb > (64ull << 20)) break;  // path state stays within ~13 GB
            spbs.push_back(spb);
        }
        const uint32_t spb_def = default_batch(p.spp, WH);
        if (std::find(spbs.begin(), spbs.end(), spb_def) == spbs.end()) spbs.push_back(spb_def);
        b_refill = std::min<uint32_t>(64u, std::max<uint32_t>(1u, traversal_refill()));
        b_grab = kGrab;
        // the best refill threshold depends on the reservation size: the
        // reservation is swept before and after it
        const std::vector<uint32_t> grabs = {64u, 128u, 256u, 512u, 1024u};
        std::vector<uint32_t> refills = {8u, 16u, 32u, 48u};
        if (std::find(refills.begin(), refills.end(), b_refill) == refills.end()) refills.push_back(b_refill);
        // the bounce from which k_tail finishes the paths: never (0), or from
        // bounce 2 up to the first bounces after Russian roulette starts (few
        // paths per frame favour an early tail: at 1 spp of the 1080p scene
        // bounce 3 beats bounce 5 by 8 %)
        std::vector<uint32_t> tails = {0u};
        for (uint32_t tb = 2u; tb < std::min<uint32_t>(p.path_depth + 1u, p.rr_min_bounces + 4u); ++tb)
            tails.push_back(tb);
        sweep(spbs, b_spb, 0);
        sweep(grabs, b_grab, 2);
        sweep(refills, b_refill, 1);
        // the camera rays' reservation, then the other bounces' own: camera
        // rays favour larger reservations than the scattered later bounces
        // (at 16 spp of the 1080p scene 512 against 256 for the later
        // bounces; one reservation for all of them is 6 % slower)
        const std::vector<uint32_t> grab0s = {256u, 512u, 1024u};
        uint32_t best_grab0 = b_grab;
        sweep(grab0s, best_grab0, 4);
        b_grab0 = best_grab0;
        sweep(grabs, b_grab, 2);
        uint32_t best_tail = 0;
        sweep(tails, best_tail, 3);
        b_tail = best_tail;
        // the tail kernel: two lanes per path while few paths remain, or one
        // (the tail bounce is swept again with the kernel picked)
        uint32_t best_pairs = 1;
        sweep({1u, 0u}, best_pairs, 6);
        b_pairs = best_pairs;
        sweep(tails, best_tail, 3);
        b_tail = best_tail;
        // the shadow rays' reservation (0: the same; at 16 spp of the 1080p
        // scene 64 against 256 saves 2 % of the frame, at 1 spp 4 %)
        const std::vector<uint32_t> grabss = {0u, 64u, 128u};
        uint32_t best_grabs = 0;
        sweep(grabss, best_grabs, 5);
        b_grabs = best_grabs;
        r->tune_spb = b_spb;
        r->tune_refill = b_refill;
        r->tune_grab = b_grab;
        r->tune_tail = b_tail;
        r->tune_grab0 = b_grab0;
        r->tune_grabs = b_grabs;
        r->tune_pairs = b_pairs;
        out->tail_pairs = b_pairs;
        out->tail_bounce = b_tail;
        out->grab0 = b_grab0;
        out->grab_shadow = b_grabs;
        out->samples_per_batch = b_spb;
        out->refill = b_refill;
        out->grab = b_grab;
        out->frame_ms = t_best;
        out->default_ms = t_def;
        out->trials = trials;
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

extern "C" int kf_renderer_set_schedule(kf_renderer* r, uint32_t samples_per_batch, uint32_t refill, uint32_t grab,
                                        uint32_t tail_bounce, uint32_t grab0, uint32_t grab_shadow) {
    if (!r) {
        set_error("kf_renderer_set_schedule: null renderer");
        return -1;
    }
    r->tune_spb = samples_per_batch;
    r->tune_refill = refill;
    r->tune_grab = grab;
    r->tune_tail = tail_bounce;
    r->tune_grab0 = grab0;
    r->tune_grabs = grab_shadow;
    r->tune_pairs = 1;
    return 0;
}

extern "C" int kf_renderer_set_tail_pairs(kf_renderer* r, uint32_t tail_pairs) {
    if (!r) {
        set_error("kf_renderer_set_tail_pairs: null renderer");
        return -1;
    }
    if (tail_pairs > 1) {
        set_error("kf_renderer_set_tail_pairs: tail_pairs must be 0 or 1");
        return -1;
    }
    r->tune_pairs = tail_pairs;
    return 0;
}

// Single-channel FP32 image of the rgba target: (r + g + b) / 3 per pixel.
__global__ void k_to_gray32f(uint32_t n, const float4* __restrict__ rgba, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = rgba[i];
    out[i] = (c.x + c.y + c.z) / 3.0f;
}

extern "C" int kf_renderer_download_gray32f(kf_renderer* r, float* out) {
    try {
        if (!r || !out) throw Error("kf_renderer_download_gray32f: null argument");
        hipStream_t st = stream();
        uint32_t n = r->W * r->H;
        // staging shared with the BGRA8 download: 4 bytes per pixel either way
        if (!r->d_bgra) KF_CHECK(hipMalloc((void**)&r->d_bgra, 4 * (size_t)n));
        k_to_gray32f<<<(n + TB - 1) / TB, TB, 0, st>>>(n, r->d_rgba, (float*)r->d_bgra);
        KF_CHECK(hipGetLastError());
        KF_CHECK(hipMemcpyAsync(out, r->d_bgra, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
        KF_CHECK(hipStreamSynchronize(st));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

extern "C" int kf_renderer_download_bgra8(kf_renderer* r, uint8_t* out) {
    try {
        if (!r || !out) throw Error("kf_renderer_download_bgra8: null argument");
        hipStream_t st = stream();
        uint32_t n = r->W * r->H;
        static double* d_thr = nullptr;
        if (!d_thr) {
            double thr[255];
            for (int k = 0; k < 255; ++k) {
                double sv = (k + 0.5) / 255.0;
                thr[k] = sv <= 0.04045 ? sv / 12.92 : std::pow((sv + 0.055) / 1.055, 2.4);
            }
            KF_CHECK(hipMalloc((void**)&d_thr, sizeof(thr)));
            KF_CHECK(hipMemcpy(d_thr, thr, sizeof(thr), hipMemcpyHostToDevice));
        }
        if (!r->d_bgra) KF_CHECK(hipMalloc((void**)&r->d_bgra, 4 * (size_t)n));
        k_to_bgra8<<<(n + TB - 1) / TB, TB, 0, st>>>(n, r->d_rgba, d_thr, (uchar4*)r->d_bgra);
        KF_CHECK(hipGetLastError());
        KF_CHECK(hipMemcpyAsync(out, r->d_bgra, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
        KF_CHECK(hipStreamSynchronize(st));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}
