// bvh.hip — GPU LBVH builder for the BLASes and the instance TLAS (gfx950).
//
// Replaces the driver's acceleration-structure builds that the reference
// requests in RayTracer::createBottomLevelAS / buildBlas (rt.cpp:142-370) and
// RayTracer::buildTlas / updateTlas (rt.cpp:372-495).
//
// One batched build handles every geometry at once (segments): primitives of
// segment s get 64-bit keys (s << 30 | morton30(centroid)), one radix sort
// orders all of them, the Karras (2012) hierarchy is emitted per segment and
// a bottom-up refit fills the child boxes.
//   k_seg_bounds   centroid bounds per segment (ordered-int atomics)
//   k_morton       30-bit Morton codes, key = seg << 30 | code
//   radix sort     8-bit digits: per-tile histogram, one-block scan, stable
//                  scatter with wave64 ballot multi-split ranks in LDS
//   k_karras       internal nodes (ranges, splits, child refs, parents)
//   k_refit        leaves walk up; the second arrival at a node unions it
#include "kf_internal.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace kf {

#define KF_DEV_INLINE __device__ __forceinline__

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kDigitBits = 8;
constexpr int kDigits = 1 << kDigitBits;

__device__ __forceinline__ int f2ord(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float ord2f(int o) { return __int_as_float(o >= 0 ? o : o ^ 0x7fffffff); }

__global__ void k_fill_int(int* p, int v, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Centroid bounds start empty: lo = ord(+inf), hi = ord(-inf) per segment.
__global__ void k_seg_init(uint32_t n_seg, int* __restrict__ sb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 6 * n_seg) return;
    sb[i] = (i % 6) < 3 ? 0x7f800000 : (int)(0xff800000u ^ 0x7fffffffu);
}

__global__ void k_seg_bounds(uint32_t n, const float4* __restrict__ lo, const float4* __restrict__ hi,
                             const uint32_t* __restrict__ seg, int* __restrict__ sb /* 6 per seg */) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = lo[i], b = hi[i];
    if (!(a.x <= b.x)) return;  // empty box
    uint32_t s = seg[i];
    float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
    for (int k = 0; k < 3; ++k) {
        atomicMin(&sb[6 * s + k], f2ord(c[k]));
        atomicMax(&sb[6 * s + 3 + k], f2ord(c[k]));
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton(uint32_t n, const float4* __restrict__ lo, const float4* __restrict__ hi,
                         const uint32_t* __restrict__ seg, const int* __restrict__ sb,
                         unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s = seg[i];
    float4 a = lo[i], b = hi[i];
    uint32_t q[3] = {0, 0, 0};
    if (a.x <= b.x) {
        float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
        for (int k = 0; k < 3; ++k) {
            float l = ord2f(sb[6 * s + k]), h = ord2f(sb[6 * s + 3 + k]);
            float e = h - l;
            float u = e > 0.0f ? (c[k] - l) / e : 0.5f;
            u = fminf(fmaxf(u, 0.0f), 1.0f);
            q[k] = min((uint32_t)(u * 1024.0f), 1023u);
        }
    }
    uint32_t m = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    keys[i] = ((unsigned long long)s << 30) | m;
    vals[i] = i;
}

// ---------------------------------------------------------------- radix sort
__global__ __launch_bounds__(kSortThreads) void k_rs_hist(uint32_t n, const unsigned long long* __restrict__ keys,
                                                          int shift, uint32_t n_tiles, uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[kDigits];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kDigits - 1)], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * n_tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of m entries with one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t m, uint32_t* __restrict__ h) {
    __shared__ uint32_t part[1024];
    uint32_t per = (m + 1023) / 1024;
    uint32_t b = threadIdx.x * per, e = min(b + per, m);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += h[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (uint32_t)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (uint32_t i = b; i < e; ++i) {
        uint32_t v = h[i];
        h[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kSortThreads) void k_rs_scatter(uint32_t n, const unsigned long long* __restrict__ kin,
                                                             const uint32_t* __restrict__ vin, int shift,
                                                             uint32_t n_tiles, const uint32_t* __restrict__ hist,
                                                             unsigned long long* __restrict__ kout,
                                                             uint32_t* __restrict__ vout) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t gbase[kDigits];
    __shared__ uint32_t run[kDigits];
    __shared__ uint32_t wcnt[kWaves][kDigits];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    gbase[t] = hist[t * n_tiles + blockIdx.x];
    run[t] = 0;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        for (int k = 0; k < kWaves; ++k) wcnt[k][t] = 0;
        __syncthreads();
        uint32_t i = base + r * kSortThreads + t;
        bool valid = i < n;
        unsigned long long key = valid ? kin[i] : 0ull;
        uint32_t d = (uint32_t)(key >> shift) & (kDigits - 1);
        unsigned long long mask = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            bool bit = (d >> b) & 1u;
            unsigned long long m = __ballot(bit);
            mask &= bit ? m : ~m;
        }
        uint32_t rank = __popcll(mask & lt);
        if (valid && rank == 0) wcnt[w][d] = __popcll(mask);
        __syncthreads();
        {
            uint32_t s = run[t];
            for (int k = 0; k < kWaves; ++k) {
                uint32_t c = wcnt[k][t];
                wcnt[k][t] = s;
                s += c;
            }
            run[t] = s;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = gbase[d] + wcnt[w][d] + rank;
            if (pos < n) {
                kout[pos] = key;
                vout[pos] = vin[i];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- Karras
__device__ __forceinline__ int delta(const unsigned long long* k, int n, int a, int b) {
    if (b < 0 || b >= n) return -1;
    unsigned long long x = k[a] ^ k[b];
    if (x) return __clzll(x);
    return 64 + __clz((uint32_t)(a ^ b));
}

__global__ void k_karras(uint32_t n, const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ order,
                         const uint32_t* __restrict__ seg_off, const uint32_t* __restrict__ node_off,
                         Node* __restrict__ nodes, int* __restrict__ parent_node, int* __restrict__ parent_leaf,
                         int leaf_is_slot) {
    uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= n) return;
    uint32_t s = (uint32_t)(keys[gi] >> 30);
    uint32_t so = seg_off[s];
    int ns = (int)(seg_off[s + 1] - so);
    int i = (int)(gi - so);
    if (i >= ns - 1) return;
    const unsigned long long* k = keys + so;
    int d = (delta(k, ns, i, i + 1) - delta(k, ns, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(k, ns, i, i - d);
    int lmax = 2;
    while (delta(k, ns, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(k, ns, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(k, ns, i, j);
    int sp = 0;
    int div = 2;
    int t = (l + div - 1) / div;
    while (true) {
        if (delta(k, ns, i, i + (sp + t) * d) > dnode) sp += t;
        if (t == 1) break;
        div <<= 1;
        t = (l + div - 1) / div;
    }
    int gamma = i + sp * d + min(d, 0);
    int nb = (int)node_off[s];
    int me = nb + i;
    int lref, rref;
    if (min(i, j) == gamma) {
        int leaf = (int)so + gamma;
        lref = ~(leaf_is_slot ? leaf : (int)order[leaf]);
        parent_leaf[leaf] = me;
    } else {
        lref = nb + gamma;
        parent_node[nb + gamma] = me;
    }
    if (max(i, j) == gamma + 1) {
        int leaf = (int)so + gamma + 1;
        rref = ~(leaf_is_slot ? leaf : (int)order[leaf]);
        parent_leaf[leaf] = me;
    } else {
        rref = nb + gamma + 1;
        parent_node[nb + gamma + 1] = me;
    }
    nodes[me].d = make_int4(lref, rref, 0, 0);
    if (i == 0) parent_node[me] = -1;
}

__device__ __forceinline__ void write_child(Node* nd, bool left, const float lo[3], const float hi[3]) {
    float* a = left ? &nd->a.x : &nd->b.x;
    a[0] = lo[0];
    a[1] = hi[0];
    a[2] = lo[1];
    a[3] = hi[1];
    float* c = left ? &nd->c.x : &nd->c.z;
    c[0] = lo[2];
    c[1] = hi[2];
}

__global__ void k_refit(uint32_t n, uint32_t n_nodes, const uint32_t* __restrict__ order, const float4* __restrict__ plo,
                        const float4* __restrict__ phi, const int* __restrict__ parent_node,
                        const int* __restrict__ parent_leaf, const unsigned long long* __restrict__ keys,
                        const uint32_t* __restrict__ seg_off, Node* nodes, int* flags, int leaf_is_slot) {
    uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= n) return;
    uint32_t s = (uint32_t)(keys[gi] >> 30);
    if (seg_off[s + 1] - seg_off[s] < 2) return;  // single-leaf segment: no internal node
    uint32_t prim = order[gi];
    float4 a = plo[prim], b = phi[prim];
    float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
    int child = ~(leaf_is_slot ? (int)gi : (int)prim);
    int p = parent_leaf[gi];
    for (int step = 0; p >= 0 && (uint32_t)p < n_nodes && step < 4096; ++step) {
        Node* nd = &nodes[p];
        bool left = nd->d.x == child;
        write_child(nd, left, lo, hi);
        int old = __hip_atomic_fetch_add(&flags[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;
        float4 na = nd->a, nb = nd->b, nc = nd->c;
        lo[0] = fminf(na.x, nb.x);
        hi[0] = fmaxf(na.y, nb.y);
        lo[1] = fminf(na.z, nb.z);
        hi[1] = fmaxf(na.w, nb.w);
        lo[2] = fminf(nc.x, nc.z);
        hi[2] = fmaxf(nc.y, nc.w);
        child = p;
        p = parent_node[p];
    }
}

__global__ void k_seg_root(uint32_t n_seg, const uint32_t* __restrict__ seg_off, const uint32_t* __restrict__ node_off,
                           const uint32_t* __restrict__ order, int* __restrict__ root, int leaf_is_slot) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_seg) return;
    uint32_t ns = seg_off[s + 1] - seg_off[s];
    if (ns == 0) root[s] = kEmptyRef;
    else if (ns == 1) root[s] = ~(leaf_is_slot ? (int)seg_off[s] : (int)order[seg_off[s]]);
    else root[s] = (int)node_off[s];
}

template <class T> T* dalloc(size_t n) {
    T* p = nullptr;
    if (n) KF_CHECK(hipMallocAsync((void**)&p, n * sizeof(T), stream()));
    return p;
}
template <class T> void dfree(T* p) {
    if (p) KF_CHECK(hipFreeAsync(p, stream()));
}

inline uint32_t blocks(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

// ---------------------------------------------------------------- BLAS inputs
__global__ void k_tri_geom(uint32_t n_tri, uint32_t n_geom, const uint32_t* __restrict__ toff,
                           uint32_t* __restrict__ tri_geom) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tri) return;
    uint32_t lo = 0, hi = n_geom;  // largest g with toff[g] <= i
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (toff[mid] <= i) lo = mid; else hi = mid;
    }
    tri_geom[i] = lo;
}

__global__ void k_tri_bounds(uint32_t n_tri, const float4* __restrict__ verts, const uint32_t* __restrict__ idx,
                             const uint32_t* __restrict__ voff, const uint32_t* __restrict__ tri_geom,
                             float4* __restrict__ lo, float4* __restrict__ hi) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tri) return;
    uint32_t vb = voff[tri_geom[i]];
    float3 p[3];
    for (int k = 0; k < 3; ++k) {
        float4 v = verts[3 * (size_t)(vb + idx[3 * (size_t)i + k])];
        p[k] = make_float3(v.x, v.y, v.z);
    }
    lo[i] = make_float4(fminf(p[0].x, fminf(p[1].x, p[2].x)), fminf(p[0].y, fminf(p[1].y, p[2].y)),
                        fminf(p[0].z, fminf(p[1].z, p[2].z)), 0.f);
    hi[i] = make_float4(fmaxf(p[0].x, fmaxf(p[1].x, p[2].x)), fmaxf(p[0].y, fmaxf(p[1].y, p[2].y)),
                        fmaxf(p[0].z, fmaxf(p[1].z, p[2].z)), 0.f);
}

// Triangle slot s, by axis: tri[3s + a] = (v0[a], v1[a], v2[a], w), w the
// local prim id bits on line 0 and 0 on lines 1-2.  Line a minus the ray
// origin's axis a is what a node's low planes subtract too, so the traversal
// step shares those differences between its node and triangle tests.
KF_DEV_INLINE void put_tri_slot(float4* __restrict__ tri, size_t s, const float3 p[3], uint32_t prim) {
    tri[3 * s + 0] = make_float4(p[0].x, p[1].x, p[2].x, __uint_as_float(prim));
    tri[3 * s + 1] = make_float4(p[0].y, p[1].y, p[2].y, 0.f);
    tri[3 * s + 2] = make_float4(p[0].z, p[1].z, p[2].z, 0.f);
}
// Box of triangle slot s.
KF_DEV_INLINE void tri_slot_box(const float4* __restrict__ tri, size_t s, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        const float4 l = tri[3 * s + a];
        lo[a] = fminf(l.x, fminf(l.y, l.z));
        hi[a] = fmaxf(l.x, fmaxf(l.y, l.z));
    }
}

// Triangle slots in BVH leaf order.
__global__ void k_tri_reorder(uint32_t n_tri, const uint32_t* __restrict__ order, const float4* __restrict__ verts,
                              const uint32_t* __restrict__ idx, const uint32_t* __restrict__ voff,
                              const uint32_t* __restrict__ toff, const uint32_t* __restrict__ tri_geom,
                              float4* __restrict__ tri) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_tri) return;
    uint32_t i = order[s];
    uint32_t g = tri_geom[i];
    uint32_t vb = voff[g];
    float3 p[3];
    for (int k = 0; k < 3; ++k) {
        const float4 v = verts[3 * (size_t)(vb + idx[3 * (size_t)i + k])];
        p[k] = make_float3(v.x, v.y, v.z);
    }
    put_tri_slot(tri, s, p, i - toff[g]);
}

// Object-space bounds of each BLAS (union of the root's children, or the
// single triangle).
__global__ void k_blas_box(uint32_t n_geom, const int* __restrict__ root, const Node* __restrict__ nodes,
                           const float4* __restrict__ tri, float4* __restrict__ glo, float4* __restrict__ ghi) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_geom) return;
    int r = root[g];
    float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0);
    if (r == kEmptyRef) {
    } else if (r < 0) {
        float l[3], h[3];
        tri_slot_box(tri, (size_t)(~r), l, h);
        lo = make_float4(l[0], l[1], l[2], 0);
        hi = make_float4(h[0], h[1], h[2], 0);
    } else {
        Node n = nodes[r];
        lo = make_float4(fminf(n.a.x, n.b.x), fminf(n.a.z, n.b.z), fminf(n.c.x, n.c.z), 0);
        hi = make_float4(fmaxf(n.a.y, n.b.y), fmaxf(n.a.w, n.b.w), fmaxf(n.c.y, n.c.w), 0);
    }
    glo[g] = lo;
    ghi[g] = hi;
}

// Instance records + padded world AABBs (corners transformed in double).
// root: BVH4 root ref of each geometry (node index, ~triangle slot for a
// BLAS of one triangle, kEmptyRef); blas4: the BVH4 nodes it indexes; mesh:
// (voff, toff) of the arrays each geometry's shading reads; toff: each
// geometry's own triangle offset.
__global__ void k_instances(uint32_t n, const kf_instance* __restrict__ src, const float* __restrict__ w2o,
                            uint32_t n_geom, const uint32_t* __restrict__ gflags, const int* __restrict__ root,
                            const Node4* __restrict__ blas4, uint32_t n_blas4,
                            const float4* __restrict__ glo, const float4* __restrict__ ghi,
                            const uint2* __restrict__ mesh, const uint32_t* __restrict__ toff,
                            const uint32_t* __restrict__ inst_env, Inst* __restrict__ out,
                            float4* __restrict__ ilo, float4* __restrict__ ihi, uint32_t* __restrict__ iseg) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const kf_instance& in = src[i];
    uint32_t g = in.geometry_index;
    Inst rec;
    for (int r = 0; r < 3; ++r) rec.w2o[r] = make_float4(w2o[12 * i + 4 * r], w2o[12 * i + 4 * r + 1],
                                                         w2o[12 * i + 4 * r + 2], w2o[12 * i + 4 * r + 3]);
    const int r4 = g < n_geom ? root[g] : kEmptyRef;
    bool valid = g < n_geom && !(gflags[g] & KF_GEOM_HIDDEN) && r4 != kEmptyRef && (r4 < 0 || (uint32_t)r4 < n_blas4);
    int flags = (valid ? kInstValid : 0) | ((g < n_geom && (gflags[g] & KF_GEOM_OPAQUE)) ? kInstOpaque : 0);
    rec.info = make_int4(valid ? r4 : kEmptyRef, (int)g, flags, 0);
    rec.shade = g < n_geom ? make_int4((int)mesh[g].x, (int)mesh[g].y, (int)toff[g], 0) : make_int4(0, 0, 0, 0);
    // the BLAS root node, as the traversal's instance step tests it
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    rec.lox = rec.hix = rec.loy = rec.hiy = rec.loz = rec.hiz = z;
    rec.ref = make_int4((int)kEmptyItem, (int)kEmptyItem, (int)kEmptyItem, (int)kEmptyItem);
    if (valid && r4 >= 0) {
        const Node4 nd = blas4[r4];
        rec.lox = nd.lox; rec.hix = nd.hix; rec.loy = nd.loy; rec.hiy = nd.hiy; rec.loz = nd.loz; rec.hiz = nd.hiz;
        rec.ref = nd.ref;
    } else if (valid) {  // one triangle: its box (the geometry's) as the only child
        const float4 a = glo[g], b = ghi[g];
        rec.lox.x = a.x; rec.hix.x = b.x; rec.loy.x = a.y; rec.hiy.x = b.y; rec.loz.x = a.z; rec.hiz.x = b.z;
        rec.ref.x = (int)((3u << 30) | (uint32_t)~r4);  // triangle item
    }
    out[i] = rec;
    float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0);
    if (valid) {
        float4 a = glo[g], b = ghi[g];
        double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
        const float* t = in.transform;
        for (int c = 0; c < 8; ++c) {
            double p[3] = {(c & 1) ? b.x : a.x, (c & 2) ? b.y : a.y, (c & 4) ? b.z : a.z};
            for (int r = 0; r < 3; ++r) {
                double v = (double)t[r] * p[0] + (double)t[4 + r] * p[1] + (double)t[8 + r] * p[2] + (double)t[12 + r];
                l[r] = fmin(l[r], v);
                h[r] = fmax(h[r], v);
            }
        }
        double ext = fmax(h[0] - l[0], fmax(h[1] - l[1], h[2] - l[2]));
        float fl[3], fh[3];
        for (int r = 0; r < 3; ++r) {
            double pad = 1e-5 * (ext + fabs(l[r]) + fabs(h[r])) + 1e-6;
            fl[r] = nextafterf((float)(l[r] - pad), -INFINITY);
            fh[r] = nextafterf((float)(h[r] + pad), INFINITY);
        }
        lo = make_float4(fl[0], fl[1], fl[2], 0);
        hi = make_float4(fh[0], fh[1], fh[2], 0);
    }
    ilo[i] = lo;
    ihi[i] = hi;
    iseg[i] = inst_env ? inst_env[i] : 0u;  // the LBVH segment is the environment
}

// ---------------------------------------------------------------- BVH4
KF_DEV_INLINE float half_area(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// One item of the collapse: (BVH2 node b, BVH4 slot w) writes BVH4 node w
// and queues its inner children for the next level.  err is set on any index
// outside the arrays (malformed input), and nothing is written out of bounds.
KF_DEV_INLINE void collapse_item(const Node* __restrict__ b2, uint32_t n_b2, int2 it, int2* __restrict__ out,
                                 uint32_t* out_count, uint32_t out_cap, Node4* __restrict__ wide,
                                 uint32_t* wide_count, uint32_t wide_cap, uint32_t inner_kind, int* err) {
    if ((uint32_t)it.x >= n_b2 || (uint32_t)it.y >= wide_cap) {
        *err = 1;
        return;
    }
    float lo[4][3], hi[4][3];
    int ref[4];
    auto take = [&](int node, int s0, int s1) {
        const Node n = b2[node];
        lo[s0][0] = n.a.x; hi[s0][0] = n.a.y; lo[s0][1] = n.a.z; hi[s0][1] = n.a.w; lo[s0][2] = n.c.x; hi[s0][2] = n.c.y;
        lo[s1][0] = n.b.x; hi[s1][0] = n.b.y; lo[s1][1] = n.b.z; hi[s1][1] = n.b.w; lo[s1][2] = n.c.z; hi[s1][2] = n.c.w;
        ref[s0] = n.d.x;
        ref[s1] = n.d.y;
    };
    take(it.x, 0, 1);
    int cnt = 2;
    while (cnt < 4) {
        int best = -1;
        float ba = -1.0f;
        for (int j = 0; j < cnt; ++j)
            if (ref[j] >= 0) {
                const float a = half_area(lo[j], hi[j]);
                if (a > ba) {
                    ba = a;
                    best = j;
                }
            }
        if (best < 0) break;
        if ((uint32_t)ref[best] >= n_b2) {
            *err = 1;
            return;
        }
        take(ref[best], best, cnt);
        ++cnt;
    }
    int k = 0;
    for (int j = 0; j < cnt; ++j) k += ref[j] >= 0 ? 1 : 0;
    const uint32_t base = k ? atomicAdd(wide_count, (uint32_t)k) : 0u;
    if (base + (uint32_t)k > wide_cap) {
        *err = 1;
        return;
    }
    float blo[3][4], bhi[3][4];
    int wref[4];
    int m = 0;
    for (int j = 0; j < 4; ++j) {
        const bool used = j < cnt;
        for (int a = 0; a < 3; ++a) {
            blo[a][j] = used ? lo[j][a] : 0.0f;
            bhi[a][j] = used ? hi[j][a] : 0.0f;
        }
        // child refs are stored as traversal items: kind (inner_kind for an
        // inner node, inner_kind + 2 for a leaf) in bits 31:30, index below
        if (!used) {
            wref[j] = (int)kEmptyItem;
        } else if (ref[j] >= 0) {
            wref[j] = (int)((inner_kind << 30) | (base + m));
            const uint32_t q = atomicAdd(out_count, 1u);
            if (q < out_cap) out[q] = make_int2(ref[j], (int)(base + m));
            else *err = 1;
            ++m;
        } else {
            wref[j] = (int)(((inner_kind + 2u) << 30) | (uint32_t)~ref[j]);
        }
    }
    Node4 W;
    W.lox = make_float4(blo[0][0], blo[0][1], blo[0][2], blo[0][3]);
    W.hix = make_float4(bhi[0][0], bhi[0][1], bhi[0][2], bhi[0][3]);
    W.loy = make_float4(blo[1][0], blo[1][1], blo[1][2], blo[1][3]);
    W.hiy = make_float4(bhi[1][0], bhi[1][1], bhi[1][2], bhi[1][3]);
    W.loz = make_float4(blo[2][0], blo[2][1], blo[2][2], blo[2][3]);
    W.hiz = make_float4(bhi[2][0], bhi[2][1], bhi[2][2], bhi[2][3]);
    W.ref = make_int4(wref[0], wref[1], wref[2], wref[3]);
    // the BVH2 ref each child was taken from (a refit recomputes the child
    // boxes from these); the traversal never reads this line
    W.pad = make_int4(0 < cnt ? ref[0] : kEmptyRef, 1 < cnt ? ref[1] : kEmptyRef, 2 < cnt ? ref[2] : kEmptyRef,
                      3 < cnt ? ref[3] : kEmptyRef);
    wide[it.y] = W;
}

// One level of the collapse over many items (BLAS builds).
__global__ void k_collapse(const Node* __restrict__ b2, uint32_t n_b2, const int2* __restrict__ in, uint32_t n_in,
                           int2* __restrict__ out, uint32_t* __restrict__ out_count, uint32_t out_cap,
                           Node4* __restrict__ wide, uint32_t* __restrict__ wide_count, uint32_t wide_cap,
                           uint32_t inner_kind, int* __restrict__ err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_in) return;
    collapse_item(b2, n_b2, in[i], out, out_count, out_cap, wide, wide_count, wide_cap, inner_kind, err);
}

// The whole collapse of a forest with no host round trip (TLAS updates,
// every frame): block e collapses tree e, level after level.  Tree e has the
// n_e = seg_off[e+1] - seg_off[e] leaves of its LBVH segment; for n_e >= 2
// its n_e - 1 BVH2 nodes start at node_off[e], with the root first, and its
// BVH4 takes slots [node_off[e], node_off[e] + n_e - 1) of `wide` (a BVH4
// never has more nodes than its BVH2), root first; qa / qb hold its level
// queues in the same index range.  cnt + 2e: [0] next free BVH4 slot,
// [1] next-level queue length (scratch).
constexpr int kTreeThreads = 1024;
__global__ __launch_bounds__(kTreeThreads) void k_collapse_forest(const Node* __restrict__ b2, uint32_t n_b2,
                                                                  const uint32_t* __restrict__ seg_off,
                                                                  const uint32_t* __restrict__ node_off, int2* qa_all,
                                                                  int2* qb_all, Node4* __restrict__ wide,
                                                                  uint32_t inner_kind, uint32_t* cnt_all, int* err) {
    const uint32_t e = blockIdx.x;
    const uint32_t n_e = seg_off[e + 1] - seg_off[e];
    if (n_e < 2u) return;  // a single instance or none: the root is a leaf or empty
    const uint32_t base = node_off[e], cap_e = n_e - 1u;
    if (base + cap_e > n_b2) {
        if (threadIdx.x == 0) *err = 1;
        return;
    }
    int2* qa = qa_all + base;
    int2* qb = qb_all + base;
    uint32_t* cnt = cnt_all + 2 * e;
    __shared__ uint32_t n_in;
    if (threadIdx.x == 0) {
        n_in = 1u;
        cnt[0] = base + 1u;
        cnt[1] = 0u;
        qa[0] = make_int2((int)base, (int)base);
    }
    __syncthreads();
    for (uint32_t level = 0; n_in > 0u; ++level) {
        if (level > cap_e + 1u) {  // deeper than a well-formed tree can be
            if (threadIdx.x == 0) *err = 1;
            break;
        }
        for (uint32_t i = threadIdx.x; i < n_in; i += blockDim.x)
            collapse_item(b2, n_b2, qa[i], qb, &cnt[1], cap_e, wide, &cnt[0], base + cap_e, inner_kind, err);
        __syncthreads();
        if (threadIdx.x == 0) {
            n_in = min(cnt[1], cap_e);
            cnt[1] = 0u;
        }
        __syncthreads();
        int2* t = qa;
        qa = qb;
        qb = t;
    }
}

}  // namespace

void lbvh_build(uint32_t n, const float4* d_lo, const float4* d_hi, const uint32_t* d_seg, uint32_t n_seg,
                const uint32_t* d_seg_off, const uint32_t* d_node_off, Node* d_nodes, uint32_t* d_order, int* d_root,
                bool leaf_is_slot, hipStream_t st) {
    const uint32_t TB = 256;
    if (n == 0) {
        k_seg_root<<<blocks(n_seg, TB), TB, 0, st>>>(n_seg, d_seg_off, d_node_off, d_order, d_root, leaf_is_slot);
        return;
    }
    int* sb = dalloc<int>(6 * (size_t)n_seg);
    unsigned long long* k0 = dalloc<unsigned long long>(n);
    unsigned long long* k1 = dalloc<unsigned long long>(n);
    uint32_t* v0 = dalloc<uint32_t>(n);
    uint32_t* v1 = dalloc<uint32_t>(n);
    int* parent_node = dalloc<int>(std::max<uint32_t>(n, 1));
    int* parent_leaf = dalloc<int>(n);
    int* flags = dalloc<int>(std::max<uint32_t>(n, 1));
    uint32_t n_tiles = blocks(n, kSortTile);
    uint32_t* hist = dalloc<uint32_t>((size_t)n_tiles * kDigits);

    k_seg_init<<<blocks(6 * n_seg, TB), TB, 0, st>>>(n_seg, sb);
    k_seg_bounds<<<blocks(n, TB), TB, 0, st>>>(n, d_lo, d_hi, d_seg, sb);
    k_morton<<<blocks(n, TB), TB, 0, st>>>(n, d_lo, d_hi, d_seg, sb, k0, v0);

    int seg_bits = 0;
    while ((1u << seg_bits) < n_seg) ++seg_bits;
    int total_bits = 30 + seg_bits;
    for (int shift = 0; shift < total_bits; shift += kDigitBits) {
        k_rs_hist<<<n_tiles, kSortThreads, 0, st>>>(n, k0, shift, n_tiles, hist);
        k_rs_scan<<<1, 1024, 0, st>>>(n_tiles * kDigits, hist);
        k_rs_scatter<<<n_tiles, kSortThreads, 0, st>>>(n, k0, v0, shift, n_tiles, hist, k1, v1);
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    KF_CHECK(hipMemcpyAsync(d_order, v0, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    k_fill_int<<<blocks(n, TB), TB, 0, st>>>(flags, 0, n);
    k_fill_int<<<blocks(n, TB), TB, 0, st>>>(parent_leaf, -1, n);
    k_fill_int<<<blocks(n, TB), TB, 0, st>>>(parent_node, -1, n);
    k_karras<<<blocks(n, TB), TB, 0, st>>>(n, k0, v0, d_seg_off, d_node_off, d_nodes, parent_node, parent_leaf,
                                           leaf_is_slot);
    k_refit<<<blocks(n, TB), TB, 0, st>>>(n, n, v0, d_lo, d_hi, parent_node, parent_leaf, k0, d_seg_off, d_nodes, flags,
                                          leaf_is_slot);
    k_seg_root<<<blocks(n_seg, TB), TB, 0, st>>>(n_seg, d_seg_off, d_node_off, v0, d_root, leaf_is_slot);
    KF_CHECK(hipGetLastError());
    dfree(sb); dfree(k0); dfree(k1); dfree(v0); dfree(v1);
    dfree(parent_node); dfree(parent_leaf); dfree(flags); dfree(hist);
}

uint32_t collapse_bvh4(const Node* d_b2, uint32_t n_b2, const std::vector<int>& roots2, std::vector<int>& roots4,
                       Node4* d_out, uint32_t cap, uint32_t inner_kind, hipStream_t st) {
    const uint32_t TB = 256;
    roots4.assign(roots2.size(), kEmptyRef);
    std::vector<int2> items;
    for (size_t g = 0; g < roots2.size(); ++g) {
        const int r = roots2[g];
        if (r >= 0 && r != kEmptyRef) {
            roots4[g] = (int)items.size();
            items.push_back(make_int2(r, (int)items.size()));
        } else {
            roots4[g] = r;  // single leaf or empty
        }
    }
    uint32_t n_in = (uint32_t)items.size();
    if (n_in == 0) return 0;
    if (n_in > cap || n_b2 == 0) throw Error("collapse_bvh4: inconsistent BVH2");
    const uint32_t qcap = std::max<uint32_t>(n_b2, 1);
    int2* qa = dalloc<int2>(qcap);
    int2* qb = dalloc<int2>(qcap);
    uint32_t* cnt = dalloc<uint32_t>(2);  // [0] next queue, [1] BVH4 nodes
    int* err = dalloc<int>(1);
    uint32_t h2[2] = {0, n_in};
    int zero = 0;
    KF_CHECK(hipMemcpyAsync(qa, items.data(), n_in * sizeof(int2), hipMemcpyHostToDevice, st));
    KF_CHECK(hipMemcpyAsync(cnt, h2, sizeof(h2), hipMemcpyHostToDevice, st));
    KF_CHECK(hipMemcpyAsync(err, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    // Every level consumes at least one BVH2 level, so n_b2 + 1 levels bound
    // a well-formed tree.
    for (uint32_t level = 0; n_in > 0; ++level) {
        if (level > n_b2 + 1) throw Error("collapse_bvh4: BVH2 deeper than its node count");
        k_collapse<<<blocks(n_in, TB), TB, 0, st>>>(d_b2, n_b2, qa, n_in, qb, cnt, qcap, d_out, cnt + 1, cap,
                                                    inner_kind, err);
        KF_CHECK(hipGetLastError());
        uint32_t next = 0;
        KF_CHECK(hipMemcpyAsync(&next, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        KF_CHECK(hipMemsetAsync(cnt, 0, sizeof(uint32_t), st));
        KF_CHECK(hipStreamSynchronize(st));
        n_in = std::min(next, qcap);
        std::swap(qa, qb);
    }
    int h_err = 0;
    KF_CHECK(hipMemcpyAsync(h2, cnt, sizeof(h2), hipMemcpyDeviceToHost, st));
    KF_CHECK(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, st));
    KF_CHECK(hipStreamSynchronize(st));
    dfree(qa); dfree(qb); dfree(cnt); dfree(err);
    if (h_err) throw Error("collapse_bvh4: malformed BVH2");
    return h2[1];
}

void build_blas(Scene& s) {
    hipStream_t st = stream();
    const uint32_t TB = 256;
    uint32_t n = s.n_tri;
    s.h_node_off.assign(s.n_geom, 0);
    uint32_t acc = 0;
    for (uint32_t g = 0; g < s.n_geom; ++g) {
        s.h_node_off[g] = acc;
        uint32_t ng = s.h_toff[g + 1] - s.h_toff[g];
        acc += ng > 1 ? ng - 1 : 0;
    }
    s.n_blas_nodes = acc;
    dfree(s.d_blas);
    dfree(s.d_tri);
    dfree(s.d_blas_root);
    dfree(s.d_tri_geom);
    s.d_blas = dalloc<Node>(std::max<uint32_t>(acc, 1));
    s.d_tri = dalloc<float4>(3 * (size_t)std::max<uint32_t>(n, 1));
    s.d_blas_root = dalloc<int>(std::max<uint32_t>(s.n_geom, 1));
    s.d_tri_geom = dalloc<uint32_t>(std::max<uint32_t>(n, 1));
    uint32_t* d_node_off = dalloc<uint32_t>(std::max<uint32_t>(s.n_geom, 1));
    if (s.n_geom)
        KF_CHECK(hipMemcpyAsync(d_node_off, s.h_node_off.data(), s.n_geom * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    float4* lo = dalloc<float4>(std::max<uint32_t>(n, 1));
    float4* hi = dalloc<float4>(std::max<uint32_t>(n, 1));
    uint32_t* order = dalloc<uint32_t>(std::max<uint32_t>(n, 1));
    if (n) {
        k_tri_geom<<<blocks(n, TB), TB, 0, st>>>(n, s.n_geom, s.d_toff, s.d_tri_geom);
        k_tri_bounds<<<blocks(n, TB), TB, 0, st>>>(n, s.d_verts, s.d_idx, s.d_voff, s.d_tri_geom, lo, hi);
    }
    lbvh_build(n, lo, hi, s.d_tri_geom, s.n_geom, s.d_toff, d_node_off, s.d_blas, order, s.d_blas_root, true, st);
    if (n)
        k_tri_reorder<<<blocks(n, TB), TB, 0, st>>>(n, order, s.d_verts, s.d_idx, s.d_voff, s.d_toff, s.d_tri_geom,
                                                    s.d_tri);
    KF_CHECK(hipGetLastError());
    s.h_blas_root.resize(s.n_geom);
    if (s.n_geom)
        KF_CHECK(hipMemcpyAsync(s.h_blas_root.data(), s.d_blas_root, s.n_geom * sizeof(int), hipMemcpyDeviceToHost, st));
    dfree(d_node_off); dfree(lo); dfree(hi); dfree(order);
    KF_CHECK(hipStreamSynchronize(st));
    // BVH4 for the traversal: geometry g's nodes in [node_off[g], node_off[g]
    // + n_g - 1), root first, so roots are known by construction and a refit
    // of one geometry touches only its own range
    dfree(s.d_blas4);
    dfree(s.d_blas_root4);
    dfree(s.d_node_off);
    dfree(s.d_blas_cnt);
    const uint32_t cap4 = std::max<uint32_t>(acc, 1);
    s.d_blas4 = dalloc<Node4>(cap4);
    s.d_node_off = dalloc<uint32_t>(std::max<uint32_t>(s.n_geom, 1));
    s.d_blas_cnt = dalloc<uint32_t>(2 * (size_t)std::max<uint32_t>(s.n_geom, 1));
    if (s.n_geom)
        KF_CHECK(hipMemcpyAsync(s.d_node_off, s.h_node_off.data(), s.n_geom * sizeof(uint32_t), hipMemcpyHostToDevice,
                                st));
    int* err = dalloc<int>(1);
    KF_CHECK(hipMemsetAsync(err, 0, sizeof(int), st));
    if (acc) {
        int2* qa = dalloc<int2>(acc);
        int2* qb = dalloc<int2>(acc);
        k_collapse_forest<<<s.n_geom, kTreeThreads, 0, st>>>(s.d_blas, acc, s.d_toff, s.d_node_off, qa, qb, s.d_blas4,
                                                               1u, s.d_blas_cnt, err);
        KF_CHECK(hipGetLastError());
        dfree(qa);
        dfree(qb);
    }
    s.h_blas_root4.assign(s.n_geom, kEmptyRef);
    for (uint32_t g = 0; g < s.n_geom; ++g) {
        const uint32_t ng = s.h_toff[g + 1] - s.h_toff[g];
        s.h_blas_root4[g] = ng >= 2 ? (int)s.h_node_off[g] : (ng == 1 ? ~(int)s.h_toff[g] : kEmptyRef);
    }
    s.n_blas4 = cap4;  // index bound; used nodes per geometry are counted in d_blas_cnt
    s.d_blas_root4 = dalloc<int>(std::max<uint32_t>(s.n_geom, 1));
    if (s.n_geom)
        KF_CHECK(hipMemcpyAsync(s.d_blas_root4, s.h_blas_root4.data(), s.n_geom * sizeof(int), hipMemcpyHostToDevice, st));
    int h_err = 0;
    KF_CHECK(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, st));
    KF_CHECK(hipStreamSynchronize(st));
    dfree(err);
    if (h_err) throw Error("BLAS build: malformed BVH2");
    upload_trav_roots(s);
    upload_mesh_sources(s);
}

// d_trav_root4[g] = the BVH4 root of the geometry whose BVH g traverses.
void upload_trav_roots(Scene& s) {
    if (s.h_canon.size() != s.n_geom) {
        s.h_canon.resize(s.n_geom);
        for (uint32_t g = 0; g < s.n_geom; ++g) s.h_canon[g] = g;
    }
    std::vector<int> roots(std::max<uint32_t>(s.n_geom, 1), kEmptyRef);
    for (uint32_t g = 0; g < s.n_geom; ++g) roots[g] = s.h_blas_root4[s.h_canon[g]];
    if (!s.d_trav_root4) s.d_trav_root4 = dalloc<int>(std::max<uint32_t>(s.n_geom, 1));
    KF_CHECK(hipMemcpy(s.d_trav_root4, roots.data(), roots.size() * sizeof(int), hipMemcpyHostToDevice));
}

// d_mesh[g] = (voff, toff) of the geometry whose arrays g's shading reads.
void upload_mesh_sources(Scene& s) {
    if (s.h_mesh.size() != s.n_geom) {
        s.h_mesh.resize(s.n_geom);
        for (uint32_t g = 0; g < s.n_geom; ++g) s.h_mesh[g] = g;
    }
    std::vector<uint2> src(std::max<uint32_t>(s.n_geom, 1), make_uint2(0u, 0u));
    for (uint32_t g = 0; g < s.n_geom; ++g) src[g] = make_uint2(s.h_voff[s.h_mesh[g]], s.h_toff[s.h_mesh[g]]);
    if (!s.d_mesh) s.d_mesh = dalloc<uint2>(std::max<uint32_t>(s.n_geom, 1));
    KF_CHECK(hipMemcpy(s.d_mesh, src.data(), src.size() * sizeof(uint2), hipMemcpyHostToDevice));
}

namespace {
// Groups of geometries equal under eq, each led by its first member, found
// through a hash of the bytes key(g) feeds: lead[g] = the leader of g's group.
template <class Key, class Eq>
void group_equal(uint32_t n, std::vector<uint32_t>& lead, Key key, Eq eq) {
    lead.resize(n);
    std::unordered_map<uint64_t, std::vector<uint32_t>> seen;
    for (uint32_t g = 0; g < n; ++g) {
        lead[g] = g;
        uint64_t h = 1469598103934665603ull;  // FNV-1a
        key(g, [&](const void* p, size_t bytes) {
            const uint8_t* b = (const uint8_t*)p;
            for (size_t i = 0; i < bytes; ++i) h = (h ^ b[i]) * 1099511628211ull;
        });
        auto& list = seen[h];
        for (uint32_t c : list)
            if (eq(c, g)) {
                lead[g] = c;
                break;
            }
        if (lead[g] == g) list.push_back(g);
    }
}
// g leaves its group: the next member leads the rest.
void leave_group(std::vector<uint32_t>& lead, uint32_t g) {
    if (lead[g] == g) {
        uint32_t m = UINT32_MAX;
        for (uint32_t d = 0; d < lead.size(); ++d)
            if (d != g && lead[d] == g) {
                if (m == UINT32_MAX) m = d;
                lead[d] = m;
            }
    }
    lead[g] = g;
}
}  // namespace

void share_geometry(Scene& s, const kf_vertex* verts, const uint32_t* idx, bool separate) {
    const uint32_t n = s.n_geom;
    s.h_canon.resize(n);
    s.h_mesh.resize(n);
    for (uint32_t g = 0; g < n; ++g) s.h_canon[g] = s.h_mesh[g] = g;
    if (separate) return;
    auto ntri = [&](uint32_t g) { return (size_t)(s.h_toff[g + 1] - s.h_toff[g]); };
    auto nvert = [&](uint32_t g) { return (size_t)(s.h_voff[g + 1] - s.h_voff[g]); };
    auto corner = [&](uint32_t g, size_t t, int k) -> const float* {
        return verts[s.h_voff[g] + idx[3 * (s.h_toff[g] + t) + k]].pos;
    };
    // the BVH: the positions of every triangle corner, in order
    group_equal(
        n, s.h_canon,
        [&](uint32_t g, auto feed) {
            const size_t nt = ntri(g);
            feed(&nt, sizeof(nt));
            for (size_t t = 0; t < nt; ++t)
                for (int k = 0; k < 3; ++k) feed(corner(g, t, k), 12);
        },
        [&](uint32_t c, uint32_t g) {
            if (ntri(c) != ntri(g)) return false;
            for (size_t t = 0; t < ntri(g); ++t)
                for (int k = 0; k < 3; ++k)
                    if (std::memcmp(corner(c, t, k), corner(g, t, k), 12)) return false;
            return true;
        });
    // shading: the vertex records and the indices
    group_equal(
        n, s.h_mesh,
        [&](uint32_t g, auto feed) {
            const size_t nv = nvert(g), nt = ntri(g);
            feed(&nv, sizeof(nv));
            feed(&nt, sizeof(nt));
            if (nv) feed(verts + s.h_voff[g], nv * sizeof(kf_vertex));
            if (nt) feed(idx + 3 * (size_t)s.h_toff[g], 3 * nt * sizeof(uint32_t));
        },
        [&](uint32_t c, uint32_t g) {
            return nvert(c) == nvert(g) && ntri(c) == ntri(g) &&
                   !std::memcmp(verts + s.h_voff[c], verts + s.h_voff[g], nvert(g) * sizeof(kf_vertex)) &&
                   !std::memcmp(idx + 3 * (size_t)s.h_toff[c], idx + 3 * (size_t)s.h_toff[g],
                                3 * ntri(g) * sizeof(uint32_t));
        });
}

void unshare_geometry(Scene& s, uint32_t g) {
    if (g >= s.n_geom) return;
    if (s.h_canon.size() == s.n_geom) leave_group(s.h_canon, g);
    if (s.h_mesh.size() == s.n_geom) leave_group(s.h_mesh, g);
    upload_trav_roots(s);
    upload_mesh_sources(s);
}

// ---------------------------------------------------------------- refit
// Refit of one geometry after its vertices moved (same triangles): the
// triangle slots take the new positions, the BVH2 boxes are recomputed
// bottom-up (the second child to arrive at a node carries on, as in the
// build), then every BVH4 child box is the box of the BVH2 ref it was
// collapsed from.  The trees keep their shape.
__global__ void k_geom_slots(uint32_t g, const float4* __restrict__ verts, const uint32_t* __restrict__ idx,
                             const uint32_t* __restrict__ voff, const uint32_t* __restrict__ toff, float4* tri,
                             float4* __restrict__ lo, float4* __restrict__ hi) {
    const uint32_t t0 = toff[g], n = toff[g + 1] - t0;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const size_t slot = t0 + j;
    const uint32_t prim = __float_as_uint(tri[3 * slot].w);
    if (prim >= n) return;  // not a slot of this geometry (malformed)
    const size_t i = t0 + prim;
    float3 p[3];
    for (int k = 0; k < 3; ++k) {
        const float4 v = verts[3 * (size_t)(voff[g] + idx[3 * i + k])];
        p[k] = make_float3(v.x, v.y, v.z);
    }
    put_tri_slot(tri, slot, p, prim);
    lo[j] = make_float4(fminf(p[0].x, fminf(p[1].x, p[2].x)), fminf(p[0].y, fminf(p[1].y, p[2].y)),
                        fminf(p[0].z, fminf(p[1].z, p[2].z)), 0.f);
    hi[j] = make_float4(fmaxf(p[0].x, fmaxf(p[1].x, p[2].x)), fmaxf(p[0].y, fmaxf(p[1].y, p[2].y)),
                        fmaxf(p[0].z, fmaxf(p[1].z, p[2].z)), 0.f);
}

__global__ void k_geom_parents(uint32_t g, const uint32_t* __restrict__ toff, const uint32_t* __restrict__ node_off,
                               const Node* __restrict__ nodes, uint32_t n_nodes, int* __restrict__ parent_node,
                               int* __restrict__ parent_leaf, int* __restrict__ flags) {
    const uint32_t t0 = toff[g], n = toff[g + 1] - t0, nb = node_off[g];
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < 2 || j >= n - 1 || nb + j >= n_nodes) return;
    const int me = (int)(nb + j);
    const int4 d = nodes[me].d;
    for (int c = 0; c < 2; ++c) {
        const int r = c ? d.y : d.x;
        if (r >= 0) {
            if ((uint32_t)r >= nb && (uint32_t)r < nb + n - 1) parent_node[r - nb] = me;
        } else if ((uint32_t)~r >= t0 && (uint32_t)~r < t0 + n) {
            parent_leaf[~r - t0] = me;
        }
    }
    flags[j] = 0;
}

__global__ void k_geom_refit2(uint32_t g, const uint32_t* __restrict__ toff, const uint32_t* __restrict__ node_off,
                              uint32_t n_nodes, const float4* __restrict__ plo, const float4* __restrict__ phi,
                              const int* __restrict__ parent_node, const int* __restrict__ parent_leaf, Node* nodes,
                              int* flags) {
    const uint32_t t0 = toff[g], n = toff[g + 1] - t0, nb = node_off[g];
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < 2 || j >= n) return;
    float4 a = plo[j], b = phi[j];
    float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
    int child = ~(int)(t0 + j);
    int p = parent_leaf[j];
    for (int step = 0; p >= (int)nb && (uint32_t)p < nb + n - 1 && (uint32_t)p < n_nodes && step < 4096; ++step) {
        Node* nd = &nodes[p];
        const bool left = nd->d.x == child;
        write_child(nd, left, lo, hi);
        const int old = __hip_atomic_fetch_add(&flags[p - nb], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;
        const float4 na = nd->a, nbx = nd->b, nc = nd->c;
        lo[0] = fminf(na.x, nbx.x);
        hi[0] = fmaxf(na.y, nbx.y);
        lo[1] = fminf(na.z, nbx.z);
        hi[1] = fmaxf(na.w, nbx.w);
        lo[2] = fminf(nc.x, nc.z);
        hi[2] = fmaxf(nc.y, nc.w);
        child = p;
        p = parent_node[p - nb];
    }
}

__global__ void k_geom_refit4(uint32_t g, const uint32_t* __restrict__ toff, const uint32_t* __restrict__ node_off,
                              const uint32_t* __restrict__ cnt4, const Node* __restrict__ nodes, uint32_t n_nodes,
                              const float4* __restrict__ tri, Node4* wide) {
    const uint32_t t0 = toff[g], n = toff[g + 1] - t0, nb = node_off[g];
    if (n < 2) return;
    const uint32_t used = min(cnt4[2 * g] - nb, n - 1);
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= used) return;
    Node4& W = wide[nb + j];
    const int4 pr = W.pad;
    float blo[3][4], bhi[3][4];
    for (int c = 0; c < 4; ++c) {
        const int r = c == 0 ? pr.x : (c == 1 ? pr.y : (c == 2 ? pr.z : pr.w));
        float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};
        if (r == kEmptyRef) {
        } else if (r >= 0) {
            if ((uint32_t)r < n_nodes) {
                const Node m = nodes[r];
                lo[0] = fminf(m.a.x, m.b.x); hi[0] = fmaxf(m.a.y, m.b.y);
                lo[1] = fminf(m.a.z, m.b.z); hi[1] = fmaxf(m.a.w, m.b.w);
                lo[2] = fminf(m.c.x, m.c.z); hi[2] = fmaxf(m.c.y, m.c.w);
            }
        } else if ((uint32_t)~r >= t0 && (uint32_t)~r < t0 + n) {
            tri_slot_box(tri, (size_t)~r, lo, hi);
        }
        for (int a = 0; a < 3; ++a) {
            blo[a][c] = lo[a];
            bhi[a][c] = hi[a];
        }
    }
    W.lox = make_float4(blo[0][0], blo[0][1], blo[0][2], blo[0][3]);
    W.hix = make_float4(bhi[0][0], bhi[0][1], bhi[0][2], bhi[0][3]);
    W.loy = make_float4(blo[1][0], blo[1][1], blo[1][2], blo[1][3]);
    W.hiy = make_float4(bhi[1][0], bhi[1][1], bhi[1][2], bhi[1][3]);
    W.loz = make_float4(blo[2][0], blo[2][1], blo[2][2], blo[2][3]);
    W.hiz = make_float4(bhi[2][0], bhi[2][1], bhi[2][2], bhi[2][3]);
}

__global__ void k_tlas_aux(uint32_t n, uint32_t* aux) {
    aux[0] = 0u;  // segment offsets {0, n}
    aux[1] = n;
    aux[2] = 0u;  // node offset
}

// Root refs of each environment's BVH4, known by construction: an LBVH of
// n_e >= 2 leaves has its root at its first node and the collapse puts the
// BVH4 root at the first slot of the environment's range; one instance is a
// leaf root, none an empty one.
void env_roots(const std::vector<uint32_t>& env_off, const std::vector<uint32_t>& node_off, std::vector<int>& roots) {
    const size_t ne = env_off.size() - 1;
    roots.assign(ne, kEmptyRef);
    for (size_t e = 0; e < ne; ++e) {
        const uint32_t n_e = env_off[e + 1] - env_off[e];
        roots[e] = n_e >= 2 ? (int)node_off[e] : (n_e == 1 ? ~(int)env_off[e] : kEmptyRef);
    }
}

// TLAS (re)build, once per instance update: instance records + world boxes,
// an LBVH over the instances of each environment (one segment each), the
// BVH4 collapse of every environment's tree in one launch.  Nothing here
// waits for the GPU: roots are known by construction (env_roots) and the
// host inputs go through pinned staging.  One environment: its segment and
// node offsets are written on the device (the instance count may change);
// several: they are fixed at scene creation (the count may not).
void build_tlas(Scene& s, const kf_instance* h_inst, uint32_t n) {
    hipStream_t st = stream();
    const uint32_t TB = 256;
    const bool multi = s.n_env > 1;
    if (multi && n != s.h_env_off.back())
        throw Error("kf_scene_update_instances: a scene with environments keeps its instance count");
    if (h_inst != s.h_inst.data()) s.h_inst.assign(h_inst, h_inst + n);
    if (n != s.n_inst || !s.d_inst) {
        dfree(s.d_inst); dfree(s.d_tlas); dfree(s.d_inst_src); dfree(s.d_ilo); dfree(s.d_ihi);
        s.d_inst = dalloc<Inst>(std::max<uint32_t>(n, 1));
        s.d_ilo = dalloc<float4>(std::max<uint32_t>(n, 1));
        s.d_ihi = dalloc<float4>(std::max<uint32_t>(n, 1));
        s.d_tlas = dalloc<Node>(std::max<uint32_t>(n, 1));
        s.d_inst_src = dalloc<kf_instance>(std::max<uint32_t>(n, 1));
        s.n_inst = n;
    }
    if (!s.d_gflags) {  // static per scene
        s.d_gflags = dalloc<uint32_t>(std::max<uint32_t>(s.n_geom, 1));
        if (s.n_geom)
            KF_CHECK(hipMemcpy(s.d_gflags, s.h_gflags.data(), s.n_geom * sizeof(uint32_t), hipMemcpyHostToDevice));
        s.d_tlas_aux = dalloc<uint32_t>(8);
        s.d_tlas_err = dalloc<int>(1);
        KF_CHECK(hipMemsetAsync(s.d_tlas_err, 0, sizeof(int), st));
        KF_CHECK(hipEventCreateWithFlags(&s.stage_ev, hipEventDisableTiming));
    }
    // host inputs -> pinned staging -> device
    const size_t w2o_bytes = 12 * sizeof(float) * (size_t)n, inst_bytes = sizeof(kf_instance) * (size_t)n;
    if (s.stage_pending) KF_CHECK(hipEventSynchronize(s.stage_ev));  // the previous update's copies are done
    s.stage_pending = false;
    if (w2o_bytes + inst_bytes > s.stage_cap) {
        if (s.h_stage) KF_CHECK(hipHostFree(s.h_stage));
        s.stage_cap = std::max<size_t>(w2o_bytes + inst_bytes, 4096);
        KF_CHECK(hipHostMalloc((void**)&s.h_stage, s.stage_cap, hipHostMallocDefault));
    }
    float* w2o = (float*)s.h_stage;
    for (uint32_t i = 0; i < n; ++i) kf_world_to_object(h_inst[i].transform, &w2o[12 * i]);
    if (n) std::memcpy(s.h_stage + w2o_bytes, h_inst, inst_bytes);
    float* d_w2o = dalloc<float>(std::max<size_t>(12 * (size_t)n, 1));
    if (n) {
        KF_CHECK(hipMemcpyAsync(d_w2o, s.h_stage, w2o_bytes, hipMemcpyHostToDevice, st));
        KF_CHECK(hipMemcpyAsync(s.d_inst_src, s.h_stage + w2o_bytes, inst_bytes, hipMemcpyHostToDevice, st));
        KF_CHECK(hipEventRecord(s.stage_ev, st));
        s.stage_pending = true;
    }
    float4* glo = dalloc<float4>(std::max<uint32_t>(s.n_geom, 1));
    float4* ghi = dalloc<float4>(std::max<uint32_t>(s.n_geom, 1));
    if (s.n_geom) k_blas_box<<<blocks(s.n_geom, TB), TB, 0, st>>>(s.n_geom, s.d_blas_root, s.d_blas, s.d_tri, glo, ghi);
    float4* ilo = s.d_ilo;  // kept: the renderer culls instances per pixel tile with them
    float4* ihi = s.d_ihi;
    uint32_t* iseg = dalloc<uint32_t>(std::max<uint32_t>(n, 1));
    if (n)
        k_instances<<<blocks(n, TB), TB, 0, st>>>(n, s.d_inst_src, d_w2o, s.n_geom, s.d_gflags, s.d_trav_root4,
                                                  s.d_blas4, s.n_blas4, glo, ghi, s.d_mesh, s.d_toff,
                                                  multi ? s.d_inst_env : nullptr, s.d_inst, ilo, ihi, iseg);
    const uint32_t n_seg = multi ? s.n_env : 1u;
    const uint32_t* seg_off = multi ? s.d_env_off : s.d_tlas_aux;
    const uint32_t* node_off = multi ? s.d_env_node_off : s.d_tlas_aux + 2;
    uint32_t* cnt = multi ? s.d_env_cnt : s.d_tlas_aux + 3;
    if (!multi) k_tlas_aux<<<1, 1, 0, st>>>(n, s.d_tlas_aux);
    uint32_t* order = dalloc<uint32_t>(std::max<uint32_t>(n, 1));
    int* d_root = dalloc<int>(n_seg);
    lbvh_build(n, ilo, ihi, iseg, n_seg, seg_off, node_off, s.d_tlas, order, d_root, false, st);
    // BVH2 nodes: n_e - 1 per environment of n_e >= 2 instances
    uint32_t n_tlas2 = 0;
    if (multi)
        n_tlas2 = s.h_env_node_off.back();
    else
        n_tlas2 = n > 1 ? n - 1 : 0;
    std::vector<int> roots;
    if (multi)
        env_roots(s.h_env_off, s.h_env_node_off, roots);
    else
        env_roots({0u, n}, {0u}, roots);
    s.tlas_root = roots[0];
    if (!s.d_tlas4 || s.tlas4_cap < std::max<uint32_t>(n_tlas2, 1)) {
        dfree(s.d_tlas4);
        s.tlas4_cap = std::max<uint32_t>(n_tlas2, 1);
        s.d_tlas4 = dalloc<Node4>(s.tlas4_cap);
    }
    const uint32_t qcap = std::max<uint32_t>(n_tlas2, 1);
    int2* qa = dalloc<int2>(qcap);
    int2* qb = dalloc<int2>(qcap);
    if (n_tlas2)
        k_collapse_forest<<<n_seg, kTreeThreads, 0, st>>>(s.d_tlas, n_tlas2, seg_off, node_off, qa, qb, s.d_tlas4, 0u,
                                                          cnt, s.d_tlas_err);
    KF_CHECK(hipGetLastError());
    s.tlas_root4 = roots[0];  // BVH4 and BVH2 roots coincide by construction
    s.n_tlas4 = n_tlas2 ? s.tlas4_cap : 0;  // index bound; the node count stays on the device
    dfree(d_w2o); dfree(glo); dfree(ghi); dfree(iseg);
    dfree(order); dfree(d_root); dfree(qa); dfree(qb);
}

// Geometry g's vertices were replaced on the device (same count, same
// triangles): refit its BLAS on the GPU, no host round trip.
void refit_blas(Scene& s, uint32_t g) {
    hipStream_t st = stream();
    const uint32_t TB = 256;
    const uint32_t n = s.h_toff[g + 1] - s.h_toff[g];
    if (n == 0) return;
    float4* lo = dalloc<float4>(n);
    float4* hi = dalloc<float4>(n);
    k_geom_slots<<<blocks(n, TB), TB, 0, st>>>(g, s.d_verts, s.d_idx, s.d_voff, s.d_toff, s.d_tri, lo, hi);
    if (n >= 2) {
        int* parent_node = dalloc<int>(n - 1);
        int* parent_leaf = dalloc<int>(n);
        int* flags = dalloc<int>(n - 1);
        k_fill_int<<<blocks(n - 1, TB), TB, 0, st>>>(parent_node, -1, n - 1);
        k_fill_int<<<blocks(n, TB), TB, 0, st>>>(parent_leaf, -1, n);
        k_geom_parents<<<blocks(n - 1, TB), TB, 0, st>>>(g, s.d_toff, s.d_node_off, s.d_blas, s.n_blas_nodes, parent_node,
                                                        parent_leaf, flags);
        k_geom_refit2<<<blocks(n, TB), TB, 0, st>>>(g, s.d_toff, s.d_node_off, s.n_blas_nodes, lo, hi, parent_node,
                                                   parent_leaf, s.d_blas, flags);
        k_geom_refit4<<<blocks(n - 1, TB), TB, 0, st>>>(g, s.d_toff, s.d_node_off, s.d_blas_cnt, s.d_blas,
                                                       s.n_blas_nodes, s.d_tri, s.d_blas4);
        dfree(parent_node);
        dfree(parent_leaf);
        dfree(flags);
    }
    KF_CHECK(hipGetLastError());
    dfree(lo);
    dfree(hi);
}

}  // namespace kf
