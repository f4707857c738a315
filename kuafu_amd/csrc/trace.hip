// trace.hip — closest-hit / any-hit queries on fixed ray batches.
//
// The traceRayEXT calls of PathTrace.rgen:97-107 (path rays) and
// PathTrace.rchit:186-197 (shadow rays), exposed on their own so that the
// traversal can be checked and measured on fixed ray batches.
#include "kf_internal.h"
#include "traverse.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace kf {
namespace {

constexpr int TB = 256;
static_assert(TB == (int)kRingThreads, "traversal kernels index their LDS stack ring by a 256-thread block");

// Counters of a trace call: [0] node records, [1] triangle tests, [2] stack
// overflow / malformed tree, [4..11] work-chunk counters of the persistent grid.
constexpr int kTraceCnt = 12;

// Fixed batches trace environment 0 (the whole TLAS of a scene without
// environments).
KF_DEV void ray_of(const SceneView& S, const kf_ray& r, uint32_t q, TRay& t) {
    t.o = mk(r.org[0], r.org[1], r.org[2]);
    t.d = mk(r.dir[0], r.dir[1], r.dir[2]);
    t.tmin = r.tmin;
    t.tmax = r.tmax;
    t.root = env_root_item(S, 0u);
    t.alpha_seed = 0u;
    t.payload = q;
}

template <bool COUNT>
__global__ __launch_bounds__(TB) void k_closest(SceneView S, const kf_ray* __restrict__ rays, uint32_t n,
                                                kf_hit* __restrict__ hits, unsigned long long* __restrict__ cnt,
                                                uint32_t refill) {
    uint32_t ovf = 0;
    TCount tot{0, 0, 0, 0};
    auto load = [&](uint32_t q, TRay& t) __attribute__((always_inline)) { ray_of(S, rays[q], q, t); };
    auto store = [&](uint32_t q, THit b, bool) __attribute__((always_inline)) {
        kf_hit o;
        o.t = b.t;
        o.inst = b.inst;
        o.prim = b.prim;
        o.b1 = b.b1;
        o.b2 = b.b2;
        o.back = b.back;
        hits[q] = o;
    };
    trace_persistent<false, COUNT>(S, n, cnt + 4, refill, kGrab, false, load, store, ovf, tot);
    if (COUNT) {
        atomicAdd(&cnt[0], (unsigned long long)tot.nodes);
        atomicAdd(&cnt[1], (unsigned long long)tot.tris);
    }
    if (ovf) atomicAdd(&cnt[2], 1ull);
}

__global__ __launch_bounds__(TB) void k_any(SceneView S, const kf_ray* __restrict__ rays, uint32_t n,
                                            uint8_t* __restrict__ occ, unsigned long long* __restrict__ cnt,
                                            uint32_t refill) {
    uint32_t ovf = 0;
    TCount tot{0, 0, 0, 0};
    auto load = [&](uint32_t q, TRay& t) __attribute__((always_inline)) { ray_of(S, rays[q], q, t); };
    auto store = [&](uint32_t q, THit, bool found) __attribute__((always_inline)) { occ[q] = found ? 1 : 0; };
    trace_persistent<true, false>(S, n, cnt + 4, refill, kGrab, false, load, store, ovf, tot);
    if (ovf) atomicAdd(&cnt[2], 1ull);
}

// Persistent grid: exactly the blocks of `kernel` that are resident at once.
template <class Kernel>
unsigned trace_grid(Kernel kernel) {
    int dev = 0, cus = 256, nb = 0;
    KF_CHECK(hipGetDevice(&dev));
    KF_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    KF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, TB, 0));
    return (unsigned)std::max(nb, 1) * (unsigned)cus;
}

// Streaming copy for the achievable-bandwidth figure: each block copies one
// contiguous 64 KiB chunk, each thread four 16-byte vectors per round with all
// loads issued before the stores, through nontemporal loads and stores (the
// data is touched once).  Measured on MI355X against grid-stride variants and
// hipMemcpy device-to-device: 5.4-5.7 TB/s (read + write) on 1-4 GiB, where
// grid-stride copies reach 4.0-5.2 and hipMemcpy 4.6-5.1.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr size_t kCopyChunk = 4096;  // 16-byte vectors per block
__global__ __launch_bounds__(TB) void k_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst, size_t n) {
    const size_t b0 = (size_t)blockIdx.x * kCopyChunk;
    const size_t b1 = b0 + kCopyChunk < n ? b0 + kCopyChunk : n;
    size_t i = b0 + threadIdx.x;
    for (; i + 3 * TB < b1; i += 4 * TB) {
        u32x4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(&src[i + u * TB]);
#pragma unroll
        for (int u = 0; u < 4; ++u) __builtin_nontemporal_store(v[u], &dst[i + u * TB]);
    }
    for (; i < b1; i += TB) dst[i] = src[i];
}

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" int kf_bench_copy(uint64_t bytes, int iters, float* gbps) {
    try {
        if (!gbps || bytes < 16 || iters < 1) throw Error("kf_bench_copy: bad argument");
        hipStream_t st = stream();
        size_t n = bytes / 16;
        u32x4 *a = nullptr, *b = nullptr;
        KF_CHECK(hipMalloc((void**)&a, n * 16));
        KF_CHECK(hipMalloc((void**)&b, n * 16));
        KF_CHECK(hipMemsetAsync(a, 1, n * 16, st));
        const unsigned grid = (unsigned)((n + kCopyChunk - 1) / kCopyChunk);
        k_copy<<<grid, TB, 0, st>>>(a, b, n);  // warm-up
        hipEvent_t e0, e1;
        KF_CHECK(hipEventCreate(&e0));
        KF_CHECK(hipEventCreate(&e1));
        KF_CHECK(hipEventRecord(e0, st));
        for (int i = 0; i < iters; ++i) k_copy<<<grid, TB, 0, st>>>(a, b, n);
        KF_CHECK(hipEventRecord(e1, st));
        KF_CHECK(hipEventSynchronize(e1));
        KF_CHECK(hipGetLastError());
        float ms = 0.f;
        KF_CHECK(hipEventElapsedTime(&ms, e0, e1));
        *gbps = (float)(2.0 * n * 16 * iters / (ms * 1e-3) / 1e9);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        KF_CHECK(hipFree(a));
        KF_CHECK(hipFree(b));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

namespace {
unsigned long long* trace_counters() {
    static unsigned long long* c = nullptr;
    if (!c) KF_CHECK(hipMalloc((void**)&c, kTraceCnt * sizeof(unsigned long long)));
    return c;
}
}  // namespace

extern "C" int kf_trace_closest(kf_scene* s, const kf_ray* rays, uint32_t n, kf_hit* hits) {
    try {
        if (!s || (n && (!rays || !hits))) throw Error("kf_trace_closest: null argument");
        if (n == 0) return 0;
        hipStream_t st = stream();
        kf_ray* d_r = nullptr;
        kf_hit* d_h = nullptr;
        KF_CHECK(hipMalloc((void**)&d_r, n * sizeof(kf_ray)));
        KF_CHECK(hipMalloc((void**)&d_h, n * sizeof(kf_hit)));
        unsigned long long* c = trace_counters();
        KF_CHECK(hipMemsetAsync(c, 0, kTraceCnt * sizeof(unsigned long long), st));
        KF_CHECK(hipMemcpyAsync(d_r, rays, n * sizeof(kf_ray), hipMemcpyHostToDevice, st));
        k_closest<false><<<trace_grid(k_closest<false>), TB, 0, st>>>(s->view(), d_r, n, d_h, c, traversal_refill());
        KF_CHECK(hipGetLastError());
        unsigned long long h[kTraceCnt];
        KF_CHECK(hipMemcpyAsync(hits, d_h, n * sizeof(kf_hit), hipMemcpyDeviceToHost, st));
        KF_CHECK(hipMemcpyAsync(h, c, sizeof(h), hipMemcpyDeviceToHost, st));
        KF_CHECK(hipStreamSynchronize(st));
        KF_CHECK(hipFree(d_r));
        KF_CHECK(hipFree(d_h));
        if (h[2]) throw Error("kf_trace_closest: traversal stack overflow");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

extern "C" int kf_trace_any(kf_scene* s, const kf_ray* rays, uint32_t n, uint8_t* occluded) {
    try {
        if (!s || (n && (!rays || !occluded))) throw Error("kf_trace_any: null argument");
        if (n == 0) return 0;
        hipStream_t st = stream();
        kf_ray* d_r = nullptr;
        uint8_t* d_o = nullptr;
        KF_CHECK(hipMalloc((void**)&d_r, n * sizeof(kf_ray)));
        KF_CHECK(hipMalloc((void**)&d_o, n));
        unsigned long long* c = trace_counters();
        KF_CHECK(hipMemsetAsync(c, 0, kTraceCnt * sizeof(unsigned long long), st));
        KF_CHECK(hipMemcpyAsync(d_r, rays, n * sizeof(kf_ray), hipMemcpyHostToDevice, st));
        k_any<<<trace_grid(k_any), TB, 0, st>>>(s->view(), d_r, n, d_o, c, traversal_refill());
        KF_CHECK(hipGetLastError());
        unsigned long long h[kTraceCnt];
        KF_CHECK(hipMemcpyAsync(occluded, d_o, n, hipMemcpyDeviceToHost, st));
        KF_CHECK(hipMemcpyAsync(h, c, sizeof(h), hipMemcpyDeviceToHost, st));
        KF_CHECK(hipStreamSynchronize(st));
        KF_CHECK(hipFree(d_r));
        KF_CHECK(hipFree(d_o));
        if (h[2]) throw Error("kf_trace_any: traversal stack overflow");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

// ------------------------------------------------------------ no-alloc variant
// Fixed device scratch of kf_trace_closest_noalloc: the trace counters, then
// staging for up to KF_TRACE_NOALLOC_MAX_RAYS rays and their hits.  Each GPU
// holds its own copy from module load until the process ends.  In host code
// the name `scratch` is only a host-side stand-in: the device address comes
// from hipGetSymbolAddress, once per device, and only that address is used.
constexpr uint32_t kNoAllocMaxRays = KF_TRACE_NOALLOC_MAX_RAYS;
constexpr size_t kNoAllocCntBytes = 256;
constexpr size_t kNoAllocRayBytes = (size_t)kNoAllocMaxRays * sizeof(kf_ray);
constexpr size_t kNoAllocBytes = kNoAllocCntBytes + kNoAllocRayBytes + (size_t)kNoAllocMaxRays * sizeof(kf_hit);
static_assert(kTraceCnt * sizeof(unsigned long long) <= kNoAllocCntBytes, "trace counters fit their slot");
static_assert(kNoAllocRayBytes % 256 == 0, "hit staging stays 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on first use).
unsigned char* noalloc_scratch() {
    int dev = -1;
    KF_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("kf_trace_closest_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(std::string("kf_trace_closest_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = (unsigned char*)p;
    }
    return g_scratch_dev[dev];
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}
}  // namespace

namespace kf {
// Called by kf_init with the chosen device current.
void trace_noalloc_setup() { (void)noalloc_scratch(); }
}  // namespace kf

extern "C" int kf_trace_closest_noalloc(kf_scene* s, const kf_ray* rays, uint32_t n, kf_hit* hits) {
    try {
        if (!s || (n && (!rays || !hits))) throw Error("kf_trace_closest_noalloc: null argument");
        if (n > kNoAllocMaxRays)
            throw Error("kf_trace_closest_noalloc: " + std::to_string(n) + " rays exceed the scratch's maximum of " +
                        std::to_string(kNoAllocMaxRays) + " (use kf_trace_closest)");
        if (n == 0) return 0;
        hipStream_t st = stream();
        unsigned char* base = noalloc_scratch();
        unsigned long long* c = (unsigned long long*)base;
        kf_ray* d_r = (kf_ray*)(base + kNoAllocCntBytes);
        kf_hit* d_h = (kf_hit*)(base + kNoAllocCntBytes + kNoAllocRayBytes);
        k_zero_u64<<<1, 64, 0, st>>>(c, kTraceCnt);  // the last call's counters stay in the scratch
        KF_CHECK(hipMemcpyAsync(d_r, rays, n * sizeof(kf_ray), hipMemcpyHostToDevice, st));
        k_closest<false><<<trace_grid(k_closest<false>), TB, 0, st>>>(s->view(), d_r, n, d_h, c, traversal_refill());
        KF_CHECK(hipGetLastError());
        unsigned long long h[kTraceCnt];
        KF_CHECK(hipMemcpyAsync(hits, d_h, n * sizeof(kf_hit), hipMemcpyDeviceToHost, st));
        KF_CHECK(hipMemcpyAsync(h, c, sizeof(h), hipMemcpyDeviceToHost, st));
        KF_CHECK(hipStreamSynchronize(st));
        if (h[2]) throw Error("kf_trace_closest_noalloc: traversal stack overflow");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

extern "C" int kf_trace_closest_device(kf_scene* s, const kf_ray* d_rays, uint32_t n, kf_hit* d_hits,
                                       float* kernel_ms) {
    try {
        if (!s) throw Error("kf_trace_closest_device: null scene");
        if (n == 0) return 0;
        hipStream_t st = stream();
        unsigned long long* c = trace_counters();
        KF_CHECK(hipMemsetAsync(c, 0, kTraceCnt * sizeof(unsigned long long), st));
        hipEvent_t e0, e1;
        KF_CHECK(hipEventCreate(&e0));
        KF_CHECK(hipEventCreate(&e1));
        KF_CHECK(hipEventRecord(e0, st));
        k_closest<false><<<trace_grid(k_closest<false>), TB, 0, st>>>(s->view(), d_rays, n, d_hits, c, traversal_refill());
        KF_CHECK(hipEventRecord(e1, st));
        KF_CHECK(hipGetLastError());
        KF_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        KF_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (kernel_ms) *kernel_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}
