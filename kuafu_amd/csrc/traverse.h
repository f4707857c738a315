// traverse.h — stack-based two-level BVH traversal with the watertight
// ray/triangle test (gfx950 device code).
//
// Replaces the fixed-function traversal behind traceRayEXT in
// PathTrace.rgen:97-107 (closest hit, tMin 0.001, tMax 1e4) and
// PathTrace.rchit:186-197 (shadow rays: opaque, terminate on first hit).
// Semantics fixed by this build (the reference leaves them to the driver):
//   * watertight test of Woo, Benthin & Wald (JCGT 2013), double-precision
//     fallback when an edge function is exactly zero;
//   * accept t in [tmin, tmax]; closest hit = smallest t, ties broken by the
//     smaller (instance, primitive) pair, so the result does not depend on the
//     BVH or the visiting order;
//   * hit kind front-facing iff dot(d_obj, cross(v1 - v0, v2 - v0)) < 0;
//   * the object-space ray is (W2O * (o,1), W2O * (d,0)), t is shared.
#pragma once
#include "kf_internal.h"
#include "kf_math.h"

namespace kf {

constexpr int kStack = 64;
constexpr uint32_t kMaxSteps = 1u << 22;  // bound on loop iterations (malformed-tree guard)
// Slab margin: each computed slab distance (b - o) * rcp(d) carries three
// roundings, one of them the 1-ulp hardware reciprocal (2 units of 2^-24),
// so |relative error| <= gamma_4; near <= far * (1 + 2 gamma_5) then never
// rejects a box the exact ray meets (Ize, "Robust BVH ray traversal", 2013).
constexpr float kUnit = 5.96046448e-8f;  // 2^-24
constexpr float kGamma5 = 5.0f * kUnit / (1.0f - 5.0f * kUnit);
constexpr float kSlabScale = 1.0f + 2.0f * kGamma5;

// The shear axes kx, ky, kz are kept as lane masks (k0: axis 0, k1: axis
// 1, neither: axis 2), which the compiler holds in scalar registers.
struct Pre {
    v3 o, invd;  // invd: culling only (hardware reciprocal)
    bool x0, x1, y0, y1, z0, z1;
    float Sx, Sy, Sz;
};

KF_DEV float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
KF_DEV float comp2(v3 a, bool k0, bool k1) { return k0 ? a.x : (k1 ? a.y : a.z); }
// Component-wise choice.  (c ? a : b on structure lvalues compiles to a copy
// from a chosen address, which forces the traversal state into memory.)
KF_DEV v3 sel3(bool c, v3 a, v3 b) { return mk(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }
KF_DEV v3 rcp3(v3 d) { return mk(__builtin_amdgcn_rcpf(d.x), __builtin_amdgcn_rcpf(d.y), __builtin_amdgcn_rcpf(d.z)); }

// Shear constants of the watertight test: kz = dominant axis, (kx, ky)
// swapped when d[kz] < 0 to keep the winding; Sz = 1 / d[kz] correctly
// rounded, Sx = d[kx] * Sz, Sy = d[ky] * Sz.  In two parts so that a caller
// can share the division: precompute_axes returns d[kz], precompute_scale
// takes its correctly rounded reciprocal.  The caller sets invd.
// kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2), kx = kz + 1
// and ky = kz + 2 (mod 3), swapped when d[kz] < 0; the flags are written as
// logic on the three comparisons (lane masks, no integer axes).
KF_DEV float precompute_axes(v3 o, v3 d, Pre& p) {
    p.o = o;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const bool xy = ax >= ay, xz = ax >= az, yz = ay >= az;
    const bool z0 = xy && xz, z1 = !xy && yz, z2 = !z0 && !z1;
    const float dz = z0 ? d.x : (z1 ? d.y : d.z);
    const bool sw = dz < 0.0f;
    p.z0 = z0;
    p.z1 = z1;
    p.x0 = sw ? z1 : z2;  // the (swapped) x axis is 0
    p.x1 = sw ? z2 : z0;  // ... is 1
    p.y0 = sw ? z2 : z1;  // the (swapped) y axis is 0
    p.y1 = sw ? z0 : z2;  // ... is 1
    return dz;
}
KF_DEV void precompute_scale(v3 d, float rcp_dz, Pre& p) {
    p.Sz = rcp_dz;
    p.Sx = comp2(d, p.x0, p.x1) * p.Sz;
    p.Sy = comp2(d, p.y0, p.y1) * p.Sz;
}

struct TriHit {
    float t, b1, b2;
    uint32_t back;
};

// Watertight ray/triangle test without early exits (the caller runs it for
// every lane of a wave) on the vertices relative to the ray origin, in two
// parts around its one division 1 / det, so that a caller can share that
// division: intersect_begin returns the edge
// functions, det and the scaled distance; intersect_end takes the correctly
// rounded 1 / det.  en gates the double-precision fallback so that lanes
// holding other records never take it.
struct TriEdge {
    float U, V, W, det, Tn;
    bool outside;
};
KF_DEV TriEdge intersect_begin(const Pre& p, v3 A, v3 B, v3 C, bool en) {  // A, B, C: vertices - p.o
    float Akz = comp2(A, p.z0, p.z1), Bkz = comp2(B, p.z0, p.z1), Ckz = comp2(C, p.z0, p.z1);
    float Ax = comp2(A, p.x0, p.x1) - p.Sx * Akz;
    float Ay = comp2(A, p.y0, p.y1) - p.Sy * Akz;
    float Bx = comp2(B, p.x0, p.x1) - p.Sx * Bkz;
    float By = comp2(B, p.y0, p.y1) - p.Sy * Bkz;
    float Cx = comp2(C, p.x0, p.x1) - p.Sx * Ckz;
    float Cy = comp2(C, p.y0, p.y1) - p.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (en && (U == 0.0f || V == 0.0f || W == 0.0f)) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    TriEdge e;
    e.outside = (U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f);
    e.det = U + V + W;
    const float Az = p.Sz * Akz;
    const float Bz = p.Sz * Bkz;
    const float Cz = p.Sz * Ckz;
    e.Tn = U * Az + V * Bz + W * Cz;
    e.U = U;
    e.V = V;
    e.W = W;
    return e;
}
KF_DEV bool intersect_end(const TriEdge& e, float rcp_det, float tmin, float tmax, TriHit& h) {
    const float t = e.Tn * rcp_det;
    h.t = t;
    h.b1 = e.V * rcp_det;
    h.b2 = e.W * rcp_det;
    h.back = e.det < 0.0f ? 1u : 0u;
    return !e.outside && e.det != 0.0f && t >= tmin && t <= tmax;
}

KF_DEV float alpha_coin(uint32_t seed, uint32_t inst, uint32_t prim) {
    uint32_t h = tea(seed ^ (inst * 0x9e3779b9u), prim);
    return float(h & 0x00FFFFFFu) / float(0x01000000);
}

struct THit {
    float t, b1, b2;
    uint32_t inst, prim, back;
};

struct TCount {
    uint32_t nodes, tris, insts;  // BVH4 node records, triangle tests, instance records
    uint32_t tlas;                // of the node records, TLAS ones
};

KF_DEV v3 xform_point(const Inst& I, v3 v) {
    return mk(I.w2o[0].x * v.x + I.w2o[0].y * v.y + I.w2o[0].z * v.z + I.w2o[0].w,
              I.w2o[1].x * v.x + I.w2o[1].y * v.y + I.w2o[1].z * v.z + I.w2o[1].w,
              I.w2o[2].x * v.x + I.w2o[2].y * v.y + I.w2o[2].z * v.z + I.w2o[2].w);
}
KF_DEV v3 xform_dir(const Inst& I, v3 v) {
    return mk(I.w2o[0].x * v.x + I.w2o[0].y * v.y + I.w2o[0].z * v.z,
              I.w2o[1].x * v.x + I.w2o[1].y * v.y + I.w2o[1].z * v.z,
              I.w2o[2].x * v.x + I.w2o[2].y * v.y + I.w2o[2].z * v.z);
}

// ---------------------------------------------------------------- traversal
// The traversal is a state machine over 32-bit items: bits 31:30 the kind,
// bits 29:0 an index into that kind's array.  Every step fetches exactly one
// record (BVH4 node, instance record, triangle) and processes it, so all
// lanes of a wave wait on one memory round trip per step whatever kind of
// record each lane is on.  An instance record carries its BLAS root node:
// its step transforms the ray and tests the root's children at once, so
// entering an instance costs no step of its own.  The stack keeps its top
// in a register; the entry below it is read with every record fetch, so a
// pop costs no extra round trip.
enum : uint32_t { kTlasNode = 0u, kBlasNode = 1u, kInstance = 2u, kTriangle = 3u };
constexpr uint32_t kIdxMask = 0x3fffffffu;  // scenes are limited to 2^30 records of each kind
constexpr uint32_t kNoPrim = 0x7fffffffu;   // best.prim before the first hit

KF_DEV uint32_t make_item(uint32_t kind, uint32_t idx) { return (kind << 30) | (idx & kIdxMask); }

// Record arrays of the traversal.  Where a lane chooses between them by the
// kind of its record, each field passes through an empty asm first: that
// keeps the choice a register select, which the optimiser otherwise turns
// into a dynamically indexed private-memory copy of the whole structure.
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) vf4 gf4;  // vf4 in the global address space

template <class T>
KF_DEV T opaque_sgpr(T v) {
    asm volatile("" : "+s"(v));
    return v;
}

__device__ const float4 kZeroRecord[8] = {};  // at least the lines a step reads of any record

struct TravScene {
    const float4 *tlas, *blas, *inst, *tri;
    uint32_t n_tlas, n_blas, n_inst, n_tri;
    const float4* zero;  // 128 zero bytes, read in place of an out-of-range record
    const kf_material* mats;
    const uint32_t *matidx, *toff;
    KF_DEV explicit TravScene(const SceneView& S)
        : tlas((const float4*)S.tlas), blas((const float4*)S.blas), inst((const float4*)S.inst), tri(S.tri),
          n_tlas(S.n_tlas_nodes), n_blas(S.n_blas_nodes), n_inst(S.n_inst), n_tri(S.n_tri),
          zero(kZeroRecord), mats(S.mats), matidx(S.matidx), toff(S.toff) {}
};

// ------------------------------------------------------------- stacks
// A lane's traversal stack below its register-cached top.  Entry i is the
// i-th oldest; sp counts the entries.
//   peek(sp): entry sp - 1 (read speculatively every step; any in-range
//             address when sp == 0);  push(sp, v): store v as entry sp;
//   popped(sp): entry sp has just become the register top (sp is the new
//             count).
struct PrivStack {  // everything in private memory
    static constexpr int kCap = kStack;
    uint32_t a[kStack];
    KF_DEV uint32_t peek(int sp) const { return a[sp > 0 ? sp - 1 : 0]; }
    KF_DEV void push(int sp, uint32_t v, uint32_t& overflow) {
        if (sp < kStack) a[sp] = v;
        else overflow = 1;
    }
    KF_DEV void popped(int) {}
};

// The newest kRing entries in an LDS ring, one column per thread of the
// block (slot-major, so a wave's accesses hit 64 distinct banks whatever
// slots its lanes use); older entries spill to private memory.  Private
// stores dirty whole cache lines that end up written back to HBM, so the
// ring keeps the common shallow stack off the memory system entirely.
// The ring pointer is typed as an LDS (address space 3) pointer and the spill
// array lives outside the struct: with both in one struct the whole struct
// stayed in scratch, so every step reloaded the ring pointer from scratch and
// reached the ring through flat loads and stores (which wait on every
// outstanding global load as well).
constexpr int kRing = 16;                 // power of two
constexpr uint32_t kRingThreads = 256;    // block size of kernels using RingStack
typedef __attribute__((address_space(3))) uint32_t lds_u32;
struct RingStack {
    static constexpr int kCap = kRing + kStack;
    lds_u32* ring;    // this thread's column: slot k at ring[k * kRingThreads]
    uint32_t* spill;  // kStack entries in private memory
    KF_DEV lds_u32& slot(int i) const { return ring[(uint32_t)(i & (kRing - 1)) * kRingThreads]; }
    KF_DEV uint32_t peek(int sp) const { return slot(sp - 1); }
    KF_DEV void push(int sp, uint32_t v, uint32_t& overflow) {
        if (sp >= kCap) {
            overflow = 1;
            return;
        }
        if (sp >= kRing) spill[sp - kRing] = slot(sp);  // evict the entry that shares the slot
        slot(sp) = v;
    }
    KF_DEV void popped(int sp) {
        if (sp >= kRing) slot(sp) = spill[sp - kRing];  // bring the next older entry back into the ring
    }
};

static_assert(RingStack::kCap < 256 && PrivStack::kCap < 256, "stack depth in the low byte of Trav::spx");

// One query: ray, t range, the TLAS root item it starts from (its
// environment's; kEmptyItem: nothing to hit), alpha coin seed and an opaque
// payload for the caller (where to write the result).
struct TRay {
    v3 o, d;
    float tmin, tmax;
    uint32_t root, alpha_seed, payload;
};

KF_DEV uint32_t root_item(int root) {
    return root >= 0 ? make_item(kTlasNode, (uint32_t)root) : make_item(kInstance, (uint32_t)~root);
}

// Root item of environment e's TLAS (the whole TLAS for a scene without
// environments); kEmptyItem when the environment has no instance.
KF_DEV uint32_t env_root_item(const SceneView& S, uint32_t e) {
    if (S.n_env > 1) return S.env_root[e < S.n_env ? e : 0u];
    return S.tlas_root == kEmptyRef ? kEmptyItem : root_item(S.tlas_root);
}

// ANY: shadow-ray query (opaque, stop at the first hit in [tmin, tmax]).
// Closest-hit queries with use_alpha run the any-hit alpha test
// (PathTrace.rahit:30-48) for non-opaque instances, with an order-independent
// coin (alpha_coin).  MIXED (with ANY false): each lane's query is a shadow
// query or a closest-hit one (any_q), so that lanes of one wave trace both
// kinds with the same steps.
template <bool ANY, bool COUNT, bool MIXED = false>
struct Trav {
    v3 o, d;         // world-space ray
    Pre obj;         // object-space ray of the current instance (its axis
                     // flags are kept in the masks below instead)
    // Shear axes of the object-space ray as wave-wide lane masks (bit l:
    // lane l's flag), updated with scalar instructions when lanes enter an
    // instance and read back as per-lane conditions at no cost; as per-lane
    // flags their update took about 30 vector instructions per step.
    uint64_t mx0, mx1, my0, my1, mz0, mz1;
    float tmin;
    THit best;
    uint32_t cur_inst;  // instance whose BLAS the object-space ray belongs to
    uint32_t cur;    // item processed by the next step
    uint32_t top;    // stack top, cached in a register (valid when has_top)
    uint32_t spx;    // bits 7:0 the entries in the stack below the cached
                     // top, bits 31:8 the steps taken (malformed-tree guard)
    bool has_top, found, use_alpha;
    bool any_q;      // MIXED: this lane's query is a shadow query
    uint32_t alpha_seed;
    TCount cnt;

    // Benign state for a lane without a query (it then reads only the zero
    // record until start() gives it one).
    KF_DEV void reset() {
        o = d = mk(0.0f, 0.0f, 0.0f);
        obj.o = obj.invd = o;
        obj.x0 = obj.y1 = true;  // kx = 0, ky = 1, kz = 2
        obj.x1 = obj.y0 = obj.z0 = obj.z1 = false;
        mx0 = my1 = ~0ull;
        mx1 = my0 = mz0 = mz1 = 0ull;
        obj.Sx = obj.Sy = obj.Sz = 0.0f;
        tmin = 0.001f;  // the path tracer's constant, so that its kernels can fold it
        best.t = best.b1 = best.b2 = 0.0f;
        best.inst = ~0u;
        best.prim = kNoPrim;  // closest hit: bit 31 holds the hit kind while tracing
        best.back = 0u;
        cur_inst = 0u;
        cur = top = 0u;
        spx = 0u;
        has_top = found = false;
        use_alpha = false;
        any_q = false;
        alpha_seed = 0u;
        cnt.nodes = cnt.tris = cnt.insts = cnt.tlas = 0u;
    }

    // Lanes with f take a new query r starting at item root (the scene is
    // not empty); the others keep their state.  Written as selects so that
    // refilling some lanes does not copy the state of all of them.
    KF_DEV void start(bool f, const TRay& r, uint32_t root, bool use_alpha_) {
        o = sel3(f, r.o, o);
        d = sel3(f, r.d, d);
        tmin = f ? r.tmin : tmin;
        best.t = f ? r.tmax : best.t;
        best.inst = f ? ~0u : best.inst;
        best.prim = f ? kNoPrim : best.prim;
        best.b1 = f ? 0.0f : best.b1;
        best.b2 = f ? 0.0f : best.b2;
        best.back = f ? 0u : best.back;
        cur = f ? root : cur;
        spx = f ? 0u : spx;
        has_top = f ? false : has_top;
        found = f ? false : found;
        use_alpha = f ? use_alpha_ : use_alpha;
        alpha_seed = f ? r.alpha_seed : alpha_seed;
        cnt.nodes = f ? 0u : cnt.nodes;
        cnt.tlas = f ? 0u : cnt.tlas;
        cnt.tris = f ? 0u : cnt.tris;
        cnt.insts = f ? 0u : cnt.insts;
    }

    // The query's result: closest hits keep the hit kind in bit 31 of
    // best.prim while tracing.
    KF_DEV THit result() const {
        THit r = best;
        const bool hit = !ANY && found;
        r.prim = hit ? (best.prim & 0x7fffffffu) : ~0u;
        r.back = hit ? (best.prim >> 31) : 0u;
        return r;
    }

    // One record; returns false when the query has finished.  stk: kStack
    // private entries owned by this query (a separate array so that the
    // scalar state above stays in registers).
    //
    // The entry below the cached top is loaded speculatively together with
    // the record, so a pop needs no memory access of its own: every step
    // costs the wave one memory round trip.  Each branch only decides an
    // action; the stack is updated once at the end.
    // live: the lane holds a query.  Lanes without one read the zero record
    // and have no side effects beyond their own (discarded) state.
    // SKIP: skip the instance and triangle evaluations when no lane of the
    // wave needs them (for waves that often run with few queries: shadow
    // rays, one-thread traversals; the full waves of the closest-hit launch
    // are 1 % faster without the branches).
    template <bool SKIP, class Stack>
    KF_DEV bool step(const TravScene& S, Stack& stk, uint32_t& overflow, bool live) {
        enum { kNext, kPush, kPop, kDone };
        const uint32_t kind = cur >> 30, idx = cur & kIdxMask;
        // All operands of the per-lane choices are evaluated unconditionally,
        // so the choices stay selects (an asm cannot be speculated).
        const uint32_t n_tlas = opaque_sgpr(S.n_tlas), n_blas = opaque_sgpr(S.n_blas);
        const uint32_t n_inst = opaque_sgpr(S.n_inst), n_tri = opaque_sgpr(S.n_tri);
        const float4 *p_tlas = opaque_sgpr(S.tlas), *p_blas = opaque_sgpr(S.blas);
        const float4 *p_inst = opaque_sgpr(S.inst), *p_tri = opaque_sgpr(S.tri), *p_zero = opaque_sgpr(S.zero);
        const uint32_t lim = kind == kTlasNode ? n_tlas : kind == kBlasNode ? n_blas : kind == kInstance ? n_inst : n_tri;
        // Out-of-range record or step budget spent (malformed tree): read the
        // zero record instead and finish the query with the overflow flag.
        // Kept free of branches: an early exit here makes the compiler copy
        // the whole traversal state on every step.
        int sp = (int)(spx & 0xffu);
        const uint32_t steps = (spx & ~0xffu) + 0x100u;  // this step counted, in bits 31:8
        const bool bad = live && (idx >= lim || steps > (kMaxSteps << 8));
        const bool on = live && !bad;
        const float4* base = kind == kTlasNode ? p_tlas : kind == kBlasNode ? p_blas : kind == kInstance ? p_inst : p_tri;
        const bool node_kind = kind <= kBlasNode;
        const float4* P = on ? base + (node_kind ? 8u : (kind == kTriangle ? 3u : 12u)) * (size_t)idx : p_zero;
        // Ten 16-byte global loads issued together: lines 0-6 of a BVH4 node
        // (128 B) or of an instance record's root node, lines 7-9 of an
        // instance record (its transform; other lanes read the zero record
        // there).  A triangle's x, y and z lines (48 B) go where a node's
        // low x, y and z planes go (r0, r2, r4), and its other loads repeat
        // its own lines instead of reading past the record.  The empty asm needs all
        // results, which keeps the compiler from sinking parts of the fetch
        // into the kind-specific code below (each sunk load would cost the
        // wave another memory round trip).
        const gf4* G = (const gf4*)P;
        const bool wide = kind != kTriangle && on;
        const bool irec = kind == kInstance && on;
        const gf4* GI = irec ? G + 7 : (const gf4*)p_zero;
        const bool trec = kind == kTriangle && on;
        const vf4 r0 = G[0], r1 = G[1], r2 = G[trec ? 1 : 2], r3 = G[trec ? 2 : 3];
        const vf4 r4 = G[wide ? 4 : 2], r5 = G[wide ? 5 : 1], r6 = G[wide ? 6 : 0];  // line 7 of a node is padding
        const vf4 r7 = GI[0], r8 = GI[1], r9 = GI[2];
        const uint32_t below = stk.peek(sp);
        asm volatile("" ::"v"(r0), "v"(r1), "v"(r2), "v"(r3), "v"(r4), "v"(r5), "v"(r6), "v"(r7), "v"(r8), "v"(r9),
                     "v"(below));

        // Every lane evaluates the record as a node, an instance and a
        // triangle and keeps the result of its kind: in a full wave64 nearly
        // every step has lanes of all three kinds, so per-lane branches would
        // run all three anyway and would copy the traversal state at each
        // join.  With SKIP the instance and triangle evaluations are skipped
        // when no lane of the wave needs them (wave-uniform branches), which
        // happens once few lanes still hold queries.  An instance record is a
        // node as well (its BLAS root's children).
        const bool is_node = kind != kTriangle && on;
        const bool is_inst = irec;
        const bool is_tri = kind == kTriangle && on;
        const bool any_inst = !SKIP || __builtin_amdgcn_ballot_w64(is_inst) != 0;
        const bool any_tri = !SKIP || __builtin_amdgcn_ballot_w64(is_tri) != 0;

        // instance: its BLAS is entered with the object-space ray, which this
        // step's node test already uses for the BLAS root's children (an
        // invalid instance's record has no children, so entering it changes
        // nothing the traversal reads later).  Other lanes keep their ray.
        // A lane enters an instance or tests a triangle, never both: the two
        // correctly rounded reciprocals (1 / d[kz] of the new shear, 1 / det
        // of the triangle) share one division.
        const bool enter = is_inst;
        Pre nobj;
        v3 nd = d;
        float dz = 1.0f;
        if (any_inst) {
            Inst I;
            I.w2o[0] = make_float4(r7.x, r7.y, r7.z, r7.w);
            I.w2o[1] = make_float4(r8.x, r8.y, r8.z, r8.w);
            I.w2o[2] = make_float4(r9.x, r9.y, r9.z, r9.w);
            nd = xform_dir(I, d);
            dz = precompute_axes(xform_point(I, o), nd, nobj);
            obj.o = sel3(enter, nobj.o, obj.o);
            // lanes outside the exec mask are not in E and keep their bits
            const uint64_t E = __builtin_amdgcn_ballot_w64(enter);
            mx0 = (__builtin_amdgcn_ballot_w64(nobj.x0) & E) | (mx0 & ~E);
            mx1 = (__builtin_amdgcn_ballot_w64(nobj.x1) & E) | (mx1 & ~E);
            my0 = (__builtin_amdgcn_ballot_w64(nobj.y0) & E) | (my0 & ~E);
            my1 = (__builtin_amdgcn_ballot_w64(nobj.y1) & E) | (my1 & ~E);
            mz0 = (__builtin_amdgcn_ballot_w64(nobj.z0) & E) | (mz0 & ~E);
            mz1 = (__builtin_amdgcn_ballot_w64(nobj.z1) & E) | (mz1 & ~E);
            cur_inst = enter ? idx : cur_inst;
        }
        // one set of hardware reciprocals (culling only) for the ray this
        // step's node test uses fresh: the world ray of a TLAS node, the new
        // object-space ray of an instance (BLAS nodes reuse obj.invd)
        const bool world = kind == kTlasNode;
        const v3 rinv = rcp3(sel3(world, d, nd));
        obj.invd = sel3(enter, rinv, obj.invd);

        // The ray this step's node test uses: the world ray for a TLAS node,
        // the object-space ray (new for an instance) otherwise.  Its low-plane
        // differences are, on a triangle's lines, the triangle's vertices
        // minus the object-space origin.
        const v3 ro = sel3(world, o, obj.o), ri = sel3(world, rinv, obj.invd);
        const float dlx[4] = {r0.x - ro.x, r0.y - ro.x, r0.z - ro.x, r0.w - ro.x};
        const float dly[4] = {r2.x - ro.y, r2.y - ro.y, r2.z - ro.y, r2.w - ro.y};
        const float dlz[4] = {r4.x - ro.z, r4.y - ro.z, r4.z - ro.z, r4.w - ro.z};

        // triangle, against the current instance's object-space ray (lanes
        // entering an instance discard it)
        TriEdge te;
        te.det = 1.0f;
        if (any_tri) {
            Pre pt = obj;
            pt.x0 = __builtin_amdgcn_inverse_ballot_w64(mx0);
            pt.x1 = __builtin_amdgcn_inverse_ballot_w64(mx1);
            pt.y0 = __builtin_amdgcn_inverse_ballot_w64(my0);
            pt.y1 = __builtin_amdgcn_inverse_ballot_w64(my1);
            pt.z0 = __builtin_amdgcn_inverse_ballot_w64(mz0);
            pt.z1 = __builtin_amdgcn_inverse_ballot_w64(mz1);
            te = intersect_begin(pt, mk(dlx[0], dly[0], dlz[0]), mk(dlx[1], dly[1], dlz[1]),
                                 mk(dlx[2], dly[2], dlz[2]), is_tri);
        }
        TriHit h;
        bool tri_hit = false;
        if (any_tri || any_inst) {
            const float q = 1.0f / (enter ? dz : te.det);
            if (any_tri) tri_hit = intersect_end(te, q, tmin, best.t, h) && is_tri;
            if (any_inst) {
                precompute_scale(nd, q, nobj);
                obj.Sx = enter ? nobj.Sx : obj.Sx;
                obj.Sy = enter ? nobj.Sy : obj.Sy;
                obj.Sz = enter ? nobj.Sz : obj.Sz;
            }
        }
        const uint32_t prim = __float_as_uint(r0.w);
        bool take = tri_hit;
        if (!ANY) {
            const bool closer = h.t < best.t || cur_inst < best.inst ||
                                (cur_inst == best.inst && prim < (best.prim & 0x7fffffffu));
            take = take && ((MIXED && any_q) || closer);
            // scenes with non-opaque geometry only: the instance's flags and
            // geometry are read here, for candidate hits
            if (take && use_alpha) {
                const float4 info = S.inst[12u * (size_t)cur_inst + 10u];  // Inst::info
                if (!(__float_as_int(info.z) & kInstOpaque)) {
                    const float a = S.mats[S.matidx[S.toff[__float_as_uint(info.y)] + prim]].alpha;
                    take = !(a == 0.0f || alpha_coin(alpha_seed, cur_inst, prim) > a);
                }
            }
        }
        best.t = take ? h.t : best.t;
        if (!ANY) {
            best.b1 = take ? h.b1 : best.b1;
            best.b2 = take ? h.b2 : best.b2;
            best.inst = take ? cur_inst : best.inst;
            best.prim = take ? (prim | (h.back << 31)) : best.prim;
        }
        found = found || take;

        // node: four child boxes (Node4 layout: lox, hix, loy, hiy, loz, hiz,
        // ref), against the world ray (TLAS) or the object-space ray (BLAS
        // nodes, and instance records, whose ray was set above)
        const float hix[4] = {r1.x, r1.y, r1.z, r1.w};
        const float hiy[4] = {r3.x, r3.y, r3.z, r3.w};
        const float hiz[4] = {r5.x, r5.y, r5.z, r5.w};
        uint32_t citem[4] = {__float_as_uint(r6.x), __float_as_uint(r6.y), __float_as_uint(r6.z), __float_as_uint(r6.w)};
        float key[4];
        int nhit = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float l0x = dlx[c] * ri.x, l1x = (hix[c] - ro.x) * ri.x;
            const float l0y = dly[c] * ri.y, l1y = (hiy[c] - ro.y) * ri.y;
            const float l0z = dlz[c] * ri.z, l1z = (hiz[c] - ro.z) * ri.z;
            const float tn = fmaxf(fmaxf(fminf(l0x, l1x), fminf(l0y, l1y)), fmaxf(fminf(l0z, l1z), tmin));
            const float tf = fminf(fminf(fmaxf(l0x, l1x), fmaxf(l0y, l1y)), fminf(fmaxf(l0z, l1z), best.t));
            const bool h = tn <= tf * kSlabScale && citem[c] != kEmptyItem;
            // Nearer child first; any-hit queries take leaf children
            // (triangles, instances: kinds 2 and 3) before inner ones, which
            // finds occluders sooner.
            key[c] = h ? (((ANY || (MIXED && any_q)) && (citem[c] >> 31)) ? -INFINITY : tn) : INFINITY;
            nhit += h ? 1 : 0;
        }
        // sorting network on (key, item), ascending
        auto cx = [&](int i, int j) {
            const bool sw = key[j] < key[i];
            const float ki = key[i], kj = key[j];
            const uint32_t ii = citem[i], ij = citem[j];
            key[i] = sw ? kj : ki;
            key[j] = sw ? ki : kj;
            citem[i] = sw ? ij : ii;
            citem[j] = sw ? ii : ij;
        };
        cx(0, 1);
        cx(2, 3);
        cx(0, 2);
        cx(1, 3);
        cx(1, 2);

        if (COUNT) {
            cnt.nodes += (node_kind && on) ? 1u : 0u;
            cnt.tlas += (on && kind == kTlasNode) ? 1u : 0u;
            cnt.tris += is_tri ? 1u : 0u;
            cnt.insts += is_inst ? 1u : 0u;
        }
        const int node_act = nhit >= 2 ? kPush : (nhit == 1 ? kNext : kPop);
        int act = is_node ? node_act : (((ANY || (MIXED && any_q)) && take) ? kDone : kPop);
        const uint32_t next = citem[0];
        if (bad) {
            overflow = 1;
            act = kDone;
        }

        // Stack update, resolved with selects.  A node with k >= 2 hit
        // children pushes the k - 1 farther ones, farthest first, so the
        // nearest of them ends up as the cached top.
        const bool push = act == kPush, popping = act == kPop;
        if (!live) act = kDone;
        const bool alive = act != kDone && (!popping || has_top);
        const uint32_t popped_item = top;
#pragma unroll
        for (int j = 3; j >= 1; --j) {
            const bool pj = push && j < nhit;
            if (pj && has_top) stk.push(sp, top, overflow);
            sp = (pj && has_top && sp < Stack::kCap) ? sp + 1 : sp;
            top = pj ? citem[j] : top;
            has_top = has_top || pj;
        }
        cur = popping ? popped_item : next;
        const int sp_in = sp;
        sp = (popping && sp_in > 0) ? sp_in - 1 : sp_in;
        top = popping ? below : top;
        has_top = popping ? sp_in > 0 : has_top;
        if (popping && sp_in > 0) stk.popped(sp_in - 1);
        spx = steps | (uint32_t)sp;
        return alive;
    }
};

// Whole query for one lane (used where rays are generated inside a kernel).
template <bool ANY, bool COUNT>
KF_DEV bool traverse(const SceneView& S, uint32_t root, v3 o, v3 d, float tmin, float tmax, bool use_alpha,
                     uint32_t alpha_seed, THit& best, TCount& cnt, uint32_t& overflow) {
    const TravScene TS(S);
    Trav<ANY, COUNT> T;
    PrivStack stk;
    T.reset();
    TRay r;
    r.o = o;
    r.d = d;
    r.tmin = tmin;
    r.tmax = tmax;
    r.root = root;
    r.alpha_seed = alpha_seed;
    r.payload = 0u;
    T.start(true, r, root, use_alpha);
    if (root != kEmptyItem)
        while (T.template step<true>(TS, stk, overflow, true)) {
        }
    best = T.result();
    cnt.nodes += T.cnt.nodes;
    cnt.tlas += T.cnt.tlas;
    cnt.tris += T.cnt.tris;
    cnt.insts += T.cnt.insts;
    return T.found;
}

// ------------------------------------------------------- persistent driver
// Queries [0, n) are split into 8 contiguous chunks, one per XCD (workgroups
// are dispatched to the 8 XCDs round-robin, so block b runs on XCD b % 8 and
// neighbouring queries share that XCD's L2).  A wave reserves `grab` queries
// at a time from its chunk's counter and refills its idle lanes from that
// reserve once at least `refill` lanes are idle (or all are), so no lane
// waits for the slowest ray of a fixed group.  When its chunk is used up the
// wave moves on to the next one; a plain load of the counter tells a used-up
// chunk without an atomic, which keeps the end of a launch (every wave
// passing over every chunk) off the contended counters.
//   ctr[8]: per-chunk counters, zero before the launch.
//   load(q, TRay&) fills the ray of query q and its payload;
//   store(payload, best, found) writes a finished query.
constexpr uint32_t kGrab = 128;  // default reservation

KF_DEV uint64_t chunk_begin(uint32_t n, uint32_t c) { return (uint64_t)n * c / 8u; }

template <bool ANY, bool COUNT, class Load, class Store>
KF_DEV void trace_persistent(const SceneView& S, uint32_t n, unsigned long long* __restrict__ ctr, uint32_t refill,
                             uint32_t grab, bool use_alpha, Load load, Store store, uint32_t& overflow,
                             TCount& total) {
    const TravScene TS(S);
    Trav<ANY, COUNT> T;
    T.reset();
    __shared__ uint32_t ring[kRing * kRingThreads];
    uint32_t spill[kStack];
    RingStack stk;
    stk.ring = (lds_u32*)(ring + threadIdx.x);
    stk.spill = spill;
    bool live = false;
    uint32_t payload = 0;
    uint32_t chunk = blockIdx.x & 7u;  // wave-uniform from here on
    uint32_t moved = 0;                // chunks found used up
    uint32_t res_q = 0, res_n = 0;     // reserved queries [res_q, res_q + res_n)
    while (true) {
        const unsigned long long idle = __ballot(!live);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if ((res_n > 0u || moved < 8u) && (nidle >= refill || nidle == 64u)) {
            bool fresh = false;
            TRay r;  // the ray this lane fetches, if any
            r.o = r.d = mk(0.0f, 0.0f, 0.0f);
            r.tmin = 0.001f;
            r.tmax = 0.0f;
            r.root = kEmptyItem;
            r.alpha_seed = r.payload = 0u;
            // idle lanes below this one (mbcnt: no per-lane mask to keep)
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            const bool wants = !live;
            uint32_t got = 0;
            while (got < nidle && (res_n > 0u || moved < 8u)) {
                if (res_n == 0u) {
                    const uint64_t c0 = chunk_begin(n, chunk), len = chunk_begin(n, chunk + 1u) - c0;
                    unsigned long long base = 0;
                    if (__builtin_amdgcn_mbcnt_lo(~0u, 0u) == 0u) {  // lane 0
                        base = __hip_atomic_load(&ctr[chunk], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (base < len) base = atomicAdd(&ctr[chunk], (unsigned long long)grab);
                    }
                    base = __shfl(base, 0);
                    if (base < len) {
                        res_q = (uint32_t)(c0 + base);
                        res_n = (uint32_t)min((unsigned long long)grab, len - base);
                    }
                    if (base + grab >= len) {
                        chunk = (chunk + 1u) & 7u;
                        ++moved;
                    }
                    continue;
                }
                const uint32_t take = min(nidle - got, res_n);
                if (wants && rank >= got && rank < got + take) {
                    load(res_q + (rank - got), r);
                    fresh = true;
                }
                got += take;
                res_q += take;
                res_n -= take;
            }
            if (__ballot(fresh) != 0ull) {  // wave-uniform
                T.start(fresh, r, r.root, use_alpha);
                payload = fresh ? r.payload : payload;
                // An empty t range (tmax < tmin, or NaN) or an environment
                // without instances has nothing to trace: finish it at once
                // as a miss.
                const bool empty = fresh && (!(r.tmin <= r.tmax) || r.root == kEmptyItem);
                if (empty) store(r.payload, T.result(), false);
                live = live || (fresh && !empty);
            }
        }
        if (__ballot(live) == 0ull && moved >= 8u && res_n == 0u) break;
        // The step runs even when no lane holds a query (idle lanes read the
        // zero record; this happens only while a wave runs out of work):
        // a path that skipped it would make every step pay for merging the
        // two paths' registers, and the refill state lives only in its branch.
        const bool alive = T.template step<ANY>(TS, stk, overflow, live);
        if (live && !alive) {
            live = false;
            store(payload, T.result(), T.found);
            if (COUNT) {
                total.nodes += T.cnt.nodes;
                total.tlas += T.cnt.tlas;
                total.tris += T.cnt.tris;
                total.insts += T.cnt.insts;
            }
        }
    }
}

}  // namespace kf
