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
constexpr int kPopRef = (int)0x80000000;      // "pop the stack"
constexpr int kExitBlasRef = (int)0x80000001; // marker: leave the current instance
constexpr float kGamma3 = 3.0f * 5.96046448e-8f / (1.0f - 3.0f * 5.96046448e-8f);

struct Pre {
    v3 o, d, invd;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

KF_DEV float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

KF_DEV void precompute(v3 o, v3 d, Pre& p) {
    p.o = o;
    p.d = d;
    p.invd = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(d, kz);
    if (dz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    p.kx = kx;
    p.ky = ky;
    p.kz = kz;
    p.Sx = comp(d, kx) / dz;
    p.Sy = comp(d, ky) / dz;
    p.Sz = 1.0f / dz;
}

struct TriHit {
    float t, b1, b2;
    uint32_t back;
};

KF_DEV bool intersect(const Pre& p, float4 v0, float4 v1, float4 v2, float tmin, float tmax, TriHit& h) {
    v3 A = mk(v0.x, v0.y, v0.z) - p.o;
    v3 B = mk(v1.x, v1.y, v1.z) - p.o;
    v3 C = mk(v2.x, v2.y, v2.z) - p.o;
    float Akz = comp(A, p.kz), Bkz = comp(B, p.kz), Ckz = comp(C, p.kz);
    float Ax = comp(A, p.kx) - p.Sx * Akz;
    float Ay = comp(A, p.ky) - p.Sy * Akz;
    float Bx = comp(B, p.kx) - p.Sx * Bkz;
    float By = comp(B, p.ky) - p.Sy * Bkz;
    float Cx = comp(C, p.kx) - p.Sx * Ckz;
    float Cy = comp(C, p.ky) - p.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = p.Sz * Akz;
    float Bz = p.Sz * Bkz;
    float Cz = p.Sz * Ckz;
    float Tn = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = Tn * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    h.t = t;
    h.b1 = V * rcp;
    h.b2 = W * rcp;
    h.back = det < 0.0f ? 1u : 0u;
    return true;
}

// Conservative slab test of the two child boxes of a node (Ize 2013).
KF_DEV void slab2(const Node& n, v3 o, v3 invd, float tmin, float tmax, bool& hl, bool& hr, float& tl, float& tr) {
    const float s = 1.0f + 2.0f * kGamma3;
    float l0x = (n.a.x - o.x) * invd.x, l1x = (n.a.y - o.x) * invd.x;
    float l0y = (n.a.z - o.y) * invd.y, l1y = (n.a.w - o.y) * invd.y;
    float l0z = (n.c.x - o.z) * invd.z, l1z = (n.c.y - o.z) * invd.z;
    float r0x = (n.b.x - o.x) * invd.x, r1x = (n.b.y - o.x) * invd.x;
    float r0y = (n.b.z - o.y) * invd.y, r1y = (n.b.w - o.y) * invd.y;
    float r0z = (n.c.z - o.z) * invd.z, r1z = (n.c.w - o.z) * invd.z;
    float ln = fmaxf(fmaxf(fminf(l0x, l1x), fminf(l0y, l1y)), fmaxf(fminf(l0z, l1z), tmin));
    float lf = fminf(fminf(fmaxf(l0x, l1x), fmaxf(l0y, l1y)), fminf(fmaxf(l0z, l1z), tmax));
    float rn = fmaxf(fmaxf(fminf(r0x, r1x), fminf(r0y, r1y)), fmaxf(fminf(r0z, r1z), tmin));
    float rf = fminf(fminf(fmaxf(r0x, r1x), fmaxf(r0y, r1y)), fminf(fmaxf(r0z, r1z), tmax));
    hl = ln <= lf * s;
    hr = rn <= rf * s;
    tl = ln;
    tr = rn;
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
    uint32_t nodes, tris;
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

// ANY: shadow-ray query (opaque, stop at the first hit in [tmin, tmax]).
// ALPHA: run the any-hit alpha test (PathTrace.rahit:30-48) for non-opaque
// instances, with an order-independent coin (alpha_coin).
template <bool ANY, bool COUNT>
KF_DEV bool traverse(const SceneView& S, v3 o, v3 d, float tmin, float tmax, bool use_alpha, uint32_t alpha_seed,
                     THit& best, TCount& cnt, uint32_t& overflow) {
    best.t = tmax;
    best.inst = ~0u;
    best.prim = ~0u;
    best.b1 = 0.0f;
    best.b2 = 0.0f;
    best.back = 0u;
    if (S.tlas_root == kEmptyRef) return false;
    v3 winvd = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    Pre obj;
    v3 co = o, cinvd = winvd;
    bool in_blas = false;
    uint32_t cur_inst = 0;
    int cur_flags = 0;
    uint32_t cur_geom = 0;
    int stack[kStack];
    int sp = 0;
    int ref = S.tlas_root;
    bool found = false;

    auto test_tri = [&](int slot) -> bool {
        if (COUNT) cnt.tris++;
        if ((uint32_t)slot >= S.n_tri) {
            overflow = 1;
            return false;
        }
        const float4* t = S.tri + 3 * (size_t)slot;
        float4 v0 = t[0], v1 = t[1], v2 = t[2];
        TriHit h;
        if (!intersect(obj, v0, v1, v2, tmin, best.t, h)) return false;
        uint32_t prim = __float_as_uint(v0.w);
        if (!ANY) {
            if (!(h.t < best.t || cur_inst < best.inst || (cur_inst == best.inst && prim < best.prim))) return false;
            if (use_alpha && !(cur_flags & kInstOpaque)) {
                float a = S.mats[S.matidx[S.toff[cur_geom] + prim]].alpha;
                if (a == 0.0f || alpha_coin(alpha_seed, cur_inst, prim) > a) return false;
            }
        }
        best.t = h.t;
        best.b1 = h.b1;
        best.b2 = h.b2;
        best.back = h.back;
        best.inst = cur_inst;
        best.prim = prim;
        return true;
    };

    uint32_t steps = 0;
    while (true) {
        if (++steps > kMaxSteps) {
            overflow = 1;
            break;
        }
        if (ref == kPopRef) {
            if (sp == 0) break;
            ref = stack[--sp];
            if (ref == kExitBlasRef) {
                in_blas = false;
                co = o;
                cinvd = winvd;
                ref = kPopRef;
            }
            continue;
        }
        if (ref >= 0) {
            if (COUNT) cnt.nodes++;
            if ((uint32_t)ref >= (in_blas ? S.n_blas_nodes : S.n_tlas_nodes)) {
                overflow = 1;
                ref = kPopRef;
                continue;
            }
            Node n = in_blas ? S.blas[ref] : S.tlas[ref];
            bool hl, hr;
            float tl, tr;
            slab2(n, co, cinvd, tmin, best.t, hl, hr, tl, tr);
            int lref = n.d.x, rref = n.d.y;
            if (in_blas) {
                if (hl && lref < 0) {
                    hl = false;
                    if (test_tri(~lref)) {
                        found = true;
                        if (ANY) return true;
                    }
                }
                if (hr && rref < 0) {
                    hr = false;
                    if (test_tri(~rref)) {
                        found = true;
                        if (ANY) return true;
                    }
                }
            }
            if (hl && hr) {
                int nearr = tl <= tr ? lref : rref;
                int farr = tl <= tr ? rref : lref;
                if (sp < kStack) stack[sp++] = farr; else overflow = 1;
                ref = nearr;
            } else if (hl) {
                ref = lref;
            } else if (hr) {
                ref = rref;
            } else {
                ref = kPopRef;
            }
            continue;
        }
        // leaf
        if (!in_blas) {
            uint32_t ii = (uint32_t)(~ref);
            ref = kPopRef;
            if (ii >= S.n_inst) {
                overflow = 1;
                continue;
            }
            Inst I = S.inst[ii];
            if (!(I.info.z & kInstValid)) continue;
            v3 oo = xform_point(I, o);
            v3 od = xform_dir(I, d);
            precompute(oo, od, obj);
            cur_inst = ii;
            cur_flags = I.info.z;
            cur_geom = (uint32_t)I.info.y;
            int root = I.info.x;
            if (root < 0) {  // single-triangle BLAS
                if (test_tri(~root)) {
                    found = true;
                    if (ANY) return true;
                }
                continue;
            }
            if (sp < kStack) stack[sp++] = kExitBlasRef; else { overflow = 1; continue; }
            in_blas = true;
            co = obj.o;
            cinvd = obj.invd;
            ref = root;
        } else {
            if (test_tri(~ref)) {
                found = true;
                if (ANY) return true;
            }
            ref = kPopRef;
        }
    }
    return found;
}

}  // namespace kf
