// oracle.cpp — CPU oracle of Kuafu's path-tracing hot path (TEST INFRASTRUCTURE).
//
// Restates resources/shaders/PathTrace.rgen, PathTrace.rchit, PathTrace.rmiss,
// PathTraceShadow.rmiss, PathTrace.rahit and base/*.glsl of the reference on
// the CPU, templated on the arithmetic type:
//   float  — the per-sample parity oracle for the HIP kernels (same op order,
//            own deterministic sin/cos/exp2 from ko_math.h);
//   double — the convergence oracle (libm math).
// Traversal semantics (which the reference leaves to the Vulkan driver) are
// fixed here: watertight ray/triangle test (Woo, Benthin, Wald 2013), closest
// hit = smallest t with ties broken by (instance, primitive), t in [tmin, tmax].
// The oracle's own BVH (binned SAH) is pinned against brute force by tests.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).
#include "oracle.h"
#include "ko_math.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <omp.h>

using namespace ko;

namespace {

constexpr float kPi = 3.141592f;  // Random.glsl:1 (M_PI)
constexpr float kGamma3 = 3.0f * 5.96046448e-8f / (1.0f - 3.0f * 5.96046448e-8f);

// ------------------------------------------------------------------ sRGB
struct SrgbLut {
    float dec[256];        // sRGB byte -> linear (IEC 61966-2-1)
    double enc_thr[255];   // linear thresholds between consecutive codes
    SrgbLut() {
        for (int i = 0; i < 256; ++i) {
            double c = i / 255.0;
            double l = c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
            dec[i] = (float)l;
        }
        for (int k = 0; k < 255; ++k) {
            // code k+1 is chosen when encode(x)*255 >= k + 0.5
            double s = (k + 0.5) / 255.0;
            enc_thr[k] = s <= 0.04045 ? s / 12.92 : std::pow((s + 0.055) / 1.055, 2.4);
        }
    }
};
const SrgbLut& srgb() { static SrgbLut l; return l; }

uint8_t encode_srgb8(float x) {
    if (!(x > 0.0f)) return 0;  // NaN and <= 0 -> 0
    const SrgbLut& l = srgb();
    int lo = 0, hi = 255;       // number of thresholds <= x
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if ((double)x >= l.enc_thr[mid]) lo = mid + 1; else hi = mid;
    }
    return (uint8_t)lo;
}
uint8_t encode_unorm8(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 1.0f) return 255;
    return (uint8_t)(int)rintf(x * 255.0f);
}

// ------------------------------------------------------------------ BVH
struct Box {
    float lo[3], hi[3];
    Box() { for (int i = 0; i < 3; ++i) { lo[i] = INFINITY; hi[i] = -INFINITY; } }
    void grow(const float* p) { for (int i = 0; i < 3; ++i) { lo[i] = std::min(lo[i], p[i]); hi[i] = std::max(hi[i], p[i]); } }
    void grow(const Box& b) { for (int i = 0; i < 3; ++i) { lo[i] = std::min(lo[i], b.lo[i]); hi[i] = std::max(hi[i], b.hi[i]); } }
    float area() const {
        float d[3];
        for (int i = 0; i < 3; ++i) d[i] = std::max(0.0f, hi[i] - lo[i]);
        return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
    }
};
struct BNode {
    Box box;
    int left = -1, right = -1;  // internal: children; leaf: left = -1
    int first = 0, count = 0;   // leaf: range into item order
};
struct Bvh {
    std::vector<BNode> nodes;
    std::vector<int> order;
};

// Binned SAH, 16 bins, leaves of <= 4 items.
void build_bvh(const std::vector<Box>& boxes, Bvh& out) {
    int n = (int)boxes.size();
    out.nodes.clear();
    out.order.resize(n);
    for (int i = 0; i < n; ++i) out.order[i] = i;
    if (n == 0) return;
    std::vector<float> cen(3 * (size_t)n);
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) cen[3 * i + a] = 0.5f * (boxes[i].lo[a] + boxes[i].hi[a]);
    struct Task { int node, first, count; };
    std::vector<Task> stack;
    out.nodes.push_back(BNode());
    stack.push_back({0, 0, n});
    while (!stack.empty()) {
        Task t = stack.back();
        stack.pop_back();
        BNode& nd = out.nodes[t.node];
        Box b, cb;
        for (int i = t.first; i < t.first + t.count; ++i) {
            b.grow(boxes[out.order[i]]);
            cb.grow(&cen[3 * out.order[i]]);
        }
        nd.box = b;
        nd.first = t.first;
        nd.count = t.count;
        if (t.count <= 4) continue;
        int best_axis = -1, best_split = -1;
        float best_cost = INFINITY;
        const int NB = 16;
        for (int a = 0; a < 3; ++a) {
            float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0)) continue;
            Box bb[NB];
            int bc[NB] = {0};
            for (int i = t.first; i < t.first + t.count; ++i) {
                int it = out.order[i];
                int k = (int)((cen[3 * it + a] - cb.lo[a]) / ext * NB);
                k = std::min(std::max(k, 0), NB - 1);
                bb[k].grow(boxes[it]);
                bc[k]++;
            }
            for (int s = 1; s < NB; ++s) {
                Box l, r;
                int nl = 0, nr = 0;
                for (int k = 0; k < s; ++k) { if (bc[k]) { l.grow(bb[k]); nl += bc[k]; } }
                for (int k = s; k < NB; ++k) { if (bc[k]) { r.grow(bb[k]); nr += bc[k]; } }
                if (!nl || !nr) continue;
                float cost = l.area() * nl + r.area() * nr;
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = s; }
            }
        }
        int mid;
        if (best_axis < 0) {
            mid = t.first + t.count / 2;  // all centroids equal: split by count
        } else {
            float ext = cb.hi[best_axis] - cb.lo[best_axis];
            auto it = std::partition(out.order.begin() + t.first, out.order.begin() + t.first + t.count,
                                     [&](int i) {
                                         int k = (int)((cen[3 * i + best_axis] - cb.lo[best_axis]) / ext * NB);
                                         k = std::min(std::max(k, 0), NB - 1);
                                         return k < best_split;
                                     });
            mid = (int)(it - out.order.begin());
            if (mid == t.first || mid == t.first + t.count) mid = t.first + t.count / 2;
        }
        int l = (int)out.nodes.size();
        out.nodes.push_back(BNode());
        out.nodes.push_back(BNode());
        out.nodes[t.node].left = l;
        out.nodes[t.node].right = l + 1;
        out.nodes[t.node].count = 0;
        stack.push_back({l, t.first, mid - t.first});
        stack.push_back({l + 1, mid, t.first + t.count - mid});
    }
}

// ------------------------------------------------------------------ scene
struct Tex { uint32_t w = 0, h = 0; std::vector<uint8_t> px; };

}  // namespace

struct ko_scene {
    uint32_t ngeom = 0;
    std::vector<uint32_t> voff, toff, gflags;
    std::vector<kf_vertex> verts;
    std::vector<uint32_t> idx, matidx;
    std::vector<kf_instance> inst;
    std::vector<float> w2o;  // 12 per instance
    std::vector<kf_material> mats;
    std::vector<Tex> tex;
    kf_dir_light dl;
    kf_point_lights pl;
    kf_active_lights al;
    uint32_t env_kind = 0;
    Tex env[6];
    float env_grad[8] = {0};
    std::vector<Bvh> blas;       // per geometry, items = local triangles
    std::vector<Box> inst_box;   // world AABB per instance (padded)
    std::vector<int> inst_valid;
    Bvh tlas;
};

extern "C" void ko_world_to_object(const float t[16], float out[12]) {
    // m[r][c] = t[c*4 + r]; inverse of the 3x3 part by cofactors in double,
    // translation -M^-1 * t, rounded to float.
    double m[3][3], tr[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) m[r][c] = t[c * 4 + r];
        tr[r] = t[3 * 4 + r];
    }
    double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    double inv[3][3];
    inv[0][0] = c00 / det;
    inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    inv[1][0] = c01 / det;
    inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    inv[2][0] = c02 / det;
    inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) out[r * 4 + c] = (float)inv[r][c];
        out[r * 4 + 3] = (float)(-(inv[r][0] * tr[0] + inv[r][1] * tr[1] + inv[r][2] * tr[2]));
    }
}

namespace {

// ------------------------------------------------------------------ rays
template <class T> struct Ray {
    V3<T> o, d;
    T tmin, tmax;
};
template <class T> struct Pre {
    V3<T> o, d, invd;
    int kx, ky, kz;
    T Sx, Sy, Sz;
};
template <class T> T tabs(T x) { return x < 0 ? -x : x; }

template <class T> Pre<T> precompute(V3<T> o, V3<T> d) {
    Pre<T> p;
    p.o = o;
    p.d = d;
    p.invd = V3<T>(T(1) / d.x, T(1) / d.y, T(1) / d.z);
    T ax = tabs(d.x), ay = tabs(d.y), az = tabs(d.z);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (d[kz] < 0) std::swap(kx, ky);
    p.kx = kx; p.ky = ky; p.kz = kz;
    // Sz correctly rounded; Sx, Sy as products with it (any per-ray shear
    // constants keep the test watertight, the kernels use these).
    p.Sz = T(1) / d[kz];
    p.Sx = d[kx] * p.Sz;
    p.Sy = d[ky] * p.Sz;
    return p;
}

template <class T> struct TriHit { T t, b1, b2; uint32_t back; };

// Watertight ray/triangle test (Woo, Benthin, Wald 2013, "Watertight
// Ray/Triangle Intersection", JCGT 2(1)).  Front-facing iff
// dot(d, cross(v1 - v0, v2 - v0)) < 0, which is det > 0 below.
template <class T>
bool intersect(const Pre<T>& p, const float* v0, const float* v1, const float* v2, T tmin, T tmax, TriHit<T>& h) {
    V3<T> A = V3<T>(T(v0[0]), T(v0[1]), T(v0[2])) - p.o;
    V3<T> B = V3<T>(T(v1[0]), T(v1[1]), T(v1[2])) - p.o;
    V3<T> C = V3<T>(T(v2[0]), T(v2[1]), T(v2[2])) - p.o;
    T Ax = A[p.kx] - p.Sx * A[p.kz];
    T Ay = A[p.ky] - p.Sy * A[p.kz];
    T Bx = B[p.kx] - p.Sx * B[p.kz];
    T By = B[p.ky] - p.Sy * B[p.kz];
    T Cx = C[p.kx] - p.Sx * C[p.kz];
    T Cy = C[p.ky] - p.Sy * C[p.kz];
    T U = Cx * By - Cy * Bx;
    T V = Ax * Cy - Ay * Cx;
    T W = Bx * Ay - By * Ax;
    if (sizeof(T) == 4 && (U == T(0) || V == T(0) || W == T(0))) {
        U = (T)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (T)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (T)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < T(0) || V < T(0) || W < T(0)) && (U > T(0) || V > T(0) || W > T(0))) return false;
    T det = U + V + W;
    if (det == T(0)) return false;
    T Az = p.Sz * A[p.kz];
    T Bz = p.Sz * B[p.kz];
    T Cz = p.Sz * C[p.kz];
    T Tn = U * Az + V * Bz + W * Cz;
    T rcp = T(1) / det;
    T t = Tn * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    h.t = t;
    h.b1 = V * rcp;
    h.b2 = W * rcp;
    h.back = det < T(0) ? 1u : 0u;
    return true;
}

// Conservative slab test (Ize 2013: scale the far distance by 1 + 2*gamma(3)).
template <class T> bool slab(const Pre<T>& p, const Box& b, T tmin, T tmax) {
    T tn = tmin, tf = tmax;
    for (int a = 0; a < 3; ++a) {
        T t0 = (T(b.lo[a]) - p.o[a]) * p.invd[a];
        T t1 = (T(b.hi[a]) - p.o[a]) * p.invd[a];
        T lo = std::fmin(t0, t1), hi = std::fmax(t0, t1);
        tn = std::fmax(tn, lo);
        tf = std::fmin(tf, hi);
    }
    return tn <= tf * (T(1) + T(2) * T(kGamma3));
}

template <class T> struct Hit {
    T t;
    uint32_t inst, prim;
    T b1, b2;
    uint32_t back;
};

// closest-hit tie rule: smaller t, then smaller (instance, primitive)
template <class T> inline bool better(T t, uint32_t i, uint32_t p, const Hit<T>& h) {
    if (t < h.t) return true;
    if (t > h.t) return false;
    return i < h.inst || (i == h.inst && p < h.prim);
}

inline float alpha_coin(uint32_t seed, uint32_t inst, uint32_t prim) {
    // Deterministic stand-in for rahit:44 `rnd(ray.seed) > alpha` that does
    // not depend on the order in which candidates are visited.
    uint32_t h = tea(seed ^ (inst * 0x9e3779b9u), prim);
    return float(h & 0x00FFFFFFu) / float(0x01000000);
}

template <class T> V3<T> xform_point(const float* w, V3<T> v) {
    return V3<T>(T(w[0]) * v.x + T(w[1]) * v.y + T(w[2]) * v.z + T(w[3]),
                 T(w[4]) * v.x + T(w[5]) * v.y + T(w[6]) * v.z + T(w[7]),
                 T(w[8]) * v.x + T(w[9]) * v.y + T(w[10]) * v.z + T(w[11]));
}
template <class T> V3<T> xform_dir(const float* w, V3<T> v) {
    return V3<T>(T(w[0]) * v.x + T(w[1]) * v.y + T(w[2]) * v.z,
                 T(w[4]) * v.x + T(w[5]) * v.y + T(w[6]) * v.z,
                 T(w[8]) * v.x + T(w[9]) * v.y + T(w[10]) * v.z);
}

const float* tri_vertex(const ko_scene& s, uint32_t g, uint32_t prim, int k) {
    uint32_t vi = s.idx[3 * (size_t)(s.toff[g] + prim) + k];
    return s.verts[s.voff[g] + vi].pos;
}

// any_hit: shadow rays (opaque, terminate on first hit: rchit:182-197).
// alpha_seed: used for non-opaque geometry on path rays only.
template <class T>
bool trace_blas(const ko_scene& s, uint32_t inst, const Pre<T>& p, T tmin, Hit<T>& best, bool any,
                bool use_alpha, uint32_t alpha_seed, bool brute) {
    uint32_t g = s.inst[inst].geometry_index;
    if (s.gflags[g] & KF_GEOM_HIDDEN) return false;
    bool opaque = (s.gflags[g] & KF_GEOM_OPAQUE) != 0;
    uint32_t ntri = s.toff[g + 1] - s.toff[g];
    auto test = [&](uint32_t prim) -> bool {
        TriHit<T> h;
        if (!intersect(p, tri_vertex(s, g, prim, 0), tri_vertex(s, g, prim, 1), tri_vertex(s, g, prim, 2), tmin,
                       best.t, h))
            return false;
        if (!any && !better(h.t, inst, prim, best)) return false;
        if (use_alpha && !opaque) {
            float a = s.mats[s.matidx[s.toff[g] + prim]].alpha;
            if (a == 0.0f || alpha_coin(alpha_seed, inst, prim) > a) return false;
        }
        best.t = h.t; best.inst = inst; best.prim = prim; best.b1 = h.b1; best.b2 = h.b2; best.back = h.back;
        return true;
    };
    if (brute) {
        for (uint32_t prim = 0; prim < ntri; ++prim)
            if (test(prim) && any) return true;
        return false;
    }
    const Bvh& b = s.blas[g];
    if (b.nodes.empty()) return false;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BNode& nd = b.nodes[stack[--sp]];
        if (!slab(p, nd.box, tmin, best.t)) continue;
        if (nd.left < 0) {
            for (int i = nd.first; i < nd.first + nd.count; ++i)
                if (test((uint32_t)b.order[i]) && any) return true;
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    return false;
}

template <class T>
bool trace(const ko_scene& s, V3<T> o, V3<T> d, T tmin, T tmax, Hit<T>& best, bool any, bool use_alpha,
           uint32_t alpha_seed, bool brute) {
    best.t = tmax;
    best.inst = ~0u;
    best.prim = ~0u;
    best.b1 = best.b2 = 0;
    best.back = 0;
    Pre<T> wp = precompute(o, d);
    auto visit = [&](uint32_t i) -> bool {
        Pre<T> p = precompute(xform_point(&s.w2o[12 * i], o), xform_dir(&s.w2o[12 * i], d));
        return trace_blas(s, i, p, tmin, best, any, use_alpha, alpha_seed, brute);
    };
    if (brute) {
        for (uint32_t i = 0; i < s.inst.size(); ++i)
            if (visit(i) && any) return true;
        return best.inst != ~0u;
    }
    if (s.tlas.nodes.empty()) return false;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BNode& nd = s.tlas.nodes[stack[--sp]];
        if (!slab(wp, nd.box, tmin, best.t)) continue;
        if (nd.left < 0) {
            for (int k = nd.first; k < nd.first + nd.count; ++k) {
                uint32_t i = (uint32_t)s.tlas.order[k];
                if (!s.inst_valid[i]) continue;
                if (!slab(wp, s.inst_box[i], tmin, best.t)) continue;
                if (visit(i) && any) return true;
            }
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    return best.inst != ~0u;
}

// ------------------------------------------------------------------ textures
// Sampler of vkCore.hpp:619-637: linear filter, repeat, one mip level; sRGB
// RGBA8 texels are decoded before filtering.  Bilinear at texel centres.
template <class T> void tex_sample(const Tex& t, T u, T v, bool repeat, T out[4]) {
    const SrgbLut& L = srgb();
    if (t.w == 0 || t.h == 0) { out[0] = out[1] = out[2] = out[3] = 0; return; }
    T x = u * T(t.w) - T(0.5), y = v * T(t.h) - T(0.5);
    T fx0 = std::floor(x), fy0 = std::floor(y);
    T fx = x - fx0, fy = y - fy0;
    long x0 = (long)fx0, y0 = (long)fy0;
    auto wrap = [&](long i, long n) -> long {
        if (repeat) { long r = i % n; return r < 0 ? r + n : r; }
        return i < 0 ? 0 : (i >= n ? n - 1 : i);
    };
    long xs[2] = {wrap(x0, t.w), wrap(x0 + 1, t.w)};
    long ys[2] = {wrap(y0, t.h), wrap(y0 + 1, t.h)};
    T c[2][2][4];
    for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
            const uint8_t* px = &t.px[4 * ((size_t)ys[j] * t.w + xs[i])];
            for (int k = 0; k < 3; ++k) c[j][i][k] = T(L.dec[px[k]]);
            c[j][i][3] = T(px[3]) / T(255);
        }
    for (int k = 0; k < 4; ++k) {
        T a = c[0][0][k] * (T(1) - fx) + c[0][1][k] * fx;
        T b = c[1][0][k] * (T(1) - fx) + c[1][1][k] * fx;
        out[k] = a * (T(1) - fy) + b * fy;
    }
}

// Cube map lookup: Vulkan face selection, clamp-to-edge, per-face bilinear.
template <class T> V3<T> cube_sample(const ko_scene& s, V3<T> r) {
    T ax = tabs(r.x), ay = tabs(r.y), az = tabs(r.z);
    int face;
    T sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (r.x >= 0) { face = 0; sc = -r.z; tc = -r.y; } else { face = 1; sc = r.z; tc = -r.y; }
    } else if (ay >= az) {
        ma = ay;
        if (r.y >= 0) { face = 2; sc = r.x; tc = r.z; } else { face = 3; sc = r.x; tc = -r.z; }
    } else {
        ma = az;
        if (r.z >= 0) { face = 4; sc = r.x; tc = -r.y; } else { face = 5; sc = -r.x; tc = -r.y; }
    }
    T u = T(0.5) * (sc / ma + T(1)), v = T(0.5) * (tc / ma + T(1));
    T o[4];
    tex_sample(s.env[face], u, v, false, o);
    return V3<T>(o[0], o[1], o[2]);
}

// Miss shader (rmiss:12-31): env lookup direction (-d.y, d.z, -d.x).
template <class T> V3<T> miss_emission(const ko_scene& s, const kf_push_constants& pc, V3<T> d) {
    if (pc.use_env_map) {
        V3<T> dir(-d.y, d.z, -d.x);
        if (s.env_kind == KF_ENV_CUBEMAP) return cube_sample(s, dir);
        if (s.env_kind == KF_ENV_GRADIENT) {
            T t = gclamp(T(0.5) * dir.y + T(0.5), T(0), T(1));
            V3<T> lo(T(s.env_grad[0]), T(s.env_grad[1]), T(s.env_grad[2]));
            V3<T> hi(T(s.env_grad[4]), T(s.env_grad[5]), T(s.env_grad[6]));
            return lo * (T(1) - t) + hi * t;
        }
        return V3<T>(T(0));
    }
    return V3<T>(T(pc.clear_color[0]), T(pc.clear_color[1]), T(pc.clear_color[2])) * T(pc.clear_color[3]);
}

// ------------------------------------------------------------------ BSDF pieces
template <class T> T schlick(T cosine, T ior) {  // Random.glsl:68-73 (pow(x,5) as x^2*x^2*x)
    T r0 = (T(1) - ior) / (T(1) + ior);
    r0 *= r0;
    T x = T(1) - cosine;
    T x2 = x * x;
    return r0 + (T(1) - r0) * (x2 * x2 * x);
}
template <class T> T ggx_d(T NdotH, T a2) {  // Random.glsl:75-79
    T d = gmax(NdotH * NdotH * (a2 - T(1)) + T(1), T(1e-6f));
    return a2 / (d * d * T(kPi));
}
template <class T> T g1(T x, T a2) {  // Random.glsl:117-121
    return (T(2) * x) / (x + Fn<T>::sqrt(a2 + (T(1) - a2) * (x * x)));
}
// GeometricShadowing(NdotV, NdotL, a2) = G1(NdotL) * G1(NdotV)  (Random.glsl:123-127)
template <class T> T geom_shadow(T NdotV, T NdotL, T a2) { return g1(NdotL, a2) * g1(NdotV, a2); }
template <class T> V3<T> schlick_fresnel(V3<T> f0, T lDotH) {  // Random.glsl:133-136
    T e = Fn<T>::exp2((T(-5.55473f) * lDotH - T(6.98316f)) * lDotH);
    return f0 + (V3<T>(T(1)) - f0) * e;
}
template <class T> V3<T> perpendicular(V3<T> u) {  // Random.glsl:58-65
    V3<T> a(tabs(u.x), tabs(u.y), tabs(u.z));
    uint32_t xm = ((a.x - a.y) < 0 && (a.x - a.z) < 0) ? 1u : 0u;
    uint32_t ym = (a.y - a.z) < 0 ? (1u ^ xm) : 0u;
    uint32_t zm = 1u ^ (xm | ym);
    return cross(u, V3<T>(T(xm), T(ym), T(zm)));
}
template <class T> V3<T> sample_ggx(uint32_t& seed, T a2, V3<T> N) {  // Random.glsl:83-101
    T rx = T(rnd(seed));
    T ry = T(rnd(seed));
    V3<T> B = perpendicular(N);
    V3<T> Tg = cross(B, N);
    T cosT = Fn<T>::sqrt(gclamp((T(1) - rx) / ((a2 - T(1)) * rx + T(1)), T(0), T(1)));
    T sinT = Fn<T>::sqrt(gclamp(T(1) - cosT * cosT, T(0), T(1)));
    T phi = ry * T(kPi) * T(2);
    return Tg * (sinT * Fn<T>::cos(phi)) + B * (sinT * Fn<T>::sin(phi)) + N * cosT;
}
template <class T> V3<T> local_to_world(V3<T> d, V3<T> n) {  // Sampling.glsl:6-23
    V3<T> t;
    if (tabs(n.x) > tabs(n.y))
        t = V3<T>(n.z, T(0), -n.x) / Fn<T>::sqrt(n.x * n.x + n.z * n.z);
    else
        t = V3<T>(T(0), -n.z, n.y) / Fn<T>::sqrt(n.y * n.y + n.z * n.z);
    V3<T> b = cross(n, t);
    return d.x * t + d.y * b + d.z * n;
}
template <class T> V3<T> cosine_hemisphere(uint32_t& seed, V3<T> n) {  // Sampling.glsl:25-35
    T u0 = T(rnd(seed));
    T u1 = T(rnd(seed));
    T sq = Fn<T>::sqrt(T(1) - u1);
    T ang = T(2.0f * kPi) * u0;
    V3<T> dir(Fn<T>::cos(ang) * sq, Fn<T>::sin(ang) * sq, Fn<T>::sqrt(u1));
    return local_to_world(dir, n);
}
template <class T> V3<T> uniform_sphere(uint32_t& seed) {  // Sampling.glsl:72-83
    V3<T> p;
    do {
        T a = T(rnd(seed)), b = T(rnd(seed)), c = T(rnd(seed));
        p = V3<T>(a, b, c) * T(2) - V3<T>(T(1));
    } while (dot(p, p) >= T(1));
    return p;
}
template <class T> void disk(uint32_t& seed, T& px, T& py) {  // Sampling.glsl:124-134
    do {
        T a = T(rnd(seed)), b = T(rnd(seed));
        px = T(2) * a - T(1);
        py = T(2) * b - T(1);
    } while (px * px + py * py >= T(1));
}

// ------------------------------------------------------------------ shading
template <class T> struct Shade {
    V3<T> N, V, worldPos;
    T f, a2;
    V3<T> diffuseColor, specularColor, transmissionColor;
    bool inside;
};

// calcDirectContribution (rchit:103-171).
template <class T> V3<T> direct(const Shade<T>& sh, V3<T> L, V3<T> lightEmission, uint32_t& seed) {
    const V3<T>& N = sh.N;
    const V3<T>& V = sh.V;
    V3<T> weight(T(0));
    if (!all_zero(sh.transmissionColor)) {
        if (sh.inside) {
            weight = V3<T>(T(0));
        } else {
            T NdotV0 = dot(N, V);
            V3<T> refr = refract(-V, N, T(1) / sh.f);
            T reflectProb = !all_zero(refr) ? schlick(NdotV0, sh.f) : T(1);
            if (T(rnd(seed)) <= reflectProb) {
                V3<T> H = normalize(L + V);
                T NdotL = dot(N, L);
                T NdotH = dot(N, H);
                T HdotV = dot(H, V);
                T NdotV = gmax(dot(N, V), T(1e-6f));
                T LdotH = dot(L, H);
                T D = ggx_d(NdotH, sh.a2);
                T G = geom_shadow(NdotL, NdotV, sh.a2);
                V3<T> F = schlick_fresnel(sh.transmissionColor, LdotH);
                weight = D * F * G * HdotV / NdotH * NdotV;
            } else {
                weight = V3<T>(T(0));
            }
        }
    } else {
        V3<T> H = normalize(L + V);
        T NdotL = dot(N, L);
        T NdotH = dot(N, H);
        T HdotV = dot(H, V);
        T NdotV = gmax(dot(N, V), T(1e-6f));
        T LdotH = dot(L, H);
        T D = ggx_d(NdotH, sh.a2);
        T G = geom_shadow(NdotL, NdotV, sh.a2);
        V3<T> F = schlick_fresnel(sh.specularColor, LdotH);
        T dl = length(sh.diffuseColor);
        T sl = length(sh.specularColor);
        T pD = dl / (dl + sl);
        if (dl == T(0) && sl == T(0)) pD = T(0.5);
        V3<T> dw = sh.diffuseColor * V3<T>(NdotL);
        V3<T> sw = D * F * G * HdotV / NdotH * NdotV;
        weight = T(rnd(seed)) < pD ? dw * pD : sw * (T(1) - pD);
    }
    return lightEmission * weight;
}

struct Counts { uint64_t primary = 0, extension = 0, shadow = 0; };

// traceShadowRay (rchit:175-200).
template <class T>
V3<T> shadow_ray(const ko_scene& s, const Shade<T>& sh, V3<T> L, T maxDist, V3<T> lightEmission, uint32_t& seed,
                 Counts& cnt) {
    bool shadowed = true;
    T NdotL = dot(sh.N, L);
    if (NdotL > T(0)) {
        Hit<T> h;
        cnt.shadow++;
        shadowed = trace(s, sh.worldPos, L, T(0.001f), maxDist, h, true, false, 0u, false);
    }
    return shadowed ? V3<T>(T(0)) : direct(sh, L, lightEmission, seed);
}

template <class T> V3<T> dir_light(const ko_scene& s, const Shade<T>& sh, V3<T> rayDir, uint32_t& seed, Counts& c) {
    // traceDirectionalLight (rchit:204-223)
    const kf_dir_light& d = s.dl;
    V3<T> em = V3<T>(T(d.rgbs[0]), T(d.rgbs[1]), T(d.rgbs[2])) * T(d.rgbs[3]);
    if (all_zero(em)) return V3<T>(T(0));
    V3<T> L = -V3<T>(T(d.direction[0]), T(d.direction[1]), T(d.direction[2]));
    (void)rayDir;
    if (d.direction[3] != 0.0f) {
        T a = T(rnd(seed)), b = T(rnd(seed)), cc = T(rnd(seed));
        L = normalize(L + T(d.direction[3]) * V3<T>(a, b, cc));
    }
    return shadow_ray(s, sh, L, T(1e6f), em, seed, c);
}

template <class T> V3<T> point_lights(const ko_scene& s, const Shade<T>& sh, uint32_t& seed, Counts& c) {
    // tracePointLights (rchit:227-258)
    V3<T> ret(T(0));
    for (int i = 0; i < KF_MAX_POINT_LIGHTS; ++i) {
        if (!(s.pl.rgbs[i][3] > 0.0f)) continue;
        V3<T> rgb(T(s.pl.rgbs[i][0]), T(s.pl.rgbs[i][1]), T(s.pl.rgbs[i][2]));
        T lum = length(rgb * T(s.pl.rgbs[i][3]));
        if (lum == T(0)) continue;
        V3<T> lpos(T(s.pl.posr[i][0]), T(s.pl.posr[i][1]), T(s.pl.posr[i][2]));
        if (s.pl.posr[i][3] != 0.0f) {
            V3<T> pert = uniform_sphere<T>(seed);
            lpos = lpos + T(s.pl.posr[i][3]) * normalize(pert);
        }
        V3<T> L = lpos - sh.worldPos;
        T d = length(L);
        L = normalize(L);
        V3<T> em = rgb * T(s.pl.rgbs[i][3]) / d / d;
        ret = ret + shadow_ray(s, sh, L, d, em, seed, c);
    }
    return ret;
}

template <class T> V3<T> active_lights(const ko_scene& s, const Shade<T>& sh, uint32_t& seed, Counts& c) {
    // traceActiveLights (rchit:262-316)
    V3<T> ret(T(0));
    const kf_active_lights& a = s.al;
    for (int i = 0; i < KF_MAX_ACTIVE_LIGHTS; ++i) {
        if (!(a.front[i][3] > 0.0f)) continue;
        V3<T> rgb(T(a.rgbs[i][0]), T(a.rgbs[i][1]), T(a.rgbs[i][2]));
        T lum = length(rgb * T(a.rgbs[i][3]));
        if (lum == T(0)) continue;
        T softness = T(a.sftp[i][0]);
        V3<T> lpos(T(a.position[i][0]), T(a.position[i][1]), T(a.position[i][2]));
        if (softness != T(0)) {
            V3<T> pert = uniform_sphere<T>(seed);
            lpos = lpos + softness * normalize(pert);
        }
        V3<T> L = lpos - sh.worldPos;
        T d = length(L);
        L = normalize(L);
        T fov = T(a.sftp[i][1]);
        V3<T> ldir = normalize(V3<T>(T(a.front[i][0]), T(a.front[i][1]), T(a.front[i][2])));
        T half = gclamp(fov, T(0), T(kPi)) / T(2);
        T cs = dot(ldir, -L);
        if (cs > Fn<T>::cos(half)) {
            int texID = (int)a.sftp[i][2];
            V3<T> color = rgb;
            if (texID >= 0 && texID < (int)s.tex.size()) {
                // texCoord = proj * view * vec4(worldPos, 1): (proj * view) first.
                float pv[16];
                for (int cc = 0; cc < 4; ++cc)
                    for (int r = 0; r < 4; ++r) {
                        float acc = a.proj[i][0 * 4 + r] * a.view[i][cc * 4 + 0] + a.proj[i][1 * 4 + r] * a.view[i][cc * 4 + 1] +
                                    a.proj[i][2 * 4 + r] * a.view[i][cc * 4 + 2] + a.proj[i][3 * 4 + r] * a.view[i][cc * 4 + 3];
                        pv[cc * 4 + r] = acc;
                    }
                T v4[4] = {sh.worldPos.x, sh.worldPos.y, sh.worldPos.z, T(1)};
                T tc[4];
                mat4_mul_vec4(pv, v4, tc);
                T u = (tc[0] / tc[3]) * T(0.5) + T(0.5);
                T v = (tc[1] / tc[3]) * T(0.5) + T(0.5);
                T o[4];
                tex_sample(s.tex[texID], u, v, true, o);
                color = color * V3<T>(o[0], o[1], o[2]);
            }
            V3<T> em = color * T(a.rgbs[i][3]) / d / d;
            ret = ret + shadow_ray(s, sh, L, d, em, seed, c);
        }
    }
    return ret;
}

template <class T> struct Payload {
    V3<T> dir, albedo, normal, emission, origin, weight, shadow;
    uint32_t seed;
    uint32_t depth;
};

// closest hit (rchit:322-469) for a hit, or miss (rmiss:12-31).
template <class T>
void shade_segment(const ko_scene& s, const kf_push_constants& pc, const Hit<T>& h, bool hit, Payload<T>& ray,
                   Counts& cnt) {
    uint32_t maxDepth = pc.path_depth;
    if (!hit) {
        ray.emission = miss_emission(s, pc, ray.dir);
        ray.depth = maxDepth + 1;
        return;
    }
    uint32_t g = s.inst[h.inst].geometry_index;
    size_t tri = (size_t)s.toff[g] + h.prim;
    const kf_vertex& v0 = s.verts[s.voff[g] + s.idx[3 * tri + 0]];
    const kf_vertex& v1 = s.verts[s.voff[g] + s.idx[3 * tri + 1]];
    const kf_vertex& v2 = s.verts[s.voff[g] + s.idx[3 * tri + 2]];
    T bx = T(1) - h.b1 - h.b2, by = h.b1, bz = h.b2;
    V3<T> n0(T(v0.normal[0]), T(v0.normal[1]), T(v0.normal[2]));
    V3<T> n1(T(v1.normal[0]), T(v1.normal[1]), T(v1.normal[2]));
    V3<T> n2(T(v2.normal[0]), T(v2.normal[1]), T(v2.normal[2]));
    V3<T> ln = n0 * bx + n1 * by + n2 * bz;
    const float* w = &s.w2o[12 * h.inst];
    // localNormal * gl_WorldToObjectEXT: component c = sum_r n[r] * W2O[r][c]
    V3<T> wn(ln.x * T(w[0]) + ln.y * T(w[4]) + ln.z * T(w[8]),
             ln.x * T(w[1]) + ln.y * T(w[5]) + ln.z * T(w[9]),
             ln.x * T(w[2]) + ln.y * T(w[6]) + ln.z * T(w[10]));
    V3<T> N = normalize(wn);
    V3<T> worldPos = ray.origin + ray.dir * h.t;
    T uvx = T(v0.uv[0]) * bx + T(v1.uv[0]) * by + T(v2.uv[0]) * bz;
    T uvy = T(v0.uv[1]) * bx + T(v1.uv[1]) * by + T(v2.uv[1]) * bz;
    const kf_material& m = s.mats[s.matidx[tri]];

    V3<T> emission = V3<T>(T(m.emission[0]), T(m.emission[1]), T(m.emission[2])) * T(m.emission[3]);
    if (!all_zero(emission)) {
        ray.depth = maxDepth + 1;
        ray.emission = emission;
        return;
    }
    auto texel = [&](int id, T out[4]) { tex_sample(s.tex[id], uvx, uvy, true, out); };
    auto valid_tex = [&](int id) { return id >= 0 && id < (int)s.tex.size(); };
    V3<T> base(T(m.diffuse[0]), T(m.diffuse[1]), T(m.diffuse[2]));
    if (valid_tex(m.diffuse_tex)) { T o[4]; texel(m.diffuse_tex, o); base = V3<T>(o[0], o[1], o[2]); }
    base = base / T(kPi);
    T metallic = T(m.metallic);
    if (valid_tex(m.metallic_tex)) { T o[4]; texel(m.metallic_tex, o); metallic = o[0]; }
    T roughness = T(m.roughness);
    if (valid_tex(m.roughness_tex)) { T o[4]; texel(m.roughness_tex, o); roughness = o[0]; }
    T a2 = roughness * roughness;
    T transmission = T(m.transmission);
    if (valid_tex(m.transmission_tex)) { T o[4]; texel(m.transmission_tex, o); transmission = o[0]; }

    T f = gmax(T(m.ior), T(1e-5f));
    T diffuse_weight = (T(1) - gclamp(metallic, T(0), T(1))) * (T(1) - gclamp(transmission, T(0), T(1)));
    T final_transmission = gclamp(transmission, T(0), T(1)) * (T(1) - gclamp(metallic, T(0), T(1)));
    T specular_weight = T(1) - final_transmission;
    V3<T> weight(T(0));
    V3<T> V = normalize(-ray.dir);
    bool inside = h.back != 0;
    N = inside ? -N : N;
    V3<T> L(T(0));
    T NdotV = dot(N, V);
    V3<T> diffuseColor = diffuse_weight * base;
    V3<T> specularColor = specular_weight * (base * metallic + (V3<T>(T(m.specular) * T(0.08f))) * (T(1) - metallic));
    V3<T> transmissionColor = transmission * base;
    T dl = length(diffuseColor), sl = length(specularColor);
    T pD = dl / (dl + sl);
    if (dl == T(0) && sl == T(0)) {
        if (all_zero(base)) {
            if (diffuse_weight == T(1)) pD = T(1);
            else if (diffuse_weight == T(0)) pD = T(0);
            else pD = T(0.5);
        } else {
            pD = T(0);
        }
    } else {
        pD *= specular_weight;
    }
    bool chooseDiffuse = T(rnd(ray.seed)) < pD;
    if (chooseDiffuse) {
        L = cosine_hemisphere(ray.seed, N);
        T NdotL = gclamp(dot(N, L), T(0), T(1));
        weight = T(kPi) * diffuseColor * NdotL / pD;
    } else {
        if (final_transmission == T(0)) {
            V3<T> H = sample_ggx(ray.seed, a2, N);
            T HdotV = dot(H, V);
            L = T(2) * HdotV * H - V;
            T NoV = gmax(dot(N, V), T(1e-7f));
            T NoL = dot(N, L);
            T NoH = gmax(dot(N, H), T(1e-7f));
            T VoH = gmax(dot(V, H), T(1e-7f));
            if (NoL >= T(0)) {
                T G = geom_shadow(NoV, NoL, a2);
                V3<T> F = schlick_fresnel(specularColor, VoH);
                weight = T(kPi) * F * G * VoH / (NoH * NoV * (T(1) - pD));
            } else {
                weight = V3<T>(T(0));
            }
        } else {
            T ior = inside ? T(1) / f : f;
            T dd = inside ? NdotV * ior : NdotV;
            V3<T> refr = refract(-V, N, T(1) / f);
            T reflectProb = !all_zero(refr) ? schlick(dd, f) : T(1);
            if (T(rnd(ray.seed)) >= reflectProb) {
                L = refr;
                weight = T(kPi) * transmissionColor / (T(1) - pD);
            } else {
                L = reflect(-V, N);
                weight = T(kPi) * transmissionColor / (T(1) - pD);
            }
        }
    }
    Shade<T> sh;
    sh.N = N; sh.V = V; sh.worldPos = worldPos; sh.f = f; sh.a2 = a2;
    sh.diffuseColor = diffuseColor; sh.specularColor = specularColor; sh.transmissionColor = transmissionColor;
    sh.inside = inside;
    V3<T> d_ = dir_light(s, sh, ray.dir, ray.seed, cnt);
    V3<T> p_ = point_lights(s, sh, ray.seed, cnt);
    V3<T> a_ = active_lights(s, sh, ray.seed, cnt);
    ray.shadow = d_ + p_ + a_;
    ray.origin = worldPos;
    ray.dir = L;
    ray.emission = V3<T>(T(0));
    ray.weight = weight;
    ray.albedo = base;
    ray.normal = N;
}

// One pixel of PathTrace.rgen:19-167 (all samples).  Seeds: clockARB() is
// replaced by S0 = tea(seed, frameCount): jitter stream tea(pixel, S0),
// sample i stream tea(pixel, S0 + 1 + i).
template <class T>
void render_pixel(const ko_scene& s, const kf_camera_ubo& cam, const kf_push_constants& pc, uint32_t S0, uint32_t x,
                  uint32_t y, uint32_t W, uint32_t H, float* rgba, float* albedo_out, float* normal_out,
                  float* samples, Counts& cnt) {
    uint32_t pixel = y * W + x;
    uint32_t seed = tea(pixel, S0);
    V3<T> colors(T(0)), albedo(T(0)), normal(T(0));
    uint32_t maxDepth = pc.path_depth;
    for (uint32_t i = 0; i < pc.spp; ++i) {
        Payload<T> ray;
        ray.seed = tea(pixel, S0 + 1u + i);
        const uint32_t sample_seed = ray.seed;
        T r0 = T(rnd(seed));
        T r1 = T(rnd(seed));
        T posx = T(float(x)) + r0, posy = T(float(y)) + r1;
        T npx = posx / T(float(W)), npy = posy / T(float(H));
        T dx = npx * T(2) - T(1), dy = npy * T(2) - T(1);
        T aperture = T(cam.position[3]);
        T focus = T(cam.front[3]);
        T ox, oy;
        disk(ray.seed, ox, oy);
        T half = aperture / T(2);
        ox = half * ox;
        oy = half * oy;
        T tv[4] = {dx, dy, T(1), T(1)};
        T target[4];
        mat4_mul_vec4(cam.proj_inverse, tv, target);
        T org[4], dir4[4];
        if (aperture > T(0)) {
            T o4[4] = {ox, oy, T(0), T(1)};
            mat4_mul_vec4(cam.view_inverse, o4, org);
            V3<T> tgt = V3<T>(target[0], target[1], target[2]) * focus - V3<T>(ox, oy, T(0));
            V3<T> nd = normalize(tgt);
            T d4[4] = {nd.x, nd.y, nd.z, T(0)};
            mat4_mul_vec4(cam.view_inverse, d4, dir4);
        } else {
            T o4[4] = {T(0), T(0), T(0), T(1)};
            mat4_mul_vec4(cam.view_inverse, o4, org);
            V3<T> nd = normalize(V3<T>(target[0], target[1], target[2]));
            T d4[4] = {nd.x, nd.y, nd.z, T(0)};
            mat4_mul_vec4(cam.view_inverse, d4, dir4);
        }
        ray.dir = V3<T>(dir4[0], dir4[1], dir4[2]);
        ray.origin = V3<T>(org[0], org[1], org[2]);
        ray.weight = V3<T>(T(0));
        ray.emission = V3<T>(T(1));
        ray.albedo = V3<T>(T(0));
        ray.normal = V3<T>(T(0));
        ray.shadow = V3<T>(T(0));
        V3<T> weight(T(1)), color(T(0));
        for (ray.depth = 0; ray.depth <= maxDepth; ++ray.depth) {
            if (ray.depth == 0) cnt.primary++; else cnt.extension++;
            Hit<T> h;
            // alpha coin seed: tea(initial seed of the sample's stream, depth),
            // independent of how many numbers the path has drawn so far
            bool hit = trace(s, ray.origin, ray.dir, T(0.001f), T(10000.0f), h, false, true,
                             tea(sample_seed, ray.depth), false);
            shade_segment(s, pc, h, hit, ray, cnt);
            color = color + ray.emission * weight;
            weight = weight * ray.weight;
            color = color + ray.shadow * weight;
            ray.shadow = V3<T>(T(0));
            if (i == 0 && ray.depth == 0) {
                albedo = ray.albedo;
                normal = ray.normal;
            }
            if (all_zero(weight)) break;
            if (pc.russian_roulette && ray.depth >= pc.rr_min_bounces) {
                T p = gmax(weight.x, gmax(weight.y, weight.z));
                T rr = T(rnd(ray.seed));
                if (rr > p) break;
                weight = weight * (T(1) / p);
            }
        }
        colors = colors + color;
        if (samples) {
            float* o = samples + 3 * ((size_t)i * W * H + pixel);
            o[0] = (float)color.x; o[1] = (float)color.y; o[2] = (float)color.z;
        }
    }
    V3<T> fc = colors / T(float(pc.spp));
    float* px = rgba + 4 * (size_t)pixel;
    if (pc.frame_count <= 0) {
        px[0] = (float)fc.x; px[1] = (float)fc.y; px[2] = (float)fc.z; px[3] = 1.0f;
    } else {
        // mix(old, new, a) = old * (1 - a) + new * a, a = 1 / (frameCount + 1)
        float a = 1.0f / float(pc.frame_count + 1);
        T ta = T(a);
        V3<T> old = V3<T>(T(px[0]), T(px[1]), T(px[2]));
        V3<T> m = old * (T(1) - ta) + fc * ta;
        px[0] = (float)m.x; px[1] = (float)m.y; px[2] = (float)m.z; px[3] = 1.0f;
    }
    float* al = albedo_out + 4 * (size_t)pixel;
    al[0] = (float)albedo.x; al[1] = (float)albedo.y; al[2] = (float)albedo.z; al[3] = 1.0f;
    float* nn = normal_out + 4 * (size_t)pixel;
    nn[0] = (float)normal.x; nn[1] = (float)normal.y; nn[2] = (float)normal.z; nn[3] = 1.0f;
}

}  // namespace

namespace {
template <class T> V3<T> v3_at(const double* p) { return V3<T>(T(p[0]), T(p[1]), T(p[2])); }
template <class T> void put3(double* o, V3<T> v) { o[0] = (double)v.x; o[1] = (double)v.y; o[2] = (double)v.z; }

template <class T> int shading_term(int term, const double* in, uint32_t n, double* out) {
    static const int ins[] = {2, 2, 2, 3, 2, 7, 6, 6, 3};
    static const int outs[] = {1, 1, 1, 1, 1, 3, 3, 3, 3};
    if (term < 0 || term > KO_PERPENDICULAR) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        const double* a = in + (size_t)i * ins[term];
        double* o = out + (size_t)i * outs[term];
        switch (term) {
        case KO_SCHLICK: o[0] = (double)schlick(T(a[0]), T(a[1])); break;
        case KO_GGX_D: o[0] = (double)ggx_d(T(a[0]), T(a[1])); break;
        case KO_G1: o[0] = (double)g1(T(a[0]), T(a[1])); break;
        case KO_GEOM_SHADOW: o[0] = (double)geom_shadow(T(a[0]), T(a[1]), T(a[2])); break;
        case KO_SCHLICK_FRESNEL: o[0] = (double)schlick_fresnel(V3<T>(T(a[0])), T(a[1])).x; break;
        case KO_REFRACT: put3(o, refract(v3_at<T>(a), v3_at<T>(a + 3), T(a[6]))); break;
        case KO_REFLECT: put3(o, reflect(v3_at<T>(a), v3_at<T>(a + 3))); break;
        case KO_LOCAL_TO_WORLD: put3(o, local_to_world(v3_at<T>(a), v3_at<T>(a + 3))); break;
        case KO_PERPENDICULAR: put3(o, perpendicular(v3_at<T>(a))); break;
        }
    }
    return 0;
}

template <class T> int sample_n(int sampler, uint32_t& seed, const double* param, uint32_t n, double* out) {
    for (uint32_t i = 0; i < n; ++i) {
        switch (sampler) {
        case KO_SAMPLE_COSINE: put3(out + 3 * (size_t)i, cosine_hemisphere(seed, v3_at<T>(param))); break;
        case KO_SAMPLE_GGX: put3(out + 3 * (size_t)i, sample_ggx(seed, T(param[0]), v3_at<T>(param + 1))); break;
        case KO_SAMPLE_SPHERE: put3(out + 3 * (size_t)i, uniform_sphere<T>(seed)); break;
        case KO_SAMPLE_DISK: {
            T x, y;
            disk(seed, x, y);
            out[2 * (size_t)i] = (double)x;
            out[2 * (size_t)i + 1] = (double)y;
            break;
        }
        default: return -1;
        }
    }
    return 0;
}

template <class T> void direct_one(const double* a, uint32_t& seed, double* out) {
    Shade<T> sh;
    sh.N = v3_at<T>(a);
    sh.V = v3_at<T>(a + 3);
    V3<T> L = v3_at<T>(a + 6);
    sh.f = T(a[9]);
    sh.a2 = T(a[10]);
    sh.diffuseColor = v3_at<T>(a + 11);
    sh.specularColor = v3_at<T>(a + 14);
    sh.transmissionColor = v3_at<T>(a + 17);
    sh.inside = a[20] != 0.0;
    sh.worldPos = V3<T>(T(0));
    put3(out, direct(sh, L, v3_at<T>(a + 21), seed));
}
}  // namespace

// ======================================================================== C ABI
extern "C" {

void ko_tea_batch(const uint32_t* a, const uint32_t* b, uint32_t n, uint32_t* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = tea(a[i], b[i]);
}
void ko_rnd_sequence(uint32_t seed, uint32_t n, float* out, uint32_t* final_state) {
    for (uint32_t i = 0; i < n; ++i) out[i] = rnd(seed);
    if (final_state) *final_state = seed;
}
void ko_sincos_batch(const float* x, uint32_t n, float* s, float* c) {
    for (uint32_t i = 0; i < n; ++i) sincos_f(x[i], &s[i], &c[i]);
}
void ko_exp2_batch(const float* x, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = exp2_f(x[i]);
}

int ko_intersect_triangle(const float org[3], const float dir[3], float tmin, float tmax, const float v0[3],
                          const float v1[3], const float v2[3], float* t, float* b1, float* b2, uint32_t* back) {
    Pre<float> p = precompute(V3<float>(org[0], org[1], org[2]), V3<float>(dir[0], dir[1], dir[2]));
    TriHit<float> h;
    if (!intersect(p, v0, v1, v2, tmin, tmax, h)) return 0;
    *t = h.t; *b1 = h.b1; *b2 = h.b2; *back = h.back;
    return 1;
}

ko_scene* ko_scene_create(const kf_scene_desc* d) {
    ko_scene* s = new ko_scene();
    s->ngeom = d->num_geometries;
    s->voff.assign(d->geom_vertex_offset, d->geom_vertex_offset + d->num_geometries + 1);
    s->toff.assign(d->geom_tri_offset, d->geom_tri_offset + d->num_geometries + 1);
    s->gflags.assign(d->geom_flags, d->geom_flags + d->num_geometries);
    uint32_t nv = s->voff[s->ngeom], nt = s->toff[s->ngeom];
    s->verts.assign(d->vertices, d->vertices + nv);
    s->idx.assign(d->indices, d->indices + 3 * (size_t)nt);
    s->matidx.assign(d->mat_index, d->mat_index + nt);
    s->inst.assign(d->instances, d->instances + d->num_instances);
    s->mats.assign(d->materials, d->materials + d->num_materials);
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        Tex t;
        t.w = d->textures[i].width;
        t.h = d->textures[i].height;
        t.px.assign(d->textures[i].rgba8, d->textures[i].rgba8 + 4 * (size_t)t.w * t.h);
        s->tex.push_back(std::move(t));
    }
    std::memset(&s->dl, 0, sizeof(s->dl));
    std::memset(&s->pl, 0, sizeof(s->pl));
    std::memset(&s->al, 0, sizeof(s->al));
    if (d->dir_light) s->dl = *d->dir_light;
    if (d->point_lights) s->pl = *d->point_lights;
    if (d->active_lights) s->al = *d->active_lights;
    s->env_kind = d->env_kind;
    if (d->env_kind == KF_ENV_CUBEMAP && d->env_faces)
        for (int f = 0; f < 6; ++f) {
            s->env[f].w = d->env_faces[f].width;
            s->env[f].h = d->env_faces[f].height;
            s->env[f].px.assign(d->env_faces[f].rgba8, d->env_faces[f].rgba8 + 4 * (size_t)s->env[f].w * s->env[f].h);
        }
    for (int i = 0; i < 8; ++i) s->env_grad[i] = d->env_gradient[i];

    // BLAS per geometry (object space)
    s->blas.resize(s->ngeom);
    std::vector<Box> geom_box(s->ngeom);
    for (uint32_t g = 0; g < s->ngeom; ++g) {
        uint32_t n = s->toff[g + 1] - s->toff[g];
        std::vector<Box> boxes(n);
        for (uint32_t p = 0; p < n; ++p)
            for (int k = 0; k < 3; ++k) boxes[p].grow(tri_vertex(*s, g, p, k));
        for (auto& b : boxes) geom_box[g].grow(b);
        build_bvh(boxes, s->blas[g]);
    }
    // instances: world->object, padded world AABB, TLAS
    uint32_t ni = d->num_instances;
    s->w2o.resize(12 * (size_t)ni);
    s->inst_box.resize(ni);
    s->inst_valid.resize(ni);
    for (uint32_t i = 0; i < ni; ++i) {
        ko_world_to_object(s->inst[i].transform, &s->w2o[12 * i]);
        uint32_t g = s->inst[i].geometry_index;
        const Box& gb = geom_box[g];
        bool valid = g < s->ngeom && !(s->gflags[g] & KF_GEOM_HIDDEN) && s->toff[g + 1] > s->toff[g];
        s->inst_valid[i] = valid;
        Box wb;
        if (valid) {
            const float* t = s->inst[i].transform;
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int c = 0; c < 8; ++c) {
                double p[3] = {(c & 1) ? gb.hi[0] : gb.lo[0], (c & 2) ? gb.hi[1] : gb.lo[1], (c & 4) ? gb.hi[2] : gb.lo[2]};
                for (int r = 0; r < 3; ++r) {
                    double v = (double)t[0 * 4 + r] * p[0] + (double)t[1 * 4 + r] * p[1] + (double)t[2 * 4 + r] * p[2] +
                               (double)t[3 * 4 + r];
                    lo[r] = std::min(lo[r], v);
                    hi[r] = std::max(hi[r], v);
                }
            }
            double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
            for (int r = 0; r < 3; ++r) {
                double pad = 1e-5 * (ext + std::fabs(lo[r]) + std::fabs(hi[r])) + 1e-6;
                wb.lo[r] = std::nextafter((float)(lo[r] - pad), -INFINITY);
                wb.hi[r] = std::nextafter((float)(hi[r] + pad), INFINITY);
            }
        }
        s->inst_box[i] = wb;
    }
    std::vector<Box> ib(ni);
    for (uint32_t i = 0; i < ni; ++i) {
        ib[i] = s->inst_box[i];
        if (!s->inst_valid[i]) { for (int r = 0; r < 3; ++r) { ib[i].lo[r] = 0; ib[i].hi[r] = 0; } }
    }
    build_bvh(ib, s->tlas);
    return s;
}

void ko_scene_destroy(ko_scene* s) { delete s; }

static void set_threads(int threads) {
    if (threads > 0) omp_set_num_threads(threads);
}

void ko_trace_closest(const ko_scene* s, const kf_ray* rays, uint32_t n, kf_hit* hits, int brute, int threads) {
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < (long)n; ++i) {
        const kf_ray& r = rays[i];
        Hit<float> h;
        bool ok = trace(*s, V3<float>(r.org[0], r.org[1], r.org[2]), V3<float>(r.dir[0], r.dir[1], r.dir[2]), r.tmin,
                        r.tmax, h, false, false, 0u, brute != 0);
        kf_hit& o = hits[i];
        if (ok) { o.t = h.t; o.inst = h.inst; o.prim = h.prim; o.b1 = h.b1; o.b2 = h.b2; o.back = h.back; }
        else { o.t = r.tmax; o.inst = ~0u; o.prim = ~0u; o.b1 = 0; o.b2 = 0; o.back = 0; }
    }
}

void ko_trace_any(const ko_scene* s, const kf_ray* rays, uint32_t n, uint8_t* occ, int brute, int threads) {
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < (long)n; ++i) {
        const kf_ray& r = rays[i];
        Hit<float> h;
        occ[i] = trace(*s, V3<float>(r.org[0], r.org[1], r.org[2]), V3<float>(r.dir[0], r.dir[1], r.dir[2]), r.tmin,
                       r.tmax, h, true, false, 0u, brute != 0) ? 1 : 0;
    }
}

int ko_render(const ko_scene* s, const kf_camera_ubo* cam, const kf_push_constants* pc, uint32_t seed,
              uint32_t width, uint32_t height, uint32_t y0, uint32_t y1, int precision, int threads, float* rgba,
              float* albedo, float* normal, float* samples, uint64_t* counts) {
    return ko_render_region(s, cam, pc, seed, width, height, 0, width, y0, y1, precision, threads, rgba, albedo,
                            normal, samples, counts);
}

int ko_render_region(const ko_scene* s, const kf_camera_ubo* cam, const kf_push_constants* pc, uint32_t seed,
                     uint32_t width, uint32_t height, uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1,
                     int precision, int threads, float* rgba, float* albedo, float* normal, float* samples,
                     uint64_t* counts) {
    if (!s || !cam || !pc || !rgba || !albedo || !normal || y1 > height || y0 > y1 || x1 > width || x0 > x1)
        return -1;
    if (precision != 32 && precision != 64) return -1;
    set_threads(threads);
    uint32_t S0 = tea(seed, (uint32_t)pc->frame_count);
    uint64_t c0 = 0, c1 = 0, c2 = 0;
    const uint32_t rw = x1 - x0;
    long npx = (long)(y1 - y0) * rw;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : c0, c1, c2)
    for (long k = 0; k < npx; ++k) {
        uint32_t y = y0 + (uint32_t)(k / rw), x = x0 + (uint32_t)(k % rw);
        Counts c;
        if (precision == 32)
            render_pixel<float>(*s, *cam, *pc, S0, x, y, width, height, rgba, albedo, normal, samples, c);
        else
            render_pixel<double>(*s, *cam, *pc, S0, x, y, width, height, rgba, albedo, normal, samples, c);
        c0 += c.primary; c1 += c.extension; c2 += c.shadow;
    }
    if (counts) { counts[0] = c0; counts[1] = c1; counts[2] = c2; }
    return 0;
}

double ko_relmse(const float* x, const float* y, uint32_t n, uint32_t sx, uint32_t sy) {
    double acc = 0;
    for (uint32_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            double a = x[(size_t)i * sx + c], b = y[(size_t)i * sy + c];
            acc += (a - b) * (a - b) / (b * b + 1e-2);
        }
    return acc / (3.0 * n);
}

int ko_shading_term(int term, int precision, const double* in, uint32_t n, double* out) {
    if (precision == 32) return shading_term<float>(term, in, n, out);
    if (precision == 64) return shading_term<double>(term, in, n, out);
    return -1;
}

int ko_sample(int sampler, int precision, uint32_t seed, const double* param, uint32_t n, double* out,
              uint64_t* draws) {
    const uint32_t start = seed;
    int rc = precision == 32 ? sample_n<float>(sampler, seed, param, n, out)
                             : precision == 64 ? sample_n<double>(sampler, seed, param, n, out) : -1;
    if (rc == 0 && draws) {
        // Count LCG steps from the start state to the final one (the LCG has
        // period 2^32, so the count is exact below that).
        uint64_t k = 0;
        for (uint32_t s = start; s != seed; ++k) lcg(s);
        *draws = k;
    }
    return rc;
}

int ko_direct(int precision, const double* shade, uint32_t* seed, double* out) {
    if (!shade || !seed || !out) return -1;
    if (precision == 32) direct_one<float>(shade, *seed, out);
    else if (precision == 64) direct_one<double>(shade, *seed, out);
    else return -1;
    return 0;
}

void ko_to_bgra8_srgb(const float* rgba, uint32_t n, uint8_t* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = rgba + 4 * (size_t)i;
        out[4 * i + 0] = encode_srgb8(p[2]);
        out[4 * i + 1] = encode_srgb8(p[1]);
        out[4 * i + 2] = encode_srgb8(p[0]);
        out[4 * i + 3] = encode_unorm8(p[3]);
    }
}

}  // extern "C"
