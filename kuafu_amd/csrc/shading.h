// shading.h — closest-hit / miss / NEE shading of PathTrace.rchit and
// PathTrace.rmiss, restated for gfx950 with the operation order of the
// GLSL source (compiled with -ffp-contract=off).
#pragma once
#include "kf_internal.h"
#include "kf_math.h"

namespace kf {

// Light data uploaded per frame (Scene::uploadLightBuffers, scene.cpp:252-292)
// plus the compact list of NEE slots in evaluation order: directional light,
// point lights 0..31, active lights 0..7 (rchit:457-460).
constexpr int kMaxSlots = 1 + KF_MAX_POINT_LIGHTS + KF_MAX_ACTIVE_LIGHTS;
struct DevLights {
    kf_dir_light dl;
    kf_point_lights pl;
    kf_active_lights al;
    float pv[KF_MAX_ACTIVE_LIGHTS][16];  // proj * view per active light
    int n_slots;
    int kind[kMaxSlots];  // 0 directional, 1 point, 2 active
    int index[kMaxSlots];
};

// ------------------------------------------------------------------ textures
// Sampler of vkCore.hpp:619-637 (linear, repeat, single mip): RGBA8 sRGB
// texels decoded through a LUT, bilinear at texel centres.
KF_DEV float4 tex_sample(const SceneView& S, const uint8_t* base, uint4 info, float u, float v, bool repeat) {
    uint32_t w = info.y, h = info.z;
    if (w == 0 || h == 0) return make_float4(0, 0, 0, 0);
    float x = u * float(w) - 0.5f, y = v * float(h) - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float fx = x - fx0, fy = y - fy0;
    long x0 = (long)fx0, y0 = (long)fy0;
    long xs[2], ys[2];
    for (int k = 0; k < 2; ++k) {
        long xi = x0 + k, yi = y0 + k;
        if (repeat) {
            long r = xi % (long)w;
            xs[k] = r < 0 ? r + w : r;
            r = yi % (long)h;
            ys[k] = r < 0 ? r + h : r;
        } else {
            xs[k] = xi < 0 ? 0 : (xi >= (long)w ? (long)w - 1 : xi);
            ys[k] = yi < 0 ? 0 : (yi >= (long)h ? (long)h - 1 : yi);
        }
    }
    float c[2][2][4];
    for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
            const uint8_t* px = base + info.x + 4 * ((size_t)ys[j] * w + xs[i]);
            uchar4 t = *(const uchar4*)px;
            c[j][i][0] = S.srgb_lut[t.x];
            c[j][i][1] = S.srgb_lut[t.y];
            c[j][i][2] = S.srgb_lut[t.z];
            c[j][i][3] = float(t.w) / 255.0f;
        }
    float o[4];
    for (int k = 0; k < 4; ++k) {
        float a = c[0][0][k] * (1.0f - fx) + c[0][1][k] * fx;
        float b = c[1][0][k] * (1.0f - fx) + c[1][1][k] * fx;
        o[k] = a * (1.0f - fy) + b * fy;
    }
    return make_float4(o[0], o[1], o[2], o[3]);
}

// Vulkan cube-map face selection, clamp-to-edge, per-face bilinear.
KF_DEV v3 cube_sample(const SceneView& S, v3 r) {
    float ax = fabsf(r.x), ay = fabsf(r.y), az = fabsf(r.z);
    int face;
    float sc, tc, ma;
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
    float u = 0.5f * (sc / ma + 1.0f), v = 0.5f * (tc / ma + 1.0f);
    float4 o = tex_sample(S, S.tex_data, S.env_info[face], u, v, false);
    return mk(o.x, o.y, o.z);
}

// Miss shader (rmiss:12-31): lookup direction (-d.y, d.z, -d.x).
KF_DEV v3 miss_emission(const SceneView& S, const kf_push_constants& pc, v3 d) {
    if (pc.use_env_map) {
        v3 dir = mk(-d.y, d.z, -d.x);
        if (S.env_kind == KF_ENV_CUBEMAP) return cube_sample(S, dir);
        if (S.env_kind == KF_ENV_GRADIENT) {
            float t = gclamp(0.5f * dir.y + 0.5f, 0.0f, 1.0f);
            v3 lo = mk(S.env_grad[0], S.env_grad[1], S.env_grad[2]);
            v3 hi = mk(S.env_grad[4], S.env_grad[5], S.env_grad[6]);
            return lo * (1.0f - t) + hi * t;
        }
        return splat(0.0f);
    }
    return mk(pc.clear_color[0], pc.clear_color[1], pc.clear_color[2]) * pc.clear_color[3];
}

// ------------------------------------------------------------------ BSDF
KF_DEV float schlick(float cosine, float ior) {  // Random.glsl:68-73, pow(x,5) = x^2 x^2 x
    float r0 = (1.0f - ior) / (1.0f + ior);
    r0 *= r0;
    float x = 1.0f - cosine;
    float x2 = x * x;
    return r0 + (1.0f - r0) * (x2 * x2 * x);
}
KF_DEV float ggx_d(float NdotH, float a2) {  // Random.glsl:75-79
    float d = gmax(NdotH * NdotH * (a2 - 1.0f) + 1.0f, 1e-6f);
    return a2 / (d * d * kPi);
}
KF_DEV float g1(float x, float a2) { return (2.0f * x) / (x + sqrtf(a2 + (1.0f - a2) * (x * x))); }
KF_DEV float geom_shadow(float NdotV, float NdotL, float a2) { return g1(NdotL, a2) * g1(NdotV, a2); }
KF_DEV v3 schlick_fresnel(v3 f0, float lDotH) {
    float e = exp2_k((-5.55473f * lDotH - 6.98316f) * lDotH);
    return f0 + (splat(1.0f) - f0) * e;
}
KF_DEV v3 perpendicular(v3 u) {
    v3 a = mk(fabsf(u.x), fabsf(u.y), fabsf(u.z));
    uint32_t xm = ((a.x - a.y) < 0 && (a.x - a.z) < 0) ? 1u : 0u;
    uint32_t ym = (a.y - a.z) < 0 ? (1u ^ xm) : 0u;
    uint32_t zm = 1u ^ (xm | ym);
    return cross(u, mk(float(xm), float(ym), float(zm)));
}
KF_DEV v3 sample_ggx(uint32_t& seed, float a2, v3 N) {
    float rx = rnd(seed);
    float ry = rnd(seed);
    v3 B = perpendicular(N);
    v3 T = cross(B, N);
    float cosT = sqrtf(gclamp((1.0f - rx) / ((a2 - 1.0f) * rx + 1.0f), 0.0f, 1.0f));
    float sinT = sqrtf(gclamp(1.0f - cosT * cosT, 0.0f, 1.0f));
    float phi = ry * kPi * 2.0f;
    float sp, cp;
    sincos_k(phi, &sp, &cp);
    return T * (sinT * cp) + B * (sinT * sp) + N * cosT;
}
KF_DEV v3 local_to_world(v3 d, v3 n) {
    v3 t;
    if (fabsf(n.x) > fabsf(n.y))
        t = mk(n.z, 0.0f, -n.x) / sqrtf(n.x * n.x + n.z * n.z);
    else
        t = mk(0.0f, -n.z, n.y) / sqrtf(n.y * n.y + n.z * n.z);
    v3 b = cross(n, t);
    return d.x * t + d.y * b + d.z * n;
}
KF_DEV v3 cosine_hemisphere(uint32_t& seed, v3 n) {
    float u0 = rnd(seed);
    float u1 = rnd(seed);
    float sq = sqrtf(1.0f - u1);
    float s, c;
    sincos_k((2.0f * kPi) * u0, &s, &c);
    return local_to_world(mk(c * sq, s * sq, sqrtf(u1)), n);
}
KF_DEV v3 uniform_sphere(uint32_t& seed) {
    v3 p;
    do {
        float a = rnd(seed), b = rnd(seed), c = rnd(seed);
        p = mk(a, b, c) * 2.0f - splat(1.0f);
    } while (dot(p, p) >= 1.0f);
    return p;
}

struct ShadeCtx {
    v3 N, V, worldPos;
    float f, a2;
    v3 diffuseColor, specularColor, transmissionColor;
    bool inside;
};

// calcDirectContribution (rchit:103-171) split at its single random draw:
// the result is (u OP theta) ? ca : cb with OP '<=' (mode 1) or '<'
// (mode 2); mode 0 consumes no draw and yields cb.
KF_DEV void direct_split(const ShadeCtx& sh, v3 L, v3 em, v3& ca, v3& cb, float& theta, int& mode) {
    const v3 N = sh.N, V = sh.V;
    if (!all_zero(sh.transmissionColor)) {
        if (sh.inside) {
            mode = 0;
            theta = 0.0f;
            ca = em * splat(0.0f);
            cb = ca;
            return;
        }
        float NdotV0 = dot(N, V);
        v3 refr = refract(-V, N, 1.0f / sh.f);
        float reflectProb = !all_zero(refr) ? schlick(NdotV0, sh.f) : 1.0f;
        v3 H = normalize(L + V);
        float NdotL = dot(N, L);
        float NdotH = dot(N, H);
        float HdotV = dot(H, V);
        float NdotV = gmax(dot(N, V), 1e-6f);
        float LdotH = dot(L, H);
        float D = ggx_d(NdotH, sh.a2);
        float G = geom_shadow(NdotL, NdotV, sh.a2);
        v3 F = schlick_fresnel(sh.transmissionColor, LdotH);
        v3 w = D * F * G * HdotV / NdotH * NdotV;
        ca = em * w;
        cb = em * splat(0.0f);
        theta = reflectProb;
        mode = 1;
        return;
    }
    v3 H = normalize(L + V);
    float NdotL = dot(N, L);
    float NdotH = dot(N, H);
    float HdotV = dot(H, V);
    float NdotV = gmax(dot(N, V), 1e-6f);
    float LdotH = dot(L, H);
    float D = ggx_d(NdotH, sh.a2);
    float G = geom_shadow(NdotL, NdotV, sh.a2);
    v3 F = schlick_fresnel(sh.specularColor, LdotH);
    float dl = length(sh.diffuseColor);
    float sl = length(sh.specularColor);
    float pD = dl / (dl + sl);
    if (dl == 0.0f && sl == 0.0f) pD = 0.5f;
    v3 dw = sh.diffuseColor * splat(NdotL);
    v3 sw = D * F * G * HdotV / NdotH * NdotV;
    ca = em * (dw * pD);
    cb = em * (sw * (1.0f - pD));
    theta = pD;
    mode = 2;
}

KF_DEV v3 direct_resolve(v3 ca, v3 cb, float theta, int mode, uint32_t& seed) {
    if (mode == 0) return cb;
    float u = rnd(seed);
    bool a = mode == 1 ? (u <= theta) : (u < theta);
    return a ? ca : cb;
}

}  // namespace kf
