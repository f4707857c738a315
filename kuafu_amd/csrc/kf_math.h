// kf_math.h — device arithmetic of the path tracer (gfx950).
//
// The GLSL shaders of the reference (resources/shaders/base/Random.glsl,
// Sampling.glsl, PathTrace.rchit) are restated with a fixed operation order.
// All HIP sources are compiled with -ffp-contract=off, so a*b+c is two
// roundings unless written as fmaf; + - * / sqrt are IEEE-exact on gfx950
// (HIP's default correctly-rounded division and square root).  sin/cos/exp2
// are this file's own: Cody-Waite reduction plus Taylor polynomials in fmaf
// Horner form, so results do not depend on the device math library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define KF_DEV __device__ __forceinline__

namespace kf {

constexpr float kPi = 3.141592f;  // Random.glsl:1 (M_PI)

// ------------------------------------------------------------------ RNG
// tea: Random.glsl:7-21.
KF_DEV uint32_t tea(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
#pragma unroll
    for (int n = 0; n < 16; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
// lcg / rnd: Random.glsl:26-39.
KF_DEV float rnd(uint32_t& s) {
    s = 1664525u * s + 1013904223u;
    return float(s & 0x00FFFFFFu) / float(0x01000000);
}

// ------------------------------------------------------------------ sin/cos/exp2
constexpr float kTwoOverPi = 0.636619772367581343f;
constexpr float kPio2Hi = 1.57079637050628662109375f;
constexpr float kPio2Lo = -4.37113900018624283e-8f;
constexpr float kS1 = (float)(-1.0 / 6.0);
constexpr float kS2 = (float)(1.0 / 120.0);
constexpr float kS3 = (float)(-1.0 / 5040.0);
constexpr float kS4 = (float)(1.0 / 362880.0);
constexpr float kC1 = -0.5f;
constexpr float kC2 = (float)(1.0 / 24.0);
constexpr float kC3 = (float)(-1.0 / 720.0);
constexpr float kC4 = (float)(1.0 / 40320.0);
constexpr float kC5 = (float)(-1.0 / 3628800.0);

KF_DEV void sincos_k(float x, float* s, float* c) {
    float k = rintf(x * kTwoOverPi);
    float r = fmaf(-k, kPio2Hi, x);
    r = fmaf(-k, kPio2Lo, r);
    float r2 = r * r;
    float ps = fmaf(r2, kS4, kS3);
    ps = fmaf(r2, ps, kS2);
    ps = fmaf(r2, ps, kS1);
    float sp = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, kC5, kC4);
    pc = fmaf(r2, pc, kC3);
    pc = fmaf(r2, pc, kC2);
    pc = fmaf(r2, pc, kC1);
    float cp = fmaf(r2, pc, 1.0f);
    int q = ((int)k) & 3;
    float ss = (q & 1) ? cp : sp;
    float cc = (q & 1) ? sp : cp;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}
KF_DEV float sin_k(float x) { float s, c; sincos_k(x, &s, &c); return s; }
KF_DEV float cos_k(float x) { float s, c; sincos_k(x, &s, &c); return c; }

constexpr float kE1 = (float)0.69314718055994530942;
constexpr float kE2 = (float)0.24022650695910071233;
constexpr float kE3 = (float)0.05550410866482157995;
constexpr float kE4 = (float)0.00961812910762847716;
constexpr float kE5 = (float)0.00133335581464284434;
constexpr float kE6 = (float)0.00015403530393381609;
constexpr float kE7 = (float)0.00001525273380405984;
KF_DEV float exp2_k(float y) {
    if (y != y) return y;
    if (y < -150.0f) return 0.0f;
    if (y > 128.0f) return __builtin_inff();
    float n = rintf(y);
    float f = y - n;
    float p = fmaf(f, kE7, kE6);
    p = fmaf(f, p, kE5);
    p = fmaf(f, p, kE4);
    p = fmaf(f, p, kE3);
    p = fmaf(f, p, kE2);
    p = fmaf(f, p, kE1);
    p = fmaf(f, p, 1.0f);
    return ldexpf(p, (int)n);
}

// GLSL min/max/clamp.
KF_DEV float gmin(float x, float y) { return y < x ? y : x; }
KF_DEV float gmax(float x, float y) { return x < y ? y : x; }
KF_DEV float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

// ------------------------------------------------------------------ vec3
struct v3 {
    float x, y, z;
};
KF_DEV v3 mk(float a, float b, float c) { return v3{a, b, c}; }
KF_DEV v3 splat(float a) { return v3{a, a, a}; }
KF_DEV v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
KF_DEV v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
KF_DEV v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
KF_DEV v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
KF_DEV v3 operator*(float s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
KF_DEV v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
KF_DEV v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
KF_DEV bool all_zero(v3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
KF_DEV float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KF_DEV v3 cross(v3 a, v3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
KF_DEV float length(v3 a) { return sqrtf(dot(a, a)); }
KF_DEV v3 normalize(v3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return a * inv;
}
KF_DEV v3 reflect(v3 I, v3 N) { return I - (2.0f * dot(N, I)) * N; }
KF_DEV v3 refract(v3 I, v3 N, float eta) {
    float d = dot(N, I);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return splat(0.0f);
    return eta * I - (eta * d + sqrtf(k)) * N;
}
KF_DEV v3 xyz(float4 a) { return {a.x, a.y, a.z}; }
KF_DEV float4 f4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// mat4 (column-major) times vec4, summed in column order.
KF_DEV void mat4_mul(const float* m, const float v[4], float out[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
        out[r] = m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1] + m[2 * 4 + r] * v[2] + m[3 * 4 + r] * v[3];
}

}  // namespace kf
