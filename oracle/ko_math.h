// ko_math.h — oracle (test infrastructure only; never linked into the product).
//
// Scalar/vector arithmetic of the reference's GLSL path tracer, restated for
// the CPU.  The FP32 flavour uses this file's own sin/cos/exp2 (Cody–Waite
// reduction + Taylor polynomials, evaluated with fmaf), so that a GPU kernel
// written with the same operations in the same order produces the same bits.
// The FP64 flavour uses libm and is the convergence oracle.
//
// Everything here must be compiled with -ffp-contract=off: the op order below
// IS the specification.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace ko {

// ---------------------------------------------------------------- RNG
// tea: Random.glsl:7-21 (TEA, 16 rounds, delta 0x9e3779b9).
inline uint32_t tea(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
    for (int n = 0; n < 16; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
// lcg / rnd: Random.glsl:26-39 (Numerical Recipes LCG, 24-bit mantissa draw).
inline uint32_t lcg(uint32_t& s) {
    s = 1664525u * s + 1013904223u;
    return s & 0x00FFFFFFu;
}
inline float rnd(uint32_t& s) { return float(lcg(s)) / float(0x01000000); }

// ---------------------------------------------------------------- FP32 math
// sin/cos: reduce by k = rint(x * 2/pi) with a two-part pi/2 (fmaf keeps each
// step exactly rounded), then Taylor polynomials on [-pi/4, pi/4] (truncation
// error < 3e-9, far below half an ulp) in Horner form with fmaf.
constexpr float kTwoOverPi = 0.636619772367581343f;
constexpr float kPio2Hi = 1.57079637050628662109375f;      // float(pi/2)
constexpr float kPio2Lo = -4.37113900018624283e-8f;         // pi/2 - kPio2Hi
constexpr float kS1 = (float)(-1.0 / 6.0);
constexpr float kS2 = (float)(1.0 / 120.0);
constexpr float kS3 = (float)(-1.0 / 5040.0);
constexpr float kS4 = (float)(1.0 / 362880.0);
constexpr float kC1 = -0.5f;
constexpr float kC2 = (float)(1.0 / 24.0);
constexpr float kC3 = (float)(-1.0 / 720.0);
constexpr float kC4 = (float)(1.0 / 40320.0);
constexpr float kC5 = (float)(-1.0 / 3628800.0);

inline void sincos_f(float x, float* s, float* c) {
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
    switch (q) {
        case 0: *s = sp; *c = cp; break;
        case 1: *s = cp; *c = -sp; break;
        case 2: *s = -sp; *c = -cp; break;
        default: *s = -cp; *c = sp; break;
    }
}
inline float sin_f(float x) { float s, c; sincos_f(x, &s, &c); return s; }
inline float cos_f(float x) { float s, c; sincos_f(x, &s, &c); return c; }

// exp2: y = n + f with n = rint(y), |f| <= 1/2 (exact split); 2^f by the
// degree-7 Taylor polynomial of e^(f ln 2) (truncation < 6e-9); scale by 2^n.
constexpr float kE1 = (float)0.69314718055994530942;
constexpr float kE2 = (float)0.24022650695910071233;
constexpr float kE3 = (float)0.05550410866482157995;
constexpr float kE4 = (float)0.00961812910762847716;
constexpr float kE5 = (float)0.00133335581464284434;
constexpr float kE6 = (float)0.00015403530393381609;
constexpr float kE7 = (float)0.00001525273380405984;
inline float exp2_f(float y) {
    if (y != y) return y;
    if (y < -150.0f) return 0.0f;
    if (y > 128.0f) return INFINITY;
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

// Precision traits: FP32 uses the functions above, FP64 uses libm.
template <class T> struct Fn;
template <> struct Fn<float> {
    static float sin(float x) { return sin_f(x); }
    static float cos(float x) { return cos_f(x); }
    static float exp2(float x) { return exp2_f(x); }
    static float sqrt(float x) { return sqrtf(x); }
};
template <> struct Fn<double> {
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double exp2(double x) { return std::exp2(x); }
    static double sqrt(double x) { return std::sqrt(x); }
};

// GLSL min/max/clamp (min(x,y) = y < x ? y : x as in the GLSL spec).
template <class T> inline T gmin(T x, T y) { return y < x ? y : x; }
template <class T> inline T gmax(T x, T y) { return x < y ? y : x; }
template <class T> inline T gclamp(T x, T lo, T hi) { return gmin(gmax(x, lo), hi); }

// ---------------------------------------------------------------- vectors
template <class T> struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a) : x(a), y(a), z(a) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline bool all_zero(V3<T> a) { return a.x == 0 && a.y == 0 && a.z == 0; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// GLSL cross(x, y) = (x1 y2 - y1 x2, x2 y0 - y2 x0, x0 y1 - y0 x1).
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> inline T length(V3<T> a) { return Fn<T>::sqrt(dot(a, a)); }
// normalize(v) = v * (1 / sqrt(dot(v, v))).
template <class T> inline V3<T> normalize(V3<T> a) {
    T inv = T(1) / Fn<T>::sqrt(dot(a, a));
    return a * inv;
}
// reflect / refract: GLSL definitions.
template <class T> inline V3<T> reflect(V3<T> I, V3<T> N) { return I - (T(2) * dot(N, I)) * N; }
template <class T> inline V3<T> refract(V3<T> I, V3<T> N, T eta) {
    T d = dot(N, I);
    T k = T(1) - eta * eta * (T(1) - d * d);
    if (k < T(0)) return V3<T>(T(0));
    return eta * I - (eta * d + Fn<T>::sqrt(k)) * N;
}
template <class T> inline T maxc(V3<T> a) { return gmax(a.x, gmax(a.y, a.z)); }

// mat4 (column-major float[16]) times vec4, summed in column order.
template <class T> inline void mat4_mul_vec4(const float* m, const T v[4], T out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = T(m[0 * 4 + r]) * v[0] + T(m[1 * 4 + r]) * v[1] + T(m[2 * 4 + r]) * v[2] + T(m[3 * 4 + r]) * v[3];
}

}  // namespace ko
