// kuafu_glm.hpp — the glm types the kuafu:: API uses in its signatures.
//
// The reference API passes glm::vec3 / glm::vec4 / glm::mat4 (stdafx.hpp:13-17).
// When glm is installed it is used as is; otherwise this header provides the
// subset the API and its callers use, with glm's memory layout (column-major
// mat4 of vec4 columns) and glm's formulas for lookAt / perspective /
// translate / rotate / scale / inverse.
#pragma once

#if __has_include(<glm/glm.hpp>)
#define GLM_ENABLE_EXPERIMENTAL
#include <glm/glm.hpp>
#include <glm/gtc/matrix_transform.hpp>
#include <glm/gtc/type_ptr.hpp>
#include <glm/gtx/hash.hpp>
#else
#include <cmath>
#include <cstddef>
#include <functional>

namespace glm {

struct vec2 {
    float x = 0, y = 0;
    vec2() = default;
    explicit vec2(float s) : x(s), y(s) {}
    vec2(float a, float b) : x(a), y(b) {}
    float& operator[](int i) { return (&x)[i]; }
    const float& operator[](int i) const { return (&x)[i]; }
    bool operator==(const vec2& o) const { return x == o.x && y == o.y; }
    bool operator!=(const vec2& o) const { return !(*this == o); }
};

struct vec3 {
    float x = 0, y = 0, z = 0;
    vec3() = default;
    explicit vec3(float s) : x(s), y(s), z(s) {}
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return (&x)[i]; }
    const float& operator[](int i) const { return (&x)[i]; }
    vec3& operator+=(const vec3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    vec3& operator-=(const vec3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    vec3& operator*=(float s) { x *= s; y *= s; z *= s; return *this; }
    bool operator==(const vec3& o) const { return x == o.x && y == o.y && z == o.z; }
    bool operator!=(const vec3& o) const { return !(*this == o); }
};

struct vec4 {
    float x = 0, y = 0, z = 0, w = 0;
    vec4() = default;
    explicit vec4(float s) : x(s), y(s), z(s), w(s) {}
    vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    vec4(const vec3& v, float d) : x(v.x), y(v.y), z(v.z), w(d) {}
    float& operator[](int i) { return (&x)[i]; }
    const float& operator[](int i) const { return (&x)[i]; }
    bool operator==(const vec4& o) const { return x == o.x && y == o.y && z == o.z && w == o.w; }
    bool operator!=(const vec4& o) const { return !(*this == o); }
};

inline vec3 operator+(vec3 a, const vec3& b) { return a += b; }
inline vec3 operator-(vec3 a, const vec3& b) { return a -= b; }
inline vec3 operator-(const vec3& a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(vec3 a, float s) { return a *= s; }
inline vec3 operator*(float s, vec3 a) { return a *= s; }
inline vec3 operator*(const vec3& a, const vec3& b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3 operator/(const vec3& a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec2 operator+(const vec2& a, const vec2& b) { return {a.x + b.x, a.y + b.y}; }
inline vec2 operator*(const vec2& a, float s) { return {a.x * s, a.y * s}; }
inline vec4 operator+(const vec4& a, const vec4& b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline vec4 operator*(const vec4& a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }

inline float dot(const vec3& a, const vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(const vec3& a, const vec3& b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float length(const vec3& a) { return std::sqrt(dot(a, a)); }
inline vec3 normalize(const vec3& a) { return a * (1.0f / std::sqrt(dot(a, a))); }

struct mat4 {
    vec4 c[4];
    mat4() : mat4(1.0f) {}
    explicit mat4(float d) {
        for (int i = 0; i < 4; ++i) c[i] = vec4(0.0f);
        c[0].x = c[1].y = c[2].z = c[3].w = d;
    }
    mat4(const vec4& a, const vec4& b, const vec4& cc, const vec4& d) { c[0] = a; c[1] = b; c[2] = cc; c[3] = d; }
    vec4& operator[](int i) { return c[i]; }
    const vec4& operator[](int i) const { return c[i]; }
    bool operator==(const mat4& o) const { return c[0] == o.c[0] && c[1] == o.c[1] && c[2] == o.c[2] && c[3] == o.c[3]; }
};

inline vec4 operator*(const mat4& m, const vec4& v) {
    vec4 r;
    for (int i = 0; i < 4; ++i) r[i] = m[0][i] * v[0] + m[1][i] * v[1] + m[2][i] * v[2] + m[3][i] * v[3];
    return r;
}
inline mat4 operator*(const mat4& a, const mat4& b) {
    mat4 r(0.0f);
    for (int j = 0; j < 4; ++j) r[j] = a * b[j];
    return r;
}
inline mat4 transpose(const mat4& m) {
    mat4 r(0.0f);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r[i][j] = m[j][i];
    return r;
}
mat4 inverse(const mat4& m);

template <class T> constexpr T pi() { return T(3.14159265358979323846264338327950288); }
inline float radians(float deg) { return deg * pi<float>() / 180.0f; }

inline mat4 translate(const mat4& m, const vec3& v) {
    mat4 r = m;
    r[3] = m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3];
    return r;
}
inline mat4 scale(const mat4& m, const vec3& v) {
    mat4 r = m;
    r[0] = m[0] * v.x;
    r[1] = m[1] * v.y;
    r[2] = m[2] * v.z;
    return r;
}
mat4 rotate(const mat4& m, float angle, const vec3& axis);
mat4 lookAt(const vec3& eye, const vec3& center, const vec3& up);
mat4 perspective(float fovy, float aspect, float zNear, float zFar);
inline const float* value_ptr(const mat4& m) { return &m.c[0].x; }
inline float* value_ptr(mat4& m) { return &m.c[0].x; }

}  // namespace glm

namespace std {
template <> struct hash<glm::vec3> {
    size_t operator()(const glm::vec3& v) const {
        size_t h = std::hash<float>()(v.x);
        h ^= std::hash<float>()(v.y) + 0x9e3779b9 + (h << 6) + (h >> 2);
        h ^= std::hash<float>()(v.z) + 0x9e3779b9 + (h << 6) + (h >> 2);
        return h;
    }
};
template <> struct hash<glm::vec2> {
    size_t operator()(const glm::vec2& v) const {
        size_t h = std::hash<float>()(v.x);
        return h ^ (std::hash<float>()(v.y) + 0x9e3779b9 + (h << 6) + (h >> 2));
    }
};
}  // namespace std
#endif
