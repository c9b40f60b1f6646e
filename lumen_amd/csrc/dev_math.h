// dev_math.h — device vector math, RNG and sampling helpers (gfx950).
// Semantics follow the reference GLSL (src/shaders/utils.glsl); every
// function cites the lines it restates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lt.h"

#define LT_DEV __device__ __forceinline__

namespace ltd {

constexpr float PI = 3.14159265359f;       // utils.glsl:4
constexpr float TWO_PI = 6.28318530718f;   // utils.glsl:5
constexpr float INV_PI = 1.0f / PI;        // utils.glsl:6
constexpr float EPS = 0.001f;              // utils.glsl:8

struct f2 {
    float x, y;
};
struct f3 {
    float x, y, z;
    LT_DEV float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
struct f4 {
    float x, y, z, w;
};

LT_DEV f3 mk3(float a) { return {a, a, a}; }
LT_DEV f3 mk3(float a, float b, float c) { return {a, b, c}; }
LT_DEV f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LT_DEV f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LT_DEV f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
LT_DEV f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
LT_DEV f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
LT_DEV f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
LT_DEV f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
LT_DEV f3& operator+=(f3& a, f3 b) { a = a + b; return a; }
LT_DEV f3& operator*=(f3& a, f3 b) { a = a * b; return a; }
LT_DEV f3& operator*=(f3& a, float s) { a = a * s; return a; }
LT_DEV f3& operator/=(f3& a, float s) { a = a / s; return a; }
LT_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LT_DEV f3 cross(f3 a, f3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
LT_DEV float length(f3 a) { return sqrtf(dot(a, a)); }
LT_DEV f3 normalize(f3 a) { return a * (1.0f / length(a)); }
LT_DEV f3 reflect(f3 i, f3 n) { return i - 2.0f * dot(n, i) * n; }
LT_DEV f3 vsqrt(f3 a) { return {sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)}; }
LT_DEV f3 mix(f3 a, f3 b, float t) { return a * (1.0f - t) + b * t; }
LT_DEV float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
LT_DEV float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
LT_DEV float luminance(f3 c) { return dot(c, mk3(0.2126f, 0.7152f, 0.0722f)); }  // utils.glsl:99
LT_DEV f3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

// column-major 4x4 times vec4
LT_DEV f4 mul4(const float* m, f4 v) {
    return {m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w, m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w,
            m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w, m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w};
}

// ---- pcg4d (Jarzynski & Olano 2020), utils.glsl:121-154 ----
struct RNG {
    uint32_t x, y, z, w;
};
LT_DEV uint32_t pcg4d_x(RNG v) {
    // only the x lane of pcg4d's output is consumed by rand()
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    v.x ^= v.x >> 16u;
    v.y ^= v.y >> 16u;
    v.z ^= v.z >> 16u;
    v.w ^= v.w >> 16u;
    return v.x + v.y * v.w;
}
LT_DEV float uint_to_float(uint32_t x) { return __uint_as_float(0x3f800000u | (x >> 9)) - 1.0f; }
LT_DEV float rnd(RNG& s) {
    s.w++;
    return uint_to_float(pcg4d_x(s));
}

// ---- geometry helpers (utils.glsl) ----
// utils.glsl:73-84
LT_DEV f3 offset_ray(f3 p, f3 n) {
    const float origin = 1.0f / 32.0f;
    const float float_scale = 1.0f / 65536.0f;
    const float int_scale = 256.0f;
    int ox = (int)(int_scale * n.x), oy = (int)(int_scale * n.y), oz = (int)(int_scale * n.z);
    float px = __int_as_float(__float_as_int(p.x) + ((p.x < 0) ? -ox : ox));
    float py = __int_as_float(__float_as_int(p.y) + ((p.y < 0) ? -oy : oy));
    float pz = __int_as_float(__float_as_int(p.z) + ((p.z < 0) ? -oz : oz));
    return {fabsf(p.x) < origin ? p.x + float_scale * n.x : px, fabsf(p.y) < origin ? p.y + float_scale * n.y : py,
            fabsf(p.z) < origin ? p.z + float_scale * n.z : pz};
}
// utils.glsl:90-93
LT_DEV f3 offset_ray2(f3 p, f3 n) { return p + (2.0f / 65536.0f) * n; }
// utils.glsl:185-191
LT_DEV void branchless_onb(f3 n, f3& b1, f3& b2) {
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    b1 = {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    b2 = {b, sign + n.y * n.y * a, -n.y};
}
LT_DEV f3 to_world(f3 v, f3 T, f3 B, f3 N) { return v.x * T + v.y * B + v.z * N; }
LT_DEV f3 to_local(f3 v, f3 T, f3 B, f3 N) { return {dot(v, T), dot(v, B), dot(v, N)}; }
// utils.glsl:215-229
LT_DEV f2 concentric_sample_disk(float rx, float ry) {
    float ox = 2.0f * rx - 1.0f, oy = 2.0f * ry - 1.0f;
    if (ox == 0.0f && oy == 0.0f) return {0.0f, 0.0f};
    float theta, r;
    if (fabsf(ox) > fabsf(oy)) {
        r = ox;
        theta = 0.25f * PI * oy / ox;
    } else {
        r = oy;
        theta = PI * (0.5f - 0.25f * ox / oy);
    }
    float s, c;
    sincosf(theta, &s, &c);
    return {r * c, r * s};
}
// sampling_commons.glsl:351-361 (SAMPLING_MODE 1)
LT_DEV f3 sample_hemisphere(float rx, float ry) {
    f2 d = concentric_sample_disk(rx, ry);
    float z = sqrtf(fmaxf(0.0f, 1.0f - (d.x * d.x + d.y * d.y)));
    return {d.x, d.y, z};
}

}  // namespace ltd
