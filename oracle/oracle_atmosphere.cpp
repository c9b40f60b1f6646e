// oracle_atmosphere.cpp — TEST INFRASTRUCTURE: CPU restatement of the sky of
// scenes with a directional light, shade_atmosphere (commons.glsl:156-168)
// over integrate_scattering and its helpers (atmosphere/atmosphere.glsl:
// 40-205, Felix Westin's single-scattering model): planet and atmosphere
// spheres (:45-67), Rayleigh and Schlick-Mie phase functions (:71-81),
// density profiles (:85-105), optical depth by 8 midpoint samples
// (:112-132), transmittance (:135-139) and the 64-sample power-law march
// (:143-205).  Every operation in the shaders' order, compiled without
// contraction.
#include "oracle.h"

namespace ora {
namespace {

constexpr float kPi = 3.14159265359f;           // atmosphere.glsl:27
constexpr float kPlanetR = 6371000.0f;          // :29
constexpr float kAtmH = 100000.0f;              // :31
constexpr float kRayleighH = kAtmH * 0.08f;     // :32
constexpr float kMieH = kAtmH * 0.012f;         // :33
constexpr float kExposure = 20.0f;              // :42

const vec3 kCRayleigh = v3(5.802f * 1e-6f, 13.558f * 1e-6f, 33.100f * 1e-6f);  // :37
const vec3 kCMie = v3(3.996f * 1e-6f, 3.996f * 1e-6f, 3.996f * 1e-6f);         // :38
const vec3 kCOzone = v3(0.650f * 1e-6f, 1.881f * 1e-6f, 0.085f * 1e-6f);       // :39

vec2 sphere_intersection(vec3 start, vec3 dir, vec3 centre, float radius) {  // :45-59
    start = start - centre;
    const float a = dot(dir, dir);
    const float b = 2.0f * dot(start, dir);
    const float c = dot(start, start) - radius * radius;
    float d = b * b - 4.0f * a * c;
    if (d < 0.0f) return vec2{-1.0f, -1.0f};
    d = sqrtf(d);
    return vec2{(-b - d) / (2.0f * a), (-b + d) / (2.0f * a)};
}
vec2 planet_intersection(vec3 start, vec3 dir) {  // :60-63
    return sphere_intersection(start, dir, v3(0.0f, -kPlanetR * 1.00001f, 0.0f), kPlanetR);
}
vec2 atmosphere_intersection(vec3 start, vec3 dir) {  // :64-67
    return sphere_intersection(start, dir, v3(0.0f, -kPlanetR, 0.0f), kPlanetR + kAtmH);
}
float atmosphere_height(vec3 p) { return length(p - v3(0.0f, -kPlanetR, 0.0f)) - kPlanetR; }  // :85-88
vec3 atmosphere_density(float h) {                                                           // :89-105
    return v3(expf(-fmaxf(0.0f, h / kRayleighH)), expf(-fmaxf(0.0f, h / kMieH)),
              fmaxf(0.0f, 1.0f - fabsf(h - 25000.0f) / 15000.0f));
}
vec3 integrate_optical_depth(vec3 start, vec3 dir) {  // :112-132
    const float len = atmosphere_intersection(start, dir).y;
    const float step = len / 8.0f;
    vec3 od = v3(0.0f);
    for (int i = 0; i < 8; i++) {
        const vec3 p = start + dir * ((float)i + 0.5f) * step;
        od += atmosphere_density(atmosphere_height(p)) * step;
    }
    return od;
}
vec3 absorb(vec3 od) {  // :135-139
    const vec3 e = kCRayleigh * od.x + kCMie * od.y * 1.1f + kCOzone * od.z;
    return v3(expf(-e.x), expf(-e.y), expf(-e.z));
}

}  // namespace

float phase_rayleigh(float costh) { return 3.0f * (1.0f + costh * costh) / (16.0f * kPi); }  // :71-74
float phase_mie(float costh, float g) {                                                       // :75-81
    g = fminf(g, 0.9381f);
    const float k = 1.55f * g - 0.55f * g * g * g;
    const float kc = k * costh;
    return (1.0f - k * k) / ((4.0f * kPi) * (1.0f - kc) * (1.0f - kc));
}

vec3 integrate_scattering(vec3 start, vec3 dir, float len, vec3 ldir, vec3 Lc) {  // :143-205
    const float h0 = atmosphere_height(start);
    const float expo = 1.0f + clampf(1.0f - h0 / kAtmH, 0.0f, 1.0f) * 8.0f;
    const vec2 is = atmosphere_intersection(start, dir);
    len = fminf(len, is.y);
    if (is.x > 0.0f) {
        start = start + dir * is.x;
        len -= is.x;
    }
    const float costh = dot(dir, ldir);
    const float pr = phase_rayleigh(costh);
    const float pm = phase_mie(costh, 0.85f);
    vec3 od = v3(0.0f), ray = v3(0.0f), mie = v3(0.0f);
    float prev = 0.0f;
    for (int i = 0; i < 64; i++) {
        const float t = powf((float)i / 64.0f, expo) * len;
        const float step = t - prev;
        const vec3 p = start + dir * t;
        const vec3 dens = atmosphere_density(atmosphere_height(p));
        od += dens * step;
        const vec3 tv = absorb(od);
        const vec3 tl = absorb(integrate_optical_depth(p, ldir));
        ray += tv * tl * pr * dens.x * step;
        mie += tv * tl * pm * dens.y * step;
        prev = t;
    }
    return (ray * kCRayleigh + mie * kCMie) * Lc * kExposure;
}

vec3 shade_atmosphere(const lt_scene_desc& s, vec3 sky_col, vec3 o, vec3 d, float len) {  // commons.glsl:156-168
    if (s.dir_light_idx == -1) return sky_col;
    const lt_light& L = s.lights[s.dir_light_idx];
    const vec3 ldir = -normalize(load3(L.to) - load3(L.pos));
    const vec2 pi = planet_intersection(o, d);
    if (pi.x > 0.0f) len = fminf(len, pi.x);
    return integrate_scattering(o, d, len, ldir, load3(L.L));
}

}  // namespace ora
