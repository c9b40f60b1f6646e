// dev_atmosphere.h — sky radiance of scenes with a directional light.
//
// Missed path rays of such scenes see single scattering in a planetary
// atmosphere (shade_atmosphere, commons.glsl:156-168): Rayleigh, Mie and an
// ozone layer with exponential / tent density profiles over a planet of
// radius 6,371 km whose surface lies just below the origin, marched along
// the view ray in 64 samples spaced by a power law that crowds them near the
// viewer, each sample attenuated toward the light by an 8-sample
// optical-depth march (the model of Felix Westin's MIT-licensed atmosphere
// shader that the reference includes, shaders/atmosphere/atmosphere.glsl:
// 40-205).  The constants and the order of the sums are the model's; every
// product is rounded on its own (no fused multiply-add), so that the CPU
// oracle's restatement, compiled with -ffp-contract=off, gives the same
// values up to the last bits of expf / powf / sqrtf.  Scenes without a
// directional light get the constant sky colour.
#pragma once
#include "context.h"
#include "dev_math.h"

namespace ltd {
namespace atm {

constexpr float kPi = 3.14159265359f;
constexpr float kPlanetR = 6371000.0f;
constexpr float kAtmH = 100000.0f;
constexpr float kRayleighH = kAtmH * 0.08f;
constexpr float kMieH = kAtmH * 0.012f;
constexpr float kExposure = 20.0f;

// the model's coefficients (per metre): Rayleigh, Mie, ozone
LT_DEV f3 c_rayleigh() { return f3{5.802f * 1e-6f, 13.558f * 1e-6f, 33.100f * 1e-6f}; }
LT_DEV f3 c_mie() { return f3{3.996f * 1e-6f, 3.996f * 1e-6f, 3.996f * 1e-6f}; }
LT_DEV f3 c_ozone() { return f3{0.650f * 1e-6f, 1.881f * 1e-6f, 0.085f * 1e-6f}; }

LT_DEV float dot_nc(f3 a, f3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

// ray-sphere distances (near, far), (-1, -1) on a miss
LT_DEV f2 sphere_hit(f3 o, f3 d, f3 c, float r) {
#pragma clang fp contract(off)
    o = f3{o.x - c.x, o.y - c.y, o.z - c.z};
    const float a = dot_nc(d, d);
    const float b = 2.0f * dot_nc(o, d);
    const float cc = dot_nc(o, o) - r * r;
    float disc = b * b - 4.0f * a * cc;
    if (disc < 0.0f) return f2{-1.0f, -1.0f};
    disc = sqrtf(disc);
    return f2{(-b - disc) / (2.0f * a), (-b + disc) / (2.0f * a)};
}
// the planet surface sits 63.71 m below the origin (its centre scaled by 1.00001)
LT_DEV f2 planet_hit(f3 o, f3 d) { return sphere_hit(o, d, f3{0.0f, -kPlanetR * 1.00001f, 0.0f}, kPlanetR); }
LT_DEV f2 shell_hit(f3 o, f3 d) { return sphere_hit(o, d, f3{0.0f, -kPlanetR, 0.0f}, kPlanetR + kAtmH); }

LT_DEV float height(f3 p) {
#pragma clang fp contract(off)
    const f3 q{p.x, p.y + kPlanetR, p.z};
    return sqrtf(dot_nc(q, q)) - kPlanetR;
}
// densities (Rayleigh, Mie, ozone) at height h
LT_DEV f3 density(float h) {
#pragma clang fp contract(off)
    return f3{expf(-fmaxf(0.0f, h / kRayleighH)), expf(-fmaxf(0.0f, h / kMieH)),
              fmaxf(0.0f, 1.0f - fabsf(h - 25000.0f) / 15000.0f)};
}
// transmittance of an optical depth (Mie absorbs 10 % beyond its scattering)
LT_DEV f3 absorb(f3 od) {
#pragma clang fp contract(off)
    const f3 r = c_rayleigh(), m = c_mie(), z = c_ozone();
    return f3{expf(-(od.x * r.x + od.y * m.x * 1.1f + od.z * z.x)), expf(-(od.x * r.y + od.y * m.y * 1.1f + od.z * z.y)),
              expf(-(od.x * r.z + od.y * m.z * 1.1f + od.z * z.z))};
}
// optical depth from p to the top of the atmosphere along d: 8 midpoint samples
LT_DEV f3 optical_depth(f3 p, f3 d) {
#pragma clang fp contract(off)
    const float len = shell_hit(p, d).y;
    const float step = len / 8.0f;
    f3 od{0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 8; i++) {
        const float s = (float)i + 0.5f;  // p + (d * s) * step, rounded in that order
        const f3 q{p.x + d.x * s * step, p.y + d.y * s * step, p.z + d.z * s * step};
        const f3 dn = density(height(q));
        od = f3{od.x + dn.x * step, od.y + dn.y * step, od.z + dn.z * step};
    }
    return od;
}

// in-scattered radiance along (o, d) up to len, lit from direction ldir
LT_DEV f3 scatter(f3 o, f3 d, float len, f3 ldir, f3 Lc) {
#pragma clang fp contract(off)
    const float h0 = height(o);
    const float expo = 1.0f + fminf(fmaxf(1.0f - h0 / kAtmH, 0.0f), 1.0f) * 8.0f;
    const f2 sh = shell_hit(o, d);
    len = fminf(len, sh.y);
    if (sh.x > 0.0f) {
        o = f3{o.x + d.x * sh.x, o.y + d.y * sh.x, o.z + d.z * sh.x};
        len -= sh.x;
    }
    const float cth = dot_nc(d, ldir);
    const float ph_r = 3.0f * (1.0f + cth * cth) / (16.0f * kPi);
    const float g = fminf(0.85f, 0.9381f);
    const float k = 1.55f * g - 0.55f * g * g * g;
    const float kc = k * cth;
    const float ph_m = (1.0f - k * k) / ((4.0f * kPi) * (1.0f - kc) * (1.0f - kc));
    f3 od{0.0f, 0.0f, 0.0f}, ray{0.0f, 0.0f, 0.0f}, mie{0.0f, 0.0f, 0.0f};
    float prev = 0.0f;
    for (int i = 0; i < 64; i++) {
        const float t = powf((float)i / 64.0f, expo) * len;
        const float step = t - prev;
        const f3 q{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
        const f3 dn = density(height(q));
        od = f3{od.x + dn.x * step, od.y + dn.y * step, od.z + dn.z * step};
        const f3 tv = absorb(od);
        const f3 tl = absorb(optical_depth(q, ldir));
        const f3 tt{tv.x * tl.x, tv.y * tl.y, tv.z * tl.z};
        ray = f3{ray.x + tt.x * ph_r * dn.x * step, ray.y + tt.y * ph_r * dn.x * step, ray.z + tt.z * ph_r * dn.x * step};
        mie = f3{mie.x + tt.x * ph_m * dn.y * step, mie.y + tt.y * ph_m * dn.y * step, mie.z + tt.z * ph_m * dn.y * step};
        prev = t;
    }
    const f3 r = c_rayleigh(), m = c_mie();
    return f3{(ray.x * r.x + mie.x * m.x) * Lc.x * kExposure, (ray.y * r.y + mie.y * m.y) * Lc.y * kExposure,
              (ray.z * r.z + mie.z * m.z) * Lc.z * kExposure};
}

}  // namespace atm

// shade_atmosphere (commons.glsl:156-168): the sky seen by a ray that left
// the scene from `o` along `d` (ray length `len`)
LT_DEV f3 sky_radiance(const SceneView& sv, f3 sky_col, f3 o, f3 d, float len) {
    if (sv.dir_light_idx < 0) return sky_col;
    const lt_light& L = sv.lights[sv.dir_light_idx];
    const f3 v{L.to[0] - L.pos[0], L.to[1] - L.pos[1], L.to[2] - L.pos[2]};
    const float inv = 1.0f / sqrtf(atm::dot_nc(v, v));
    const f3 ldir{-(v.x * inv), -(v.y * inv), -(v.z * inv)};
    const f2 ph = atm::planet_hit(o, d);
    if (ph.x > 0.0f) len = fminf(len, ph.x);
    return atm::scatter(o, d, len, ldir, f3{L.L[0], L.L[1], L.L[2]});
}

}  // namespace ltd
