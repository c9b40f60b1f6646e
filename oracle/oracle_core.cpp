// oracle_core.cpp — RNG, sampling helpers, BSDFs and light sampling restated
// from the reference GLSL (TEST INFRASTRUCTURE, see oracle.h).
#include "oracle.h"

namespace ora {

// ---------------------------------------------------------------- RNG ----
// pcg4d, Jarzynski & Olano 2020 — utils.glsl:121-133
uvec4 pcg4d(uvec4 v) {
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
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    return v;
}

// utils.glsl:136
float uint_to_float(uint32_t x) {
    uint32_t b = 0x3f800000u | (x >> 9);
    f32 f;
    memcpy(&f, &b, 4);
    return f - 1.0f;
}

// utils.glsl:145-148 (pre-increment of the counter word)
float rand(uvec4& s) {
    s.w++;
    return uint_to_float(pcg4d(s).x);
}

// ------------------------------------------------------------ geometry ----
static inline int f2i(f32 f) { int i; memcpy(&i, &f, 4); return i; }
static inline f32 i2f(int i) { f32 f; memcpy(&f, &i, 4); return f; }

// utils.glsl:73-84 (Wächter & Binder, Ray Tracing Gems ch. 6)
vec3 offset_ray(vec3 p, vec3 n) {
    const float origin = 1.0f / 32.0f;
    const float float_scale = 1.0f / 65536.0f;
    const float int_scale = 256.0f;
    int oi_x = (int)(int_scale * n.x), oi_y = (int)(int_scale * n.y), oi_z = (int)(int_scale * n.z);
    vec3 p_i = {i2f(f2i(p.x) + ((p.x < 0) ? -oi_x : oi_x)), i2f(f2i(p.y) + ((p.y < 0) ? -oi_y : oi_y)),
                i2f(f2i(p.z) + ((p.z < 0) ? -oi_z : oi_z))};
    return {fabsf(p.x) < origin ? p.x + float_scale * n.x : p_i.x,
            fabsf(p.y) < origin ? p.y + float_scale * n.y : p_i.y,
            fabsf(p.z) < origin ? p.z + float_scale * n.z : p_i.z};
}

// utils.glsl:90-93
vec3 offset_ray2(vec3 p, vec3 n) {
    const float float_scale = 2.0f / 65536.0f;
    return p + float_scale * n;
}

// utils.glsl:185-191 (Duff et al. 2017)
void branchless_onb(vec3 n, vec3& b1, vec3& b2) {
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    b1 = {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    b2 = {b, sign + n.y * n.y * a, -n.y};
}
static inline vec3 to_world(vec3 v, vec3 T, vec3 B, vec3 N) { return v.x * T + v.y * B + v.z * N; }
static inline vec3 to_local(vec3 v, vec3 T, vec3 B, vec3 N) { return {dot(v, T), dot(v, B), dot(v, N)}; }

// utils.glsl:215-229 (Shirley & Chiu)
vec2 concentric_sample_disk(vec2 r) {
    vec2 o = {2.0f * r.x - 1.0f, 2.0f * r.y - 1.0f};
    if (o.x == 0.0f && o.y == 0.0f) return {0.0f, 0.0f};
    float theta, rad;
    if (fabsf(o.x) > fabsf(o.y)) {
        rad = o.x;
        theta = 0.25f * PI * o.y / o.x;
    } else {
        rad = o.y;
        theta = PI * (0.5f - 0.25f * o.x / o.y);
    }
    return {rad * cosf(theta), rad * sinf(theta)};
}

// sampling_commons.glsl:351-361 with SAMPLING_MODE 1 (Malley's method)
vec3 sample_hemisphere(vec2 xi) {
    vec2 d = concentric_sample_disk(xi);
    float z = sqrtf(fmaxf(0.0f, 1.0f - (d.x * d.x + d.y * d.y)));
    return {d.x, d.y, z};
}

// sampling_commons.glsl:331-343 (explicit normal)
vec3 sample_hemisphere_n(vec2 xi, vec3 n, float& phi) {
    phi = TWO_PI * xi.x;
    vec3 T, B;
    branchless_onb(n, T, B);
    vec2 d = concentric_sample_disk(xi);
    float z = sqrtf(fmaxf(0.0f, 1.0f - (d.x * d.x + d.y * d.y)));
    return to_world(v3(d.x, d.y, z), T, B, n);
}

// ------------------------------------------------------- BSDF helpers ----
static inline bool has_prop(uint32_t props, uint32_t f) { return (props & f) != 0; }
static inline float sgnf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// sampling_commons.glsl:18-19
static inline bool eff_delta(float a) { return a <= 0.0064f; }
static inline bool eff_delta2(vec2 a) { return fminf(a.x, a.y) <= 0.0064f; }

// sampling_commons.glsl:22-41
static bool refract_(vec3 n_s, vec3 wo, bool ff, float eta, uint32_t mode, vec3& wi, vec3& f, float& inv_eta) {
    wi = v3(0.0f);
    f = v3(0.0f);
    float cos_i = dot(n_s, wo);
    inv_eta = ff ? 1.0f / eta : eta;
    const float sin2_t = inv_eta * inv_eta * (1.0f - cos_i * cos_i);
    if (sin2_t >= 1.0f) {
        wi = reflect(-wo, n_s);
    } else {
        const float cos_t = sqrtf(1.0f - sin2_t);
        wi = -inv_eta * wo + (inv_eta * cos_i - cos_t) * n_s;
    }
    f = mode == 1 ? v3(inv_eta * inv_eta) : v3(1.0f);
    return true;
}

// sampling_commons.glsl:48-65
float fresnel_dielectric(float cos_i, float eta, bool ff) {
    cos_i = clampf(cos_i, -1.0f, 1.0f);
    if (!ff) eta = 1.0f / eta;
    if (cos_i < 0.0f) {
        eta = 1.0f / eta;
        cos_i = -cos_i;
    }
    float sin2_i = 1.0f - cos_i * cos_i;
    float sin2_t = sin2_i / (eta * eta);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = sqrtf(1.0f - sin2_t);
    float r_par = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (r_par * r_par + r_perp * r_perp);
}

// sampling_commons.glsl:68-82 (Lagarde's memo)
float fresnel_conductor(float cos_i, float eta, float k) {
    float cos_sqr = cos_i * cos_i;
    float sin_sqr = fmaxf(1.0f - cos_sqr, 0.0f);
    float sin_4 = sin_sqr * sin_sqr;
    float inner = eta * eta - k * k - sin_sqr;
    float a2b2 = sqrtf(fmaxf(inner * inner + 4.0f * eta * eta * k * k, 0.0f));
    float a = sqrtf(fmaxf((a2b2 + inner) * 0.5f, 0.0f));
    float rs = ((a2b2 + cos_sqr) - (2.0f * a * cos_i)) / ((a2b2 + cos_sqr) + (2.0f * a * cos_i));
    float rp = ((cos_sqr * a2b2 + sin_4) - (2.0f * a * cos_i * sin_sqr)) /
               ((cos_sqr * a2b2 + sin_4) + (2.0f * a * cos_i * sin_sqr));
    return 0.5f * (rs + rs * rp);
}
static vec3 fresnel_conductor3(float c, vec3 eta, vec3 k) {
    return {fresnel_conductor(c, eta.x, k.x), fresnel_conductor(c, eta.y, k.y), fresnel_conductor(c, eta.z, k.z)};
}

// sampling_commons.glsl:84-92
static float fresnel_schlick(float f0, float f90, float ns) {
    return f0 + (f90 - f0) * powf(fmaxf(1.0f - ns, 0.0f), 5.0f);
}
static vec3 fresnel_schlick3(vec3 f0, vec3 f90, float ns) {
    return f0 + (f90 - f0) * powf(fmaxf(1.0f - ns, 0.0f), 5.0f);
}
static float eta_to_schlick_R0(float eta) {
    float v = (eta - 1.0f) / (eta + 1.0f);
    return v * v;
}
// sampling_commons.glsl:102-110
static float disney_fresnel_s(vec3 wi, vec3 wo, float roughness, float& f_wi, float& f_wo) {
    vec3 h = normalize(wi + wo);
    float wo_dot_h = dot(wo, h);
    float fd90 = 0.5f + 2.0f * wo_dot_h * wo_dot_h * roughness;
    float fd0 = 1.0f;
    f_wi = fresnel_schlick(fd0, fd90, wi.z);
    f_wo = fresnel_schlick(fd0, fd90, wo.z);
    return f_wi * f_wo;
}
// sampling_commons.glsl:119-122
static vec3 calc_tint(vec3 albedo) {
    float lum = luminance(albedo);
    return lum > 0.0f ? albedo / lum : v3(1.0f);
}
// sampling_commons.glsl:125-139
static vec3 disney_fresnel_m(const Material& m, vec3 wo, vec3 h, vec3 wi, float eta) {
    (void)wi;
    float wo_dot_h = dot(wo, h);
    vec3 R0 = eta_to_schlick_R0(eta) * mix(v3(1.0f), calc_tint(m.albedo), m.specular_tint);
    R0 = mix(R0, m.albedo, m.metallic);
    float fr_d = fresnel_dielectric(wo_dot_h, eta, true);
    vec3 fr_m = fresnel_schlick3(R0, v3(1.0f), wo_dot_h);
    return mix(v3(fr_d), fr_m, m.metallic);
}

// ------------------------------------------------------ microfacets ----
// microfacet_commons.glsl:5-12
static float smith_lambda_iso(float a2, float c) {
    if (c == 0.0f) return 0.0f;
    float c2 = c * c;
    float t2 = fmaxf(1.0f - c2, 0.0f) / c2;
    return 0.5f * (sqrtf(1.0f + a2 * t2) - 1.0f);
}
// microfacet_commons.glsl:13-23
static float smith_lambda_aniso(vec3 w, vec2 alpha) {
    float c2 = w.z * w.z;
    float s2 = fmaxf(1.0f - c2, 0.0f);
    float t2 = s2 / c2;
    if (std::isinf(t2)) return 0.0f;
    vec2 cp2 = s2 == 0.0f ? v2(1.0f, 0.0f)
                          : v2(clampf(w.x * w.x, 0.0f, 1.0f) / s2, clampf(w.y * w.y, 0.0f, 1.0f) / s2);
    float a2 = cp2.x * (alpha.x * alpha.x) + cp2.y * (alpha.y * alpha.y);
    return 0.5f * (sqrtf(1.0f + a2 * t2) - 1.0f);
}
static float G1_aniso(vec3 w, vec2 a) { return 1.0f / (1.0f + smith_lambda_aniso(w, a)); }
static float G_corr_iso(float alpha, vec3 wo, vec3 wi) {
    float a2 = alpha * alpha;
    return 1.0f / (1.0f + smith_lambda_iso(a2, wo.z) + smith_lambda_iso(a2, wi.z));
}
static float G_corr_aniso(vec2 a, vec3 wo, vec3 wi) {
    return 1.0f / (1.0f + smith_lambda_aniso(wo, a) + smith_lambda_aniso(wi, a));
}
// microfacet_commons.glsl:36-51
float D_GGX_anisotropic(vec2 alpha, vec3 h) {
    float c2 = h.z * h.z;
    float s2 = fmaxf(1.0f - c2, 0.0f);
    float t2 = s2 / c2;
    if (std::isinf(t2)) return 0.0f;
    float c4 = c2 * c2;
    if (c4 < 1e-16f) return 0.0f;
    vec2 p2 = s2 == 0.0f ? v2(1.0f, 0.0f)
                         : v2(clampf(h.x * h.x, 0.0f, 1.0f) / s2, clampf(h.y * h.y, 0.0f, 1.0f) / s2);
    vec2 a2 = {p2.x / (alpha.x * alpha.x), p2.y / (alpha.y * alpha.y)};
    float e = t2 * (a2.x + a2.y);
    return 1.0f / (PI * alpha.x * alpha.y * c4 * (1.0f + e) * (1.0f + e));
}
// microfacet_commons.glsl:54-57
float D_GGX_isotropic(float a2, float c) {
    float d = ((c * a2 - c) * c + 1.0f);
    return a2 / (d * d * PI);
}
// microfacet_commons.glsl:60-64
static float G1_iso(float a2, float c) {
    float c2 = c * c;
    float t2 = fmaxf(1.0f - c2, 0.0f) / c2;
    return 2.0f / (1.0f + sqrtf(1.0f + a2 * t2));
}
// microfacet_commons.glsl:67-76
static float vndf_pdf_iso(float alpha, vec3 wo, vec3 h, float& D) {
    D = 0.0f;
    if (wo.z <= 0.0f) return 0.0f;
    float a2 = alpha * alpha;
    float G1 = G1_iso(a2, wo.z);
    D = D_GGX_isotropic(a2, h.z);
    return G1 * D * fmaxf(0.0f, dot(wo, h)) / fabsf(wo.z);
}
// microfacet_commons.glsl:84-92
static float vndf_pdf_aniso(vec2 alpha, vec3 wo, vec3 h, float& D) {
    D = 0.0f;
    if (wo.z <= 0.0f) return 0.0f;
    float G1 = G1_aniso(wo, alpha);
    D = D_GGX_anisotropic(alpha, h);
    return G1 * D * fmaxf(0.0f, dot(wo, h)) / fabsf(wo.z);
}
// microfacet_commons.glsl:99-125 (Dupuy & Benyoub spherical caps)
static vec3 sample_vndf(vec2 alpha, vec3 wo, vec2 xi) {
    vec3 wh = normalize(v3(alpha.x * wo.x, alpha.y * wo.y, wo.z));
    float phi = 2.0f * PI * xi.x;
    float z = ((1.0f - xi.y) * (1.0f + wh.z)) - wh.z;
    float st = sqrtf(clampf(1.0f - z * z, 0.0f, 1.0f));
    float x = st * cosf(phi);
    float y = st * sinf(phi);
    vec3 nh = v3(x, y, z) + wh;
    return normalize(v3(alpha.x * nh.x, alpha.y * nh.y, fmaxf(0.0f, nh.z)));
}

// ------------------------------------------------------------ diffuse ----
// diffuse.glsl:10-21 (Lambertian)
static vec3 sample_diffuse(const Material& m, vec3 wo, vec3& wi, float& pdf_w, float& cos_t, vec2 xi) {
    wi = sample_hemisphere(xi);
    cos_t = wi.z;
    pdf_w = cos_t * INV_PI;
    if (fminf(wi.z, wo.z) <= 0.0f) return v3(0.0f);
    return m.albedo * INV_PI;
}
// diffuse.glsl:58-65
static vec3 eval_diffuse(const Material& m, vec3 wo, vec3 wi, float& pdf_w, float& pdf_rev_w) {
    pdf_w = 0.0f;
    pdf_rev_w = 0.0f;
    if (fminf(wi.z, wo.z) <= 0.0f) return v3(0.0f);
    pdf_w = wi.z * INV_PI;
    pdf_rev_w = pdf_w;
    return m.albedo * INV_PI;
}
static float diffuse_pdf(vec3 wo, vec3 wi) {
    if (fminf(wi.z, wo.z) <= 0.0f) return 0.0f;
    return wi.z * INV_PI;
}

// ------------------------------------------------------ mirror, glass ----
// mirror.glsl:3-15
static vec3 sample_mirror(vec3 n, vec3 wo, vec3& wi, float& pdf_w, float& cos_t) {
    if (dot(wo, n) <= 0.0f) {
        pdf_w = 0.0f;
        cos_t = 0.0f;
        wi = v3(0.0f);
        return v3(0.0f);
    }
    wi = reflect(-wo, n);
    cos_t = dot(wi, n);
    pdf_w = 1.0f;
    return v3(1.0f) / cos_t;
}
// glass.glsl:4-17
static vec3 sample_glass(const Material& m, vec3 n, vec3 wo, vec3& wi, float& pdf_w, float& cos_t, uint32_t mode,
                         bool ff) {
    wi = v3(0.0f);
    pdf_w = 0.0f;
    cos_t = 0.0f;
    vec3 f;
    float ie;
    refract_(n, wo, ff, m.ior, mode, wi, f, ie);
    cos_t = dot(n, wi);
    pdf_w = 1.0f;
    f = f / fabsf(cos_t);
    return f;
}

// --------------------------------------------------------- dielectric ----
static float thin_rough(float ior, float r) { return clampf((0.65f * ior - 0.35f) * r, 0.0f, 1.0f); }

// dielectric.glsl:10-108
static vec3 sample_dielectric(const Material& m, vec3 wo, vec3& wi, uint32_t mode, bool ff, float& pdf_w,
                              float& cos_t, vec2 xi) {
    wi = v3(0.0f);
    pdf_w = 0.0f;
    cos_t = 0.0f;
    if (wo.z <= 0.0f) return v3(0.0f);
    float rough = m.thin == 1 ? thin_rough(m.ior, m.roughness) : m.roughness;
    float alpha = rough * rough;
    bool has_r = has_prop(m.bsdf_props, LT_BSDF_FLAG_REFLECTION);
    bool has_t = has_prop(m.bsdf_props, LT_BSDF_FLAG_TRANSMISSION);
    if ((m.ior == 1.0f && m.thin == 0) || eff_delta(alpha)) {
        float F = fresnel_dielectric(wo.z, m.ior, ff);
        if (!has_r && !has_t) return v3(0.0f);
        float pr = F, pt = 1.0f - F;
        if (!has_r) pr = 0.0f;
        if (!has_t) pt = 0.0f;
        bool is_refl = ((pr + pt) * xi.x) < pr;
        vec3 f;
        if (is_refl) {
            wi = v3(-wo.x, -wo.y, wo.z);
            cos_t = wi.z;
            f = v3(F) / fabsf(cos_t);
            pdf_w = pr / (pr + pt);
        } else {
            float ie;
            refract_(v3(0.0f, 0.0f, 1.0f), wo, ff, m.ior, mode, wi, f, ie);
            cos_t = wi.z;
            f = f * (1.0f - F) / fabsf(cos_t);
            pdf_w = pt / (pr + pt);
        }
        return f;
    }
    if (!has_r && !has_t) return v3(0.0f);
    float D;
    vec3 h = sample_vndf(v2(alpha, alpha), wo, xi);
    pdf_w = vndf_pdf_iso(alpha, wo, h, D);
    float F = fresnel_dielectric(dot(wo, h), m.ior, ff);
    float pr = has_r ? F : 0.0f;
    float pt = has_t ? (1.0f - F) : 0.0f;
    bool is_refl = ((pr + pt) * xi.x) < pr;
    vec3 f;
    if (is_refl) {
        wi = reflect(-wo, h);
        if (wo.z * wi.z < 0.0f) return v3(0.0f);
        pdf_w = pdf_w * (pr / (pr + pt)) / (4.0f * fabsf(dot(wo, h)));
        f = v3(0.25f * D * F * G_corr_iso(alpha, wo, wi) / (wi.z * wo.z));
    } else {
        vec3 base;
        float ie = 0.0f;
        bool tir = false;
        if (m.thin == 1) {
            wi = reflect(-wo, h);
            base = vsqrt(m.albedo);
            ie = m.ior;
            f = v3(0.0f);  // GLSL leaves f unset on this branch (dielectric.glsl:88-92); zero here
        } else {
            tir = !refract_(h, wo, ff, m.ior, mode, wi, f, ie);
            base = m.albedo;
        }
        if (wo.z * wi.z > 0.0f || wi.z > 0.0f || tir) return v3(0.0f);
        float jd = dot(wi, h) + dot(wo, h) * ie;
        jd = jd * jd;
        float jac = fabsf(dot(wi, h)) / jd;
        pdf_w = pdf_w * (pt / (pr + pt)) * jac;
        f = base * f * (1.0f - F) * D * G_corr_iso(alpha, wo, wi) *
            fabsf(dot(wi, h) * dot(wo, h) / (wi.z * wo.z * jd));
    }
    cos_t = wi.z;
    return f;
}

// dielectric.glsl:110-189.  The refraction branch of the reference declares a
// second `vec3 f` that shadows the returned one (dielectric.glsl:171), so the
// value returned for rough transmission is the uninitialised outer f; this
// restatement returns 0 there, the value a zero-initialised outer f gives.
static vec3 eval_dielectric(const Material& m, vec3 wo, vec3 wi, float& pdf_w, float& pdf_rev_w, bool ff,
                            uint32_t mode, bool rev) {
    (void)mode;
    pdf_w = 0.0f;
    pdf_rev_w = 0.0f;
    float rough = m.thin == 1 ? thin_rough(m.ior, m.roughness) : m.roughness;
    float alpha = rough * rough;
    if (alpha == 0.0f || m.ior == 1.0f) return v3(0.0f);
    bool has_r = has_prop(m.bsdf_props, LT_BSDF_FLAG_REFLECTION);
    bool has_t = has_prop(m.bsdf_props, LT_BSDF_FLAG_TRANSMISSION);
    if (!has_r && !has_t) return v3(0.0f);
    float eta = 1.0f;
    bool is_refl = wi.z * wo.z > 0.0f;
    if (!is_refl) eta = ff ? m.ior : 1.0f / m.ior;
    vec3 h = normalize(wo + wi * eta);
    h = h * sgnf(h.z);
    if (wi.z == 0.0f || wo.z == 0.0f || dot(h, h) == 0.0f) return v3(0.0f);
    if (dot(wi, h) * wi.z < 0.0f || dot(wo, h) * wo.z < 0.0f) return v3(0.0f);
    float F = fresnel_dielectric(dot(wo, h), m.ior, ff);
    float pr = has_r ? F : 0.0f;
    float pt = has_t ? (1.0f - F) : 0.0f;
    float D;
    if (rev) pdf_rev_w = vndf_pdf_iso(alpha, wi, -h, D);
    pdf_w = vndf_pdf_iso(alpha, wo, h, D);
    float G = G_corr_iso(alpha, wo, wi);
    vec3 f = v3(0.0f);
    if (is_refl) {
        float jac = 1.0f / (4.0f * fabsf(dot(wo, h)));
        float prr = pr / (pr + pt);
        pdf_w = pdf_w * jac * prr;
        if (rev) {
            float jr = 1.0f / (4.0f * fabsf(dot(wi, h)));
            pdf_rev_w = pdf_rev_w * jr * prr;
        }
        f = v3(0.25f * D * G * F / fabsf(wo.z * wi.z));
    } else {
        float jd = dot(wi, h) + dot(wo, h) / eta;
        jd = jd * jd;
        float jac = fabsf(dot(wi, h)) / jd;
        float prt = pt / (pr + pt);
        pdf_w = pdf_w * jac * prt;
        if (rev) {
            vec3 hr = -h;
            float jdr = eta * dot(wi, hr) + dot(wo, hr);
            jdr *= jdr;
            float jr = fabsf(dot(wo, hr)) / jdr;
            pdf_rev_w = pdf_rev_w * jr * prt;
        }
    }
    return f;
}

// dielectric.glsl:191-240
static float dielectric_pdf(const Material& m, vec3 wo, vec3 wi, bool ff) {
    float rough = m.thin == 1 ? thin_rough(m.ior, m.roughness) : m.roughness;
    float alpha = rough * rough;
    if (alpha == 0.0f || m.ior == 1.0f) return 0.0f;
    bool has_r = has_prop(m.bsdf_props, LT_BSDF_FLAG_REFLECTION);
    bool has_t = has_prop(m.bsdf_props, LT_BSDF_FLAG_TRANSMISSION);
    if (!has_r && !has_t) return 0.0f;
    bool is_refl = wi.z * wo.z > 0.0f;
    float eta = 1.0f;
    if (!is_refl) eta = ff ? m.ior : 1.0f / m.ior;
    vec3 h = normalize(wo + wi * eta);
    h = h * sgnf(h.z);
    if (wi.z == 0.0f || wo.z == 0.0f || dot(h, h) == 0.0f) return 0.0f;
    if (dot(wi, h) * wi.z < 0.0f || dot(wo, h) * wo.z < 0.0f) return 0.0f;
    float F = fresnel_dielectric(dot(wo, h), m.ior, ff);
    float pr = has_r ? F : 0.0f;
    float pt = has_t ? (1.0f - F) : 0.0f;
    float D;
    float pdf = vndf_pdf_iso(alpha, wo, h, D);
    if (is_refl) {
        pdf = pdf * (1.0f / (4.0f * fabsf(dot(wo, h)))) * (pr / (pr + pt));
    } else {
        float jd = dot(wi, h) + dot(wo, h) / eta;
        jd = jd * jd;
        pdf = pdf * (fabsf(dot(wi, h)) / jd) * (pt / (pr + pt));
    }
    return pdf;
}

// ---------------------------------------------------------- conductor ----
// conductor.glsl:6-43
static vec3 sample_conductor(const Material& m, vec3 wo, vec3& wi, float& pdf_w, float& cos_t, vec2 xi) {
    wi = v3(0.0f);
    pdf_w = 0.0f;
    cos_t = 0.0f;
    if (wo.z <= 0.0f) return v3(0.0f);
    float alpha = m.roughness * m.roughness;
    if (!has_prop(m.bsdf_props, LT_BSDF_FLAG_REFLECTION)) return v3(0.0f);
    if (eff_delta(alpha)) {
        wi = v3(-wo.x, -wo.y, wo.z);
        pdf_w = 1.0f;
        cos_t = wi.z;
        vec3 F = fresnel_conductor3(cos_t, m.albedo, m.k);
        return F / fabsf(cos_t);
    }
    float D;
    vec3 h = sample_vndf(v2(alpha, alpha), wo, xi);
    pdf_w = vndf_pdf_iso(alpha, wo, h, D);
    vec3 F = fresnel_conductor3(dot(wo, h), m.albedo, m.k);
    wi = reflect(-wo, h);
    if (wo.z * wi.z < 0.0f) return v3(0.0f);
    pdf_w /= (4.0f * dot(wo, h));
    cos_t = wi.z;
    return 0.25f * D * F * G_corr_iso(alpha, wo, wi) / (wi.z * wo.z);
}
// conductor.glsl:45-72
static vec3 eval_conductor(const Material& m, vec3 wo, vec3 wi, float& pdf_w, float& pdf_rev_w, bool rev) {
    pdf_w = 0.0f;
    pdf_rev_w = 0.0f;
    float alpha = m.roughness * m.roughness;
    if (eff_delta(alpha)) return v3(0.0f);
    if (wo.z * wi.z < 0.0f) return v3(0.0f);
    if (wo.z == 0.0f || wi.z == 0.0f) return v3(0.0f);
    vec3 h = normalize(wo + wi);
    h = h * sgnf(h.z);
    float jac = 1.0f / (4.0f * dot(wo, h));
    float D, Dr;
    pdf_w = vndf_pdf_iso(alpha, wo, h, D) * jac;
    if (rev) pdf_rev_w = vndf_pdf_iso(alpha, wi, h, Dr) * jac;
    vec3 F = fresnel_conductor3(dot(wo, h), m.albedo, m.k);
    return 0.25f * D * F * G_corr_iso(alpha, wo, wi) / (wi.z * wo.z);
}
// conductor.glsl:74-90
static float conductor_pdf(const Material& m, vec3 wo, vec3 wi) {
    float alpha = m.roughness * m.roughness;
    if (eff_delta(alpha)) return 0.0f;
    if (wo.z * wi.z < 0.0f) return 0.0f;
    if (wo.z == 0.0f || wi.z == 0.0f) return 0.0f;
    vec3 h = normalize(wo + wi);
    h = h * sgnf(h.z);
    float D;
    return vndf_pdf_iso(alpha, wo, h, D) / (4.0f * dot(wo, h));
}

// ---------------------------------------------------------- principled ----
// principled.glsl:20-24
static vec2 calc_aniso(float roughness, float aniso) {
    float aspect = sqrtf(1.0f - 0.9f * aniso);
    float r2 = roughness * roughness;
    return {fmaxf(0.001f, r2 / aspect), fmaxf(0.001f, r2 * aspect)};
}
// principled.glsl:26-48
static void init_probs(const Material& m, float Fd, bool ff, float& ps, float& pd, float& pc, float& pt) {
    float brdf_w = (1.0f - m.spec_trans) * (1.0f - m.metallic);
    float bsdf_w = (1.0f - m.metallic) * m.spec_trans;
    ps = ff ? (1.0f - bsdf_w * (1.0f - Fd)) : Fd;
    pt = ff ? (bsdf_w * (1.0f - Fd)) : (1.0f - Fd);
    pd = ff ? brdf_w : 0.0f;
    pc = ff ? 0.25f * clampf(m.clearcoat, 0.0f, 1.0f) : 0.0f;
    float norm = 1.0f / (ps + pt + pd + pc);
    ps *= norm;
    pd *= norm;
    pc *= norm;
    pt *= norm;
}
// principled.glsl:50-74
static vec3 disney_diffuse_factor(const Material& m, vec3 wo, vec3 wi) {
    vec3 h = normalize(wi + wo);
    float f_wi, f_wo;
    disney_fresnel_s(wi, wo, m.roughness, f_wi, f_wo);
    float r2 = m.roughness * m.roughness;
    float ss = 0.0f;
    float rr = 2.0f * r2 * wi.z * wi.z;
    float f_retro = rr * (f_wi + f_wo * f_wi * f_wo * (rr - 1.0f));
    float f_diff = (1.0f - 0.5f * f_wi) * (1.0f - 0.5f * f_wo);
    if (m.flatness > 0.0f) {
        float fss90 = 0.5f * rr;
        float hdw = dot(h, wi);
        hdw *= hdw;
        float f_ss = mixf(1.0f, fss90, f_wi) * mixf(1.0f, fss90, f_wo);
        ss = 1.25f * (f_ss * (1.0f / (wi.z + wo.z) - 0.5f) + 0.5f);
    }
    float s = mixf(f_diff + f_retro, ss, m.flatness);
    return m.albedo * s * INV_PI;
}
// principled.glsl:76-83
static float clearcoat_factor(const Material& m, vec3 wo, vec3 wi, vec3 h, float& D) {
    const float a2 = 0.25f * 0.25f;
    D = D_GGX_isotropic(mixf(0.1f, 0.001f, m.clearcoat_gloss), h.z);
    float F = fresnel_schlick(0.04f, 1.0f, dot(wi, h));
    float G = G1_iso(a2, wo.z) * G1_iso(a2, wi.z);
    return 0.25f * m.clearcoat * D * F * G;
}
// principled.glsl:96-135
static vec3 sample_principled_brdf(const Material& m, vec3 wo, vec3& wi, float& pdf_w, float& cos_t, vec2 xi,
                                   float eta) {
    if (!has_prop(m.bsdf_props, LT_BSDF_FLAG_REFLECTION)) return v3(0.0f);
    vec2 alpha = calc_aniso(m.roughness, m.anisotropy);
    if (eff_delta2(alpha)) {
        wi = v3(-wo.x, -wo.y, wo.z);
        pdf_w = 1.0f;
        cos_t = wi.z;
        vec3 F = disney_fresnel_m(m, wo, v3(0.0f, 0.0f, 1.0f), wi, eta);
        return F / fabsf(cos_t);
    }
    float D;
    vec3 h = sample_vndf(alpha, wo, xi);
    pdf_w = vndf_pdf_aniso(alpha, wo, h, D);
    wi = reflect(-wo, h);
    if (wo.z * wi.z < 0.0f) return v3(0.0f);
    vec3 F = disney_fresnel_m(m, wo, h, wi, eta);
    pdf_w /= (4.0f * dot(wo, h));
    cos_t = wi.z;
    return 0.25f * D * F * G_corr_aniso(alpha, wo, wi) / (wi.z * wo.z);
}
// principled.glsl:138-163.  sin_t is computed from the incoming cos_theta
// argument (0 on entry from sample_principled), as the reference does.
static vec3 sample_clearcoat(const Material& m, vec3 wo, vec3& wi, float& pdf_w, float& cos_t, vec2 xi) {
    const float a2 = 0.25f * 0.25f;
    float ct = sqrtf(fmaxf(0.0f, (1.0f - powf(a2, 1.0f - xi.x)) / (1.0f - a2)));
    float st = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    float phi = TWO_PI * xi.y;
    vec3 h = v3(st * cosf(phi), st * sinf(phi), ct);
    if (dot(h, wo) < 0.0f) h = h * -1.0f;
    wi = reflect(-wo, h);
    if (dot(wi, wo) < 0.0f) return v3(0.0f);
    float D;
    float fc = clearcoat_factor(m, wo, wi, h, D);
    pdf_w = D / (4.0f * dot(wo, h));
    cos_t = wi.z;
    return v3(fc);
}
// principled.glsl:165-176
static vec3 eval_clearcoat(const Material& m, vec3 wo, vec3 wi, float& pdf_w, float& pdf_rev_w) {
    vec3 h = normalize(wo + wi);
    float D;
    float fc = clearcoat_factor(m, wo, wi, h, D);
    pdf_w = D / (4.0f * dot(wo, h));
    pdf_rev_w = pdf_w;
    return v3(fc);
}
static float clearcoat_pdf(const Material& m, vec3 wo, vec3 wi) {
    vec3 h = normalize(wo + wi);
    float D = D_GGX_isotropic(mixf(0.1f, 0.001f, m.clearcoat_gloss), h.z);
    return D / (4.0f * dot(wo, h));
}
// principled.glsl:183-224
static vec3 eval_principled_brdf(const Material& m, vec3 wo, vec3 wi, float& pdf_w, float& pdf_rev_w, bool ff,
                                 bool rev) {
    pdf_w = 0.0f;
    pdf_rev_w = 0.0f;
    vec2 alpha = calc_aniso(m.roughness, m.anisotropy);
    if (eff_delta2(alpha)) return v3(0.0f);
    if (wo.z * wi.z < 0.0f) return v3(0.0f);
    if (wo.z == 0.0f || wi.z == 0.0f) return v3(0.0f);
    vec3 h = normalize(wo + wi);
    float jac = 1.0f / (4.0f * dot(wo, h));
    float D, Dr;
    pdf_w = vndf_pdf_aniso(alpha, wo, h, D) * jac;
    float eta = ff ? m.ior : 1.0f / m.ior;
    if (rev) pdf_rev_w = vndf_pdf_aniso(alpha, wi, h, Dr) * jac;
    vec3 F = disney_fresnel_m(m, wo, h, wi, eta);
    return 0.25f * D * F * G_corr_aniso(alpha, wo, wi) / (wi.z * wo.z);
}
// principled.glsl:226-252
static float principled_brdf_pdf(const Material& m, vec3 wo, vec3 wi) {
    vec2 alpha = calc_aniso(m.roughness, m.anisotropy);
    if (eff_delta2(alpha)) return 0.0f;
    if (wo.z * wi.z < 0.0f) return 0.0f;
    if (wo.z == 0.0f || wi.z == 0.0f) return 0.0f;
    vec3 h = normalize(wo + wi);
    h = h * sgnf(h.z);
    float D;
    return vndf_pdf_aniso(alpha, wo, h, D) / (4.0f * dot(wo, h));
}
// principled.glsl:254-287
static vec3 sample_principled(const Material& m, vec3 wo, vec3& wi, uint32_t mode, bool ff, float& pdf_w,
                              float& cos_t, vec3 xi) {
    wi = v3(0.0f);
    pdf_w = 0.0f;
    cos_t = 0.0f;
    if (wo.z <= 0.0f) return v3(0.0f);
    float ps, pd, pc, pt;
    float F = fresnel_dielectric(wo.z, m.ior, ff);
    init_probs(m, F, ff, ps, pd, pc, pt);
    float eta = ff ? m.ior : 1.0f / m.ior;
    vec3 f = v3(0.0f);
    float pl = 0.0f;
    if (xi.z < ps) {
        f = sample_principled_brdf(m, wo, wi, pdf_w, cos_t, v2(xi.x, xi.y), eta);
        pl = ps;
    } else if (xi.z > ps && xi.z <= (ps + pc)) {
        f = sample_clearcoat(m, wo, wi, pdf_w, cos_t, v2(xi.x, xi.y));
        pl = pc;
    } else if (xi.z > (ps + pc) && xi.z <= (ps + pc + pd)) {
        // principled.glsl:86-94 sample_disney_diffuse
        wi = sample_hemisphere(v2(xi.x, xi.y));
        cos_t = wi.z;
        pdf_w = cos_t * INV_PI;
        f = fminf(wi.z, wo.z) <= 0.0f ? v3(0.0f) : disney_diffuse_factor(m, wo, wi);
        pl = pd;
    } else if (pt >= 0.0f && xi.z <= (ps + pc + pd + pt)) {
        f = sample_dielectric(m, wo, wi, mode, ff, pdf_w, cos_t, v2(xi.x, xi.y));
        pl = pt;
    }
    pdf_w *= pl;
    return f;
}
// principled.glsl:289-336 (the diffuse lobe adds the running pdf_w into
// pdf_rev_w, principled.glsl:315, kept as is)
static vec3 eval_principled(const Material& m, vec3 wo, vec3 wi, float& pdf_w, float& pdf_rev_w, bool ff,
                            uint32_t mode, bool rev) {
    pdf_w = 0.0f;
    pdf_rev_w = 0.0f;
    float ps, pd, pc, pt;
    float F = fresnel_dielectric(wo.z, m.ior, ff);
    init_probs(m, F, ff, ps, pd, pc, pt);
    vec3 f = v3(0.0f);
    float pdf = 0.0f, pdf_rev = 0.0f;
    float brdf_w = (1.0f - m.spec_trans) * (1.0f - m.metallic);
    float bsdf_w = (1.0f - m.metallic) * m.spec_trans;
    if (ps > 0.0f) {
        f += eval_principled_brdf(m, wo, wi, pdf, pdf_rev, ff, rev);
        pdf *= ps;
        pdf_rev *= ps;
        pdf_w += pdf;
        pdf_rev_w += pdf_rev;
    }
    bool upper = fminf(wi.z, wo.z) > 0.0f;
    if (upper) {
        if (pd > 0.0f) {
            pdf_w += pd * diffuse_pdf(wo, wi);
            pdf_rev_w += pdf_w;
            f += brdf_w * disney_diffuse_factor(m, wo, wi);
        }
        if (pc > 0.0f) {
            f += eval_clearcoat(m, wo, wi, pdf, pdf_rev);
            pdf *= pc;
            pdf_rev *= pc;
            pdf_w += pdf;
            pdf_rev_w += pdf_rev;
        }
    }
    if (pt > 0.0f) {
        f += bsdf_w * eval_dielectric(m, wo, wi, pdf, pdf_rev, ff, mode, rev);
        pdf *= pt;
        pdf_rev *= pt;
        pdf_w += pdf;
        pdf_rev_w += pdf_rev;
    }
    return f;
}
// principled.glsl:338-359
static float principled_pdf(const Material& m, vec3 wo, vec3 wi, bool ff) {
    float pdf = 0.0f;
    float ps, pd, pc, pt;
    float F = fresnel_dielectric(wo.z, m.ior, ff);
    init_probs(m, F, ff, ps, pd, pc, pt);
    if (ps > 0.0f) pdf += ps * principled_brdf_pdf(m, wo, wi);
    bool upper = fminf(wi.z, wo.z) > 0.0f;
    if (upper) {
        if (pd > 0.0f) pdf += pd * diffuse_pdf(wo, wi);
        if (pc > 0.0f) pdf += pc * clearcoat_pdf(m, wo, wi);
    }
    if (pt > 0.0f) pdf += pt * dielectric_pdf(m, wo, wi, ff);
    return pdf;
}

// --------------------------------------------------------- dispatch ----
static vec4 tex_bilinear(const lt_texture& t, vec2 uv) {
    // repeat addressing, texel centres at (i + 0.5) / size
    float fx = uv.x * (float)t.width - 0.5f, fy = uv.y * (float)t.height - 0.5f;
    float x0f = floorf(fx), y0f = floorf(fy);
    float ax = fx - x0f, ay = fy - y0f;
    int w = (int)t.width, h = (int)t.height;
    int x0 = ((int)x0f % w + w) % w, y0 = ((int)y0f % h + h) % h;
    int x1 = (x0 + 1) % w, y1 = (y0 + 1) % h;
    const f32* a = t.data + 4 * ((size_t)y0 * w + x0);
    const f32* b = t.data + 4 * ((size_t)y0 * w + x1);
    const f32* c = t.data + 4 * ((size_t)y1 * w + x0);
    const f32* d = t.data + 4 * ((size_t)y1 * w + x1);
    vec4 r;
    float* rp = &r.x;
    for (int i = 0; i < 4; i++) {
        float top = a[i] * (1.0f - ax) + b[i] * ax;
        float bot = c[i] * (1.0f - ax) + d[i] * ax;
        rp[i] = top * (1.0f - ay) + bot * ay;
    }
    return r;
}

// bsdf_commons.glsl:16-22 (+ per-material roughness texture, an addition)
Material load_material(const lt_scene_desc& s, uint32_t idx, vec2 uv) {
    Material m = material_from(s.materials[idx]);
    if (m.texture_id > -1) {
        vec4 t = tex_bilinear(s.textures[m.texture_id], uv);
        m.albedo = m.albedo * v3(t.x, t.y, t.z);
    }
    if (s.roughness_texture_ids && s.roughness_texture_ids[idx] > -1) {
        vec4 t = tex_bilinear(s.textures[s.roughness_texture_ids[idx]], uv);
        m.roughness = t.x;
    }
    return m;
}

// bsdf_commons.glsl:26-66
float bsdf_pdf(const Material& m, vec3 n_s, vec3 wo, vec3 wi, bool ff) {
    vec3 T, B;
    branchless_onb(n_s, T, B);
    wo = to_local(wo, T, B, n_s);
    wi = to_local(wi, T, B, n_s);
    switch (m.bsdf_type) {
        case LT_BSDF_DIFFUSE: return diffuse_pdf(wo, wi);
        case LT_BSDF_MIRROR: return 0.0f;
        case LT_BSDF_GLASS: return 0.0f;
        case LT_BSDF_DIELECTRIC: return dielectric_pdf(m, wo, wi, ff);
        case LT_BSDF_CONDUCTOR: return conductor_pdf(m, wo, wi);
        case LT_BSDF_PRINCIPLED: return principled_pdf(m, wo, wi, ff);
        default: return 0.0f;
    }
}

// bsdf_commons.glsl:68-114
vec3 sample_bsdf(vec3 n_s, vec3 wo, const Material& m, uint32_t mode, bool ff, vec3& wi, float& pdf_w,
                 float& cos_t, vec3 r) {
    vec3 f = v3(0.0f);
    pdf_w = 0.0f;
    cos_t = 0.0f;
    wi = v3(0.0f);
    vec3 T, B;
    branchless_onb(n_s, T, B);
    wo = to_local(wo, T, B, n_s);
    vec2 xy = {r.x, r.y};
    switch (m.bsdf_type) {
        case LT_BSDF_DIFFUSE: f = sample_diffuse(m, wo, wi, pdf_w, cos_t, xy); break;
        case LT_BSDF_MIRROR: f = sample_mirror(v3(0.0f, 0.0f, 1.0f), wo, wi, pdf_w, cos_t); break;
        case LT_BSDF_GLASS: f = sample_glass(m, v3(0.0f, 0.0f, 1.0f), wo, wi, pdf_w, cos_t, mode, ff); break;
        case LT_BSDF_DIELECTRIC: f = sample_dielectric(m, wo, wi, mode, ff, pdf_w, cos_t, xy); break;
        case LT_BSDF_CONDUCTOR: f = sample_conductor(m, wo, wi, pdf_w, cos_t, xy); break;
        case LT_BSDF_PRINCIPLED: f = sample_principled(m, wo, wi, mode, ff, pdf_w, cos_t, r); break;
        default: break;
    }
    wi = to_world(wi, T, B, n_s);
    return f;
}

// bsdf_commons.glsl:123-167
vec3 eval_bsdf(vec3 n_s, vec3 wo, const Material& m, uint32_t mode, bool ff, vec3 wi, float& pdf_w,
               float& pdf_rev_w, bool rev) {
    pdf_w = 0.0f;
    pdf_rev_w = 0.0f;
    vec3 f = v3(0.0f);
    vec3 T, B;
    branchless_onb(n_s, T, B);
    wo = to_local(wo, T, B, n_s);
    wi = to_local(wi, T, B, n_s);
    switch (m.bsdf_type) {
        case LT_BSDF_DIFFUSE: f = eval_diffuse(m, wo, wi, pdf_w, pdf_rev_w); break;
        case LT_BSDF_MIRROR: break;
        case LT_BSDF_GLASS: break;
        case LT_BSDF_DIELECTRIC: f = eval_dielectric(m, wo, wi, pdf_w, pdf_rev_w, ff, mode, rev); break;
        case LT_BSDF_CONDUCTOR: f = eval_conductor(m, wo, wi, pdf_w, pdf_rev_w, rev); break;
        case LT_BSDF_PRINCIPLED: f = eval_principled(m, wo, wi, pdf_w, pdf_rev_w, ff, mode, rev); break;
        default: break;
    }
    return f;
}

// ------------------------------------------------------------- lights ----
static inline uint32_t light_type(uint32_t flags) { return flags & 7u; }

// commons.glsl:112-149
TriangleRecord sample_triangle(const lt_scene_desc& s, const lt_prim_mesh_info& pi, vec2 rands, uint32_t tri,
                               const f32* world, vec2& uv) {
    TriangleRecord r;
    uint32_t io = pi.index_offset + 3 * tri;
    uint32_t vo = pi.vertex_offset;
    const lt_vertex& a = s.vertices[s.indices[io + 0] + vo];
    const lt_vertex& b = s.vertices[s.indices[io + 1] + vo];
    const lt_vertex& c = s.vertices[s.indices[io + 2] + vo];
    vec3 v0 = load3(a.pos), v1 = load3(b.pos), v2p = load3(c.pos);
    vec3 n0 = load3(a.normal), n1 = load3(b.normal), n2 = load3(c.normal);
    float sr = sqrtf(rands.x);
    uv = {1.0f - sr, rands.y * sr};
    vec3 bc = {1.0f - uv.x - uv.y, uv.x, uv.y};
    // The reference transforms the edges with w = 1 (commons.glsl:136-137),
    // which only matters for a translated light; they are vectors (w = 0)
    // here, so the pdf is the world triangle's.  The normal goes through
    // transpose(inverse(world)) (commons.glsl:130,144).
    vec4 e0 = mul(world, vec4{v1.x - v0.x, v1.y - v0.y, v1.z - v0.z, 0.0f});
    vec4 e1 = mul(world, vec4{v2p.x - v0.x, v2p.y - v0.y, v2p.z - v0.z, 0.0f});
    vec3 pos = v0 * bc.x + v1 * bc.y + v2p * bc.z;
    vec3 nrm = normalize(n0 * bc.x + n1 * bc.y + n2 * bc.z);
    vec4 wp = mul(world, vec4{pos.x, pos.y, pos.z, 1.0f});
    f32 xf[24];
    if (xf_record(world, xf))
        nrm = v3(xf[12] * nrm.x + xf[13] * nrm.y + xf[14] * nrm.z, xf[15] * nrm.x + xf[16] * nrm.y + xf[17] * nrm.z,
                 xf[18] * nrm.x + xf[19] * nrm.y + xf[20] * nrm.z);
    r.n_s = normalize(nrm);
    r.triangle_pdf = 2.0f / length(cross(v3(e0.x, e0.y, e0.z), v3(e1.x, e1.y, e1.z)));
    r.pos = v3(wp.x, wp.y, wp.z);
    return r;
}

// commons.glsl:264-281 spot falloff (30 deg width, 25 deg falloff start)
static float spot_falloff(float cos_from_light) {
    const float cos_width = cosf(PI / 6.0f);
    const float cos_faloff = cosf(25.0f * PI / 180.0f);
    if (cos_from_light < cos_width) return 0.0f;
    if (cos_from_light >= cos_faloff) return 1.0f;
    float d = (cos_from_light - cos_width) / (cos_faloff - cos_width);
    return (d * d) * (d * d);
}
// per-light cone (an addition of this core; the reference fixes 30 / 25 degrees)
static float spot_falloff_c(float cos_from_light, float cos_width, float cos_faloff) {
    if (cos_from_light < cos_width) return 0.0f;
    if (cos_from_light >= cos_faloff) return 1.0f;
    float d = (cos_from_light - cos_width) / (cos_faloff - cos_width);
    return (d * d) * (d * d);
}
static float uniform_cone_pdf(float cos_max) { return 1.0f / (TWO_PI * (1.0f - cos_max)); }

// commons.glsl:224-300 (+ LT_LIGHT_POINT: isotropic point light, an addition)
vec3 sample_light_Li(const lt_scene_desc& s, vec4 rp, vec3 p, int num_lights, float& pdf_pos_w, vec3& wi,
                     float& wi_len, float& pdf_pos_a, float& cos_from_light, LightRecord& rec, vec3& n, vec3& pos,
                     float& pdf_pos_dir_w) {
    rec.light_idx = (uint32_t)(rp.x * (float)num_lights);
    const lt_light& L = s.lights[rec.light_idx];
    uint32_t lt = light_type(L.light_flags);
    vec3 Le = v3(0.0f);
    rec.flags = L.light_flags;
    pdf_pos_w = 0.0f;
    wi = v3(0.0f);
    wi_len = 0.0f;
    pdf_pos_a = 0.0f;
    cos_from_light = 0.0f;
    switch (lt) {
        case LT_LIGHT_AREA: {
            const lt_prim_mesh_info& pi = s.prim_infos[L.prim_mesh_idx];
            uint32_t mat_idx = pi.material_index;
            rec.triangle_idx = (uint32_t)(rp.y * (float)L.num_triangles);
            TriangleRecord r = sample_triangle(s, pi, v2(rp.z, rp.w), rec.triangle_idx, L.world_matrix, rec.bary);
            Material lm = load_material(s, mat_idx, rec.bary);
            wi = r.pos - p;
            float d2 = dot(wi, wi);
            wi_len = sqrtf(d2);
            wi /= wi_len;
            cos_from_light = fabsf(dot(r.n_s, -wi));
            Le = lm.emissive_factor;
            pdf_pos_a = r.triangle_pdf;
            pdf_pos_w = pdf_pos_a * d2 / cos_from_light;
            pdf_pos_dir_w = cos_from_light * INV_PI * r.triangle_pdf;
            rec.instance_idx = L.prim_mesh_idx;
            n = r.n_s;
            pos = r.pos;
        } break;
        case LT_LIGHT_SPOT: {
            wi = load3(L.pos) - p;
            float d2 = dot(wi, wi);
            wi_len = sqrtf(d2);
            wi /= wi_len;
            const vec3 ld = normalize(load3(L.to) - load3(L.pos));
            cos_from_light = dot(-wi, ld);
            float fo = s.spot_cones ? spot_falloff_c(cos_from_light, s.spot_cones[2 * rec.light_idx],
                                                     s.spot_cones[2 * rec.light_idx + 1])
                                    : spot_falloff(cos_from_light);
            pdf_pos_a = 1.0f;
            pdf_pos_w = d2;
            pdf_pos_dir_w = uniform_cone_pdf(s.spot_cones ? s.spot_cones[2 * rec.light_idx] : cosf(PI / 6.0f));
            Le = load3(L.L) * fo;
            n = -wi;
            pos = load3(L.pos);
        } break;
        case LT_LIGHT_DIRECTIONAL: {
            const vec3 dir = normalize(load3(L.pos) - load3(L.to));
            const vec3 lp = p + dir * (2.0f * L.world_radius);
            wi = lp - p;
            wi_len = length(wi);
            wi /= wi_len;
            pdf_pos_a = 1.0f;
            pdf_pos_w = 1.0f;
            pdf_pos_dir_w = INV_PI / (L.world_radius * L.world_radius);
            Le = load3(L.L);
            cos_from_light = 1.0f;
            n = -wi;
            pos = lp;
        } break;
        case LT_LIGHT_POINT: {
            wi = load3(L.pos) - p;
            float d2 = dot(wi, wi);
            wi_len = sqrtf(d2);
            wi /= wi_len;
            cos_from_light = 1.0f;
            pdf_pos_a = 1.0f;
            pdf_pos_w = d2;
            pdf_pos_dir_w = 1.0f / (4.0f * PI);
            Le = load3(L.L);
            n = -wi;
            pos = load3(L.pos);
        } break;
        default: break;
    }
    return Le;
}

}  // namespace ora
