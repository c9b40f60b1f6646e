// dev_bsdf.h — device BSDFs (Lambert, mirror, glass, dielectric, conductor,
// Disney principled) following src/shaders/bsdf_commons.glsl and
// src/shaders/bsdf/*.glsl.  Local shading frame: z = n_s.
#pragma once
#include "dev_math.h"

namespace ltd {

struct Mat {
    f3 albedo;
    float ior;
    f3 emissive;
    uint32_t bsdf_type, bsdf_props;
    f3 k;
    int32_t texture_id;
    float roughness, diffuse_trans, spec_trans, metallic, specular_tint, sheen_tint, clearcoat, clearcoat_gloss,
        sheen, subsurface, flatness, anisotropy;
    uint32_t thin;
};
static_assert(sizeof(Mat) == sizeof(lt_material), "Material layout");

LT_DEV bool has_prop(uint32_t props, uint32_t f) { return (props & f) != 0; }
LT_DEV bool is_specular(const Mat& m) { return (m.bsdf_props & LT_BSDF_FLAG_SPECULAR) != 0; }
LT_DEV float sgnf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
LT_DEV bool eff_delta(float a) { return a <= 0.0064f; }                 // sampling_commons.glsl:18
LT_DEV bool eff_delta2(f2 a) { return fminf(a.x, a.y) <= 0.0064f; }    // sampling_commons.glsl:19

// sampling_commons.glsl:22-41 (TIR reflects; always returns true)
LT_DEV void refract_(f3 n, f3 wo, bool ff, float eta, uint32_t mode, f3& wi, f3& f, float& inv_eta) {
    float cos_i = dot(n, wo);
    inv_eta = ff ? 1.0f / eta : eta;
    const float sin2_t = inv_eta * inv_eta * (1.0f - cos_i * cos_i);
    if (sin2_t >= 1.0f) {
        wi = reflect(-wo, n);
    } else {
        const float cos_t = sqrtf(1.0f - sin2_t);
        wi = -inv_eta * wo + (inv_eta * cos_i - cos_t) * n;
    }
    f = mode == 1 ? mk3(inv_eta * inv_eta) : mk3(1.0f);
}
// sampling_commons.glsl:48-65
LT_DEV float fresnel_dielectric(float cos_i, float eta, bool ff) {
    cos_i = clampf(cos_i, -1.0f, 1.0f);
    if (!ff) eta = 1.0f / eta;
    if (cos_i < 0.0f) {
        eta = 1.0f / eta;
        cos_i = -cos_i;
    }
    float sin2_t = (1.0f - cos_i * cos_i) / (eta * eta);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = sqrtf(1.0f - sin2_t);
    float rpa = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float rpe = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (rpa * rpa + rpe * rpe);
}
// sampling_commons.glsl:68-82
LT_DEV float fresnel_conductor(float cos_i, float eta, float k) {
    float c2 = cos_i * cos_i;
    float s2 = fmaxf(1.0f - c2, 0.0f);
    float s4 = s2 * s2;
    float inner = eta * eta - k * k - s2;
    float ab = sqrtf(fmaxf(inner * inner + 4.0f * eta * eta * k * k, 0.0f));
    float a = sqrtf(fmaxf((ab + inner) * 0.5f, 0.0f));
    float rs = ((ab + c2) - (2.0f * a * cos_i)) / ((ab + c2) + (2.0f * a * cos_i));
    float rp = ((c2 * ab + s4) - (2.0f * a * cos_i * s2)) / ((c2 * ab + s4) + (2.0f * a * cos_i * s2));
    return 0.5f * (rs + rs * rp);
}
LT_DEV f3 fresnel_conductor3(float c, f3 eta, f3 k) {
    return {fresnel_conductor(c, eta.x, k.x), fresnel_conductor(c, eta.y, k.y), fresnel_conductor(c, eta.z, k.z)};
}
LT_DEV float pow5m(float ns) {
    float m = fmaxf(1.0f - ns, 0.0f);
    return powf(m, 5.0f);
}
// sampling_commons.glsl:84-92
LT_DEV float fresnel_schlick(float f0, float f90, float ns) { return f0 + (f90 - f0) * pow5m(ns); }
LT_DEV f3 fresnel_schlick3(f3 f0, f3 f90, float ns) { return f0 + (f90 - f0) * pow5m(ns); }
// sampling_commons.glsl:102-110
LT_DEV void disney_fresnel_s(f3 wi, f3 wo, float roughness, float& f_wi, float& f_wo) {
    f3 h = normalize(wi + wo);
    float wh = dot(wo, h);
    float fd90 = 0.5f + 2.0f * wh * wh * roughness;
    f_wi = fresnel_schlick(1.0f, fd90, wi.z);
    f_wo = fresnel_schlick(1.0f, fd90, wo.z);
}
// sampling_commons.glsl:119-139
LT_DEV f3 disney_fresnel_m(const Mat& m, f3 wo, f3 h, float eta) {
    float wh = dot(wo, h);
    float lum = luminance(m.albedo);
    f3 tint = lum > 0.0f ? m.albedo / lum : mk3(1.0f);
    float r0 = (eta - 1.0f) / (eta + 1.0f);
    f3 R0 = (r0 * r0) * mix(mk3(1.0f), tint, m.specular_tint);
    R0 = mix(R0, m.albedo, m.metallic);
    float frd = fresnel_dielectric(wh, eta, true);
    f3 frm = fresnel_schlick3(R0, mk3(1.0f), wh);
    return mix(mk3(frd), frm, m.metallic);
}

// ---- microfacet_commons.glsl ----
LT_DEV float smith_lambda_iso(float a2, float c) {
    if (c == 0.0f) return 0.0f;
    float c2 = c * c;
    float t2 = fmaxf(1.0f - c2, 0.0f) / c2;
    return 0.5f * (sqrtf(1.0f + a2 * t2) - 1.0f);
}
LT_DEV float smith_lambda_aniso(f3 w, f2 al) {
    float c2 = w.z * w.z;
    float s2 = fmaxf(1.0f - c2, 0.0f);
    float t2 = s2 / c2;
    if (isinf(t2)) return 0.0f;
    float cx = s2 == 0.0f ? 1.0f : clampf(w.x * w.x, 0.0f, 1.0f) / s2;
    float cy = s2 == 0.0f ? 0.0f : clampf(w.y * w.y, 0.0f, 1.0f) / s2;
    float a2 = cx * (al.x * al.x) + cy * (al.y * al.y);
    return 0.5f * (sqrtf(1.0f + a2 * t2) - 1.0f);
}
LT_DEV float G_corr_iso(float alpha, f3 wo, f3 wi) {
    float a2 = alpha * alpha;
    return 1.0f / (1.0f + smith_lambda_iso(a2, wo.z) + smith_lambda_iso(a2, wi.z));
}
LT_DEV float G_corr_aniso(f2 a, f3 wo, f3 wi) {
    return 1.0f / (1.0f + smith_lambda_aniso(wo, a) + smith_lambda_aniso(wi, a));
}
LT_DEV float D_ggx_aniso(f2 al, f3 h) {
    float c2 = h.z * h.z;
    float s2 = fmaxf(1.0f - c2, 0.0f);
    float t2 = s2 / c2;
    if (isinf(t2)) return 0.0f;
    float c4 = c2 * c2;
    if (c4 < 1e-16f) return 0.0f;
    float px = s2 == 0.0f ? 1.0f : clampf(h.x * h.x, 0.0f, 1.0f) / s2;
    float py = s2 == 0.0f ? 0.0f : clampf(h.y * h.y, 0.0f, 1.0f) / s2;
    float e = t2 * (px / (al.x * al.x) + py / (al.y * al.y));
    return 1.0f / (PI * al.x * al.y * c4 * (1.0f + e) * (1.0f + e));
}
LT_DEV float D_ggx_iso(float a2, float c) {
    float d = ((c * a2 - c) * c + 1.0f);
    return a2 / (d * d * PI);
}
LT_DEV float G1_iso(float a2, float c) {
    float c2 = c * c;
    float t2 = fmaxf(1.0f - c2, 0.0f) / c2;
    return 2.0f / (1.0f + sqrtf(1.0f + a2 * t2));
}
LT_DEV float vndf_pdf_iso(float alpha, f3 wo, f3 h, float& D) {
    D = 0.0f;
    if (wo.z <= 0.0f) return 0.0f;
    float a2 = alpha * alpha;
    float G1 = G1_iso(a2, wo.z);
    D = D_ggx_iso(a2, h.z);
    return G1 * D * fmaxf(0.0f, dot(wo, h)) / fabsf(wo.z);
}
LT_DEV float vndf_pdf_aniso(f2 al, f3 wo, f3 h, float& D) {
    D = 0.0f;
    if (wo.z <= 0.0f) return 0.0f;
    float G1 = 1.0f / (1.0f + smith_lambda_aniso(wo, al));
    D = D_ggx_aniso(al, h);
    return G1 * D * fmaxf(0.0f, dot(wo, h)) / fabsf(wo.z);
}
// microfacet_commons.glsl:99-125 (Dupuy & Benyoub spherical caps)
LT_DEV f3 sample_vndf(f2 al, f3 wo, float r0, float r1) {
    f3 wh = normalize(mk3(al.x * wo.x, al.y * wo.y, wo.z));
    float phi = 2.0f * PI * r0;
    float z = ((1.0f - r1) * (1.0f + wh.z)) - wh.z;
    float st = sqrtf(clampf(1.0f - z * z, 0.0f, 1.0f));
    float s, c;
    sincosf(phi, &s, &c);
    f3 nh = mk3(st * c, st * s, z) + wh;
    return normalize(mk3(al.x * nh.x, al.y * nh.y, fmaxf(0.0f, nh.z)));
}

// ---- diffuse.glsl (Lambertian) ----
LT_DEV float diffuse_pdf(f3 wo, f3 wi) { return fminf(wi.z, wo.z) <= 0.0f ? 0.0f : wi.z * INV_PI; }

// ---- dielectric.glsl ----
LT_DEV float thin_rough(float ior, float r) { return clampf((0.65f * ior - 0.35f) * r, 0.0f, 1.0f); }

LT_DEV f3 sample_dielectric(const Mat& m, f3 wo, f3& wi, uint32_t mode, bool ff, float& pdf_w, float& cos_t,
                            float r0, float r1) {
    wi = mk3(0.0f);
    pdf_w = 0.0f;
    cos_t = 0.0f;
    if (wo.z <= 0.0f) return mk3(0.0f);
    float rough = m.thin == 1 ? thin_rough(m.ior, m.roughness) : m.roughness;
    float alpha = rough * rough;
    bool hr = has_prop(m.bsdf_props, LT_BSDF_FLAG_REFLECTION), ht = has_prop(m.bsdf_props, LT_BSDF_FLAG_TRANSMISSION);
    if ((m.ior == 1.0f && m.thin == 0) || eff_delta(alpha)) {
        float F = fresnel_dielectric(wo.z, m.ior, ff);
        if (!hr && !ht) return mk3(0.0f);
        float pr = hr ? F : 0.0f, pt = ht ? 1.0f - F : 0.0f;
        f3 f;
        if (((pr + pt) * r0) < pr) {
            wi = mk3(-wo.x, -wo.y, wo.z);
            cos_t = wi.z;
            f = mk3(F) / fabsf(cos_t);
            pdf_w = pr / (pr + pt);
        } else {
            float ie;
            refract_(mk3(0.0f, 0.0f, 1.0f), wo, ff, m.ior, mode, wi, f, ie);
            cos_t = wi.z;
            f = f * (1.0f - F) / fabsf(cos_t);
            pdf_w = pt / (pr + pt);
        }
        return f;
    }
    if (!hr && !ht) return mk3(0.0f);
    float D;
    f3 h = sample_vndf(f2{alpha, alpha}, wo, r0, r1);
    pdf_w = vndf_pdf_iso(alpha, wo, h, D);
    float F = fresnel_dielectric(dot(wo, h), m.ior, ff);
    float pr = hr ? F : 0.0f, pt = ht ? (1.0f - F) : 0.0f;
    f3 f;
    if (((pr + pt) * r0) < pr) {
        wi = reflect(-wo, h);
        if (wo.z * wi.z < 0.0f) return mk3(0.0f);
        pdf_w = pdf_w * (pr / (pr + pt)) / (4.0f * fabsf(dot(wo, h)));
        f = mk3(0.25f * D * F * G_corr_iso(alpha, wo, wi) / (wi.z * wo.z));
    } else {
        f3 base;
        float ie = 0.0f;
        if (m.thin == 1) {
            wi = reflect(-wo, h);
            base = vsqrt(m.albedo);
            ie = m.ior;
            f = mk3(0.0f);
        } else {
            refract_(h, wo, ff, m.ior, mode, wi, f, ie);
            base = m.albedo;
        }
        if (wo.z * wi.z > 0.0f || wi.z > 0.0f) return mk3(0.0f);
        float jd = dot(wi, h) + dot(wo, h) * ie;
        jd = jd * jd;
        float jac = fabsf(dot(wi, h)) / jd;
        pdf_w = pdf_w * (pt / (pr + pt)) * jac;
        f = base * f * (1.0f - F) * D * G_corr_iso(alpha, wo, wi) * fabsf(dot(wi, h) * dot(wo, h) / (wi.z * wo.z * jd));
    }
    cos_t = wi.z;
    return f;
}

// dielectric.glsl:110-189 (rough transmission returns 0: the reference returns
// the shadowed, never-assigned outer f there, dielectric.glsl:171)
LT_DEV f3 eval_dielectric(const Mat& m, f3 wo, f3 wi, float& pdf_w, float& pdf_rev_w, bool ff, bool rev) {
    pdf_w = 0.0f;
    pdf_rev_w = 0.0f;
    float rough = m.thin == 1 ? thin_rough(m.ior, m.roughness) : m.roughness;
    float alpha = rough * rough;
    if (alpha == 0.0f || m.ior == 1.0f) return mk3(0.0f);
    bool hr = has_prop(m.bsdf_props, LT_BSDF_FLAG_REFLECTION), ht = has_prop(m.bsdf_props, LT_BSDF_FLAG_TRANSMISSION);
    if (!hr && !ht) return mk3(0.0f);
    bool is_refl = wi.z * wo.z > 0.0f;
    float eta = is_refl ? 1.0f : (ff ? m.ior : 1.0f / m.ior);
    f3 h = normalize(wo + wi * eta);
    h = h * sgnf(h.z);
    if (wi.z == 0.0f || wo.z == 0.0f || dot(h, h) == 0.0f) return mk3(0.0f);
    if (dot(wi, h) * wi.z < 0.0f || dot(wo, h) * wo.z < 0.0f) return mk3(0.0f);
    float F = fresnel_dielectric(dot(wo, h), m.ior, ff);
    float pr = hr ? F : 0.0f, pt = ht ? (1.0f - F) : 0.0f;
    float D;
    if (rev) pdf_rev_w = vndf_pdf_iso(alpha, wi, -h, D);
    pdf_w = vndf_pdf_iso(alpha, wo, h, D);
    float G = G_corr_iso(alpha, wo, wi);
    f3 f = mk3(0.0f);
    if (is_refl) {
        float prr = pr / (pr + pt);
        pdf_w = pdf_w * (1.0f / (4.0f * fabsf(dot(wo, h)))) * prr;
        if (rev) pdf_rev_w = pdf_rev_w * (1.0f / (4.0f * fabsf(dot(wi, h)))) * prr;
        f = mk3(0.25f * D * G * F / fabsf(wo.z * wi.z));
    } else {
        float jd = dot(wi, h) + dot(wo, h) / eta;
        jd = jd * jd;
        float prt = pt / (pr + pt);
        pdf_w = pdf_w * (fabsf(dot(wi, h)) / jd) * prt;
        if (rev) {
            f3 hr2 = -h;
            float jdr = eta * dot(wi, hr2) + dot(wo, hr2);
            jdr *= jdr;
            pdf_rev_w = pdf_rev_w * (fabsf(dot(wo, hr2)) / jdr) * prt;
        }
    }
    return f;
}

LT_DEV float dielectric_pdf(const Mat& m, f3 wo, f3 wi, bool ff) {
    float rough = m.thin == 1 ? thin_rough(m.ior, m.roughness) : m.roughness;
    float alpha = rough * rough;
    if (alpha == 0.0f || m.ior == 1.0f) return 0.0f;
    bool hr = has_prop(m.bsdf_props, LT_BSDF_FLAG_REFLECTION), ht = has_prop(m.bsdf_props, LT_BSDF_FLAG_TRANSMISSION);
    if (!hr && !ht) return 0.0f;
    bool is_refl = wi.z * wo.z > 0.0f;
    float eta = is_refl ? 1.0f : (ff ? m.ior : 1.0f / m.ior);
    f3 h = normalize(wo + wi * eta);
    h = h * sgnf(h.z);
    if (wi.z == 0.0f || wo.z == 0.0f || dot(h, h) == 0.0f) return 0.0f;
    if (dot(wi, h) * wi.z < 0.0f || dot(wo, h) * wo.z < 0.0f) return 0.0f;
    float F = fresnel_dielectric(dot(wo, h), m.ior, ff);
    float pr = hr ? F : 0.0f, pt = ht ? (1.0f - F) : 0.0f;
    float D;
    float pdf = vndf_pdf_iso(alpha, wo, h, D);
    if (is_refl) return pdf * (1.0f / (4.0f * fabsf(dot(wo, h)))) * (pr / (pr + pt));
    float jd = dot(wi, h) + dot(wo, h) / eta;
    jd = jd * jd;
    return pdf * (fabsf(dot(wi, h)) / jd) * (pt / (pr + pt));
}

// ---- conductor.glsl ----
LT_DEV f3 sample_conductor(const Mat& m, f3 wo, f3& wi, float& pdf_w, float& cos_t, float r0, float r1) {
    wi = mk3(0.0f);
    pdf_w = 0.0f;
    cos_t = 0.0f;
    if (wo.z <= 0.0f) return mk3(0.0f);
    float alpha = m.roughness * m.roughness;
    if (!has_prop(m.bsdf_props, LT_BSDF_FLAG_REFLECTION)) return mk3(0.0f);
    if (eff_delta(alpha)) {
        wi = mk3(-wo.x, -wo.y, wo.z);
        pdf_w = 1.0f;
        cos_t = wi.z;
        return fresnel_conductor3(cos_t, m.albedo, m.k) / fabsf(cos_t);
    }
    float D;
    f3 h = sample_vndf(f2{alpha, alpha}, wo, r0, r1);
    pdf_w = vndf_pdf_iso(alpha, wo, h, D);
    f3 F = fresnel_conductor3(dot(wo, h), m.albedo, m.k);
    wi = reflect(-wo, h);
    if (wo.z * wi.z < 0.0f) return mk3(0.0f);
    pdf_w /= (4.0f * dot(wo, h));
    cos_t = wi.z;
    return 0.25f * D * F * G_corr_iso(alpha, wo, wi) / (wi.z * wo.z);
}
LT_DEV f3 eval_conductor(const Mat& m, f3 wo, f3 wi, float& pdf_w, float& pdf_rev_w, bool rev) {
    pdf_w = 0.0f;
    pdf_rev_w = 0.0f;
    float alpha = m.roughness * m.roughness;
    if (eff_delta(alpha)) return mk3(0.0f);
    if (wo.z * wi.z < 0.0f) return mk3(0.0f);
    if (wo.z == 0.0f || wi.z == 0.0f) return mk3(0.0f);
    f3 h = normalize(wo + wi);
    h = h * sgnf(h.z);
    float jac = 1.0f / (4.0f * dot(wo, h));
    float D, Dr;
    pdf_w = vndf_pdf_iso(alpha, wo, h, D) * jac;
    if (rev) pdf_rev_w = vndf_pdf_iso(alpha, wi, h, Dr) * jac;
    f3 F = fresnel_conductor3(dot(wo, h), m.albedo, m.k);
    return 0.25f * D * F * G_corr_iso(alpha, wo, wi) / (wi.z * wo.z);
}
LT_DEV float conductor_pdf(const Mat& m, f3 wo, f3 wi) {
    float alpha = m.roughness * m.roughness;
    if (eff_delta(alpha)) return 0.0f;
    if (wo.z * wi.z < 0.0f) return 0.0f;
    if (wo.z == 0.0f || wi.z == 0.0f) return 0.0f;
    f3 h = normalize(wo + wi);
    h = h * sgnf(h.z);
    float D;
    return vndf_pdf_iso(alpha, wo, h, D) / (4.0f * dot(wo, h));
}

// ---- principled.glsl ----
LT_DEV f2 calc_aniso(float roughness, float aniso) {
    float aspect = sqrtf(1.0f - 0.9f * aniso);
    float r2 = roughness * roughness;
    return {fmaxf(0.001f, r2 / aspect), fmaxf(0.001f, r2 * aspect)};
}
LT_DEV void init_probs(const Mat& m, float Fd, bool ff, float& ps, float& pd, float& pc, float& pt) {
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
LT_DEV f3 disney_diffuse_factor(const Mat& m, f3 wo, f3 wi) {
    f3 h = normalize(wi + wo);
    float f_wi, f_wo;
    disney_fresnel_s(wi, wo, m.roughness, f_wi, f_wo);
    float r2 = m.roughness * m.roughness;
    float ss = 0.0f;
    float rr = 2.0f * r2 * wi.z * wi.z;
    float f_retro = rr * (f_wi + f_wo * f_wi * f_wo * (rr - 1.0f));
    float f_diff = (1.0f - 0.5f * f_wi) * (1.0f - 0.5f * f_wo);
    if (m.flatness > 0.0f) {
        float fss90 = 0.5f * rr;
        float f_ss = mixf(1.0f, fss90, f_wi) * mixf(1.0f, fss90, f_wo);
        ss = 1.25f * (f_ss * (1.0f / (wi.z + wo.z) - 0.5f) + 0.5f);
        (void)h;
    }
    return m.albedo * mixf(f_diff + f_retro, ss, m.flatness) * INV_PI;
}
LT_DEV float clearcoat_factor(const Mat& m, f3 wo, f3 wi, f3 h, float& D) {
    const float a2 = 0.25f * 0.25f;
    D = D_ggx_iso(mixf(0.1f, 0.001f, m.clearcoat_gloss), h.z);
    float F = fresnel_schlick(0.04f, 1.0f, dot(wi, h));
    float G = G1_iso(a2, wo.z) * G1_iso(a2, wi.z);
    return 0.25f * m.clearcoat * D * F * G;
}
LT_DEV f3 sample_principled(const Mat& m, f3 wo, f3& wi, uint32_t mode, bool ff, float& pdf_w, float& cos_t,
                            float r0, float r1, float r2) {
    wi = mk3(0.0f);
    pdf_w = 0.0f;
    cos_t = 0.0f;
    if (wo.z <= 0.0f) return mk3(0.0f);
    float ps, pd, pc, pt;
    float Fd = fresnel_dielectric(wo.z, m.ior, ff);
    init_probs(m, Fd, ff, ps, pd, pc, pt);
    float eta = ff ? m.ior : 1.0f / m.ior;
    f3 f = mk3(0.0f);
    float pl = 0.0f;
    if (r2 < ps) {
        // principled.glsl:96-135
        pl = ps;
        if (has_prop(m.bsdf_props, LT_BSDF_FLAG_REFLECTION)) {
            f2 al = calc_aniso(m.roughness, m.anisotropy);
            if (eff_delta2(al)) {
                wi = mk3(-wo.x, -wo.y, wo.z);
                pdf_w = 1.0f;
                cos_t = wi.z;
                f = disney_fresnel_m(m, wo, mk3(0.0f, 0.0f, 1.0f), eta) / fabsf(cos_t);
            } else {
                float D;
                f3 h = sample_vndf(al, wo, r0, r1);
                pdf_w = vndf_pdf_aniso(al, wo, h, D);
                wi = reflect(-wo, h);
                if (wo.z * wi.z < 0.0f) {
                    f = mk3(0.0f);
                } else {
                    f3 F = disney_fresnel_m(m, wo, h, eta);
                    pdf_w /= (4.0f * dot(wo, h));
                    cos_t = wi.z;
                    f = 0.25f * D * F * G_corr_aniso(al, wo, wi) / (wi.z * wo.z);
                }
            }
        }
    } else if (r2 > ps && r2 <= (ps + pc)) {
        // principled.glsl:138-163 (sin_t from the incoming cos_theta = 0)
        pl = pc;
        const float a2 = 0.25f * 0.25f;
        float ct = sqrtf(fmaxf(0.0f, (1.0f - powf(a2, 1.0f - r0)) / (1.0f - a2)));
        float st = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
        float s, c;
        sincosf(TWO_PI * r1, &s, &c);
        f3 h = mk3(st * c, st * s, ct);
        if (dot(h, wo) < 0.0f) h = h * -1.0f;
        wi = reflect(-wo, h);
        if (dot(wi, wo) >= 0.0f) {
            float D;
            float fc = clearcoat_factor(m, wo, wi, h, D);
            pdf_w = D / (4.0f * dot(wo, h));
            cos_t = wi.z;
            f = mk3(fc);
        }
    } else if (r2 > (ps + pc) && r2 <= (ps + pc + pd)) {
        pl = pd;
        wi = sample_hemisphere(r0, r1);
        cos_t = wi.z;
        pdf_w = cos_t * INV_PI;
        f = fminf(wi.z, wo.z) <= 0.0f ? mk3(0.0f) : disney_diffuse_factor(m, wo, wi);
    } else if (pt >= 0.0f && r2 <= (ps + pc + pd + pt)) {
        pl = pt;
        f = sample_dielectric(m, wo, wi, mode, ff, pdf_w, cos_t, r0, r1);
    }
    pdf_w *= pl;
    return f;
}
LT_DEV f3 eval_principled(const Mat& m, f3 wo, f3 wi, float& pdf_w, float& pdf_rev_w, bool ff, bool rev) {
    pdf_w = 0.0f;
    pdf_rev_w = 0.0f;
    float ps, pd, pc, pt;
    float Fd = fresnel_dielectric(wo.z, m.ior, ff);
    init_probs(m, Fd, ff, ps, pd, pc, pt);
    f3 f = mk3(0.0f);
    float pdf = 0.0f, pdf_rev = 0.0f;
    float brdf_w = (1.0f - m.spec_trans) * (1.0f - m.metallic);
    float bsdf_w = (1.0f - m.metallic) * m.spec_trans;
    if (ps > 0.0f) {
        // principled.glsl:183-224
        f2 al = calc_aniso(m.roughness, m.anisotropy);
        pdf = 0.0f;
        pdf_rev = 0.0f;
        if (!eff_delta2(al) && !(wo.z * wi.z < 0.0f) && !(wo.z == 0.0f || wi.z == 0.0f)) {
            f3 h = normalize(wo + wi);
            float jac = 1.0f / (4.0f * dot(wo, h));
            float D, Dr;
            pdf = vndf_pdf_aniso(al, wo, h, D) * jac;
            float eta = ff ? m.ior : 1.0f / m.ior;
            if (rev) pdf_rev = vndf_pdf_aniso(al, wi, h, Dr) * jac;
            f3 F = disney_fresnel_m(m, wo, h, eta);
            f += 0.25f * D * F * G_corr_aniso(al, wo, wi) / (wi.z * wo.z);
        }
        pdf *= ps;
        pdf_rev *= ps;
        pdf_w += pdf;
        pdf_rev_w += pdf_rev;
    }
    if (fminf(wi.z, wo.z) > 0.0f) {
        if (pd > 0.0f) {
            pdf_w += pd * diffuse_pdf(wo, wi);
            pdf_rev_w += pdf_w;
            f += brdf_w * disney_diffuse_factor(m, wo, wi);
        }
        if (pc > 0.0f) {
            f3 h = normalize(wo + wi);
            float D;
            float fc = clearcoat_factor(m, wo, wi, h, D);
            pdf = D / (4.0f * dot(wo, h));
            pdf_rev = pdf;
            f += mk3(fc);
            pdf *= pc;
            pdf_rev *= pc;
            pdf_w += pdf;
            pdf_rev_w += pdf_rev;
        }
    }
    if (pt > 0.0f) {
        f += bsdf_w * eval_dielectric(m, wo, wi, pdf, pdf_rev, ff, rev);
        pdf *= pt;
        pdf_rev *= pt;
        pdf_w += pdf;
        pdf_rev_w += pdf_rev;
    }
    return f;
}
LT_DEV float principled_pdf(const Mat& m, f3 wo, f3 wi, bool ff) {
    float pdf = 0.0f;
    float ps, pd, pc, pt;
    float Fd = fresnel_dielectric(wo.z, m.ior, ff);
    init_probs(m, Fd, ff, ps, pd, pc, pt);
    if (ps > 0.0f) {
        f2 al = calc_aniso(m.roughness, m.anisotropy);
        if (!eff_delta2(al) && !(wo.z * wi.z < 0.0f) && !(wo.z == 0.0f || wi.z == 0.0f)) {
            f3 h = normalize(wo + wi);
            h = h * sgnf(h.z);
            float D;
            pdf += ps * (vndf_pdf_aniso(al, wo, h, D) / (4.0f * dot(wo, h)));
        }
    }
    if (fminf(wi.z, wo.z) > 0.0f) {
        if (pd > 0.0f) pdf += pd * diffuse_pdf(wo, wi);
        if (pc > 0.0f) {
            f3 h = normalize(wo + wi);
            float D = D_ggx_iso(mixf(0.1f, 0.001f, m.clearcoat_gloss), h.z);
            pdf += pc * (D / (4.0f * dot(wo, h)));
        }
    }
    if (pt > 0.0f) pdf += pt * dielectric_pdf(m, wo, wi, ff);
    return pdf;
}

// ---- bsdf_commons.glsl dispatch ----
LT_DEV f3 sample_bsdf(f3 n_s, f3 wo_w, const Mat& m, uint32_t mode, bool ff, f3& wi, float& pdf_w, float& cos_t,
                      float r0, float r1, float r2) {
    f3 f = mk3(0.0f);
    pdf_w = 0.0f;
    cos_t = 0.0f;
    wi = mk3(0.0f);
    f3 T, B;
    branchless_onb(n_s, T, B);
    f3 wo = to_local(wo_w, T, B, n_s);
    switch (m.bsdf_type) {
        case LT_BSDF_DIFFUSE:
            wi = sample_hemisphere(r0, r1);
            cos_t = wi.z;
            pdf_w = cos_t * INV_PI;
            f = fminf(wi.z, wo.z) <= 0.0f ? mk3(0.0f) : m.albedo * INV_PI;
            break;
        case LT_BSDF_MIRROR:
            if (wo.z <= 0.0f) {
                pdf_w = 0.0f;
                cos_t = 0.0f;
                wi = mk3(0.0f);
            } else {
                wi = reflect(-wo, mk3(0.0f, 0.0f, 1.0f));
                cos_t = wi.z;
                pdf_w = 1.0f;
                f = mk3(1.0f) / cos_t;
            }
            break;
        case LT_BSDF_GLASS: {
            float ie;
            refract_(mk3(0.0f, 0.0f, 1.0f), wo, ff, m.ior, mode, wi, f, ie);
            cos_t = wi.z;
            pdf_w = 1.0f;
            f = f / fabsf(cos_t);
        } break;
        case LT_BSDF_DIELECTRIC: f = sample_dielectric(m, wo, wi, mode, ff, pdf_w, cos_t, r0, r1); break;
        case LT_BSDF_CONDUCTOR: f = sample_conductor(m, wo, wi, pdf_w, cos_t, r0, r1); break;
        case LT_BSDF_PRINCIPLED: f = sample_principled(m, wo, wi, mode, ff, pdf_w, cos_t, r0, r1, r2); break;
        default: break;
    }
    wi = to_world(wi, T, B, n_s);
    return f;
}

LT_DEV f3 eval_bsdf(f3 n_s, f3 wo_w, const Mat& m, bool ff, f3 wi_w, float& pdf_w, float& pdf_rev_w, bool rev) {
    pdf_w = 0.0f;
    pdf_rev_w = 0.0f;
    f3 f = mk3(0.0f);
    f3 T, B;
    branchless_onb(n_s, T, B);
    f3 wo = to_local(wo_w, T, B, n_s), wi = to_local(wi_w, T, B, n_s);
    switch (m.bsdf_type) {
        case LT_BSDF_DIFFUSE:
            if (!(fminf(wi.z, wo.z) <= 0.0f)) {
                pdf_w = wi.z * INV_PI;
                pdf_rev_w = pdf_w;
                f = m.albedo * INV_PI;
            }
            break;
        case LT_BSDF_DIELECTRIC: f = eval_dielectric(m, wo, wi, pdf_w, pdf_rev_w, ff, rev); break;
        case LT_BSDF_CONDUCTOR: f = eval_conductor(m, wo, wi, pdf_w, pdf_rev_w, rev); break;
        case LT_BSDF_PRINCIPLED: f = eval_principled(m, wo, wi, pdf_w, pdf_rev_w, ff, rev); break;
        default: break;
    }
    return f;
}

LT_DEV float bsdf_pdf(const Mat& m, f3 n_s, f3 wo_w, f3 wi_w, bool ff) {
    f3 T, B;
    branchless_onb(n_s, T, B);
    f3 wo = to_local(wo_w, T, B, n_s), wi = to_local(wi_w, T, B, n_s);
    switch (m.bsdf_type) {
        case LT_BSDF_DIFFUSE: return diffuse_pdf(wo, wi);
        case LT_BSDF_DIELECTRIC: return dielectric_pdf(m, wo, wi, ff);
        case LT_BSDF_CONDUCTOR: return conductor_pdf(m, wo, wi);
        case LT_BSDF_PRINCIPLED: return principled_pdf(m, wo, wi, ff);
        default: return 0.0f;
    }
}

}  // namespace ltd
