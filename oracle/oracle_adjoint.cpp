// oracle_adjoint.cpp — host implementation of the differentiable Path
// estimator and its path-replay adjoint (TEST INFRASTRUCTURE, see oracle.h).
//
// No reference counterpart: the reference renders but does not
// differentiate.  The estimator is the Path integrator of oracle_path.cpp
// (path.rgen:26-104, pt_commons.glsl:3-42) with Russian roulette removed,
// averaging `spp` samples per pixel; sample k of pixel (x, y) draws its
// pcg4d stream from (x, y, k ^ seed, 0), the word the Path integrator uses
// for frame k, so for max_depth <= 5 (where path.rgen never reaches its
// roulette) one sample equals one Path frame.  The albedo texture
// `albedo_id` and the roughness texture `rough_id` of the scene are replaced
// by a variant's arrays (rgb and scalar texels, row-major).
//
// Derivatives hold the random numbers fixed and sampling detached: sampled
// directions, lobe choices, pdfs and MIS weights are constants, BSDF values
// are differentiated with respect to the texels through the bilinear fetch.
// Path-replay backpropagation (Vicini, Speierer and Jakob 2021): the first
// pass gives the sample's radiance L; the replay of the same path adds, at
// each vertex j,
//   delta . (df_j / f_j) (L - radiance gathered before j's continuation)
// for the sampled continuation and  delta . T_j dN_j  for the next-event
// terms N_j, where T_j is the throughput up to j.
//
// Supported texel derivatives: Lambertian albedo, and the principled BSDF
// with metallic, spec_trans, specular_tint, clearcoat, flatness and
// anisotropy all zero (the GGX reflection lobe plus the Disney diffuse lobe;
// other lobes contribute no derivative).  With alpha = max(0.001, r^2):
//   f = S(alpha) + albedo g(r) / pi
//   g = (1 - f_i/2)(1 - f_o/2) + rr (f_i + f_o^2 f_i (rr - 1)),  rr = 2 r^2 cos_i^2,
//   f_w = 1 + (fd90 - 1)(1 - cos_w)^5,  fd90 = 1/2 + 2 (wo.h)^2 r
//   d ln S / d alpha = d ln D / d alpha + d ln G / d alpha   (F is independent of r)
//   d ln D / d alpha = -2/alpha + 4 e / (alpha (1 + e)),  e = tan^2(h) c_h / alpha^2
//   d ln G / d alpha = -G (dL(wo) + dL(wi)),  dL(w) = alpha c_w tan^2(w) / (2 sqrt(1 + c_w alpha^2 tan^2(w)))
// where c_w is the (clamped) sum of the squared azimuth cosines of the Smith
// and GGX formulas (1 away from the pole).  A lobe with alpha <= 0.0064 is
// the reference's delta reflection: no roughness derivative.
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "oracle.h"

namespace ora {

static const float ADJ_TMIN = 0.001f;    // path.rgen:19
static const float ADJ_TMAX = 10000.0f;  // path.rgen:20

struct AdjTex {
    const float* albedo;  // rgb texels of the albedo texture (nullptr: not differentiated)
    int aw, ah, aid;
    const float* rough;   // texels of the roughness texture
    int rw, rh, rid;
};

struct Taps {  // bilinear footprint of one texture fetch
    int idx[4];
    float w[4];
};

// tex_bilinear of oracle_core.cpp (repeat addressing, texel centres at
// (i + 0.5) / size, rows blended first) on a `stride`-channel array, with
// the four taps and their weights
static float fetch(const float* data, int stride, int c, int W, int H, vec2 uv, Taps* t) {
    float fx = uv.x * (float)W - 0.5f, fy = uv.y * (float)H - 0.5f;
    float x0f = floorf(fx), y0f = floorf(fy);
    float ax = fx - x0f, ay = fy - y0f;
    int x0 = ((int)x0f % W + W) % W, y0 = ((int)y0f % H + H) % H;
    int x1 = (x0 + 1) % W, y1 = (y0 + 1) % H;
    const int i00 = y0 * W + x0, i01 = y0 * W + x1, i10 = y1 * W + x0, i11 = y1 * W + x1;
    if (t) {
        t->idx[0] = i00;
        t->idx[1] = i01;
        t->idx[2] = i10;
        t->idx[3] = i11;
        t->w[0] = (1.0f - ax) * (1.0f - ay);
        t->w[1] = ax * (1.0f - ay);
        t->w[2] = (1.0f - ax) * ay;
        t->w[3] = ax * ay;
    }
    const float top = data[stride * i00 + c] * (1.0f - ax) + data[stride * i01 + c] * ax;
    const float bot = data[stride * i10 + c] * (1.0f - ax) + data[stride * i11 + c] * ax;
    return top * (1.0f - ay) + bot * ay;
}

struct DMat {  // material at a vertex, and which of its parameters are texel-driven
    Material m;
    bool tex_albedo = false, tex_rough = false;
    Taps ta, tr;
    vec3 albedo_mult;  // the material's albedo factor (bsdf_commons.glsl:16-22)
};

// load_material of oracle_core.cpp with the variant's textures
static DMat load_dmat(const lt_scene_desc& s, const AdjTex& tx, uint32_t idx, vec2 uv) {
    DMat d;
    d.m = load_material(s, idx, uv);
    const lt_material& raw = s.materials[idx];
    d.albedo_mult = load3(raw.albedo);
    if (tx.albedo && raw.texture_id > -1 && raw.texture_id == tx.aid) {
        d.tex_albedo = true;
        vec3 t;
        t.x = fetch(tx.albedo, 3, 0, tx.aw, tx.ah, uv, &d.ta);
        t.y = fetch(tx.albedo, 3, 1, tx.aw, tx.ah, uv, nullptr);
        t.z = fetch(tx.albedo, 3, 2, tx.aw, tx.ah, uv, nullptr);
        d.m.albedo = d.albedo_mult * t;
    }
    if (tx.rough && s.roughness_texture_ids && s.roughness_texture_ids[idx] > -1 &&
        s.roughness_texture_ids[idx] == tx.rid) {
        d.tex_rough = true;
        d.m.roughness = fetch(tx.rough, 1, 0, tx.rw, tx.rh, uv, &d.tr);
    }
    return d;
}

// ---- BSDF derivatives (local frame, z = shading normal) ----
static inline float pow5(float c) { return powf(fmaxf(1.0f - c, 0.0f), 5.0f); }

// the principled diffuse factor g (principled.glsl:50-74 with flatness 0) and dg/dr
static void diffuse_g(vec3 wo, vec3 wi, float r, float& g, float& dg) {
    const vec3 h = normalize(wi + wo);
    const float wh = dot(wo, h);
    const float fd90 = 0.5f + 2.0f * wh * wh * r, dfd90 = 2.0f * wh * wh;
    const float Pi = pow5(wi.z), Po = pow5(wo.z);
    const float fi = 1.0f + (fd90 - 1.0f) * Pi, fo = 1.0f + (fd90 - 1.0f) * Po;
    const float dfi = dfd90 * Pi, dfo = dfd90 * Po;
    const float rr = 2.0f * (r * r) * wi.z * wi.z, drr = 4.0f * r * wi.z * wi.z;
    const float inner = fi + fo * fi * fo * (rr - 1.0f);
    g = (1.0f - 0.5f * fi) * (1.0f - 0.5f * fo) + rr * inner;
    const float dinner = dfi + (2.0f * fo * dfo * fi + fo * fo * dfi) * (rr - 1.0f) + fo * fo * fi * drr;
    dg = -0.5f * dfi * (1.0f - 0.5f * fo) - 0.5f * dfo * (1.0f - 0.5f * fi) + drr * inner + rr * dinner;
}

// tan^2 and the clamped azimuth factor c_w of microfacet_commons.glsl:13-51
static bool tan2_c(vec3 w, float& t2, float& c) {
    const float c2 = w.z * w.z;
    const float s2 = fmaxf(1.0f - c2, 0.0f);
    t2 = s2 / c2;
    if (std::isinf(t2)) return false;
    c = s2 == 0.0f ? 1.0f : (clampf(w.x * w.x, 0.0f, 1.0f) + clampf(w.y * w.y, 0.0f, 1.0f)) / s2;
    return true;
}

// d ln S / d r of the GGX reflection lobe at the half vector of (wo, wi)
static float spec_dlog_dr(vec3 wo, vec3 wi, float r) {
    const float alpha = fmaxf(0.001f, r * r);
    if (alpha <= 0.0064f) return 0.0f;  // delta lobe (sampling_commons.glsl:18-19)
    const float dadr = 2.0f * r;
    const vec3 h = normalize(wo + wi);
    float t2, c, dlnD = 0.0f;
    if (tan2_c(h, t2, c)) {
        const float e = t2 * c / (alpha * alpha);
        dlnD = -2.0f / alpha + 4.0f * e / (alpha * (1.0f + e));
    }
    float L[2] = {0.0f, 0.0f}, dL[2] = {0.0f, 0.0f};
    const vec3 ws[2] = {wo, wi};
    for (int i = 0; i < 2; i++) {
        if (!tan2_c(ws[i], t2, c)) continue;
        const float q = sqrtf(1.0f + c * alpha * alpha * t2);
        L[i] = 0.5f * (q - 1.0f);
        dL[i] = alpha * c * t2 / (2.0f * q);
    }
    const float G = 1.0f / (1.0f + L[0] + L[1]);
    return (dlnD - G * (dL[0] + dL[1])) * dadr;
}

// lobe probabilities of principled.glsl:26-48
static void lobe_probs(const Material& m, float wo_z, bool ff, float& ps, float& pd, float& pc, float& pt) {
    const float Fd = fresnel_dielectric(wo_z, m.ior, ff);
    const float brdf_w = (1.0f - m.spec_trans) * (1.0f - m.metallic);
    const float bsdf_w = (1.0f - m.metallic) * m.spec_trans;
    ps = ff ? (1.0f - bsdf_w * (1.0f - Fd)) : Fd;
    pt = ff ? (bsdf_w * (1.0f - Fd)) : (1.0f - Fd);
    pd = ff ? brdf_w : 0.0f;
    pc = ff ? 0.25f * clampf(m.clearcoat, 0.0f, 1.0f) : 0.0f;
    const float norm = 1.0f / (ps + pt + pd + pc);
    ps *= norm;
    pd *= norm;
    pc *= norm;
    pt *= norm;
}

struct DF {  // a BSDF value and its texel-parameter derivatives
    vec3 f;
    vec3 dfa;  // d f_c / d albedo_c
    vec3 dfr;  // d f / d roughness
};

static vec3 to_loc(vec3 v, vec3 T, vec3 B, vec3 N) { return {dot(v, T), dot(v, B), dot(v, N)}; }

// eval_bsdf plus derivatives
static DF eval_d(const DMat& d, vec3 n_s, vec3 wo_w, vec3 wi_w, bool side, float& pdf) {
    DF o;
    float pr;
    o.f = eval_bsdf(n_s, wo_w, d.m, 1, side, wi_w, pdf, pr, false);
    o.dfa = o.dfr = v3(0.0f);
    if (!(d.tex_albedo || d.tex_rough)) return o;
    vec3 T, B;
    branchless_onb(n_s, T, B);
    const vec3 wo = to_loc(wo_w, T, B, n_s), wi = to_loc(wi_w, T, B, n_s);
    if (d.m.bsdf_type == LT_BSDF_DIFFUSE) {
        if (fminf(wi.z, wo.z) > 0.0f) o.dfa = v3(INV_PI);
        return o;
    }
    if (d.m.bsdf_type != LT_BSDF_PRINCIPLED) return o;
    float ps, pd, pc, pt;
    lobe_probs(d.m, wo.z, side, ps, pd, pc, pt);
    vec3 diff = v3(0.0f);
    if (fminf(wi.z, wo.z) > 0.0f && pd > 0.0f) {
        float g, dg;
        diffuse_g(wo, wi, d.m.roughness, g, dg);
        diff = d.m.albedo * g * INV_PI;
        o.dfa = v3(g * INV_PI);
        o.dfr = d.m.albedo * dg * INV_PI;
    }
    if (ps > 0.0f && !(wo.z * wi.z < 0.0f) && wo.z != 0.0f && wi.z != 0.0f) {
        const vec3 S = o.f - diff;  // the GGX reflection term
        o.dfr += S * spec_dlog_dr(wo, wi, d.m.roughness);
    }
    return o;
}

// sample_bsdf plus derivatives of the returned lobe value
static DF sample_d(const DMat& d, vec3 n_s, vec3 wo_w, bool side, vec3& wi_w, float& pdf, float& cos_t, vec3 r3) {
    DF o;
    o.f = sample_bsdf(n_s, wo_w, d.m, 1, side, wi_w, pdf, cos_t, r3);
    o.dfa = o.dfr = v3(0.0f);
    if (!(d.tex_albedo || d.tex_rough)) return o;
    vec3 T, B;
    branchless_onb(n_s, T, B);
    const vec3 wo = to_loc(wo_w, T, B, n_s), wi = to_loc(wi_w, T, B, n_s);
    if (d.m.bsdf_type == LT_BSDF_DIFFUSE) {
        if (fminf(wi.z, wo.z) > 0.0f) o.dfa = v3(INV_PI);
        return o;
    }
    if (d.m.bsdf_type != LT_BSDF_PRINCIPLED || wo.z <= 0.0f) return o;
    float ps, pd, pc, pt;
    lobe_probs(d.m, wo.z, side, ps, pd, pc, pt);
    if (r3.z < ps) {  // GGX reflection (principled.glsl:96-135)
        o.dfr = o.f * spec_dlog_dr(wo, wi, d.m.roughness);
    } else if (r3.z > ps + pc && r3.z <= ps + pc + pd) {  // diffuse
        if (fminf(wi.z, wo.z) > 0.0f) {
            float g, dg;
            diffuse_g(wo, wi, d.m.roughness, g, dg);
            o.dfa = v3(g * INV_PI);
            o.dfr = d.m.albedo * dg * INV_PI;
        }
    }
    return o;
}

// The value of the lobe sample_d picked (with r3.z) re-evaluated with the
// material e, for the same directions: the estimator with sampling held at
// one set of texels and BSDF values taken at another, whose finite
// differences are the detached derivatives this file computes.
static vec3 lobe_value(const DMat& e, vec3 n_s, vec3 wo_w, vec3 wi_w, bool side, vec3 r3, vec3 f_sampled) {
    if (!(e.tex_albedo || e.tex_rough)) return f_sampled;
    vec3 T, B;
    branchless_onb(n_s, T, B);
    const vec3 wo = to_loc(wo_w, T, B, n_s), wi = to_loc(wi_w, T, B, n_s);
    if (e.m.bsdf_type == LT_BSDF_DIFFUSE) return fminf(wi.z, wo.z) <= 0.0f ? v3(0.0f) : e.m.albedo * INV_PI;
    if (e.m.bsdf_type != LT_BSDF_PRINCIPLED || wo.z <= 0.0f) return f_sampled;
    float ps, pd, pc, pt;
    lobe_probs(e.m, wo.z, side, ps, pd, pc, pt);
    vec3 diff = v3(0.0f);
    if (fminf(wi.z, wo.z) > 0.0f) {
        float g, dg;
        diffuse_g(wo, wi, e.m.roughness, g, dg);
        diff = e.m.albedo * g * INV_PI;
    }
    if (r3.z < ps) {
        float pdf, pr;
        const vec3 full = eval_bsdf(n_s, wo_w, e.m, 1, side, wi_w, pdf, pr, false);
        return (fminf(wi.z, wo.z) > 0.0f && pd > 0.0f) ? full - diff : full;
    }
    if (r3.z > ps + pc && r3.z <= ps + pc + pd) return diff;
    return f_sampled;
}

struct AdjParams {
    uint32_t W, H;
    int D;
    vec3 sky;
};

struct GradAcc {  // gradient accumulators of one variant
    double* ga;   // 3 per albedo texel
    double* gr;   // 1 per roughness texel
};

// scatter the texel gradients of one vertex term: wa and wr weight the
// albedo and roughness derivatives of the BSDF, channel by channel
static void scatter(const DMat& d, const GradAcc& g, vec3 wa, vec3 wr, vec3 dfa, vec3 dfr) {
    if (d.tex_albedo) {
        const vec3 q = wa * dfa * d.albedo_mult;
        for (int k = 0; k < 4; k++)
            for (int c = 0; c < 3; c++) g.ga[3 * d.ta.idx[k] + c] += (double)(d.ta.w[k] * q[c]);
    }
    if (d.tex_rough) {
        const float q = dot(wr, dfr);
        for (int k = 0; k < 4; k++) g.gr[d.tr.idx[k]] += (double)(d.tr.w[k] * q);
    }
}

static vec3 safe_inv(vec3 f) {
    return {f.x != 0.0f ? 1.0f / f.x : 0.0f, f.y != 0.0f ? 1.0f / f.y : 0.0f, f.z != 0.0f ? 1.0f / f.z : 0.0f};
}

// One sample.  With g == nullptr returns its radiance; otherwise replays the
// path with its radiance L and adjoint radiance delta, scattering gradients.
// path.rgen:26-104 without roulette, including the persistent payload a
// missed MIS ray leaves in place (ray.rmiss:13).
// With te != nullptr (forward only) sampling uses the texels tx and BSDF
// values the texels te.
static vec3 adj_sample(const lt_scene_desc& s, const Accel* a, const lt_camera& cam, const AdjParams& p,
                       const AdjTex& tx, const AdjTex* te, uint32_t x, uint32_t y, uint32_t word, vec3 L, vec3 delta,
                       const GradAcc* g) {
    uvec4 seed = {x, y, word, 0u};
    const float r0 = rand(seed), r1 = rand(seed);
    const vec2 in_uv = {((float)x + 0.5f + (r0 - 0.5f)) / (float)p.W, ((float)y + 0.5f + (r1 - 0.5f)) / (float)p.H};
    const vec2 dd = {in_uv.x * 2.0f - 1.0f, in_uv.y * 2.0f - 1.0f};
    const vec4 o4 = mul(cam.inv_view, vec4{0.0f, 0.0f, 0.0f, 1.0f});
    const vec4 tg = mul(cam.inv_projection, vec4{dd.x, dd.y, 1.0f, 1.0f});
    const vec3 tn = normalize(v3(tg.x, tg.y, tg.z));
    const vec4 d4 = mul(cam.inv_view, vec4{tn.x, tn.y, tn.z, 0.0f});
    vec3 origin = {o4.x, o4.y, o4.z}, direction = {d4.x, d4.y, d4.z};
    vec3 col = v3(0.0f), thr = v3(1.0f);
    bool last_specular = false;
    Payload pl{};
    for (int depth = 0;; depth++) {
        const Hit h = accel_closest(a, origin, direction, ADJ_TMIN, ADJ_TMAX);
        if (h.inst == 0xFFFFFFFFu) {
            col += thr * p.sky;
            break;
        }
        fill_payload(s, h, origin, direction, ADJ_TMIN, pl);
        const Payload hp = pl;
        const DMat dm = load_dmat(s, tx, hp.material_idx, hp.uv);
        const DMat de = te ? load_dmat(s, *te, hp.material_idx, hp.uv) : dm;
        if (depth == 0 || last_specular) col += thr * dm.m.emissive_factor;
        if (depth >= p.D - 1) break;
        const vec3 wo = -direction;
        vec3 n_s = hp.n_s, n_g = hp.n_g;
        bool side = true;
        if (dot(hp.n_g, wo) < 0.0f) n_g = -n_g;
        if (dot(n_g, hp.n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        origin = offset_ray(hp.pos, n_g);
        last_specular = is_specular(dm.m);
        if (!last_specular) {  // pt_commons.glsl:3-42, divided by the pick pdf (path.rgen:76)
            const float pick = 1.0f / (float)s.total_light_triangle_cnt;
            vec4 r4;
            r4.x = rand(seed);
            r4.y = rand(seed);
            r4.z = rand(seed);
            r4.w = rand(seed);
            vec3 wi, ln, lpos;
            float wi_len, pw, pa, cfl, pdw;
            LightRecord rec{};
            const vec3 Le =
                sample_light_Li(s, r4, hp.pos, (int)s.num_lights, pw, wi, wi_len, pa, cfl, rec, ln, lpos, pdw);
            const vec3 pp = offset_ray2(hp.pos, n_s);
            const float cos_x = dot(n_s, wi);
            float bpdf, epdf;
            DF fl = eval_d(dm, n_s, wo, wi, side, bpdf);
            if (te) fl.f = eval_d(de, n_s, wo, wi, side, epdf).f;
            vec3 res = v3(0.0f), ka = v3(0.0f), kb = v3(0.0f);
            bool hit_a = false, hit_b = false;
            DF fb{};
            if (!accel_any(a, pp, wi, 0.0f, wi_len - EPS) && pw > 0.0f) {
                const float mis = (rec.flags & LT_LIGHT_DELTA) ? 1.0f : 1.0f / (1.0f + bpdf / pw);
                ka = mis * fabsf(cos_x) * Le / pw;
                res += mis * fl.f * fabsf(cos_x) * Le / pw;
                hit_a = true;
            }
            if ((rec.flags & 7u) == LT_LIGHT_AREA) {
                vec3 r3;
                r3.x = rand(seed);
                r3.y = rand(seed);
                r3.z = rand(seed);
                float bp, cx;
                vec3 wb;
                fb = sample_d(dm, n_s, wo, side, wb, bp, cx, r3);
                if (te) fb.f = lobe_value(de, n_s, wo, wb, side, r3, fb.f);
                if (bp != 0.0f) {
                    const Hit h2 = accel_closest(a, pp, wb, ADJ_TMIN, ADJ_TMAX);
                    if (h2.inst != 0xFFFFFFFFu) fill_payload(s, h2, pp, wb, ADJ_TMIN, pl);
                    if (pl.triangle_idx == rec.triangle_idx && pl.instance_idx == rec.instance_idx) {
                        const float wl = length(pl.pos - hp.pos);
                        const float gg = fabsf(dot(pl.n_s, -wb)) / (wl * wl);
                        const float mis = 1.0f / (1.0f + pa / (gg * bp));
                        kb = mis * fabsf(cx) * Le / bp;
                        res += fb.f * mis * fabsf(cx) * Le / bp;
                        hit_b = true;
                    }
                }
            }
            col += thr * res / pick;
            if (g) {
                const vec3 w = delta * thr / pick;
                if (hit_a) scatter(dm, *g, w * ka, w * ka, fl.dfa, fl.dfr);
                if (hit_b) scatter(dm, *g, w * kb, w * kb, fb.dfa, fb.dfr);
            }
        }
        vec3 r3;
        r3.x = rand(seed);
        r3.y = rand(seed);
        r3.z = rand(seed);
        float pdf, cos_t;
        DF fc = sample_d(dm, n_s, wo, side, direction, pdf, cos_t, r3);
        if (pdf == 0.0f) break;
        if (te) fc.f = lobe_value(de, n_s, wo, direction, side, r3, fc.f);
        if (g) {
            const vec3 w = delta * (L - col) * safe_inv(fc.f);
            scatter(dm, *g, w, w, fc.dfa, fc.dfr);
        }
        thr *= fc.f * fabsf(cos_t) / pdf;
    }
    return col;
}

}  // namespace ora

using namespace ora;

extern "C" {

typedef struct ora_adj_params {
    uint32_t width, height;
    int32_t max_depth;
    uint32_t spp, seed;
    float sky_col[3];
    int32_t albedo_id, rough_id;  // scene texture ids replaced by the variants (-1: none)
    uint32_t aw, ah, rw, rh;      // their texel grids
} ora_adj_params;

// Forward (d_images == nullptr): images [V][H][W][3] = per-pixel mean over
// spp samples; with eval_albedo / eval_rough given, BSDF values use those
// texels while sampling uses albedo / rough.  Backward: grad_albedo [V][ah][aw][3] and grad_rough
// [V][rh][rw] (double, accumulated into) of sum(d_images * images).
// albedo [V][ah][aw][3] and rough [V][rh][rw] are the variants' texels.
// Returns wall seconds.
double ora_adjoint(const lt_scene_desc* s, void* accel, const lt_camera* cam, const ora_adj_params* q,
                   uint32_t variants, const float* albedo, const float* rough, const float* eval_albedo,
                   const float* eval_rough, float* images, const float* d_images, double* grad_albedo,
                   double* grad_rough, int threads) {
    const AdjParams p{q->width, q->height, q->max_depth, load3(q->sky_col)};
    const size_t npx = (size_t)p.W * p.H;
    const size_t na = (size_t)q->aw * q->ah, nr = (size_t)q->rw * q->rh;
    const int nt = threads < 1 ? 1 : threads;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t v = 0; v < variants; v++) {
        const AdjTex tx{q->albedo_id >= 0 ? albedo + 3 * na * v : nullptr, (int)q->aw, (int)q->ah, q->albedo_id,
                        q->rough_id >= 0 ? rough + nr * v : nullptr, (int)q->rw, (int)q->rh, q->rough_id};
        const bool frozen = (eval_albedo || eval_rough) && !d_images;
        const AdjTex te{eval_albedo && q->albedo_id >= 0 ? eval_albedo + 3 * na * v : tx.albedo, tx.aw, tx.ah, tx.aid,
                        eval_rough && q->rough_id >= 0 ? eval_rough + nr * v : tx.rough, tx.rw, tx.rh, tx.rid};
        std::atomic<uint32_t> next{0};
        std::vector<std::vector<double>> ga(nt), gr(nt);
        auto worker = [&](int tid) {
            if (d_images) {
                ga[tid].assign(3 * na, 0.0);
                gr[tid].assign(nr, 0.0);
            }
            const GradAcc acc{ga[tid].data(), gr[tid].data()};
            for (;;) {
                const uint32_t y = next.fetch_add(1);
                if (y >= p.H) break;
                for (uint32_t x = 0; x < p.W; x++) {
                    const size_t i = (size_t)y * p.W + x;
                    vec3 sum = v3(0.0f), delta = v3(0.0f);
                    if (d_images) delta = load3(d_images + 3 * (npx * v + i)) / (float)q->spp;
                    for (uint32_t k = 0; k < q->spp; k++) {
                        const uint32_t word = k ^ q->seed;
                        const vec3 L = adj_sample(*s, (const Accel*)accel, *cam, p, tx, frozen ? &te : nullptr, x,
                                                  y, word, v3(0.0f), v3(0.0f), nullptr);
                        if (std::isnan(luminance(L))) continue;  // path.rgen:102-104
                        if (d_images)
                            adj_sample(*s, (const Accel*)accel, *cam, p, tx, nullptr, x, y, word, L, delta, &acc);
                        else
                            sum += L;
                    }
                    if (!d_images) {
                        float* o = images + 3 * (npx * v + i);
                        o[0] = sum.x / (float)q->spp;
                        o[1] = sum.y / (float)q->spp;
                        o[2] = sum.z / (float)q->spp;
                    }
                }
            }
        };
        std::vector<std::thread> ts;
        for (int i = 1; i < nt; i++) ts.emplace_back(worker, i);
        worker(0);
        for (auto& t : ts) t.join();
        if (d_images)
            for (int t = 0; t < nt; t++) {
                for (size_t k = 0; k < 3 * na; k++) grad_albedo[3 * na * v + k] += ga[t][k];
                for (size_t k = 0; k < nr; k++) grad_rough[nr * v + k] += gr[t][k];
            }
    }
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// eval_bsdf of material m at world-space n_s, wo, wi and its derivatives
// with respect to albedo and roughness (as if both were texel-driven):
// out = [f.rgb, df/dalbedo.rgb, df/droughness.rgb]
void ora_bsdf_eval_d(const lt_material* m, const float* n_s, const float* wo, const float* wi, int side,
                     float* out) {
    DMat d;
    d.m = material_from(*m);
    d.tex_albedo = d.tex_rough = true;
    d.albedo_mult = v3(1.0f);
    float pdf;
    const DF r = eval_d(d, load3(n_s), load3(wo), load3(wi), side != 0, pdf);
    const vec3 v[3] = {r.f, r.dfa, r.dfr};
    for (int i = 0; i < 3; i++) {
        out[3 * i] = v[i].x;
        out[3 * i + 1] = v[i].y;
        out[3 * i + 2] = v[i].z;
    }
}

}  // extern "C"
