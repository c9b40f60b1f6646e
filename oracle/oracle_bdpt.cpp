// oracle_bdpt.cpp — the reference BDPT integrator restated on the CPU
// (TEST INFRASTRUCTURE, see oracle.h).
//
// Follows src/shaders/integrators/bdpt/bdpt.rgen:40-91 and
// src/shaders/integrators/bdpt_commons.glsl:13-641, with sample_light_Le of
// src/shaders/commons.glsl:335-406.  Reference semantics kept:
//   - the per-pixel vertex buffers start zeroed every frame (BDPT.cpp:76-77),
//     so fields a walk does not write read as 0 (an escaped camera vertex has
//     material 0, position 0 and normal 0);
//   - camera rays go through the pixel centre (no jitter, bdpt.rgen:41-42);
//   - MIS is the balance heuristic of calc_mis_weight, evaluated on copies
//     instead of the reference's write-and-restore of the shared vertices.
// Deviation: t = 1 splats are summed into a separate image and added after
// the frame (the reference's unsynchronised += races, bdpt.rgen:66-70).
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "oracle.h"

namespace ora {

struct PV {  // PathVertex, bdpt_commons.h:18-33
    vec3 dir, n_s, pos;
    vec2 uv;
    vec3 throughput;
    uint32_t light_flags, light_idx, material_idx, delta, side, mode;
    float area, pdf_fwd, pdf_rev;
};

static const float BDPT_TMIN = 0.001f;  // bdpt_commons.glsl:23
static const float BDPT_TMAX = 1e6f;    // bdpt_commons.glsl:24
static const int BDPT_MAX_DEPTH = 16;   // subpaths hold at most max_depth + 1 vertices

static inline bool lt_finite(uint32_t f) { return ((f >> 4) & 1u) != 0; }
static inline bool lt_delta(uint32_t f) { return ((f >> 5) & 1u) != 0; }
static inline float uniform_cone_pdf(float cos_max) { return 1.0f / (TWO_PI * (1.0f - cos_max)); }

static float spot_cos_width(const lt_scene_desc& s, uint32_t li) {
    return s.spot_cones ? s.spot_cones[2 * li] : cosf(PI / 6.0f);  // commons.glsl:49 cos(30 deg)
}
static float spot_cos_falloff(const lt_scene_desc& s, uint32_t li) {
    return s.spot_cones ? s.spot_cones[2 * li + 1] : cosf(25.0f * PI / 180.0f);
}
static float falloff(float c, float cw, float cf) {
    if (c < cw) return 0.0f;
    if (c >= cf) return 1.0f;
    float d = (c - cw) / (cf - cw);
    return (d * d) * (d * d);
}

// commons.glsl:48-62 / 81-95 (the spot cone pdf uses the light's cone)
static float light_pdf_dir(const lt_scene_desc& s, uint32_t li, uint32_t flags, vec3 n_s, vec3 wi) {
    switch (flags & 7u) {
        case LT_LIGHT_AREA: return fmaxf(dot(n_s, wi) / PI, 0.0f);
        case LT_LIGHT_SPOT: return uniform_cone_pdf(spot_cos_width(s, li));
        case LT_LIGHT_POINT: return 1.0f / (4.0f * PI);
        default: return 0.0f;
    }
}

// commons.glsl:64-79 (+ point lights: wi_len^2)
static float light_pdf_a_to_w(uint32_t flags, float pdf_a, float wi_len_sqr, float cos_from_light) {
    switch (flags & 7u) {
        case LT_LIGHT_AREA: return pdf_a * wi_len_sqr / cos_from_light;
        case LT_LIGHT_SPOT: return wi_len_sqr / cos_from_light;
        case LT_LIGHT_DIRECTIONAL: return 1.0f;
        case LT_LIGHT_POINT: return wi_len_sqr;
        default: return 0.0f;
    }
}

// utils.glsl:157-173
static vec4 to_local_quat(vec3 v) {
    if (v.z < -0.99999f) return {1.0f, 0.0f, 0.0f, 0.0f};
    float x = v.y, y = -v.x, z = 0.0f, w = 1.0f + v.z;
    float il = 1.0f / sqrtf(x * x + y * y + z * z + w * w);
    return {x * il, y * il, z * il, w * il};
}
static vec3 rot_quat(vec4 q, vec3 v) {
    vec3 qa = {q.x, q.y, q.z};
    return 2.0f * dot(qa, v) * qa + (q.w * q.w - dot(qa, qa)) * v + 2.0f * q.w * cross(qa, v);
}
// commons.glsl:217-222
static vec3 uniform_sample_cone(vec2 uv, float cos_max) {
    const float ct = (1.0f - uv.x) + uv.x * cos_max;
    const float st = sqrtf(1.0f - ct * ct);
    const float phi = uv.y * TWO_PI;
    return {cosf(phi) * st, sinf(phi) * st, ct};
}

// commons.glsl:335-406 (+ point lights: uniform sphere).  The reference
// writes the uninitialised local `triangle_idx` into the record
// (commons.glsl:351,360); BDPT only reads the record's flags.
vec3 sample_light_Le(const lt_scene_desc& s, vec4 rp, vec2 rd, int num_lights, int total_light, float& cos_from_light,
                     uint32_t& flags, vec3& pos, vec3& wi, vec3& n, float& pdf_pos_a, float& pdf_dir_w) {
    uint32_t li = (uint32_t)(rp.x * (float)num_lights);
    const lt_light& L = s.lights[li];
    vec3 Le = v3(0.0f);
    flags = L.light_flags;
    cos_from_light = 0.0f;
    pdf_pos_a = 0.0f;
    pdf_dir_w = 0.0f;
    pos = wi = n = v3(0.0f);
    switch (L.light_flags & 7u) {
        case LT_LIGHT_AREA: {
            const lt_prim_mesh_info& pi = s.prim_infos[L.prim_mesh_idx];
            uint32_t tri = (uint32_t)(rp.y * (float)L.num_triangles);
            vec2 bary;
            TriangleRecord r = sample_triangle(s, pi, v2(rp.z, rp.w), tri, L.world_matrix, bary);
            Material lm = load_material(s, pi.material_index, bary);
            pos = r.pos;
            float phi;
            wi = sample_hemisphere_n(rd, r.n_s, phi);
            Le = lm.emissive_factor;
            cos_from_light = fmaxf(dot(r.n_s, wi), 0.0f);
            pdf_pos_a = r.triangle_pdf;
            pdf_dir_w = dot(wi, r.n_s) / PI;
            n = r.n_s;
        } break;
        case LT_LIGHT_SPOT: {
            const float cw = spot_cos_width(s, li), cf = spot_cos_falloff(s, li);
            const vec3 ld = normalize(load3(L.to) - load3(L.pos));
            vec4 q = to_local_quat(ld);
            vec4 qi = {-q.x, -q.y, -q.z, q.w};
            wi = rot_quat(qi, uniform_sample_cone(rd, cw));
            pos = load3(L.pos);
            cos_from_light = dot(wi, ld);
            Le = load3(L.L) * falloff(cos_from_light, cw, cf);
            pdf_pos_a = 1.0f;
            pdf_dir_w = uniform_cone_pdf(cw);
            n = wi;
        } break;
        case LT_LIGHT_POINT: {
            float z = 1.0f - 2.0f * rd.x, r = sqrtf(fmaxf(0.0f, 1.0f - z * z)), phi = TWO_PI * rd.y;
            wi = {r * cosf(phi), r * sinf(phi), z};
            pos = load3(L.pos);
            cos_from_light = 1.0f;
            Le = load3(L.L);
            pdf_pos_a = 1.0f;
            pdf_dir_w = 1.0f / (4.0f * PI);
            n = wi;
        } break;
        default: break;  // directional lights are not restated (no workload uses them)
    }
    pdf_pos_a /= (float)total_light;
    return Le;
}

struct BdptCtx {
    const lt_scene_desc& s;
    const Accel* a;
    Counters* cnt;
    int max_depth;
    uint32_t W, H;
    float screen_size;
    float cam_area;
    float light_pdf_pos;  // the global of bdpt_commons.glsl:8
    const lt_camera& cam;
    bool power;  // power heuristic (beta = 2) instead of the reference's balance
};

static inline bool same_hemisphere(vec3 wi, vec3 wo, vec3 n) { return dot(wi, n) > 0.0f && dot(wo, n) > 0.0f; }

static bool occluded(BdptCtx& c, vec3 o, vec3 d, float tmax) {
    if (c.cnt) c.cnt->shadow++;
    return accel_any(c.a, o, d, 0.0f, tmax);
}

// bdpt_commons.glsl:13-197 (light: mode 0, eye: mode 1).  v[0] is vtx(-1).
static int random_walk(BdptCtx& c, PV* v, int max_depth, vec3 throughput, float pdf, uvec4& seed, bool eye) {
    if (max_depth == 0) return 0;
    int b = 0;
    vec3 ray_pos = v[0].pos;
    float pdf_fwd = pdf, pdf_rev = 0.0f;
    vec3 wi = v[0].dir;
    const bool finite_light = lt_finite(v[0].light_flags);
    for (;;) {
        const int prev = b - 1;
        if (c.cnt) c.cnt->closest++;
        Hit h = accel_closest(c.a, ray_pos, wi, BDPT_TMIN, BDPT_TMAX);
        if (h.inst == 0xFFFFFFFFu) {
            if (eye) {
                v[b + 1].throughput = throughput;
                v[b + 1].pdf_fwd = pdf_fwd;
                b++;
            }
            break;
        }
        Payload p;
        fill_payload(c.s, h, ray_pos, wi, BDPT_TMIN, p);
        vec3 wo = v[prev + 1].pos - p.pos;
        const float wo_len = length(wo);
        wo /= wo_len;
        vec3 n_s = p.n_s;
        bool side = true;
        vec3 n_g = p.n_g;
        if (dot(p.n_g, wo) < 0.0f) n_g = -n_g;
        if (dot(n_g, n_s) < 0.0f) {
            n_s = n_s * -1.0f;
            side = false;
        }
        PV& cur = v[b + 1];
        cur.pdf_fwd = pdf_fwd * fabsf(dot(wo, n_s)) / (wo_len * wo_len);
        cur.n_s = n_s;
        cur.area = p.area;
        cur.pos = p.pos;
        cur.uv = p.uv;
        cur.material_idx = p.material_idx;
        cur.throughput = throughput;
        cur.side = side ? 1u : 0u;
        cur.mode = eye ? 1u : 0u;
        const Material mat = load_material(c.s, p.material_idx, p.uv);
        const bool mat_specular = (mat.bsdf_props & LT_BSDF_FLAG_SPECULAR) == LT_BSDF_FLAG_SPECULAR;
        const bool mat_transmissive = (mat.bsdf_props & LT_BSDF_FLAG_TRANSMISSION) == LT_BSDF_FLAG_TRANSMISSION;
        cur.delta = mat_specular ? 1u : 0u;
        if (++b >= max_depth) break;
        float cos_theta;
        vec3 r3;
        r3.x = rand(seed);
        r3.y = rand(seed);
        r3.z = rand(seed);
        const vec3 f = sample_bsdf(n_s, wo, mat, eye ? 1u : 0u, side, wi, pdf_fwd, cos_theta, r3);
        const bool same_hem = same_hemisphere(wi, wo, n_s);
        if ((f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) || pdf_fwd == 0.0f || (!same_hem && !mat_transmissive)) break;
        throughput *= f * fabsf(cos_theta) / pdf_fwd;
        pdf_rev = pdf_fwd;
        if (!mat_specular) pdf_rev = bsdf_pdf(mat, n_s, wi, wo, side);
        const bool g_term = eye ? true : (prev > -1 || finite_light);
        if (g_term) pdf_rev *= fabsf(dot(v[prev + 1].n_s, wo)) / (wo_len * wo_len);
        v[prev + 1].pdf_rev = pdf_rev;
        ray_pos = offset_ray(p.pos, n_g);
    }
    return b;
}

// bdpt_commons.glsl:272-473 on copies of the two subpaths
static float calc_mis_weight(BdptCtx& c, const PV* lv0, const PV* cv0, int s, int t, const PV& sampled) {
    if (s + t == 2) return 1.0f;
    PV L[BDPT_MAX_DEPTH + 1], Cm[BDPT_MAX_DEPTH + 1];
    for (int i = 0; i < s; i++) L[i] = lv0[i];
    for (int i = 0; i < t; i++) Cm[i] = cv0[i];
    // light_vtx(s - 2) / cam_vtx(t - 2) may be read at index -1 by the
    // reference only under conditions that never hold; keep slot -1 unused
    if (s == 1) {
        L[0].pdf_fwd = sampled.pdf_fwd;
        L[0].pos = sampled.pos;
        L[0].n_s = sampled.n_s;
    }
    if (t == 1) {
        Cm[0].pdf_fwd = sampled.pdf_fwd;
        Cm[0].pos = sampled.pos;
        Cm[0].n_s = sampled.n_s;
    }
    if (t > 0) Cm[t - 1].delta = 0;
    if (s > 0) L[s - 1].delta = 0;
    auto remap0 = [](float x) { return x != 0.0f ? x : 1.0f; };
    if (t > 0) {
        if (s > 0) {
            vec3 dir = Cm[t - 1].pos - L[s - 1].pos;
            float dl = length(dir);
            dir /= dl;
            const Material mat = load_material(c.s, L[s - 1].material_idx, L[s - 1].uv);
            float pdf_rev = 0.0f;
            if (s >= 2) {
                vec3 wo = normalize(L[s - 2].pos - L[s - 1].pos);
                pdf_rev = bsdf_pdf(mat, L[s - 1].n_s, wo, dir, L[s - 1].side == 1);
                pdf_rev *= fabsf(dot(dir, Cm[t - 1].n_s)) / (dl * dl);
            } else if (s == 1) {
                if (!lt_finite(L[0].light_flags)) {
                    pdf_rev = c.light_pdf_pos;
                    pdf_rev *= fabsf(dot(dir, Cm[t - 1].n_s));
                } else {
                    pdf_rev = light_pdf_dir(c.s, L[0].light_idx, L[0].light_flags, L[0].n_s, dir);
                    pdf_rev *= fabsf(dot(dir, Cm[t - 1].n_s)) / (dl * dl);
                }
            }
            Cm[t - 1].pdf_rev = pdf_rev;
        } else {
            Cm[t - 1].pdf_rev = 1.0f / ((float)c.s.total_light_triangle_cnt * Cm[t - 1].area);
        }
    }
    if (t > 1) {
        vec3 dir = Cm[t - 2].pos - Cm[t - 1].pos;
        float dl = length(dir);
        dir /= dl;
        if (s > 0) {
            const Material mat = load_material(c.s, Cm[t - 1].material_idx, Cm[t - 1].uv);
            vec3 wo = normalize(L[s - 1].pos - Cm[t - 1].pos);
            Cm[t - 2].pdf_rev = bsdf_pdf(mat, Cm[t - 1].n_s, wo, dir, Cm[t - 1].side == 1);
            if (Cm[t - 2].pdf_rev != 0.0f) Cm[t - 2].pdf_rev *= fabsf(dot(dir, Cm[t - 2].n_s)) / (dl * dl);
        } else {
            float cx = dot(Cm[t - 1].n_s, dir), cy = dot(Cm[t - 2].n_s, dir);
            Cm[t - 2].pdf_rev = fabsf(cx * cy) / (PI * dl * dl);
        }
    }
    if (s > 0) {
        vec3 dir = L[s - 1].pos - Cm[t - 1].pos;
        float dl = length(dir);
        dir /= dl;
        if (t == 1) {
            float ct = dot(Cm[0].n_s, dir);
            float pdf = 1.0f / (Cm[0].area * c.screen_size * ct * ct * ct);
            pdf *= fabsf(dot(dir, L[s - 1].n_s)) / (dl * dl);
            L[s - 1].pdf_rev = pdf;
        } else {
            vec3 wo = normalize(Cm[t - 2].pos - Cm[t - 1].pos);
            const Material mat = load_material(c.s, Cm[t - 1].material_idx, Cm[t - 1].uv);
            L[s - 1].pdf_rev = bsdf_pdf(mat, Cm[t - 1].n_s, wo, dir, Cm[t - 1].side == 1);
            if ((s == 1 && lt_finite(L[0].light_flags)) || s > 1)
                L[s - 1].pdf_rev *= fabsf(dot(dir, L[s - 1].n_s)) / (dl * dl);
        }
    }
    if (s > 1) {
        vec3 dir = L[s - 2].pos - L[s - 1].pos;
        vec3 wo = normalize(Cm[t - 1].pos - L[s - 1].pos);
        float dl = length(dir);
        dir /= dl;
        const Material mat = load_material(c.s, L[s - 1].material_idx, L[s - 1].uv);
        L[s - 2].pdf_rev = bsdf_pdf(mat, L[s - 1].n_s, wo, dir, L[s - 1].side == 1);
        if ((s == 2 && lt_finite(L[0].light_flags)) || s > 2)
            L[s - 2].pdf_rev *= fabsf(dot(dir, L[s - 2].n_s)) / (dl * dl);
    }
    float sum_ri = 0.0f, w = 1.0f;
    for (int i = t - 1; i > 0; i--) {
        w *= remap0(Cm[i].pdf_rev) / remap0(Cm[i].pdf_fwd);
        if (Cm[i].delta == 0 && Cm[i - 1].delta == 0) sum_ri += c.power ? w * w : w;
    }
    w = 1.0f;
    for (int i = s - 1; i >= 0; i--) {
        w *= remap0(L[i].pdf_rev) / remap0(L[i].pdf_fwd);
        bool delta_prev = i > 0 ? L[i - 1].delta == 1 : lt_delta(L[0].light_flags);
        if (L[i].delta == 0 && !delta_prev) sum_ri += c.power ? w * w : w;
    }
    return 1.0f / (1.0f + sum_ri);
}

static vec3 eval_f(const lt_scene_desc& s, const PV& v, vec3 wo, vec3 wi) {
    const Material m = load_material(s, v.material_idx, v.uv);
    float p, pr;
    return eval_bsdf(v.n_s, wo, m, v.mode, v.side == 1, wi, p, pr, false);
}
static bool is_zero(vec3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

// bdpt_commons.glsl:475-533; returns the splat and its pixel (x, y) or false
static bool connect_cam(BdptCtx& c, const PV* lv, const PV* cv, int s, vec3& out, int& px, int& py) {
    PV sampled{};
    vec3 L = v3(0.0f);
    vec3 dir = cv[0].pos - lv[s - 1].pos;
    float len = length(dir);
    dir /= len;
    float cos_y = dot(dir, lv[s - 1].n_s);
    float cos_theta = dot(cv[0].n_s, -dir);
    if (cos_theta <= 0.0f) return false;
    float c3 = cos_theta * cos_theta * cos_theta;
    const float ratio = fabsf(cos_y) / (cv[0].area * c3 * len * len);
    vec3 ro = offset_ray2(lv[s - 1].pos, lv[s - 1].n_s);
    const vec3 wo = normalize(lv[s - 2].pos - lv[s - 1].pos);
    const vec3 f = eval_f(c.s, lv[s - 1], wo, dir);
    if (is_zero(f)) return false;
    if (ratio > 0.0f) {
        if (!occluded(c, ro, dir, len - EPS)) {
            sampled.pos = cv[0].pos;
            sampled.n_s = cv[0].n_s;
            L = lv[s - 1].throughput * ratio * f / c.screen_size;
        }
    }
    dir = -dir;
    vec4 tg = mul(c.cam.view, vec4{dir.x, dir.y, dir.z, 0.0f});
    tg = {tg.x / tg.z, tg.y / tg.z, tg.z / tg.z, tg.w / tg.z};
    vec4 pr = mul(c.cam.projection, tg);
    pr = {-pr.x, -pr.y, -pr.z, -pr.w};
    // ivec2() truncates toward zero: a coordinate in (-1, W) lands in [0, W)
    const float fx = 0.5f * (1.0f + pr.x) * (float)c.W - 0.5f, fy = 0.5f * (1.0f + pr.y) * (float)c.H - 0.5f;
    if (!(fx > -1.0f && fx < (float)c.W && fy > -1.0f && fy < (float)c.H) || dot(dir, cv[0].n_s) < 0.0f)
        return false;
    px = (int)fx;
    py = (int)fy;
    float mis = 1.0f;
    if (luminance(L) != 0.0f) mis = calc_mis_weight(c, lv, cv, s, 1, sampled);
    out = mis * L;
    return true;
}

// bdpt_commons.glsl:535-641
static vec3 connect(BdptCtx& c, const PV* lv, const PV* cv, int s, int t, uvec4& seed) {
    vec3 L = v3(0.0f);
    PV sampled{};
    if (s == 0) {
        const uint32_t mi = cv[t - 1].material_idx;  // 0 on an escaped vertex
        L = load3(c.s.materials[mi].emissive_factor) * cv[t - 1].throughput;
    } else if (s == 1) {
        vec4 r4;
        r4.x = rand(seed);
        r4.y = rand(seed);
        r4.z = rand(seed);
        r4.w = rand(seed);
        vec3 wi, n, pos;
        float wi_len, pdf_pos_a, cos_y, pw, pdw;
        LightRecord rec{};
        const vec3 Le = sample_light_Li(c.s, r4, cv[t - 1].pos, (int)c.s.num_lights, pw, wi, wi_len, pdf_pos_a, cos_y,
                                        rec, n, pos, pdw);
        const float cos_x = fabsf(dot(wi, cv[t - 1].n_s));
        const vec3 ro = offset_ray2(cv[t - 1].pos, cv[t - 1].n_s);
        const vec3 wo = normalize(cv[t - 2].pos - cv[t - 1].pos);
        const vec3 f = eval_f(c.s, cv[t - 1], wo, wi);
        if (!is_zero(f)) {
            if (!occluded(c, ro, wi, wi_len - EPS)) {
                const float ltc = (float)c.s.total_light_triangle_cnt;
                const float pdf_light_w = light_pdf_a_to_w(rec.flags, pdf_pos_a, wi_len * wi_len, cos_y) / ltc;
                sampled.pdf_fwd = pdf_pos_a / ltc;
                sampled.pos = pos;
                sampled.n_s = n;
                sampled.delta = lt_delta(rec.flags) ? 1u : 0u;
                L = cv[t - 1].throughput * f * cos_x * Le / pdf_light_w;
            }
        }
    } else {
        vec3 n_s = lv[s - 1].n_s, n_t = cv[t - 1].n_s;
        vec3 d = lv[s - 1].pos - cv[t - 1].pos;
        float len = length(d);
        d /= len;
        float G = dot(n_s, -d) * dot(n_t, d) / (len * len);
        if (G > 0.0f) {
            vec3 wo1 = normalize(cv[t - 2].pos - cv[t - 1].pos);
            vec3 wo2 = normalize(lv[s - 2].pos - lv[s - 1].pos);
            const vec3 b1 = eval_f(c.s, cv[t - 1], wo1, d);
            const vec3 b2 = eval_f(c.s, lv[s - 1], wo2, -d);
            if (!is_zero(b1) && !is_zero(b2)) {
                vec3 ro = offset_ray2(cv[t - 1].pos, cv[t - 1].n_s);
                if (!occluded(c, ro, d, len - EPS)) L = lv[s - 1].throughput * G * b1 * b2 * cv[t - 1].throughput;
            }
        }
    }
    if (luminance(L) != 0.0f) L *= calc_mis_weight(c, lv, cv, s, t, sampled);
    return L;
}

// bdpt.rgen:40-81: the pixel's own estimate (connections with t >= 2 and
// s = 0) and the splats of its light subpath.  lv/cv hold L+1 vertices.
struct Splat {
    int x, y;
    vec3 c;
};
static vec3 bdpt_sample(const lt_scene_desc& s, const Accel* a, const lt_camera& cam, uint32_t W, uint32_t H,
                        int max_depth, bool power, uint32_t frame_word, uint32_t x, uint32_t y,
                        std::vector<Splat>& splats, Counters* cnt) {
    uvec4 seed = {x, y, frame_word, 0u};
    BdptCtx c{s, a, cnt, max_depth, W, H, (float)(W * H), 0.0f, 0.0f, cam, power};
    // camera (bdpt.rgen:41-52): pixel centre, no jitter
    const float ux = ((float)x + 0.5f) / (float)W, uy = ((float)y + 0.5f) / (float)H;
    const vec2 d = {ux * 2.0f - 1.0f, uy * 2.0f - 1.0f};
    vec4 o4 = mul(cam.inv_view, vec4{0.0f, 0.0f, 0.0f, 1.0f});
    vec4 ai = mul(cam.inv_projection, vec4{2.0f / (float)W, 2.0f / (float)H, 0.0f, 1.0f});
    ai = {ai.x / ai.w, ai.y / ai.w, ai.z / ai.w, ai.w / ai.w};
    c.cam_area = fabsf(ai.x * ai.y);
    PV lv[BDPT_MAX_DEPTH + 2], cv[BDPT_MAX_DEPTH + 2];
    memset(lv, 0, sizeof(lv));
    memset(cv, 0, sizeof(cv));
    // light subpath (bdpt_commons.glsl:199-244)
    int n_light = 0;
    {
        vec4 rp;
        rp.x = rand(seed);
        rp.y = rand(seed);
        rp.z = rand(seed);
        rp.w = rand(seed);
        vec2 rd;
        rd.x = rand(seed);
        rd.y = rand(seed);
        float cos_theta, pdf_pos, pdf_dir;
        uint32_t flags;
        vec3 pos, wi, n;
        const vec3 Le = sample_light_Le(s, rp, rd, (int)s.num_lights, (int)s.total_light_triangle_cnt, cos_theta,
                                        flags, pos, wi, n, pdf_pos, pdf_dir);
        if (pdf_dir > 0.0f) {
            c.light_pdf_pos = pdf_pos;
            lv[0].pos = pos;
            lv[0].light_flags = flags;
            lv[0].light_idx = (uint32_t)(rp.x * (float)s.num_lights);
            lv[0].delta = 0;
            lv[0].dir = wi;
            lv[0].pdf_fwd = pdf_pos;
            lv[0].n_s = n;
            lv[0].side = 1;
            lv[0].mode = 0;
            vec3 thr = Le * cos_theta / (pdf_dir * lv[0].pdf_fwd);
            lv[0].throughput = Le;
            n_light = random_walk(c, lv, max_depth, thr, pdf_dir, seed, false) + 1;
            if (!lt_finite(flags)) lv[1].pdf_fwd = pdf_pos * fabsf(dot(wi, lv[1].n_s));
            if (lt_delta(flags)) lv[0].pdf_fwd = 0.0f;
        }
    }
    // camera subpath (bdpt_commons.glsl:246-270)
    int n_cam;
    {
        vec4 tg = mul(cam.inv_projection, vec4{d.x, d.y, 1.0f, 1.0f});
        vec3 tn = normalize(v3(tg.x, tg.y, tg.z));
        vec4 d4 = mul(cam.inv_view, vec4{tn.x, tn.y, tn.z, 0.0f});
        vec4 fw = mul(cam.inv_view, vec4{0.0f, 0.0f, 1.0f, 0.0f});
        cv[0].pos = {o4.x, o4.y, o4.z};
        cv[0].dir = {d4.x, d4.y, d4.z};
        cv[0].area = c.cam_area;
        cv[0].throughput = v3(1.0f);
        cv[0].delta = 0;
        cv[0].n_s = {-fw.x, -fw.y, -fw.z};
        cv[0].side = 1;
        float ct = dot(cv[0].dir, cv[0].n_s);
        float pdf = 1.0f / (c.cam_area * c.screen_size * ct * ct * ct);
        n_cam = random_walk(c, cv, max_depth, v3(1.0f), pdf, seed, true) + 1;
    }
    vec3 col = v3(0.0f);
    for (int t = 1; t <= n_cam; t++) {
        for (int sidx = 0; sidx <= n_light; sidx++) {
            int depth = sidx + t - 2;
            if (depth > max_depth - 1 || depth < 0 || (sidx == 1 && t == 1)) continue;
            if (t == 1) {
                vec3 sc;
                int px, py;
                if (connect_cam(c, lv, cv, sidx, sc, px, py) && luminance(sc) > 0.0f)
                    splats.push_back({px, py, sc});
            } else {
                col += connect(c, lv, cv, sidx, t, seed);
            }
        }
    }
    return col;
}

}  // namespace ora

using namespace ora;

extern "C" {

typedef struct ora_bdpt_params {
    uint32_t width, height;
    int32_t max_depth;
    uint32_t frame_start;     // first frame_num
    uint32_t frames;          // frames to accumulate
    uint32_t seed;            // XOR'ed into the frame word (bdpt.rgen:36-37 uses frame ^ time)
    uint32_t power;           // 1 = power heuristic, 0 = the reference's balance heuristic
    uint32_t x0, y0, x1, y1;  // camera-pixel window [x0,x1) x [y0,y1); x1 = 0 means the full image
} ora_bdpt_params;

// accum: RGBA32F row-major over the full image.  splat: float3 per pixel,
// persists between calls like the reference's color storage buffer; only
// the window's pixels read and clear it.  Returns wall seconds.
double ora_render_bdpt(const lt_scene_desc* s, void* accel, const lt_camera* cam, const ora_bdpt_params* p,
                       float* accum, float* splat, int threads, uint64_t* closest_rays, uint64_t* shadow_rays) {
    const uint32_t W = p->width, H = p->height;
    const uint32_t x0 = p->x0, y0 = p->y0, x1 = p->x1 ? p->x1 : W, y1 = p->y1 ? p->y1 : H;
    const uint32_t tw = (x1 - x0 + 15) / 16, th = (y1 - y0 + 15) / 16;
    const uint32_t ntiles = tw * th;
    std::atomic<uint64_t> cl{0}, sh{0};
    std::vector<float> fcol((size_t)W * H * 3, 0.0f);
    std::vector<std::vector<Splat>> tile_splats(ntiles);
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t fr = 0; fr < p->frames; fr++) {
        const uint32_t frame_num = p->frame_start + fr;
        const uint32_t word = frame_num ^ p->seed;
        std::atomic<uint32_t> next{0};
        auto worker = [&]() {
            Counters c;
            for (;;) {
                uint32_t t = next.fetch_add(1);
                if (t >= ntiles) break;
                uint32_t tx = x0 + (t % tw) * 16, ty = y0 + (t / tw) * 16;
                std::vector<Splat>& sp = tile_splats[t];
                sp.clear();
                for (uint32_t y = ty; y < std::min(ty + 16, y1); y++)
                    for (uint32_t x = tx; x < std::min(tx + 16, x1); x++) {
                        vec3 col = bdpt_sample(*s, (const Accel*)accel, *cam, W, H, p->max_depth, p->power != 0, word,
                                               x, y, sp, &c);
                        float* fc = &fcol[3 * ((size_t)y * W + x)];
                        fc[0] = col.x;
                        fc[1] = col.y;
                        fc[2] = col.z;
                    }
            }
            cl += c.closest;
            sh += c.shadow;
        };
        if (threads <= 1) {
            worker();
        } else {
            std::vector<std::thread> ts;
            for (int i = 0; i < threads; i++) ts.emplace_back(worker);
            for (auto& t : ts) t.join();
        }
        // splats in tile order, then each window pixel adds and clears its
        // entry (bdpt.rgen:66-70, 79-81), NaN guard and running mean (:82-90)
        for (uint32_t t = 0; t < ntiles; t++)
            for (const Splat& q : tile_splats[t]) {
                float* d = splat + 3 * ((size_t)q.y * W + q.x);
                d[0] += q.c.x;
                d[1] += q.c.y;
                d[2] += q.c.z;
            }
        for (uint32_t y = y0; y < y1; y++)
            for (uint32_t x = x0; x < x1; x++) {
                const size_t i = (size_t)y * W + x;
                float* sp = splat + 3 * i;
                vec3 col = v3(fcol[3 * i], fcol[3 * i + 1], fcol[3 * i + 2]) + v3(sp[0], sp[1], sp[2]);
                sp[0] = sp[1] = sp[2] = 0.0f;
                if (std::isnan(luminance(col))) continue;
                float* px = accum + 4 * i;
                if (frame_num > 0) {
                    float w = 1.0f / (float)(frame_num + 1);
                    px[0] = px[0] * (1.0f - w) + col.x * w;
                    px[1] = px[1] * (1.0f - w) + col.y * w;
                    px[2] = px[2] * (1.0f - w) + col.z * w;
                } else {
                    px[0] = col.x;
                    px[1] = col.y;
                    px[2] = col.z;
                }
                px[3] = 1.0f;
            }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (closest_rays) *closest_rays = cl.load();
    if (shadow_rays) *shadow_rays = sh.load();
    return std::chrono::duration<double>(t1 - t0).count();
}

// one emission sample (commons.glsl:335-406); out = [Le3, pos3, wi3, n3,
// cos_from_light, pdf_pos_a, pdf_dir_w, flags]
void ora_sample_light_Le(const lt_scene_desc* s, const float* rands6, float* out) {
    float cf, ppa, pdw;
    uint32_t flags;
    vec3 pos, wi, n;
    vec3 Le = sample_light_Le(*s, vec4{rands6[0], rands6[1], rands6[2], rands6[3]}, v2(rands6[4], rands6[5]),
                              (int)s->num_lights, (int)s->total_light_triangle_cnt, cf, flags, pos, wi, n, ppa, pdw);
    const vec3 v[4] = {Le, pos, wi, n};
    for (int i = 0; i < 4; i++) {
        out[3 * i] = v[i].x;
        out[3 * i + 1] = v[i].y;
        out[3 * i + 2] = v[i].z;
    }
    out[12] = cf;
    out[13] = ppa;
    out[14] = pdw;
    out[15] = (float)flags;
}

}  // extern "C"
