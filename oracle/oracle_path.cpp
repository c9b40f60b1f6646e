// oracle_path.cpp — the reference Path integrator restated on the CPU
// (TEST INFRASTRUCTURE, see oracle.h).
//
// Follows src/shaders/integrators/path/path.rgen:26-113 and
// src/shaders/integrators/pt_commons.glsl:3-47.  The closest-hit payload is a
// persistent variable as in the GLSL: a miss (ray.rmiss:13) only sets
// material_idx = -1 and leaves the other fields of the previous hit in place,
// which pt_commons.glsl:32 reads when its MIS ray misses.
#include "oracle.h"

namespace ora {

static const float PATH_TMIN = 0.001f;   // path.rgen:19
static const float PATH_TMAX = 10000.0f; // path.rgen:20
static const int RR_MIN_DEPTH = 3;       // path.rgen:21

struct PState {
    Payload pl;
    bool hit_valid;  // material_idx != -1
};

static void trace_closest(const lt_scene_desc& s, const Accel* a, vec3 o, vec3 d, float tmin, float tmax,
                          PState& st, Counters* cnt) {
    if (cnt) cnt->closest++;
    Hit h = accel_closest(a, o, d, tmin, tmax);
    if (h.inst == 0xFFFFFFFFu) {
        st.hit_valid = false;  // ray.rmiss: payload.material_idx = -1 only
        return;
    }
    st.hit_valid = true;
    fill_payload(s, h, o, d, tmin, st.pl);
}

// pt_commons.glsl:3-42
static vec3 uniform_sample_light(const lt_scene_desc& s, const Accel* a, uvec4& seed, const Material& mat, vec3 pos,
                                 bool side, vec3 n_s, vec3 wo, PState& st, Counters* cnt, float& pick_pdf) {
    vec3 res = v3(0.0f);
    vec3 wi;
    float wi_len, pdf_light_w, pdf_light_a, cos_from_light, pdf_dir;
    LightRecord rec{};
    vec3 ln, lpos;
    vec4 r4;
    r4.x = rand(seed);
    r4.y = rand(seed);
    r4.z = rand(seed);
    r4.w = rand(seed);
    const vec3 Le = sample_light_Li(s, r4, pos, (int)s.num_lights, pdf_light_w, wi, wi_len, pdf_light_a,
                                    cos_from_light, rec, ln, lpos, pdf_dir);
    // probability of the pick made above: a light uniformly (commons.glsl:227),
    // then, for an area light, one of its triangles uniformly (commons.glsl:238)
    const lt_light& picked = s.lights[rec.light_idx];
    pick_pdf = 1.0f / ((float)s.num_lights *
                       ((picked.light_flags & 7u) == LT_LIGHT_AREA ? (float)picked.num_triangles : 1.0f));
    const vec3 p = offset_ray2(pos, n_s);
    float bsdf_pdf, bsdf_rev;
    float cos_x = dot(n_s, wi);
    vec3 f = eval_bsdf(n_s, wo, mat, 1, side, wi, bsdf_pdf, bsdf_rev, false);
    if (cnt) cnt->shadow++;
    bool visible = !accel_any(a, p, wi, 0.0f, wi_len - EPS);
    if (visible && pdf_light_w > 0.0f) {
        const float mis = (rec.flags & LT_LIGHT_DELTA) ? 1.0f : 1.0f / (1.0f + bsdf_pdf / pdf_light_w);
        res += mis * f * fabsf(cos_x) * Le / pdf_light_w;
    }
    if ((rec.flags & 7u) == LT_LIGHT_AREA) {
        vec3 r3;
        r3.x = rand(seed);
        r3.y = rand(seed);
        r3.z = rand(seed);
        f = sample_bsdf(n_s, wo, mat, 1, side, wi, bsdf_pdf, cos_x, r3);
        if (bsdf_pdf != 0.0f) {
            trace_closest(s, a, p, wi, PATH_TMIN, PATH_TMAX, st, cnt);
            if (st.pl.triangle_idx == rec.triangle_idx && st.pl.instance_idx == rec.instance_idx) {
                const float wl = length(st.pl.pos - pos);
                const float g = fabsf(dot(st.pl.n_s, -wi)) / (wl * wl);
                const float mis = 1.0f / (1.0f + pdf_light_a / (g * bsdf_pdf));
                res += f * mis * fabsf(cos_x) * Le / bsdf_pdf;
            }
        }
    }
    return res;
}

// uniform_sample_light of pt_commons.glsl:3-42 at a vertex whose closest-hit
// payload is `hit` (what a missed MIS ray compares, ray.rmiss:13), for the
// other integrators that call it (SPPM's eye pass).
vec3 nee_uniform_sample_light(const lt_scene_desc& s, const Accel* a, uvec4& seed, const Material& mat, vec3 pos,
                              bool side, vec3 n_s, vec3 wo, const Payload& hit, Counters* cnt) {
    PState st{};
    st.pl = hit;
    st.hit_valid = true;
    float pick;
    return uniform_sample_light(s, a, seed, mat, pos, side, n_s, wo, st, cnt, pick);
}

// path.rgen:26-104
vec3 path_sample(const lt_scene_desc& s, const Accel* a, const lt_camera& cam, const PathParams& pp, uint32_t x,
                 uint32_t y, bool& nan_dropped, Counters* cnt) {
    uvec4 seed = {x, y, pp.frame_num, 0u};
    const vec2 pixel = {(float)x + 0.5f, (float)y + 0.5f};
    float r0 = rand(seed);
    float r1 = rand(seed);
    vec2 rands = {r0 - 0.5f, r1 - 0.5f};
    const vec2 in_uv = {(pixel.x + rands.x) / (float)pp.width, (pixel.y + rands.y) / (float)pp.height};
    vec2 d = {in_uv.x * 2.0f - 1.0f, in_uv.y * 2.0f - 1.0f};
    vec4 o4 = mul(cam.inv_view, vec4{0.0f, 0.0f, 0.0f, 1.0f});
    vec4 tg = mul(cam.inv_projection, vec4{d.x, d.y, 1.0f, 1.0f});
    vec3 tn = normalize(v3(tg.x, tg.y, tg.z));
    vec4 dir4 = mul(cam.inv_view, vec4{tn.x, tn.y, tn.z, 0.0f});
    vec3 origin = {o4.x, o4.y, o4.z};
    vec3 direction = {dir4.x, dir4.y, dir4.z};

    vec3 col = v3(0.0f);
    bool last_specular = false;
    vec3 throughput = v3(1.0f);
    PState st{};
    st.hit_valid = false;
    for (int depth = 0;; depth++) {
        trace_closest(s, a, origin, direction, PATH_TMIN, PATH_TMAX, st, cnt);
        if (!st.hit_valid) {
            if (depth > 0 || pp.direct_lighting == 1)
                col += throughput * shade_atmosphere(s, pp.sky_col, origin, direction, PATH_TMAX);  // path.rgen:50-53
            break;
        }
        const Payload hp = st.pl;
        const Material hit_mat = load_material(s, hp.material_idx, hp.uv);
        if ((depth == 0 && pp.direct_lighting == 1) || last_specular) col += throughput * hit_mat.emissive_factor;
        if (depth >= pp.max_depth - 1) break;
        const vec3 wo = -direction;
        vec3 n_s = hp.n_s;
        bool side = true;
        vec3 n_g = hp.n_g;
        if (dot(hp.n_g, wo) < 0.0f) n_g = -n_g;
        if (dot(n_g, hp.n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        origin = offset_ray(hp.pos, n_g);
        last_specular = is_specular(hit_mat);
        if (!last_specular) {
            // path.rgen:76-78: the reference divides by 1 / light_triangle_count
            // whichever light was picked (unbiased for one emitter mesh only);
            // nee_pick = 1 divides by the pick's own probability instead
            float light_pick_pdf = 1.0f / (float)s.total_light_triangle_cnt;
            if (depth > 0 || pp.direct_lighting == 1) {
                float pick_pdf;
                const vec3 nee = uniform_sample_light(s, a, seed, hit_mat, hp.pos, side, n_s, wo, st, cnt, pick_pdf);
                if (pp.nee_pick == 1) light_pick_pdf = pick_pdf;
                col += throughput * nee / light_pick_pdf;
            }
        }
        float pdf, cos_theta;
        vec3 r3;
        r3.x = rand(seed);
        r3.y = rand(seed);
        r3.z = rand(seed);
        const vec3 f = sample_bsdf(n_s, wo, hit_mat, 1, side, direction, pdf, cos_theta, r3);
        if (pdf == 0.0f) break;
        throughput *= f * fabsf(cos_theta) / pdf;
        float rr_scale = 1.0f;
        if (hit_mat.bsdf_props & LT_BSDF_FLAG_TRANSMISSION) rr_scale *= side ? 1.0f / hit_mat.ior : hit_mat.ior;
        if (depth > RR_MIN_DEPTH) {
            float rr_prob = fminf(0.95f, luminance(throughput) * rr_scale);
            if (rr_prob == 0.0f || rr_prob < rand(seed))
                break;
            else
                throughput /= rr_prob;
        }
    }
    nan_dropped = std::isnan(luminance(col));
    return col;
}

}  // namespace ora
