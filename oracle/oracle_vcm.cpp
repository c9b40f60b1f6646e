// oracle_vcm.cpp — the reference VCM integrator restated on the CPU
// (TEST INFRASTRUCTURE, see oracle.h).
//
// Follows src/shaders/integrators/vcm/vcm_light.rgen:123-166 (RAY_GUIDE = 0),
// vcm_eye.rgen:42-98 and src/shaders/integrators/vcm_commons.glsl:17-631,
// with the per-frame radius and grid of src/RayTracer/VCM.cpp:131-158.
//
// Two modes:
//   reference (corrected = 0): the shipped estimator, bug for bug — vertex
//     connections and merges keep only the last contribution
//     (vcm_commons.glsl:240,287), the photon table keeps one photon per hash
//     cell with an insertion count (:455-469), and vertex connections weight
//     the camera side with the NEE sample's reverse pdf (:231).  The table's
//     "last writer" is made deterministic: photons are inserted in light
//     path order, then vertex order.
//   corrected (corrected = 1): contributions are summed, every photon is kept
//     in the cell it falls in (bucketed by the same hash, collisions
//     rejected by the cell coordinate), and each connection uses its own
//     reverse pdf.  This is the VCM estimator of Georgiev et al. 2012; with
//     vertex merging off it is unbiased.
// Splats of the light pass are summed in light-path order before the eye
// pass reads them (the reference's unsynchronised += races,
// vcm_commons.glsl:399-404).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "oracle.h"

namespace ora {

static const float VCM_TMIN = 0.001f;   // vcm_light.rgen:54
static const float VCM_TMAX = 10000.0f;  // vcm_light.rgen:55

struct VV {  // VCMVertex (vcm_commons.h:88-103)
    vec3 wi, wo, n_s, pos;
    vec2 uv;
    vec3 thr;
    uint32_t mat, path_len;
    float area, d_vcm, d_vc, d_vm;
    uint32_t side;
};

struct Photon {  // VCMPhotonHash (vcm_commons.h:57-66)
    vec3 pos;
    float d_vm;
    vec3 wi;
    float d_vcm;
    vec3 thr;
    int count;
    vec3 nrm;
    uint32_t path_len;
};

struct VState {  // VCMState (utils.glsl:28-39)
    vec3 wi, n_s, pos, thr;
    float d_vcm, d_vc, d_vm;
};

struct Splat {
    int x, y;
    vec3 c;
};

// utils.glsl:62-65 (GLSL int products wrap like uint32)
static inline uint32_t cell_hash(int x, int y, int z, uint32_t size) {
    uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
    return h % (10u * size);
}

struct VcmCtx {
    const lt_scene_desc& s;
    const Accel* a;
    const lt_camera& cam;
    uint32_t W, H, N;  // N = W * H light paths
    int D;
    float radius, eta_vcm, eta_vc, eta_vm, screen_size, cam_area;
    int grid_res[3];
    vec3 bmin, bmax;
    int use_vc, use_vm;
    bool corrected;
    vec3 sky;
    std::vector<VV> verts;  // N * (D + 1)
    std::vector<uint32_t> cnt;
    std::vector<Photon> table;  // reference: 10 N cells
    std::vector<uint32_t> bstart, bitems;  // corrected: bucket -> photon refs (p * (D + 1) + i)
};

// vcm_commons.glsl:19-23 (the clamp's result is discarded there, so indices
// can fall outside the grid)
static inline void grid_idx(const VcmCtx& c, vec3 p, int o[3]) {
    const vec3 q = (p - c.bmin) / (c.bmax - c.bmin);
    o[0] = (int)(q.x * (float)c.grid_res[0]);
    o[1] = (int)(q.y * (float)c.grid_res[1]);
    o[2] = (int)(q.z * (float)c.grid_res[2]);
}

static bool occluded(const VcmCtx& c, vec3 o, vec3 d, float tmax, Counters* cnt) {
    if (cnt) cnt->shadow++;
    return accel_any(c.a, o, d, 0.0f, tmax);
}

// vcm_commons.glsl:24-75
static bool connect_cam(const VcmCtx& c, vec3 cam_pos, vec3 cam_nrm, vec3 n_s, vec3 payload_n_s, vec3 pos,
                        const VState& st, vec3 wo, const Material& mat, vec3& out, int& px, int& py, Counters* cnt) {
    out = v3(0.0f);
    vec3 dir = cam_pos - pos;
    float len = length(dir);
    dir /= len;
    float cos_y = dot(dir, n_s);
    float cos_theta = dot(cam_nrm, -dir);
    if (cos_theta <= 0.0f) return false;
    float c3 = cos_theta * cos_theta * cos_theta;
    const float ratio = fabsf(cos_y) / (c.cam_area * c3 * len * len);
    vec3 ro = offset_ray2(pos, n_s);
    float pdf_rev, pdf_fwd;
    const vec3 f = eval_bsdf(n_s, wo, mat, 0, dot(payload_n_s, wo) > 0.0f, dir, pdf_fwd, pdf_rev, true);
    if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) return false;
    vec3 L = v3(0.0f);
    if (ratio > 0.0f) {
        if (!occluded(c, ro, dir, len - EPS, cnt)) {
            const float w_light = (ratio / c.screen_size) * (c.eta_vm + st.d_vcm + pdf_rev * st.d_vc);
            const float mis = 1.0f / (1.0f + w_light);
            L = mis * st.thr * ratio * f / c.screen_size;
        }
    }
    dir = -dir;
    vec4 tg = mul(c.cam.view, vec4{dir.x, dir.y, dir.z, 0.0f});
    tg = {tg.x / tg.z, tg.y / tg.z, tg.z / tg.z, tg.w / tg.z};
    vec4 pr = mul(c.cam.projection, tg);
    pr = {-pr.x, -pr.y, -pr.z, -pr.w};
    // ivec2() truncates toward zero: a coordinate in (-1, W) lands in [0, W)
    const float fx = 0.5f * (1.0f + pr.x) * (float)c.W - 0.5f, fy = 0.5f * (1.0f + pr.y) * (float)c.H - 0.5f;
    if (!(fx > -1.0f && fx < (float)c.W && fy > -1.0f && fy < (float)c.H) || dot(dir, cam_nrm) < 0.0f) return false;
    px = (int)fx;
    py = (int)fy;
    out = L;
    return true;
}

// vcm_light.rgen:123-166 + vcm_commons.glsl:77-130, 302-474 for light path p
static void light_path(VcmCtx& c, uint32_t p, uint32_t light_word, std::vector<Splat>& splats, Counters* cnt) {
    const uint32_t x = p / c.H, y = p % c.H;  // pixel_idx = x * H + y (vcm_light.rgen:56)
    uvec4 seed = {x, y, light_word, 0u};
    VV* lv = &c.verts[(size_t)p * (c.D + 1)];
    c.cnt[p] = 0;
    // vcm_generate_light_sample
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
    const vec3 Le = sample_light_Le(c.s, rp, rd, (int)c.s.num_lights, (int)c.s.total_light_triangle_cnt, cos_theta, flags,
                                    pos, wi, n, pdf_pos, pdf_dir);
    if (pdf_dir <= 0.0f) return;
    VState st;
    const float pdf_emit = pdf_pos * pdf_dir, pdf_direct = pdf_pos;
    st.pos = pos;
    st.wi = wi;
    st.n_s = n;
    st.thr = Le * cos_theta / (pdf_dir * pdf_pos);
    st.d_vcm = pdf_direct / pdf_emit;
    const bool finite = ((flags >> 4) & 1u) != 0, delta = ((flags >> 5) & 1u) != 0;
    st.d_vc = delta ? 0.0f : (finite ? cos_theta : 1.0f) / (pdf_dir * pdf_pos);
    st.d_vm = st.d_vc * c.eta_vc;
    // vcm_fill_light
    const vec4 o4 = mul(c.cam.inv_view, vec4{0.0f, 0.0f, 0.0f, 1.0f});
    const vec4 fw = mul(c.cam.inv_view, vec4{0.0f, 0.0f, 1.0f, 0.0f});
    const vec3 cam_pos = {o4.x, o4.y, o4.z}, cam_nrm = {-fw.x, -fw.y, -fw.z};
    uint32_t path_idx = 0;
    for (int depth = 1;; depth++) {
        if (cnt) cnt->closest++;
        Hit h = accel_closest(c.a, st.pos, st.wi, VCM_TMIN, VCM_TMAX);
        if (h.inst == 0xFFFFFFFFu) break;
        Payload pl;
        fill_payload(c.s, h, st.pos, st.wi, VCM_TMIN, pl);
        vec3 wo = st.pos - pl.pos;
        vec3 n_s = pl.n_s, n_g = pl.n_g;
        bool side = true;
        if (dot(pl.n_g, wo) <= 0.0f) n_g = -n_g;
        if (dot(n_g, n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        float dist = length(pl.pos - st.pos);
        float dist_sqr = dist * dist;
        wo /= dist;
        const Material mat = load_material(c.s, pl.material_idx, pl.uv);
        const bool spec = (mat.bsdf_props & LT_BSDF_FLAG_SPECULAR) == LT_BSDF_FLAG_SPECULAR;
        const float cos_theta_wo = fabsf(dot(wo, n_s));
        if (depth > 1 || finite) st.d_vcm *= dist_sqr;
        st.d_vcm /= cos_theta_wo;
        st.d_vc /= cos_theta_wo;
        st.d_vm /= cos_theta_wo;
        if (!spec && (c.use_vc == 1 || c.use_vm == 1)) {
            VV& v = lv[path_idx];
            v.wi = st.wi;
            v.wo = wo;
            v.n_s = n_s;
            v.pos = pl.pos;
            v.uv = pl.uv;
            v.mat = pl.material_idx;
            v.area = pl.area;
            v.thr = st.thr;
            v.d_vcm = st.d_vcm;
            v.d_vc = st.d_vc;
            v.d_vm = st.d_vm;
            v.path_len = (uint32_t)depth + 1;
            v.side = side ? 1u : 0u;
            path_idx++;
        }
        if (depth >= c.D) break;
        if (!spec && c.use_vc == 1 && depth < c.D) {
            vec3 sc;
            int px, py;
            if (connect_cam(c, cam_pos, cam_nrm, n_s, pl.n_s, pl.pos, st, wo, mat, sc, px, py, cnt) &&
                luminance(sc) > 0.0f)
                splats.push_back({px, py, sc});
        }
        float pdf_d, cos_t;
        vec3 r3;
        r3.x = rand(seed);
        r3.y = rand(seed);
        r3.z = rand(seed);
        const vec3 f = sample_bsdf(n_s, wo, mat, 0, side, st.wi, pdf_d, cos_t, r3);
        const bool same_hem = dot(st.wi, n_s) > 0.0f && dot(wo, n_s) > 0.0f;
        const bool transm = (mat.bsdf_props & LT_BSDF_FLAG_TRANSMISSION) == LT_BSDF_FLAG_TRANSMISSION;
        if ((f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) || pdf_d == 0.0f || (!same_hem && !transm)) break;
        float pdf_rev = pdf_d;
        if (!spec) pdf_rev = bsdf_pdf(mat, n_s, st.wi, wo, side);
        const float ac = fabsf(cos_t);
        st.pos = offset_ray(pl.pos, n_g);
        if (!spec) {
            st.d_vc = (ac / pdf_d) * (c.eta_vm + st.d_vcm + pdf_rev * st.d_vc);
            st.d_vm = (ac / pdf_d) * (1.0f + st.d_vcm * c.eta_vc + pdf_rev * st.d_vm);
            st.d_vcm = 1.0f / pdf_d;
        } else {
            st.d_vcm = 0.0f;
            st.d_vc *= ac;
            st.d_vm *= ac;
        }
        st.thr *= f * ac / pdf_d;
        st.n_s = n_s;
    }
    c.cnt[p] = path_idx;
}

// vcm_commons.glsl:132-143
static vec3 light_radiance(const VcmCtx& c, const Material& mat, const VState& cs, const Payload& pl, int d) {
    if (d == 1) return mat.emissive_factor;
    const float pdf_light_pos = 1.0f / (pl.area * (float)c.s.total_light_triangle_cnt);
    const float pdf_light_dir = fabsf(dot(pl.n_s, -cs.wi)) / PI;
    const float w_camera = pdf_light_pos * cs.d_vcm +
                           ((c.use_vc == 1 || c.use_vm == 1) ? (pdf_light_pos * pdf_light_dir) * cs.d_vc : 0.0f);
    const float mis = 1.0f / (1.0f + w_camera);
    return mis * mat.emissive_factor;
}

// vcm_commons.glsl:145-199
static vec3 connect_light(const VcmCtx& c, vec3 n_s, vec3 wo, const Material& mat, bool side, const VState& cs,
                          const Payload& pl, uvec4& seed, float& pdf_rev, Counters* cnt) {
    vec4 r4;
    r4.x = rand(seed);
    r4.y = rand(seed);
    r4.z = rand(seed);
    r4.w = rand(seed);
    vec3 wi, nn, lp;
    float wi_len, pdf_pos_w, pdf_pos_a, cos_y, pdf_pos_dir_w;
    LightRecord rec{};
    const vec3 Le = sample_light_Li(c.s, r4, pl.pos, (int)c.s.num_lights, pdf_pos_w, wi, wi_len, pdf_pos_a, cos_y, rec,
                                    nn, lp, pdf_pos_dir_w);
    const float cos_x = dot(wi, n_s);
    const vec3 ro = offset_ray2(pl.pos, n_s);
    float pdf_fwd;
    const vec3 f = eval_bsdf(n_s, wo, mat, 1, side, wi, pdf_fwd, pdf_rev, true);
    vec3 res = v3(0.0f);
    if (!(f.x == 0.0f && f.y == 0.0f && f.z == 0.0f)) {
        if (!occluded(c, ro, wi, wi_len - EPS, cnt)) {
            if ((rec.flags >> 5) & 1u) pdf_fwd = 0.0f;
            const float ltc = (float)c.s.total_light_triangle_cnt;
            const float w_light = pdf_fwd / (pdf_pos_w / ltc);
            const float w_cam =
                pdf_pos_dir_w * fabsf(cos_x) / (pdf_pos_w * cos_y) * (c.eta_vm + cs.d_vcm + cs.d_vc * pdf_rev);
            const float mis = 1.0f / (1.0f + w_light + w_cam);
            if (mis > 0.0f) res = mis * fabsf(cos_x) * f * cs.thr * Le / (pdf_pos_w / ltc);
        }
    }
    return res;
}

// vcm_commons.glsl:201-247
static vec3 connect_light_vertices(const VcmCtx& c, uint32_t lp, int depth, vec3 n_s, vec3 wo, const Material& mat,
                                   bool side, const VState& cs, float nee_pdf_rev, const Payload& pl, Counters* cnt) {
    vec3 res = v3(0.0f);
    const VV* lv = &c.verts[(size_t)lp * (c.D + 1)];
    for (uint32_t i = 0; i < c.cnt[lp]; i++) {
        const VV& v = lv[i];
        const uint32_t md = v.path_len + (uint32_t)depth - 1u;
        if (md >= (uint32_t)c.D) break;
        vec3 dir = v.pos - pl.pos;
        const float len = length(dir);
        const float len_sqr = len * len;
        dir /= len;
        const float cos_cam = dot(n_s, dir), cos_light = dot(v.n_s, -dir);
        const float G = cos_light * cos_cam / len_sqr;
        if (!(G > 0.0f)) continue;
        float cam_pdf_fwd, cam_pdf_rev, l_pdf_fwd, l_pdf_rev;
        const vec3 f_cam = eval_bsdf(n_s, wo, mat, 1, side, dir, cam_pdf_fwd, cam_pdf_rev, c.corrected);
        const Material lm = load_material(c.s, v.mat, v.uv);
        const vec3 f_light = eval_bsdf(v.n_s, v.wo, lm, 0, v.side == 1, -dir, l_pdf_fwd, l_pdf_rev, true);
        if ((f_light.x == 0.0f && f_light.y == 0.0f && f_light.z == 0.0f) ||
            (f_cam.x == 0.0f && f_cam.y == 0.0f && f_cam.z == 0.0f))
            continue;
        cam_pdf_fwd *= fabsf(cos_light) / len_sqr;
        l_pdf_fwd *= fabsf(cos_cam) / len_sqr;
        const float w_light = cam_pdf_fwd * (c.eta_vm + v.d_vcm + l_pdf_rev * v.d_vc);
        const float rev = c.corrected ? cam_pdf_rev : nee_pdf_rev;
        const float w_camera = l_pdf_fwd * (c.eta_vm + cs.d_vcm + rev * cs.d_vc);
        const float mis = 1.0f / (1.0f + w_camera + w_light);
        const vec3 ro = offset_ray2(pl.pos, n_s);
        if (!occluded(c, ro, dir, len - EPS, cnt)) {
            const vec3 r = mis * G * cs.thr * v.thr * f_cam * f_light;
            if (c.corrected)
                res += r;
            else
                res = r;
        }
    }
    return res;
}

// one merge candidate (vcm_commons.glsl:268-290); false if rejected
static bool merge_one(const VcmCtx& c, const Photon& ph, int depth, vec3 n_s, vec3 wo, const Material& mat, bool side,
                      const VState& cs, const Payload& pl, float norm, vec3& out) {
    const vec3 pp = pl.pos - ph.pos;
    const float dist_sqr = dot(pp, pp);
    if (dist_sqr > c.radius * c.radius) return false;
    const uint32_t dd = ph.path_len + (uint32_t)depth - 1u;
    if (dd > (uint32_t)c.D) return false;
    float pf, pr;
    const vec3 f = eval_bsdf(n_s, wo, mat, 1, side, ph.wi, pf, pr, true);
    if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) return false;
    const float w_light = ph.d_vcm * c.eta_vc + ph.d_vm * pf;
    const float w_cam = cs.d_vcm * c.eta_vc + cs.d_vm * pr;
    const float mis = 1.0f / (1.0f + w_light + w_cam);
    const float cos_nrm = dot(ph.nrm, n_s);
    if (!(cos_nrm > EPS)) return false;
    const float w = 1.0f - sqrtf(dist_sqr) / c.radius;
    out = w * mis * (float)ph.count * ph.thr * f * cs.thr * norm * 3.0f;
    return true;
}

// Merge-query statistics of the corrected estimator (for tests): queries,
// queries with more than 64 photons inside the radius, and the largest such
// count, over the frames of the last ora_render_vcm call.
static std::atomic<uint64_t> g_mq_total{0}, g_mq_over64{0}, g_mq_max{0};

// vcm_commons.glsl:251-296
static vec3 merge(const VcmCtx& c, int depth, vec3 n_s, vec3 wo, const Material& mat, bool side, const VState& cs,
                  const Payload& pl, float norm) {
    vec3 res = v3(0.0f);
    uint64_t in_radius = 0;
    int lo[3], hi[3];
    grid_idx(c, pl.pos - v3(c.radius), lo);
    grid_idx(c, pl.pos + v3(c.radius), hi);
    for (int x = lo[0]; x <= hi[0]; x++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int z = lo[2]; z <= hi[2]; z++) {
                const uint32_t h = cell_hash(x, y, z, c.N);
                if (!c.corrected) {
                    const Photon& ph = c.table[h];
                    vec3 r;
                    if (ph.count > 0 && merge_one(c, ph, depth, n_s, wo, mat, side, cs, pl, norm, r)) res = r;
                } else {
                    for (uint32_t k = c.bstart[h]; k < c.bstart[h + 1]; k++) {
                        const uint32_t ref = c.bitems[k];
                        const VV& v = c.verts[ref];
                        int g[3];
                        grid_idx(c, v.pos, g);
                        if (g[0] != x || g[1] != y || g[2] != z) continue;  // another cell of the bucket
                        const vec3 dp = pl.pos - v.pos;
                        if (dot(dp, dp) <= c.radius * c.radius) in_radius++;
                        const Photon ph{v.pos, v.d_vm, -v.wi, v.d_vcm, v.thr, 1, v.n_s, v.path_len};
                        vec3 r;
                        if (merge_one(c, ph, depth, n_s, wo, mat, side, cs, pl, norm, r)) res += r;
                    }
                }
            }
    if (c.corrected) {
        g_mq_total++;
        if (in_radius > 64) g_mq_over64++;
        uint64_t m = g_mq_max.load();
        while (in_radius > m && !g_mq_max.compare_exchange_weak(m, in_radius)) {
        }
    }
    return res;
}

// vcm_eye.rgen:42-98 + vcm_commons.glsl:477-631 for pixel (x, y); nan reports the guard
static vec3 eye_path(const VcmCtx& c, uint32_t x, uint32_t y, uint32_t eye_word, Counters* cnt) {
    uvec4 seed = {x, y, eye_word, 0u};
    const float r0 = rand(seed), r1 = rand(seed);
    const float ux = ((float)x + 0.5f + (r0 - 0.5f)) / (float)c.W, uy = ((float)y + 0.5f + (r1 - 0.5f)) / (float)c.H;
    const vec2 d = {ux * 2.0f - 1.0f, uy * 2.0f - 1.0f};
    const vec4 o4 = mul(c.cam.inv_view, vec4{0.0f, 0.0f, 0.0f, 1.0f});
    const vec4 tg = mul(c.cam.inv_projection, vec4{d.x, d.y, 1.0f, 1.0f});
    const vec3 tn = normalize(v3(tg.x, tg.y, tg.z));
    const vec4 d4 = mul(c.cam.inv_view, vec4{tn.x, tn.y, tn.z, 0.0f});
    const vec4 fw = mul(c.cam.inv_view, vec4{0.0f, 0.0f, 1.0f, 0.0f});
    VState cs;
    cs.wi = {d4.x, d4.y, d4.z};
    cs.pos = {o4.x, o4.y, o4.z};
    cs.thr = v3(1.0f);
    cs.n_s = {-fw.x, -fw.y, -fw.z};
    const float ct = fabsf(dot(cs.n_s, cs.wi));
    cs.d_vcm = c.cam_area * c.screen_size * ct * ct * ct;
    cs.d_vc = 0.0f;
    cs.d_vm = 0.0f;
    // vcm_trace_eye
    uint32_t lp = (uint32_t)(rand(seed) * c.screen_size);
    if (lp >= c.N) lp = c.N - 1u;  // float rounding of rand * N at very large N
    vec3 col = v3(0.0f);
    const float norm = 1.0f / (PI * (c.radius * c.radius) * c.screen_size);
    for (int depth = 1;; depth++) {
        if (cnt) cnt->closest++;
        Hit h = accel_closest(c.a, cs.pos, cs.wi, VCM_TMIN, VCM_TMAX);
        if (h.inst == 0xFFFFFFFFu) {
            col += cs.thr * c.sky;
            break;
        }
        Payload pl;
        fill_payload(c.s, h, cs.pos, cs.wi, VCM_TMIN, pl);
        vec3 wo = cs.pos - pl.pos;
        float dist = length(pl.pos - cs.pos);
        float dist_sqr = dist * dist;
        wo /= dist;
        vec3 n_s = pl.n_s, n_g = pl.n_g;
        bool side = true;
        if (dot(pl.n_g, wo) < 0.0f) n_g = -n_g;
        if (dot(n_g, n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        const float cos_wo = fabsf(dot(wo, n_s));
        const Material mat = load_material(c.s, pl.material_idx, pl.uv);
        const bool spec = (mat.bsdf_props & LT_BSDF_FLAG_SPECULAR) == LT_BSDF_FLAG_SPECULAR;
        cs.d_vcm *= dist_sqr;
        cs.d_vcm /= cos_wo;
        cs.d_vc /= cos_wo;
        cs.d_vm /= cos_wo;
        if (luminance(mat.emissive_factor) > 0.0f) col += cs.thr * light_radiance(c, mat, cs, pl, depth);
        float pdf_rev = 0.0f;
        if (!spec && depth < c.D) col += connect_light(c, n_s, wo, mat, side, cs, pl, seed, pdf_rev, cnt);
        if (!spec) col += connect_light_vertices(c, lp, depth, n_s, wo, mat, side, cs, pdf_rev, pl, cnt);
        if (!spec && c.use_vm == 1) col += merge(c, depth, n_s, wo, mat, side, cs, pl, norm);
        if (depth >= c.D) break;
        float pdf_d, cos_t;
        vec3 r3;
        r3.x = rand(seed);
        r3.y = rand(seed);
        r3.z = rand(seed);
        const vec3 f = sample_bsdf(n_s, wo, mat, 1, side, cs.wi, pdf_d, cos_t, r3);
        const bool transm = (mat.bsdf_props & LT_BSDF_FLAG_TRANSMISSION) == LT_BSDF_FLAG_TRANSMISSION;
        const bool same_hem = dot(cs.wi, n_s) > 0.0f && dot(wo, n_s) > 0.0f;
        if ((f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) || pdf_d == 0.0f || (!same_hem && !transm)) break;
        pdf_rev = pdf_d;
        if (!spec) pdf_rev = bsdf_pdf(mat, n_s, cs.wi, wo, side);
        const float ac = fabsf(cos_t);
        cs.pos = offset_ray(pl.pos, n_g);
        if (!spec) {
            cs.d_vc = (ac / pdf_d) * (c.eta_vm + cs.d_vcm + pdf_rev * cs.d_vc);
            cs.d_vm = (ac / pdf_d) * (1.0f + cs.d_vcm * c.eta_vc + pdf_rev * cs.d_vm);
            cs.d_vcm = 1.0f / pdf_d;
        } else {
            cs.d_vcm = 0.0f;
            cs.d_vc *= ac;
            cs.d_vm *= ac;
        }
        cs.thr *= f * ac / pdf_d;
        cs.n_s = n_s;
    }
    return col;
}

// photon storage after the light pass (vcm_commons.glsl:455-469)
static void build_photons(VcmCtx& c) {
    const size_t cells = (size_t)10 * c.N;
    if (!c.corrected) {
        c.table.assign(cells, Photon{});
        for (uint32_t p = 0; p < c.N; p++)
            for (uint32_t i = 0; i < c.cnt[p]; i++) {
                const VV& v = c.verts[(size_t)p * (c.D + 1) + i];
                int g[3];
                grid_idx(c, v.pos, g);
                Photon& ph = c.table[cell_hash(g[0], g[1], g[2], c.N)];
                ph.pos = v.pos;
                ph.wi = -v.wi;
                ph.d_vm = v.d_vm;
                ph.d_vcm = v.d_vcm;
                ph.thr = v.thr;
                ph.nrm = v.n_s;
                ph.path_len = v.path_len;
                ph.count++;
            }
        return;
    }
    c.bstart.assign(cells + 1, 0u);
    std::vector<uint32_t> hv;
    for (uint32_t p = 0; p < c.N; p++)
        for (uint32_t i = 0; i < c.cnt[p]; i++) {
            int g[3];
            grid_idx(c, c.verts[(size_t)p * (c.D + 1) + i].pos, g);
            c.bstart[cell_hash(g[0], g[1], g[2], c.N) + 1]++;
        }
    for (size_t k = 0; k < cells; k++) c.bstart[k + 1] += c.bstart[k];
    c.bitems.assign(c.bstart[cells], 0u);
    std::vector<uint32_t> fill(c.bstart.begin(), c.bstart.end() - 1);
    for (uint32_t p = 0; p < c.N; p++)
        for (uint32_t i = 0; i < c.cnt[p]; i++) {
            const uint32_t ref = p * (uint32_t)(c.D + 1) + i;
            int g[3];
            grid_idx(c, c.verts[ref].pos, g);
            c.bitems[fill[cell_hash(g[0], g[1], g[2], c.N)]++] = ref;
        }
}

}  // namespace ora

using namespace ora;

extern "C" {

typedef struct ora_vcm_params {
    uint32_t width, height;
    int32_t max_depth;
    uint32_t frame_start, frames;
    uint32_t seed;          // per-frame random words derive from it (see ora_vcm_words)
    float radius_factor;    // VCMConfig::radius_factor (percent of the scene's bounding radius)
    float alpha;            // radius shrink exponent (VCM.cpp:141 uses 2/3)
    int32_t use_vc, use_vm;
    uint32_t corrected;     // 0: the reference estimator, 1: corrected VCM
    float sky_col[3];
} ora_vcm_params;

// Merge statistics of the last ora_render_vcm call (corrected estimator):
// out = [merge queries, queries with > 64 photons within the radius, the
// largest number of photons within the radius of one query].
void ora_vcm_merge_stats(uint64_t* out) {
    out[0] = g_mq_total.load();
    out[1] = g_mq_over64.load();
    out[2] = g_mq_max.load();
}

// Per-frame RNG words: the light pass uses `random_num` (vcm_light.rgen:57-58)
// and the eye pass `frame_num ^ random_num` (vcm_eye.rgen:35-36); random_num
// (host rand(), VCM.cpp:147) is a pcg4d hash of (seed, frame) here.
void ora_vcm_words(uint32_t seed, uint32_t frame_num, uint32_t* light_word, uint32_t* eye_word) {
    const uvec4 h = pcg4d(uvec4{seed, frame_num, 0x9E3779B9u, 0x7F4A7C15u});
    *light_word = h.x;
    *eye_word = frame_num ^ h.x;
}

// VCM.cpp:140-158: merge radius and hash-grid resolution of frame frame_num
void ora_vcm_radius_grid(const lt_scene_desc* s, float radius_factor, float alpha, uint32_t frame_num, float* radius,
                         int32_t* grid_res) {
    const vec3 mn = load3(s->scene_min), mx = load3(s->scene_max);
    const float R = length(mx - mn) * 0.5f;
    float r = R * radius_factor / 100.0f;
    r /= (float)pow((double)frame_num + 1.0, 0.5 * (1.0 - (double)alpha));
    *radius = r;
    const vec3 diam = mx - mn;
    const float max_comp = fmaxf(diam.x, fmaxf(diam.y, diam.z));
    const int base = (int)(max_comp / r);
    const float g[3] = {diam.x * (float)base / max_comp, diam.y * (float)base / max_comp,
                        diam.z * (float)base / max_comp};
    for (int k = 0; k < 3; k++) grid_res[k] = std::max((int)g[k], 1);
}

double ora_render_vcm(const lt_scene_desc* s, void* accel, const lt_camera* cam, const ora_vcm_params* p,
                      float* accum, float* splat, int threads, uint64_t* closest_rays, uint64_t* shadow_rays) {
    const uint32_t W = p->width, H = p->height, N = W * H;
    VcmCtx c{*s, (const Accel*)accel, *cam, W, H, N, p->max_depth};
    c.use_vc = p->use_vc;
    c.use_vm = p->use_vm;
    c.corrected = p->corrected != 0;
    c.sky = load3(p->sky_col);
    c.screen_size = (float)N;
    c.bmin = load3(s->scene_min);
    c.bmax = load3(s->scene_max);
    {
        vec4 ai = mul(cam->inv_projection, vec4{2.0f / (float)W, 2.0f / (float)H, 0.0f, 1.0f});
        ai = {ai.x / ai.w, ai.y / ai.w, ai.z / ai.w, ai.w / ai.w};
        c.cam_area = fabsf(ai.x * ai.y);
    }
    c.verts.assign((size_t)N * (c.D + 1), VV{});
    c.cnt.assign(N, 0u);
    g_mq_total = 0;
    g_mq_over64 = 0;
    g_mq_max = 0;
    if (threads < 1) threads = 1;
    std::atomic<uint64_t> cl{0}, sh{0};
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t chunk = 1024;
    const uint32_t nchunks = (N + chunk - 1) / chunk;
    std::vector<std::vector<Splat>> chunk_splats(nchunks);
    std::vector<float> ecol((size_t)N * 3);
    for (uint32_t fr = 0; fr < p->frames; fr++) {
        const uint32_t frame_num = p->frame_start + fr;
        uint32_t lw, ew;
        ora_vcm_words(p->seed, frame_num, &lw, &ew);
        ora_vcm_radius_grid(s, p->radius_factor, p->alpha, frame_num, &c.radius, c.grid_res);
        const float r2 = c.radius * c.radius;
        c.eta_vcm = PI * r2 * c.screen_size;
        c.eta_vc = c.use_vc == 1 ? 1.0f / c.eta_vcm : 0.0f;
        c.eta_vm = c.use_vm == 1 ? PI * r2 * c.screen_size : 0.0f;
        auto run = [&](auto&& body) {
            std::atomic<uint32_t> next{0};
            auto worker = [&]() {
                Counters k;
                for (;;) {
                    const uint32_t ch = next.fetch_add(1);
                    if (ch >= nchunks) break;
                    body(ch, k);
                }
                cl += k.closest;
                sh += k.shadow;
            };
            std::vector<std::thread> ts;
            for (int i = 1; i < threads; i++) ts.emplace_back(worker);
            worker();
            for (auto& t : ts) t.join();
        };
        // light pass
        run([&](uint32_t ch, Counters& k) {
            chunk_splats[ch].clear();
            for (uint32_t q = ch * chunk; q < std::min(N, (ch + 1) * chunk); q++)
                light_path(c, q, lw, chunk_splats[ch], &k);
        });
        for (uint32_t ch = 0; ch < nchunks; ch++)
            for (const Splat& q : chunk_splats[ch]) {
                float* d = splat + 3 * ((size_t)q.y * W + q.x);
                d[0] += q.c.x;
                d[1] += q.c.y;
                d[2] += q.c.z;
            }
        if (c.use_vm == 1) build_photons(c);
        // eye pass over pixels in row-major order
        run([&](uint32_t ch, Counters& k) {
            for (uint32_t q = ch * chunk; q < std::min(N, (ch + 1) * chunk); q++) {
                const uint32_t x = q % W, y = q / W;
                const vec3 col = eye_path(c, x, y, ew, &k);
                ecol[3 * (size_t)q] = col.x;
                ecol[3 * (size_t)q + 1] = col.y;
                ecol[3 * (size_t)q + 2] = col.z;
            }
        });
        for (uint32_t q = 0; q < N; q++) {
            vec3 col = v3(ecol[3 * (size_t)q], ecol[3 * (size_t)q + 1], ecol[3 * (size_t)q + 2]);
            if (std::isnan(luminance(col))) continue;  // vcm_eye.rgen:84-87: the splat stays
            float* sp = splat + 3 * (size_t)q;
            col += v3(sp[0], sp[1], sp[2]);
            sp[0] = sp[1] = sp[2] = 0.0f;
            float* px = accum + 4 * (size_t)q;
            if (frame_num > 0) {
                const float w = 1.0f / (float)(frame_num + 1);
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

}  // extern "C"

extern "C" {

// Checks the corrected mode's photon buckets against a brute-force radius
// search: photons of one light pass are gathered around `nq` query points
// (the photons' own positions, offset by up to the radius) through the grid
// and by scanning all photons; returns the number of queries whose sets
// differ (0 when the grid is right).
int ora_vcm_grid_selfcheck(const lt_scene_desc* s, void* accel, const lt_camera* cam, const ora_vcm_params* p,
                           int nq) {
    const uint32_t W = p->width, H = p->height, N = W * H;
    VcmCtx c{*s, (const Accel*)accel, *cam, W, H, N, p->max_depth};
    c.use_vc = 1;
    c.use_vm = 1;
    c.corrected = true;
    c.screen_size = (float)N;
    c.bmin = load3(s->scene_min);
    c.bmax = load3(s->scene_max);
    vec4 ai = mul(cam->inv_projection, vec4{2.0f / (float)W, 2.0f / (float)H, 0.0f, 1.0f});
    ai = {ai.x / ai.w, ai.y / ai.w, ai.z / ai.w, ai.w / ai.w};
    c.cam_area = fabsf(ai.x * ai.y);
    c.verts.assign((size_t)N * (c.D + 1), VV{});
    c.cnt.assign(N, 0u);
    g_mq_total = 0;
    g_mq_over64 = 0;
    g_mq_max = 0;
    uint32_t lw, ew;
    ora_vcm_words(p->seed, p->frame_start, &lw, &ew);
    ora_vcm_radius_grid(s, p->radius_factor, p->alpha, p->frame_start, &c.radius, c.grid_res);
    c.eta_vcm = PI * c.radius * c.radius * c.screen_size;
    c.eta_vc = 1.0f / c.eta_vcm;
    c.eta_vm = c.eta_vcm;
    std::vector<Splat> sp;
    for (uint32_t q = 0; q < N; q++) light_path(c, q, lw, sp, nullptr);
    build_photons(c);
    std::vector<uint32_t> all;
    for (uint32_t q = 0; q < N; q++)
        for (uint32_t i = 0; i < c.cnt[q]; i++) all.push_back(q * (uint32_t)(c.D + 1) + i);
    if (all.empty()) return -1;
    uvec4 rs = {1u, 2u, 3u, 0u};
    int bad = 0;
    for (int k = 0; k < nq; k++) {
        const vec3 base = c.verts[all[(size_t)(rand(rs) * (float)all.size()) % all.size()]].pos;
        const vec3 qp = base + c.radius * v3(rand(rs) * 2.0f - 1.0f, rand(rs) * 2.0f - 1.0f, rand(rs) * 2.0f - 1.0f);
        std::vector<uint32_t> grid, brute;
        int lo[3], hi[3];
        grid_idx(c, qp - v3(c.radius), lo);
        grid_idx(c, qp + v3(c.radius), hi);
        for (int x = lo[0]; x <= hi[0]; x++)
            for (int y = lo[1]; y <= hi[1]; y++)
                for (int z = lo[2]; z <= hi[2]; z++) {
                    const uint32_t h = cell_hash(x, y, z, c.N);
                    for (uint32_t j = c.bstart[h]; j < c.bstart[h + 1]; j++) {
                        const VV& v = c.verts[c.bitems[j]];
                        int g[3];
                        grid_idx(c, v.pos, g);
                        if (g[0] != x || g[1] != y || g[2] != z) continue;
                        const vec3 d = qp - v.pos;
                        if (dot(d, d) <= c.radius * c.radius) grid.push_back(c.bitems[j]);
                    }
                }
        for (uint32_t ref : all) {
            const vec3 d = qp - c.verts[ref].pos;
            if (dot(d, d) <= c.radius * c.radius) brute.push_back(ref);
        }
        std::sort(grid.begin(), grid.end());
        std::sort(brute.begin(), brute.end());
        if (grid != brute) bad++;
    }
    return bad;
}

}  // extern "C"
