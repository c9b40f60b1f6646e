// oracle_sppm.cpp — the reference SPPM integrator restated on the CPU
// (TEST INFRASTRUCTURE, see oracle.h).
//
// Follows src/RayTracer/SPPM.cpp:75-189 (frame structure, grid) and
// src/shaders/integrators/sppm/{sppm_eye.rgen, max.comp, min.comp,
// calc_bounds.comp, sppm_light.rgen, gather.comp, composite.comp}, with the
// state of sppm_commons.h.  Reference behaviours kept:
//   - the per-pixel state persists between frames (zeroed at frame 0 only,
//     SPPM.cpp:141): a pixel whose eye path records no diffuse surface keeps
//     gathering at the point, material and throughput of an earlier frame;
//   - eye rays go through pixel centres (sppm_eye.rgen:38-40), light and eye
//     paths of one pixel slot draw from the same pcg4d word
//     (frame_num ^ random_num, sppm_eye.rgen:30-31, sppm_light.rgen:30-31);
//   - eye roulette uses luminance(throughput) without the 0.95 cap
//     (sppm_eye.rgen:120-126), the light pass caps it (sppm_light.rgen:105);
//   - the grid spans min(p - r) .. max(p + r) over every pixel (recorded or
//     not) with cells of the largest radius (max.comp, min.comp,
//     calc_bounds.comp); indices are not clamped (the clamp's result is
//     discarded, sppm_eye.rgen:35);
//   - one photon per hash bucket with an insertion count (sppm_light.rgen:
//     85-93); the "last writer" is made deterministic: photons are inserted
//     in light-path order, then vertex order;
//   - the gather visits each grid cell's bucket once (hash collisions of
//     cells in the range count again), counts M per bucket and weighs by the
//     insertion count (gather.comp:45-83);
//   - tau restarts from zero every frame (composite.comp:52), so a frame's
//     estimate is its own photons' flux with the shrunk radius, averaged
//     over frames by the running mean.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "oracle.h"

namespace ora {

static const float SPPM_TMIN = 0.001f;   // sppm_eye.rgen:24
static const float SPPM_TMAX = 10000.0f;  // sppm_eye.rgen:25

struct SppmPixel {  // SPPMData (sppm_commons.h:27-42)
    vec3 p, wo, tau, col, phi, thr;
    uint32_t material_idx;
    vec2 uv;
    vec3 n_s;
    int M;
    float N;
    float radius;
    int path_len;
    uint32_t side;
};

struct SppmPhoton {  // PhotonHash (sppm_commons.h:44-50)
    vec3 pos, wi, thr;
    int count;
    vec3 nrm;
    uint32_t path_len;
};

struct SppmState {
    uint32_t W, H;
    std::vector<SppmPixel> px;  // row-major (the reference stores x * H + y)
};

static inline uint32_t sppm_hash(int x, int y, int z, uint32_t size) {  // utils.glsl:62-65
    uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
    return h % (10u * size);
}

struct SppmGrid {
    vec3 mn, mx;
    int res[3];
};

// get_grid_idx (sppm_eye.rgen:32-36)
static inline void sppm_cell(const SppmGrid& g, vec3 p, int o[3]) {
    const vec3 q = (p - g.mn) / (g.mx - g.mn);
    o[0] = (int)(q.x * (float)g.res[0]);
    o[1] = (int)(q.y * (float)g.res[1]);
    o[2] = (int)(q.z * (float)g.res[2]);
}

struct SppmCtx {
    const lt_scene_desc& s;
    const Accel* a;
    const lt_camera& cam;
    uint32_t W, H, N;
    int D;
    vec3 sky;
    float base_radius;
};

// sppm_eye.rgen:42-128 for pixel (x, y)
static void eye_pass(const SppmCtx& c, SppmPixel& P, uint32_t x, uint32_t y, uint32_t word, uint32_t frame_num,
                     Counters* cnt) {
    uvec4 seed = {x, y, word, 0u};
    const vec2 d = {((float)x + 0.5f) / (float)c.W * 2.0f - 1.0f, ((float)y + 0.5f) / (float)c.H * 2.0f - 1.0f};
    const vec4 o4 = mul(c.cam.inv_view, vec4{0.0f, 0.0f, 0.0f, 1.0f});
    const vec4 tg = mul(c.cam.inv_projection, vec4{d.x, d.y, 1.0f, 1.0f});
    const vec3 tn = normalize(v3(tg.x, tg.y, tg.z));
    const vec4 d4 = mul(c.cam.inv_view, vec4{tn.x, tn.y, tn.z, 0.0f});
    vec3 origin = {o4.x, o4.y, o4.z}, direction = {d4.x, d4.y, d4.z};
    vec3 thr = v3(1.0f);
    bool specular = false, recorded = false;
    Payload pl{};
    for (int depth = 0;; depth++) {
        if (cnt) cnt->closest++;
        const Hit h = accel_closest(c.a, origin, direction, SPPM_TMIN, SPPM_TMAX);
        if (h.inst == 0xFFFFFFFFu) {
            P.col += thr * c.sky;
            break;
        }
        fill_payload(c.s, h, origin, direction, SPPM_TMIN, pl);
        if (depth >= c.D - 1) break;
        const Material mat = load_material(c.s, pl.material_idx, pl.uv);
        if (depth == 0 || (specular && !recorded)) P.col += thr * mat.emissive_factor;
        const vec3 wo = -direction;
        vec3 n_s = pl.n_s, n_g = pl.n_g;
        bool side = true;
        if (dot(pl.n_g, wo) < 0.0f) n_g = -n_g;
        if (dot(n_g, pl.n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        const vec3 pos = pl.pos;
        origin = offset_ray(pos, n_g);
        specular = (mat.bsdf_props & LT_BSDF_FLAG_SPECULAR) != 0;
        if (!recorded && !specular) {
            const float light_pick_pdf = 1.0f / (float)c.s.total_light_triangle_cnt;
            P.col += thr * nee_uniform_sample_light(c.s, c.a, seed, mat, pos, side, n_s, wo, pl, cnt) /
                     light_pick_pdf;
        }
        if (!recorded && (mat.bsdf_type & LT_BSDF_DIFFUSE) != 0) {
            recorded = true;
            P.wo = wo;
            P.p = pos;
            P.material_idx = pl.material_idx;
            P.uv = pl.uv;
            P.thr = thr;
            P.n_s = n_s;
            P.path_len = depth;
            P.side = side ? 1u : 0u;
        }
        float pdf, cos_theta;
        vec3 r3;
        r3.x = rand(seed);
        r3.y = rand(seed);
        r3.z = rand(seed);
        const vec3 f = sample_bsdf(n_s, wo, mat, 1, side, direction, pdf, cos_theta, r3);
        if (pdf < 1e-20f || dot(f, f) < 1e-20f) break;
        thr *= f * fabsf(cos_theta) / pdf;
        if (depth > 3) {  // RR_MIN_DEPTH
            const float rr_prob = luminance(thr);
            if (rr_prob < rand(seed)) break;
            thr /= rr_prob;
        }
    }
    if (frame_num == 0) P.radius = c.base_radius;
}

// sppm_light.rgen:43-114 for the light path of pixel slot (x, y); photons go
// to `out` in vertex order
struct PhotonRec {
    uint32_t h;
    SppmPhoton ph;
};
static void light_pass(const SppmCtx& c, const SppmGrid& g, uint32_t x, uint32_t y, uint32_t word,
                       std::vector<PhotonRec>& out, Counters* cnt) {
    uvec4 seed = {x, y, word, 0u};
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
    const vec3 Le = sample_light_Le(c.s, rp, rd, (int)c.s.num_lights, (int)c.s.total_light_triangle_cnt, cos_theta,
                                    flags, pos, wi, n, pdf_pos, pdf_dir);
    const float abs_cos = fabsf(dot(n, wi));
    vec3 thr = (abs_cos == 0.0f || pdf_dir <= 0.0f) ? v3(0.0f) : Le * abs_cos / (pdf_dir * pdf_pos);
    Payload pl{};
    for (int depth = 0; depth < c.D; depth++) {
        if (cnt) cnt->closest++;
        const Hit h = accel_closest(c.a, pos, wi, SPPM_TMIN, SPPM_TMAX);
        if (h.inst == 0xFFFFFFFFu) break;
        fill_payload(c.s, h, pos, wi, SPPM_TMIN, pl);
        const vec3 wo = -wi;
        vec3 n_s = pl.n_s, n_g = pl.n_g;
        bool side = true;
        if (dot(pl.n_g, wo) < 0.0f) n_g = -n_g;
        if (dot(n_g, pl.n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        const Material mat = load_material(c.s, pl.material_idx, pl.uv);
        if (depth > 0) {
            int gi[3];
            sppm_cell(g, pl.pos, gi);
            PhotonRec r;
            r.h = sppm_hash(gi[0], gi[1], gi[2], c.N);
            r.ph.pos = pl.pos;
            r.ph.wi = -wi;
            r.ph.thr = thr;
            r.ph.count = 1;
            r.ph.nrm = n_s;
            r.ph.path_len = (uint32_t)depth + 1u;
            out.push_back(r);
        }
        pos = offset_ray(pl.pos, n_g);
        float pdf, cos_t;
        vec3 r3;
        r3.x = rand(seed);
        r3.y = rand(seed);
        r3.z = rand(seed);
        const vec3 f = sample_bsdf(n_s, wo, mat, 0, side, wi, pdf, cos_t, r3);
        thr *= f * fabsf(cos_t) / pdf;
        if ((f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) || pdf == 0.0f) break;
        if (depth > 3) {
            const float rr_prob = fminf(0.95f, luminance(thr));
            if (rr_prob < rand(seed)) break;
            thr /= rr_prob;
        }
    }
}

}  // namespace ora

using namespace ora;

extern "C" {

typedef struct ora_sppm_params {
    uint32_t width, height;
    int32_t max_depth;
    uint32_t frame_start, frames;
    uint32_t seed;       // per-frame random_num derives from it (ora_sppm_word)
    float base_radius;   // SPPMConfig::base_radius (SceneConfig.h:34)
    float sky_col[3];
} ora_sppm_params;

// Per-frame pcg4d word of both passes: frame_num ^ random_num, random_num (a
// host rand(), SPPM.cpp:81) being a pcg4d hash of (seed, frame).
uint32_t ora_sppm_word(uint32_t seed, uint32_t frame_num) {
    const uvec4 h = pcg4d(uvec4{seed, frame_num, 0x5A17F00Du, 0x3C6EF372u});
    return frame_num ^ h.x;
}

void* ora_sppm_create(uint32_t width, uint32_t height) {
    SppmState* st = new SppmState;
    st->W = width;
    st->H = height;
    st->px.assign((size_t)width * height, SppmPixel{});
    return st;
}
void ora_sppm_free(void* st) { delete (SppmState*)st; }

// Renders frames frame_start .. frame_start + frames - 1 into accum (RGBA32F,
// row-major); `state` carries the per-pixel SPPM data between calls (it is
// zeroed at frame 0).  grid (optional, 9 floats): the last frame's grid
// min, max and resolution.  Returns wall seconds.
double ora_render_sppm(const lt_scene_desc* s, void* accel, const lt_camera* cam, const ora_sppm_params* p,
                       void* state, float* accum, float* grid_out, int threads, uint64_t* closest_rays) {
    SppmState& S = *(SppmState*)state;
    const uint32_t W = p->width, H = p->height, N = W * H;
    SppmCtx c{*s, (const Accel*)accel, *cam, W, H, N, p->max_depth, load3(p->sky_col),
              p->base_radius < 1e-7f ? 1e-7f : p->base_radius};  // SPPM.cpp:86-88
    if (threads < 1) threads = 1;
    std::atomic<uint64_t> cl{0};
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t chunk = 1024, nchunks = (N + chunk - 1) / chunk;
    auto run = [&](auto&& body) {
        std::atomic<uint32_t> next{0};
        auto worker = [&]() {
            Counters k;
            for (;;) {
                const uint32_t ch = next.fetch_add(1);
                if (ch >= nchunks) break;
                body(ch, k);
            }
            cl += k.closest + k.shadow;
        };
        std::vector<std::thread> ts;
        for (int i = 1; i < threads; i++) ts.emplace_back(worker);
        worker();
        for (auto& t : ts) t.join();
    };
    std::vector<std::vector<PhotonRec>> chunk_ph(nchunks);
    std::vector<SppmPhoton> table;
    SppmGrid g{};
    for (uint32_t fr = 0; fr < p->frames; fr++) {
        const uint32_t frame_num = p->frame_start + fr;
        const uint32_t word = ora_sppm_word(p->seed, frame_num);
        if (frame_num == 0) std::fill(S.px.begin(), S.px.end(), SppmPixel{});  // SPPM.cpp:141
        run([&](uint32_t ch, Counters& k) {
            for (uint32_t q = ch * chunk; q < std::min(N, (ch + 1) * chunk); q++)
                eye_pass(c, S.px[q], q % W, q / W, word, frame_num, &k);
        });
        // max.comp / min.comp / calc_bounds.comp over every pixel
        vec3 mn = v3(INFINITY), mx = v3(-INFINITY);
        float max_r = -INFINITY;
        for (const SppmPixel& P : S.px) {
            const vec3 lo = P.p - v3(P.radius), hi = P.p + v3(P.radius);
            mn = {fminf(mn.x, lo.x), fminf(mn.y, lo.y), fminf(mn.z, lo.z)};
            mx = {fmaxf(mx.x, hi.x), fmaxf(mx.y, hi.y), fmaxf(mx.z, hi.z)};
            max_r = fmaxf(max_r, P.radius);
        }
        g.mn = mn;
        g.mx = mx;
        const vec3 ext = mx - mn;
        const float max_comp = fmaxf(ext.x, fmaxf(ext.y, ext.z));
        const int base = (int)(max_comp / max_r);
        if (max_comp > 0.0f && max_r > 0.0f) {
            const float fb = (float)base;
            g.res[0] = std::max((int)(fb * ext.x / max_comp), 1);
            g.res[1] = std::max((int)(fb * ext.y / max_comp), 1);
            g.res[2] = std::max((int)(fb * ext.z / max_comp), 1);
        } else {
            g.res[0] = g.res[1] = g.res[2] = 0;
        }
        // light pass, photons inserted in light-path order then vertex order
        run([&](uint32_t ch, Counters& k) {
            chunk_ph[ch].clear();
            for (uint32_t q = ch * chunk; q < std::min(N, (ch + 1) * chunk); q++)
                light_pass(c, g, q % W, q / W, word, chunk_ph[ch], &k);
        });
        table.assign((size_t)10 * N, SppmPhoton{});
        for (uint32_t ch = 0; ch < nchunks; ch++)
            for (const PhotonRec& r : chunk_ph[ch]) {
                const int count = table[r.h].count + 1;
                table[r.h] = r.ph;
                table[r.h].count = count;
            }
        // gather.comp + composite.comp
        const float size = (float)N;
        run([&](uint32_t ch, Counters&) {
            for (uint32_t q = ch * chunk; q < std::min(N, (ch + 1) * chunk); q++) {
                SppmPixel& P = S.px[q];
                float r = P.radius;
                if (r < EPS) r = EPS;
                const float r_sqr = r * r;
                int lo[3], hi[3];
                sppm_cell(g, P.p - v3(r), lo);
                sppm_cell(g, P.p + v3(r), hi);
                for (int x = lo[0]; x <= hi[0]; x++)
                    for (int y = lo[1]; y <= hi[1]; y++)
                        for (int z = lo[2]; z <= hi[2]; z++) {
                            const SppmPhoton& ph = table[sppm_hash(x, y, z, N)];
                            if (ph.count <= 0) continue;
                            const vec3 pp = P.p - ph.pos;
                            if (dot(pp, pp) > r_sqr) continue;
                            if ((uint32_t)c.D < ph.path_len + (uint32_t)P.path_len + 1u) continue;
                            const Material mat = load_material(*s, P.material_idx, P.uv);
                            float pf, pr;
                            const vec3 f = eval_bsdf(P.n_s, P.wo, mat, 1, P.side == 1u, ph.wi, pf, pr, false);
                            P.phi += (float)ph.count * ph.thr * f * P.thr;
                            P.M += 1;
                        }
                vec3 col = P.col;
                if (P.M > 0) {
                    const float gamma = 2.0f / 3.0f;
                    const float n_new = P.N + gamma * (float)P.M;
                    const float r_new = P.radius * sqrtf(n_new / (P.N + (float)P.M));
                    P.tau = (P.tau + P.phi) * r_new * r_new / (P.radius * P.radius);
                    P.N = n_new;
                    P.radius = r_new;
                    P.M = 0;
                    P.phi = v3(0.0f);
                    col += P.tau / (size * PI * r_new * r_new);
                }
                P.col = v3(0.0f);
                P.tau = v3(0.0f);
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
        });
    }
    if (grid_out) {
        const float gv[9] = {g.mn.x, g.mn.y, g.mn.z, g.mx.x, g.mx.y, g.mx.z, (float)g.res[0], (float)g.res[1],
                             (float)g.res[2]};
        memcpy(grid_out, gv, sizeof(gv));
    }
    auto t1 = std::chrono::steady_clock::now();
    if (closest_rays) *closest_rays = cl.load();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
