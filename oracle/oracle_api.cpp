// oracle_api.cpp — extern "C" hooks of the oracle for the tests and for
// bench.py's cpu_baseline leg (TEST INFRASTRUCTURE, see oracle.h).
//
// The multithreaded renderer is the CPU baseline: std::thread workers pulling
// 16x16 tiles from an atomic counter, one path sample per pixel per frame,
// running-mean accumulation exactly as path.rgen:106-112.
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "oracle.h"

using namespace ora;

extern "C" {

void ora_pcg4d(const uint32_t* in, uint32_t* out, size_t n) {
    for (size_t i = 0; i < n; i++) {
        uvec4 v = {in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3]};
        v = pcg4d(v);
        out[4 * i] = v.x;
        out[4 * i + 1] = v.y;
        out[4 * i + 2] = v.z;
        out[4 * i + 3] = v.w;
    }
}

float ora_uint_to_float(uint32_t x) { return uint_to_float(x); }

void ora_rand_seq(uint32_t x, uint32_t y, uint32_t frame, uint32_t w0, size_t n, float* out) {
    uvec4 s = {x, y, frame, w0};
    for (size_t i = 0; i < n; i++) out[i] = rand(s);
}

void ora_offset_ray(const float* p, const float* n, float* out, int which) {
    vec3 r = which == 2 ? offset_ray2(load3(p), load3(n)) : offset_ray(load3(p), load3(n));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void ora_concentric_disk(const float* r, float* out) {
    vec2 d = concentric_sample_disk({r[0], r[1]});
    out[0] = d.x;
    out[1] = d.y;
}

void ora_onb(const float* n, float* b1, float* b2) {
    vec3 t, b;
    branchless_onb(load3(n), t, b);
    memcpy(b1, &t, 12);
    memcpy(b2, &b, 12);
}

// --- BSDF hooks: one call per sample (inputs are world-space vectors) ---
void ora_sample_bsdf(const lt_material* m, const float* n_s, const float* wo, uint32_t mode, int side,
                     const float* rands, float* wi_out, float* f_out, float* pdf_out, float* cos_out) {
    Material mm;
    memcpy(&mm, m, sizeof(mm));
    vec3 wi;
    float pdf, c;
    vec3 f = sample_bsdf(load3(n_s), load3(wo), mm, mode, side != 0, wi, pdf, c, load3(rands));
    memcpy(wi_out, &wi, 12);
    memcpy(f_out, &f, 12);
    *pdf_out = pdf;
    *cos_out = c;
}

void ora_eval_bsdf(const lt_material* m, const float* n_s, const float* wo, const float* wi, uint32_t mode, int side,
                   int rev, float* f_out, float* pdf_out, float* pdf_rev_out) {
    Material mm;
    memcpy(&mm, m, sizeof(mm));
    float p, pr;
    vec3 f = eval_bsdf(load3(n_s), load3(wo), mm, mode, side != 0, load3(wi), p, pr, rev != 0);
    memcpy(f_out, &f, 12);
    *pdf_out = p;
    *pdf_rev_out = pr;
}

float ora_bsdf_pdf(const lt_material* m, const float* n_s, const float* wo, const float* wi, int side) {
    Material mm;
    memcpy(&mm, m, sizeof(mm));
    return bsdf_pdf(mm, load3(n_s), load3(wo), load3(wi), side != 0);
}

float ora_fresnel_dielectric(float c, float eta, int ff) { return fresnel_dielectric(c, eta, ff != 0); }
float ora_fresnel_conductor(float c, float eta, float k) { return fresnel_conductor(c, eta, k); }
float ora_D_ggx_iso(float a2, float c) { return D_GGX_isotropic(a2, c); }

// sample_light_Li; out = [Le(3), wi(3), wi_len, pdf_pos_w, pdf_pos_a, cos_from_light, light_idx, tri_idx]
void ora_sample_light_Li(const lt_scene_desc* s, const float* rands4, const float* p, float* out) {
    LightRecord rec{};
    vec3 wi, n, pos;
    float wl, pw, pa, cfl, pdw;
    vec4 r = {rands4[0], rands4[1], rands4[2], rands4[3]};
    vec3 Le = sample_light_Li(*s, r, load3(p), (int)s->num_lights, pw, wi, wl, pa, cfl, rec, n, pos, pdw);
    float o[12] = {Le.x, Le.y, Le.z, wi.x, wi.y, wi.z, wl, pw, pa, cfl, (float)rec.light_idx, (float)rec.triangle_idx};
    memcpy(out, o, sizeof(o));
}

// --- traversal ---
void* ora_accel_build(const lt_scene_desc* s) { return accel_build(*s); }
void ora_accel_free(void* a) { accel_free((Accel*)a); }

// mode 0: host BVH, mode 1: brute force
void ora_trace_closest(void* a, const lt_ray* rays, size_t n, lt_hit* hits, int mode, int threads) {
    auto work = [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) {
            const lt_ray& r = rays[i];
            Hit h = mode ? brute_closest((Accel*)a, load3(r.org), load3(r.dir), r.tmin, r.tmax)
                         : accel_closest((Accel*)a, load3(r.org), load3(r.dir), r.tmin, r.tmax);
            hits[i].t = h.inst == 0xFFFFFFFFu ? 0.0f : h.t;
            hits[i].prim = h.prim;
            hits[i].inst = h.inst;
            hits[i].u = h.inst == 0xFFFFFFFFu ? 0.0f : h.u;
            hits[i].v = h.inst == 0xFFFFFFFFu ? 0.0f : h.v;
        }
    };
    if (threads <= 1) {
        work(0, n);
        return;
    }
    std::vector<std::thread> ts;
    size_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) ts.emplace_back(work, b, e);
    }
    for (auto& t : ts) t.join();
}

void ora_trace_any(void* a, const lt_ray* rays, size_t n, uint8_t* occ, int mode, int threads) {
    auto work = [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) {
            const lt_ray& r = rays[i];
            occ[i] = (mode ? brute_any((Accel*)a, load3(r.org), load3(r.dir), r.tmin, r.tmax)
                           : accel_any((Accel*)a, load3(r.org), load3(r.dir), r.tmin, r.tmax))
                         ? 1
                         : 0;
        }
    };
    if (threads <= 1) {
        work(0, n);
        return;
    }
    std::vector<std::thread> ts;
    size_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) ts.emplace_back(work, b, e);
    }
    for (auto& t : ts) t.join();
}

// payload of a closest hit: out = [n_g(3), n_s(3), pos(3), uv(2), area, dist]
// shade_atmosphere of scene s for the ray (o, d) of length len: rgb into out
void ora_atmosphere(const lt_scene_desc* s, const float* sky, const float* o, const float* d, float len, float* out) {
    const vec3 r = shade_atmosphere(*s, load3(sky), load3(o), load3(d), len);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
// integrate_scattering for an explicit light direction and colour
void ora_scattering(const float* o, const float* d, float len, const float* ldir, const float* Lc, float* out) {
    const vec3 r = integrate_scattering(load3(o), load3(d), len, load3(ldir), load3(Lc));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
// phase functions: which = 0 Rayleigh, 1 Mie (asymmetry g)
float ora_phase(int which, float costh, float g) { return which == 0 ? phase_rayleigh(costh) : phase_mie(costh, g); }

void ora_payload(const lt_scene_desc* s, const lt_hit* h, const lt_ray* r, float* out) {
    Hit hh{h->t, h->u, h->v, h->prim, h->inst};
    Payload p;
    fill_payload(*s, hh, load3(r->org), load3(r->dir), r->tmin, p);
    float o[13] = {p.n_g.x, p.n_g.y, p.n_g.z, p.n_s.x, p.n_s.y, p.n_s.z, p.pos.x, p.pos.y, p.pos.z,
                   p.uv.x, p.uv.y, p.area, p.dist};
    memcpy(out, o, sizeof(o));
}

// --- Path integrator ---
typedef struct ora_path_params {
    uint32_t width, height;
    int32_t max_depth;
    float sky_col[3];
    uint32_t direct_lighting;
    uint32_t frame_start;  // first frame_num
    uint32_t frames;       // frames to accumulate
    uint32_t seed;         // XOR'ed into the frame word (0 = reference)
    uint32_t x0, y0, x1, y1;  // pixel window [x0,x1) x [y0,y1); x1 = 0 means full width/height
    uint32_t nee_pick;        // 0: the reference's light pick weight, 1: the pick's own probability
} ora_path_params;

// accum: RGBA32F, row-major over the FULL image; returns wall seconds
double ora_render_path(const lt_scene_desc* s, void* accel, const lt_camera* cam, const ora_path_params* p,
                       float* accum, int threads, uint64_t* closest_rays, uint64_t* shadow_rays) {
    const uint32_t W = p->width, H = p->height;
    const uint32_t x0 = p->x0, y0 = p->y0, x1 = p->x1 ? p->x1 : W, y1 = p->y1 ? p->y1 : H;
    const uint32_t tw = (x1 - x0 + 15) / 16, th = (y1 - y0 + 15) / 16;
    const uint32_t ntiles = tw * th;
    std::atomic<uint64_t> cl{0}, sh{0};
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t fr = 0; fr < p->frames; fr++) {
        const uint32_t frame_num = p->frame_start + fr;
        PathParams pp;
        pp.width = W;
        pp.height = H;
        pp.max_depth = p->max_depth;
        pp.nee_pick = p->nee_pick;
        pp.sky_col = load3(p->sky_col);
        pp.direct_lighting = p->direct_lighting;
        pp.frame_num = frame_num ^ p->seed;
        std::atomic<uint32_t> next{0};
        auto worker = [&]() {
            Counters c;
            for (;;) {
                uint32_t t = next.fetch_add(1);
                if (t >= ntiles) break;
                uint32_t tx = x0 + (t % tw) * 16, ty = y0 + (t / tw) * 16;
                for (uint32_t y = ty; y < std::min(ty + 16, y1); y++)
                    for (uint32_t x = tx; x < std::min(tx + 16, x1); x++) {
                        bool nan_drop;
                        vec3 col = path_sample(*s, (const Accel*)accel, *cam, pp, x, y, nan_drop, &c);
                        if (nan_drop) continue;  // path.rgen:102-104
                        float* px = accum + 4 * ((size_t)y * W + x);
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
    }
    auto t1 = std::chrono::steady_clock::now();
    if (closest_rays) *closest_rays = cl.load();
    if (shadow_rays) *shadow_rays = sh.load();
    return std::chrono::duration<double>(t1 - t0).count();
}

int ora_hardware_threads(void) { return (int)std::thread::hardware_concurrency(); }

}  // extern "C"
