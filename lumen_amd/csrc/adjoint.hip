// adjoint.hip — the differentiable Path estimator and its path-replay
// adjoint on gfx950.
//
// No reference counterpart (the reference renders but does not
// differentiate).  The estimator is the Path megakernel's per-sample logic
// (src/shaders/integrators/path/path.rgen:26-104, pt_commons.glsl:3-42)
// without roulette; the albedo and roughness textures aid / rid are read
// from per-variant texel arrays.  Derivatives hold the random numbers fixed
// and sampling detached (directions, lobe choices, pdfs and MIS weights are
// constants): path-replay backpropagation (Vicini, Speierer and Jakob 2021)
// adds, at vertex j of a replayed sample with radiance L and adjoint delta,
//   delta . (df_j / f_j) (L - radiance gathered before j's continuation)
// for the sampled continuation and delta . T_j dN_j for the next-event
// terms, scattered to the four texels of the bilinear footprint with float
// atomics.  The first pass's L comes from the forward's per-sample cache
// (one 12-B record per sample in HBM) or is recomputed.
//
// Texel derivatives: Lambertian albedo, and the principled BSDF with
// metallic, spec_trans, specular_tint, clearcoat, flatness and anisotropy
// zero (GGX reflection + Disney diffuse), alpha = max(0.001, r^2):
//   f = S(alpha) + albedo g(r) / pi,  dS/dr = S (dlnD/dalpha + dlnG/dalpha) 2r
//   dlnD/dalpha = -2/alpha + 4e/(alpha (1 + e)),  e = tan^2(h) c_h / alpha^2
//   dlnG/dalpha = -G (dL(wo) + dL(wi)),  dL = alpha c tan^2 / (2 sqrt(1 + c alpha^2 tan^2))
//   g = (1 - f_i/2)(1 - f_o/2) + rr (f_i + f_o^2 f_i (rr - 1)),  rr = 2 r^2 cos_i^2
// (c = the clamped azimuth factor of the Smith / GGX formulas; a lobe with
// alpha <= 0.0064 is the reference's delta reflection and has no roughness
// derivative).
#include <hip/hip_runtime.h>

#include "adjoint.h"
#include "context.h"
#include "dev_scene.h"
// The adjoint kernels test each hit leaf before taking another node: their
// traversals run beside a long per-lane replay state, and the leaf queue's
// deferred culling cost them more node steps than it saved in idle lanes.
#define LT_NODE_PEND_MAX 0
#include "dev_count.h"
#include "dev_trace.h"

using namespace ltd;

namespace {

constexpr float ADJ_TMIN = 0.001f;    // path.rgen:19
constexpr float ADJ_TMAX = 10000.0f;  // path.rgen:20

// bilinear footprint (repeat addressing, texel centres at (i + 0.5) / size)
struct Taps {
    uint32_t i00, i01, i10, i11;
    float ax, ay;
};

LT_DEV Taps taps_of(int W, int H, f2 uv) {
    float fx = uv.x * (float)W - 0.5f, fy = uv.y * (float)H - 0.5f;
    float x0f = floorf(fx), y0f = floorf(fy);
    Taps t;
    t.ax = fx - x0f;
    t.ay = fy - y0f;
    int x0 = ((int)x0f % W + W) % W, y0 = ((int)y0f % H + H) % H;
    int x1 = (x0 + 1) % W, y1 = (y0 + 1) % H;
    t.i00 = (uint32_t)(y0 * W + x0);
    t.i01 = (uint32_t)(y0 * W + x1);
    t.i10 = (uint32_t)(y1 * W + x0);
    t.i11 = (uint32_t)(y1 * W + x1);
    return t;
}

LT_DEV float fetch(const float* __restrict__ d, int stride, int c, const Taps& t) {
    const float top = d[stride * t.i00 + c] * (1.0f - t.ax) + d[stride * t.i01 + c] * t.ax;
    const float bot = d[stride * t.i10 + c] * (1.0f - t.ax) + d[stride * t.i11 + c] * t.ax;
    return top * (1.0f - t.ay) + bot * t.ay;
}

struct DMat {
    Mat m;
    f3 amult;  // the material's albedo factor
    Taps ta, tr;
    bool ta_on, tr_on;
};

// load_material (bsdf_commons.glsl:16-22) with the variant's texels
LT_DEV DMat load_dmat(const SceneView& sv, const AdjLaunch& L, const float* alb, const float* rgh, uint32_t idx,
                      f2 uv) {
    DMat d;
    uint32_t* dst = (uint32_t*)&d.m;
    const uint32_t* s32 = (const uint32_t*)(sv.mats + idx);
#pragma unroll
    for (int i = 0; i < 26; i++) dst[i] = s32[i];
    d.amult = d.m.albedo;
    d.ta_on = d.tr_on = false;
    if (d.m.texture_id > -1) {
        if (alb && d.m.texture_id == L.aid) {
            d.ta = taps_of((int)L.aw, (int)L.ah, uv);
            d.ta_on = true;
            d.m.albedo = d.amult * mk3(fetch(alb, 3, 0, d.ta), fetch(alb, 3, 1, d.ta), fetch(alb, 3, 2, d.ta));
        } else {
            const float4 t = tex_bilinear(sv, d.m.texture_id, uv);
            d.m.albedo = d.m.albedo * mk3(t.x, t.y, t.z);
        }
    }
    if (sv.rough) {
        const int rt = sv.rough[idx];
        if (rt > -1) {
            if (rgh && rt == L.rid) {
                d.tr = taps_of((int)L.rw, (int)L.rh, uv);
                d.tr_on = true;
                d.m.roughness = fetch(rgh, 1, 0, d.tr);
            } else {
                d.m.roughness = tex_bilinear(sv, rt, uv).x;
            }
        }
    }
    return d;
}

// ---- BSDF derivatives (local frame) ----
LT_DEV void diffuse_g(f3 wo, f3 wi, float r, float& g, float& dg) {
    const f3 h = normalize(wi + wo);
    const float wh = dot(wo, h);
    const float fd90 = 0.5f + 2.0f * wh * wh * r, dfd90 = 2.0f * wh * wh;
    const float Pi = pow5m(wi.z), Po = pow5m(wo.z);
    const float fi = 1.0f + (fd90 - 1.0f) * Pi, fo = 1.0f + (fd90 - 1.0f) * Po;
    const float dfi = dfd90 * Pi, dfo = dfd90 * Po;
    const float rr = 2.0f * (r * r) * wi.z * wi.z, drr = 4.0f * r * wi.z * wi.z;
    const float inner = fi + fo * fi * fo * (rr - 1.0f);
    g = (1.0f - 0.5f * fi) * (1.0f - 0.5f * fo) + rr * inner;
    const float dinner = dfi + (2.0f * fo * dfo * fi + fo * fo * dfi) * (rr - 1.0f) + fo * fo * fi * drr;
    dg = -0.5f * dfi * (1.0f - 0.5f * fo) - 0.5f * dfo * (1.0f - 0.5f * fi) + drr * inner + rr * dinner;
}

LT_DEV bool tan2_c(f3 w, float& t2, float& c) {
    const float c2 = w.z * w.z;
    const float s2 = fmaxf(1.0f - c2, 0.0f);
    t2 = s2 / c2;
    if (isinf(t2)) return false;
    c = s2 == 0.0f ? 1.0f : (clampf(w.x * w.x, 0.0f, 1.0f) + clampf(w.y * w.y, 0.0f, 1.0f)) / s2;
    return true;
}

LT_DEV float spec_dlog_dr(f3 wo, f3 wi, float r) {
    const float alpha = fmaxf(0.001f, r * r);
    if (alpha <= 0.0064f) return 0.0f;
    const f3 h = normalize(wo + wi);
    float t2, c, dlnD = 0.0f;
    if (tan2_c(h, t2, c)) {
        const float e = t2 * c / (alpha * alpha);
        dlnD = -2.0f / alpha + 4.0f * e / (alpha * (1.0f + e));
    }
    float Ls = 0.0f, dLs = 0.0f;
    if (tan2_c(wo, t2, c)) {
        const float q = sqrtf(1.0f + c * alpha * alpha * t2);
        Ls += 0.5f * (q - 1.0f);
        dLs += alpha * c * t2 / (2.0f * q);
    }
    if (tan2_c(wi, t2, c)) {
        const float q = sqrtf(1.0f + c * alpha * alpha * t2);
        Ls += 0.5f * (q - 1.0f);
        dLs += alpha * c * t2 / (2.0f * q);
    }
    const float G = 1.0f / (1.0f + Ls);
    return (dlnD - G * dLs) * (2.0f * r);
}

struct DF {
    f3 f, dfa, dfr;
};

// eval_bsdf plus texel derivatives
LT_DEV DF eval_d(const DMat& d, f3 n_s, f3 wo_w, f3 wi_w, bool side, float& pdf) {
    DF o;
    float pr;
    o.f = eval_bsdf(n_s, wo_w, d.m, side, wi_w, pdf, pr, false);
    o.dfa = o.dfr = mk3(0.0f);
    if (!(d.ta_on || d.tr_on)) return o;
    f3 T, B;
    branchless_onb(n_s, T, B);
    const f3 wo = to_local(wo_w, T, B, n_s), wi = to_local(wi_w, T, B, n_s);
    if (d.m.bsdf_type == LT_BSDF_DIFFUSE) {
        if (fminf(wi.z, wo.z) > 0.0f) o.dfa = mk3(INV_PI);
        return o;
    }
    if (d.m.bsdf_type != LT_BSDF_PRINCIPLED) return o;
    float ps, pd, pc, pt;
    init_probs(d.m, fresnel_dielectric(wo.z, d.m.ior, side), side, ps, pd, pc, pt);
    f3 diff = mk3(0.0f);
    if (fminf(wi.z, wo.z) > 0.0f && pd > 0.0f) {
        float g, dg;
        diffuse_g(wo, wi, d.m.roughness, g, dg);
        diff = d.m.albedo * g * INV_PI;
        o.dfa = mk3(g * INV_PI);
        o.dfr = d.m.albedo * dg * INV_PI;
    }
    if (ps > 0.0f && !(wo.z * wi.z < 0.0f) && wo.z != 0.0f && wi.z != 0.0f)
        o.dfr += (o.f - diff) * spec_dlog_dr(wo, wi, d.m.roughness);
    return o;
}

// sample_bsdf plus derivatives of the sampled lobe's value
LT_DEV DF sample_d(const DMat& d, f3 n_s, f3 wo_w, bool side, f3& wi_w, float& pdf, float& cos_t, float r0,
                   float r1, float r2) {
    DF o;
    o.f = sample_bsdf(n_s, wo_w, d.m, 1, side, wi_w, pdf, cos_t, r0, r1, r2);
    o.dfa = o.dfr = mk3(0.0f);
    if (!(d.ta_on || d.tr_on)) return o;
    f3 T, B;
    branchless_onb(n_s, T, B);
    const f3 wo = to_local(wo_w, T, B, n_s), wi = to_local(wi_w, T, B, n_s);
    if (d.m.bsdf_type == LT_BSDF_DIFFUSE) {
        if (fminf(wi.z, wo.z) > 0.0f) o.dfa = mk3(INV_PI);
        return o;
    }
    if (d.m.bsdf_type != LT_BSDF_PRINCIPLED || wo.z <= 0.0f) return o;
    float ps, pd, pc, pt;
    init_probs(d.m, fresnel_dielectric(wo.z, d.m.ior, side), side, ps, pd, pc, pt);
    if (r2 < ps) {
        o.dfr = o.f * spec_dlog_dr(wo, wi, d.m.roughness);
    } else if (r2 > ps + pc && r2 <= ps + pc + pd) {
        if (fminf(wi.z, wo.z) > 0.0f) {
            float g, dg;
            diffuse_g(wo, wi, d.m.roughness, g, dg);
            o.dfa = mk3(g * INV_PI);
            o.dfr = d.m.albedo * dg * INV_PI;
        }
    }
    return o;
}

LT_DEV bool nz(f3 a) { return a.x != 0.0f || a.y != 0.0f || a.z != 0.0f; }

// one vertex's texel gradients: qa per albedo channel, qr for roughness
LT_DEV void scatter(const DMat& d, float* __restrict__ ga, float* __restrict__ gr, f3 qa, float qr) {
    if (d.ta_on && nz(qa)) {
        const f3 q = qa * d.amult;
        const float w[4] = {(1.0f - d.ta.ax) * (1.0f - d.ta.ay), d.ta.ax * (1.0f - d.ta.ay),
                            (1.0f - d.ta.ax) * d.ta.ay, d.ta.ax * d.ta.ay};
        const uint32_t idx[4] = {d.ta.i00, d.ta.i01, d.ta.i10, d.ta.i11};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            atomicAdd(ga + 3 * idx[k] + 0, w[k] * q.x);
            atomicAdd(ga + 3 * idx[k] + 1, w[k] * q.y);
            atomicAdd(ga + 3 * idx[k] + 2, w[k] * q.z);
        }
    }
    if (d.tr_on && qr != 0.0f) {
        const float w[4] = {(1.0f - d.tr.ax) * (1.0f - d.tr.ay), d.tr.ax * (1.0f - d.tr.ay),
                            (1.0f - d.tr.ax) * d.tr.ay, d.tr.ax * d.tr.ay};
        const uint32_t idx[4] = {d.tr.i00, d.tr.i01, d.tr.i10, d.tr.i11};
#pragma unroll
        for (int k = 0; k < 4; k++) atomicAdd(gr + idx[k], w[k] * qr);
    }
}

LT_DEV f3 safe_inv(f3 f) {
    return {f.x != 0.0f ? 1.0f / f.x : 0.0f, f.y != 0.0f ? 1.0f / f.y : 0.0f, f.z != 0.0f ? 1.0f / f.z : 0.0f};
}

// A hit as the sample cache records it: instance, primitive, barycentrics
// (instance ~0: a miss), enough to rebuild the payload without a trace.  A
// record naming no point of a triangle of the scene (a miss, or a cache that
// no forward of this scene wrote) decodes as a miss, so any cache contents
// stay in bounds and give finite hit points.
LT_DEV void payload_from_record(const SceneView& sv, uint4 r, Payload& pl) {
    const float u = __uint_as_float(r.z), v = __uint_as_float(r.w);
    if (r.x >= sv.num_prims || r.y >= sv.tri_start[r.x + 1] - sv.tri_start[r.x] ||
        !(u >= -1e-3f && v >= -1e-3f && u + v <= 1.001f))
        pl.valid = false;  // ray.rmiss:13 — the other fields keep the last hit
    else
        fill_payload(sv, r.x, r.y, u, v, pl);
}

template <bool COUNT>
LT_DEV void closest(const SceneView& sv, f3 o, f3 d, const TraceStack& stk, Payload& pl, Ctr& c,
                    uint4* rec = nullptr) {
    HitRec h;
    if (COUNT) c.closest++;
    const bool hit = trace_bvh<false, COUNT>(sv, o, d, ADJ_TMIN, ADJ_TMAX, h, stk, c.nodes, c.tris);
    const uint4 r = hit ? make_uint4(h.inst, h.prim, __float_as_uint(h.u), __float_as_uint(h.v))
                        : make_uint4(0xFFFFFFFFu, 0u, 0u, 0u);
    if (rec) *rec = r;
    if (hit)
        fill_payload(sv, h.inst, h.prim, h.u, h.v, pl);
    else
        pl.valid = false;  // ray.rmiss:13 — the other fields keep the last hit
}

// The path records of one sample in the sample cache: hit records
// [k][bounce][2][item] (the path's closest hit, the NEE MIS ray's hit), a flag
// word [k][item] with 3 bits per bounce (NEE occlusion traced, occluded, MIS
// ray traced).  The forward writes them; the replay reads them instead of
// tracing, and traces what the forward did not (a light whose contribution
// was zero but whose derivative is not).
struct AdjRec {
    uint4* hits;  // nullptr: no records
    uint32_t* flags;
    size_t n;     // items
    uint32_t depth, k, i;
    LT_DEV uint4* hit(int bounce, int plane) const {
        return hits + (((size_t)k * depth + (size_t)bounce) * 2u + (size_t)plane) * n + i;
    }
};

template <bool COUNT>
LT_DEV bool occluded(const SceneView& sv, f3 o, f3 d, float tmax, const TraceStack& stk, Ctr& c) {
    HitRec h;
    if (COUNT) c.shadow++;
    return trace_bvh<true, COUNT>(sv, o, d, 0.0f, tmax, h, stk, c.snodes, c.stris);
}

// One sample of pixel (x, y).  REPLAY: the same path again with its radiance
// Lt and adjoint radiance delta, scattering texel gradients into ga / gr.
template <bool REPLAY, bool COUNT>
LT_DEV f3 adj_sample(const SceneView& sv, const AdjLaunch& L, const float* alb, const float* rgh, uint32_t x,
                     uint32_t y, uint32_t word, f3 Lt, f3 delta, float* ga, float* gr, const TraceStack& stk,
                     Ctr& cnt, const AdjRec& R) {
    const bool rec_on = R.hits != nullptr;
    uint32_t fl_word = (REPLAY && rec_on) ? R.flags[(size_t)R.k * R.n + R.i] : 0u;
    RNG seed = {x, y, word, 0u};
    const float r0 = rnd(seed);
    const float r1 = rnd(seed);
    const float ux = ((float)x + 0.5f + (r0 - 0.5f)) / (float)L.width;
    const float uy = ((float)y + 0.5f + (r1 - 0.5f)) / (float)L.height;
    const f4 o4 = mul4(L.inv_view, f4{0.0f, 0.0f, 0.0f, 1.0f});
    const f4 tg = mul4(L.inv_projection, f4{ux * 2.0f - 1.0f, uy * 2.0f - 1.0f, 1.0f, 1.0f});
    const f3 tn = normalize(mk3(tg.x, tg.y, tg.z));
    const f4 d4 = mul4(L.inv_view, f4{tn.x, tn.y, tn.z, 0.0f});
    f3 origin = {o4.x, o4.y, o4.z};
    f3 direction = {d4.x, d4.y, d4.z};
    f3 col = mk3(0.0f), thr = mk3(1.0f);
    bool last_specular = false;
    Payload pl;
    pl.valid = false;
    pl.triangle_idx = pl.instance_idx = 0xFFFFFFFFu;
    pl.pos = pl.n_s = pl.n_g = mk3(0.0f);
    const float pick = 1.0f / (float)sv.total_light_triangle_cnt;
    for (int depth = 0;; depth++) {
        if (REPLAY && rec_on)
            payload_from_record(sv, *R.hit(depth, 0), pl);
        else
            closest<COUNT>(sv, origin, direction, stk, pl, cnt, rec_on ? R.hit(depth, 0) : nullptr);
        if (!pl.valid) {
            col += thr * ld3(L.sky_col);
            break;
        }
        const Payload hp = pl;
        const DMat dm = load_dmat(sv, L, alb, rgh, hp.material_idx, hp.uv);
        if (depth == 0 || last_specular) col += thr * dm.m.emissive;
        if (depth >= L.max_depth - 1) break;
        const f3 wo = -direction;
        f3 n_s = hp.n_s, n_g = hp.n_g;
        bool side = true;
        if (dot(hp.n_g, wo) < 0.0f) n_g = -n_g;
        if (dot(n_g, hp.n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        origin = offset_ray(hp.pos, n_g);
        last_specular = is_specular(dm.m);
        f3 qa = mk3(0.0f);
        float qr = 0.0f;
        if (!last_specular) {  // pt_commons.glsl:3-42 / pick pdf (path.rgen:76)
            f4 r4;
            r4.x = rnd(seed);
            r4.y = rnd(seed);
            r4.z = rnd(seed);
            r4.w = rnd(seed);
            f3 wi;
            float wi_len, pw, pa, cfl;
            LightRecord rec;
            rec.triangle_idx = 0;
            rec.instance_idx = 0;
            const f3 Le = sample_light_Li(sv, r4, hp.pos, pw, wi, wi_len, pa, cfl, rec);
            const f3 pp = offset_ray2(hp.pos, n_s);
            const float cos_x = dot(n_s, wi);
            f3 res = mk3(0.0f);
            float bpdf;
            if (REPLAY) {
                const DF fl = eval_d(dm, n_s, wo, wi, side, bpdf);
                if (pw > 0.0f) {
                    const float mis = (rec.flags & LT_LIGHT_DELTA) ? 1.0f : 1.0f / (1.0f + bpdf / pw);
                    const f3 ka = mis * fabsf(cos_x) * Le / pw;
                    const f3 contrib = mis * fl.f * fabsf(cos_x) * Le / pw;
                    bool occ = false;
                    if (nz(contrib) || nz(fl.dfa) || nz(fl.dfr)) {
                        const uint32_t b = fl_word >> (3 * depth);
                        occ = (rec_on && (b & 1u)) ? (b & 2u) != 0u : occluded<COUNT>(sv, pp, wi, wi_len - EPS, stk, cnt);
                    }
                    if ((nz(contrib) || nz(fl.dfa) || nz(fl.dfr)) && !occ) {
                        res += contrib;
                        const f3 w = delta * thr / pick * ka;
                        qa += w * fl.dfa;
                        qr += dot(w, fl.dfr);
                    }
                }
            } else {
                float brev;
                const f3 f = eval_bsdf(n_s, wo, dm.m, side, wi, bpdf, brev, false);
                if (pw > 0.0f) {
                    const float mis = (rec.flags & LT_LIGHT_DELTA) ? 1.0f : 1.0f / (1.0f + bpdf / pw);
                    const f3 contrib = mis * f * fabsf(cos_x) * Le / pw;
                    if (nz(contrib)) {
                        const bool occ = occluded<COUNT>(sv, pp, wi, wi_len - EPS, stk, cnt);
                        fl_word |= (occ ? 3u : 1u) << (3 * depth);
                        if (!occ) res += contrib;
                    }
                }
            }
            if ((rec.flags & 7u) == LT_LIGHT_AREA) {
                const float s0 = rnd(seed), s1 = rnd(seed), s2 = rnd(seed);
                float bp, cx;
                f3 wb;
                const DF fb = REPLAY ? sample_d(dm, n_s, wo, side, wb, bp, cx, s0, s1, s2)
                                     : DF{sample_bsdf(n_s, wo, dm.m, 1, side, wb, bp, cx, s0, s1, s2), mk3(0.0f),
                                          mk3(0.0f)};
                if (bp != 0.0f) {
                    if (REPLAY && rec_on && ((fl_word >> (3 * depth)) & 4u)) {
                        payload_from_record(sv, *R.hit(depth, 1), pl);
                    } else {
                        closest<COUNT>(sv, pp, wb, stk, pl, cnt, (!REPLAY && rec_on) ? R.hit(depth, 1) : nullptr);
                        if (!REPLAY) fl_word |= 4u << (3 * depth);
                    }
                    if (pl.triangle_idx == rec.triangle_idx && pl.instance_idx == rec.instance_idx) {
                        const float wl = length(pl.pos - hp.pos);
                        const float gg = fabsf(dot(pl.n_s, -wb)) / (wl * wl);
                        const float mis = 1.0f / (1.0f + pa / (gg * bp));
                        res += fb.f * mis * fabsf(cx) * Le / bp;
                        if (REPLAY) {
                            const f3 w = delta * thr / pick * (mis * fabsf(cx) * Le / bp);
                            qa += w * fb.dfa;
                            qr += dot(w, fb.dfr);
                        }
                    }
                }
            }
            col += thr * res / pick;
        }
        const float s0 = rnd(seed), s1 = rnd(seed), s2 = rnd(seed);
        float pdf, cos_t;
        if (REPLAY) {
            const DF fc = sample_d(dm, n_s, wo, side, direction, pdf, cos_t, s0, s1, s2);
            if (pdf != 0.0f) {
                const f3 w = delta * (Lt - col) * safe_inv(fc.f);
                qa += w * fc.dfa;
                qr += dot(w, fc.dfr);
            }
            scatter(dm, ga, gr, qa, qr);
            if (pdf == 0.0f) break;
            thr *= fc.f * fabsf(cos_t) / pdf;
        } else {
            const f3 f = sample_bsdf(n_s, wo, dm.m, 1, side, direction, pdf, cos_t, s0, s1, s2);
            if (pdf == 0.0f) break;
            thr *= f * fabsf(cos_t) / pdf;
        }
    }
    if (!REPLAY && rec_on) R.flags[(size_t)R.k * R.n + R.i] = fl_word;
    return col;
}

// Three waves per SIMD (168 VGPRs): about as fast as four on the
// differentiable Cornell box and ~20 % faster than the unconstrained two.
// Four (128 VGPRs, 304-444 B of scratch per lane) gave wrong gradients in the
// replay test once the BVH8 internal children moved to the low slots, with
// no cause found in this source, so it is not used.
#ifndef LT_ADJ_WAVES
#define LT_ADJ_WAVES 3
#endif
// COUNT: the counting variant (rays traced, BVH8 nodes and triangles visited,
// summed into counts)
template <bool BACKWARD, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_ADJ_WAVES))) void k_adjoint(SceneView sv, AdjLaunch L, AdjBuffers B,
                                                         uint32_t* __restrict__ work, uint32_t* __restrict__ spill,
                                                         unsigned long long* __restrict__ counts) {
    Ctr cnt{0u, 0u, 0u, 0u, 0u, 0u};
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
    const uint32_t lane = threadIdx.x & 63u;
    const size_t na = (size_t)L.aw * L.ah, nr = (size_t)L.rw * L.rh, npx = (size_t)L.width * L.height;
    const float inv_spp = 1.0f / (float)L.spp;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(work, 1u);
        t = __shfl(t, 0);
        if ((size_t)t * 64 >= L.n_items) break;  // wave-uniform: the whole wave leaves together
        const uint32_t i = t * 64 + lane;  // n_items is a multiple of 64
        const uint32_t v = i % L.variants, p = i / L.variants;
        const uint32_t tile = p >> 6, w = p & 63u;
        const uint32_t x = (tile % L.tiles_x) * 8 + (w & 7u), y = (tile / L.tiles_x) * 8 + (w >> 3);
        if (x >= L.width || y >= L.height) continue;
        const float* alb = L.aid >= 0 ? B.albedo + 3 * na * v : nullptr;
        const float* rgh = L.rid >= 0 ? B.rough + nr * v : nullptr;
        const size_t pix = npx * v + (size_t)y * L.width + x;
        // the sample cache's path records of sample k of this item (none
        // without a cache or beyond kAdjRecMaxDepth bounces)
        auto rec_of = [&](uint32_t k) {
            if (!B.cache || L.rec_depth == 0u) return AdjRec{nullptr, nullptr, 0, 0, 0, 0};
            uint4* hits = reinterpret_cast<uint4*>(B.cache + (size_t)L.spp * 3u * L.n_items);
            uint32_t* flags = reinterpret_cast<uint32_t*>(hits + (size_t)L.spp * L.rec_depth * 2u * L.n_items);
            return AdjRec{hits, flags, (size_t)L.n_items, L.rec_depth, k, i};
        };
        if (!BACKWARD) {
            f3 sum = mk3(0.0f);
            for (uint32_t k = 0; k < L.spp; k++) {
                const f3 c = adj_sample<false, COUNT>(sv, L, alb, rgh, x, y, k ^ L.seed, mk3(0.0f), mk3(0.0f),
                                                      nullptr, nullptr, stk, cnt, rec_of(k));
                if (B.cache) {
                    float* cp = B.cache + (size_t)k * 3 * L.n_items + i;
                    cp[0] = c.x;
                    cp[L.n_items] = c.y;
                    cp[2 * (size_t)L.n_items] = c.z;
                }
                if (!isnan(luminance(c))) sum += c;  // path.rgen:102-104
            }
            B.images[3 * pix + 0] = sum.x * inv_spp;
            B.images[3 * pix + 1] = sum.y * inv_spp;
            B.images[3 * pix + 2] = sum.z * inv_spp;
        } else {
            const f3 delta = ld3(B.d_images + 3 * pix) * inv_spp;
            if (!nz(delta)) continue;
            float* ga = L.aid >= 0 ? B.g_albedo + 3 * na * v : nullptr;
            float* gr = L.rid >= 0 ? B.g_rough + nr * v : nullptr;
            for (uint32_t k = 0; k < L.spp; k++) {
                f3 c;
                if (B.cache) {
                    const float* cp = B.cache + (size_t)k * 3 * L.n_items + i;
                    c = mk3(cp[0], cp[L.n_items], cp[2 * (size_t)L.n_items]);
                } else {
                    c = adj_sample<false, COUNT>(sv, L, alb, rgh, x, y, k ^ L.seed, mk3(0.0f), mk3(0.0f), nullptr,
                                                 nullptr, stk, cnt, AdjRec{nullptr, nullptr, 0, 0, 0, 0});
                }
                if (isnan(luminance(c))) continue;
                adj_sample<true, COUNT>(sv, L, alb, rgh, x, y, k ^ L.seed, c, delta, ga, gr, stk, cnt, rec_of(k));
            }
        }
    }
    if (COUNT) flush_counts(cnt, counts);
}

__global__ void k_fill_f32(float* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

__global__ void k_zero_word(uint32_t* p) {
    if (threadIdx.x == 0) p[0] = 0u;
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

}  // namespace

lt_status launch_adjoint(lt_context* ctx, const AdjLaunch& L, const AdjBuffers& B, bool backward, uint32_t* work,
                         uint32_t* spill, uint32_t blocks, unsigned long long* counts, hipStream_t stream) {
    SceneView sv = make_scene_view(ctx);
    hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(64), 0, stream, work);
    if (counts) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, counts, (uint32_t)kCountSlots);
    if (backward) {
        const size_t na = (size_t)L.variants * L.aw * L.ah * 3, nr = (size_t)L.variants * L.rw * L.rh;
        if (L.aid >= 0 && na)
            hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)std::min<size_t>((na + 255) / 256, 4096)), dim3(256), 0,
                               stream, B.g_albedo, na);
        if (L.rid >= 0 && nr)
            hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)std::min<size_t>((nr + 255) / 256, 4096)), dim3(256), 0,
                               stream, B.g_rough, nr);
        if (counts)
            hipLaunchKernelGGL((k_adjoint<true, true>), dim3(blocks), dim3(kTraceBlock), 0, stream, sv, L, B, work, spill,
                               counts);
        else
            hipLaunchKernelGGL((k_adjoint<true, false>), dim3(blocks), dim3(kTraceBlock), 0, stream, sv, L, B, work,
                               spill, counts);
    } else {
        if (counts)
            hipLaunchKernelGGL((k_adjoint<false, true>), dim3(blocks), dim3(kTraceBlock), 0, stream, sv, L, B, work,
                               spill, counts);
        else
            hipLaunchKernelGGL((k_adjoint<false, false>), dim3(blocks), dim3(kTraceBlock), 0, stream, sv, L, B, work,
                               spill, counts);
    }
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}
