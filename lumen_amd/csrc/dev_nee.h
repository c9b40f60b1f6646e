// dev_nee.h — next-event estimation of the Path integrator
// (src/shaders/integrators/pt_commons.glsl:3-47), shared by the Path kernels
// and the SPPM eye pass (sppm_eye.rgen:98-104 calls the same
// uniform_sample_light).
#pragma once
#include "context.h"
#include "dev_count.h"
#include "dev_scene.h"
#include "dev_trace.h"
#include "path.h"

namespace ltd {

constexpr float PATH_TMIN = 0.001f;    // path.rgen:19
constexpr float PATH_TMAX = 10000.0f;  // path.rgen:20
constexpr int RR_MIN_DEPTH = 3;        // path.rgen:21

template <bool COUNT>
LT_DEV void trace_closest(const SceneView& sv, f3 o, f3 d, const TraceStack& stk, Payload& pl, Ctr& c) {
    if (COUNT) c.closest++;
    HitRec h;
    if (trace_bvh<false, COUNT>(sv, o, d, PATH_TMIN, PATH_TMAX, h, stk, c.nodes, c.tris))
        fill_payload(sv, h.inst, h.prim, h.u, h.v, pl);
    else
        pl.valid = false;  // ray.rmiss:13 — the other fields keep the last hit
}

// pt_commons.glsl:3-42 in two halves, so that a bounce draws all its random
// numbers and evaluates its BSDF before it traces anything and only this
// record, not the material and the hit, lives across the traces.  The terms
// are pre-scaled by the throughput and the light pick pdf of the bounce
// (path.rgen:76-78).
struct NeeRays {
    f3 org;             // offset_ray2(pos, n_s): both rays start here
    f3 wi, c_light;     // light sample: direction and contribution if unoccluded
    float tmax;         // wi_len - EPS; 0 = no shadow ray
    f3 wi_bsdf, c_bsdf; // area light: BSDF-sampled direction and its contribution before MIS
    f3 pos, stale_n;    // the hit: pos, and its payload normal for a missed MIS ray
    float pdf_light_a, pdf_bsdf;
    uint32_t tri, inst; // the sampled light triangle
    uint32_t flags;     // bit 0: MIS ray; bit 1: a missed MIS ray still matches (stale payload)
};

LT_DEV void nee_sample(const SceneView& sv, RNG& seed, const Mat& mat, const Payload& hp, bool side, f3 n_s, f3 wo,
                       f3 thr, uint32_t nee_pick, NeeRays& r) {
    f3 wi;
    float wi_len, pdf_light_w, pdf_light_a, cos_from_light;
    LightRecord rec;
    rec.triangle_idx = 0;
    rec.instance_idx = 0;
    f4 r4;
    r4.x = rnd(seed);
    r4.y = rnd(seed);
    r4.z = rnd(seed);
    r4.w = rnd(seed);
    const f3 Le = sample_light_Li(sv, r4, hp.pos, pdf_light_w, wi, wi_len, pdf_light_a, cos_from_light, rec);
    const f3 scale = thr / nee_pick_pdf(sv, nee_pick, rec.light_idx);
    r.org = offset_ray2(hp.pos, n_s);
    r.wi = wi;
    r.tmax = 0.0f;
    r.flags = 0u;
    float bsdf_pdf, bsdf_rev;
    float cos_x = dot(n_s, wi);
    f3 f = eval_bsdf(n_s, wo, mat, side, wi, bsdf_pdf, bsdf_rev, false);
    if (pdf_light_w > 0.0f) {
        const float mis = (rec.flags & LT_LIGHT_DELTA) ? 1.0f : 1.0f / (1.0f + bsdf_pdf / pdf_light_w);
        const f3 contrib = mis * f * fabsf(cos_x) * Le / pdf_light_w;
        // the occlusion query only matters for a non-zero contribution
        if (contrib.x != 0.0f || contrib.y != 0.0f || contrib.z != 0.0f) {
            r.c_light = scale * contrib;
            r.tmax = wi_len - EPS;
        }
    }
    if ((rec.flags & 7u) == LT_LIGHT_AREA) {
        const float r0 = rnd(seed), r1 = rnd(seed), r2 = rnd(seed);
        f = sample_bsdf(n_s, wo, mat, 1, side, wi, bsdf_pdf, cos_x, r0, r1, r2);
        if (bsdf_pdf != 0.0f) {
            r.flags = 1u;
            if (hp.triangle_idx == rec.triangle_idx && hp.instance_idx == rec.instance_idx) r.flags |= 2u;
            r.wi_bsdf = wi;
            r.c_bsdf = scale * (f * fabsf(cos_x) * Le / bsdf_pdf);
            r.pos = hp.pos;
            r.stale_n = hp.n_s;
            r.pdf_light_a = pdf_light_a;
            r.pdf_bsdf = bsdf_pdf;
            r.tri = rec.triangle_idx;
            r.inst = rec.instance_idx;
        }
    }
}

// The shadow ray and the MIS ray of nee_sample; the MIS ray's payload is the
// bounce's hit where it misses (ray.rmiss:13 leaves the IDs, pos and normal).
template <bool COUNT>
LT_DEV f3 nee_trace(const SceneView& sv, const NeeRays& r, const TraceStack& stk, Ctr& c) {
    f3 res = mk3(0.0f);
    if (r.tmax != 0.0f) {
        if (COUNT) c.shadow++;
        HitRec h;
        if (!trace_bvh<true, COUNT>(sv, r.org, r.wi, 0.0f, r.tmax, h, stk, c.snodes, c.stris)) res += r.c_light;
    }
    if (r.flags & 1u) {
        Payload pl;
        trace_closest<COUNT>(sv, r.org, r.wi_bsdf, stk, pl, c);
        const bool match = pl.valid ? (pl.triangle_idx == r.tri && pl.instance_idx == r.inst) : (r.flags & 2u) != 0u;
        if (match) {
            const f3 hpos = pl.valid ? pl.pos : r.pos, hn = pl.valid ? pl.n_s : r.stale_n;
            const float wl = length(hpos - r.pos);
            const float g = fabsf(dot(hn, -r.wi_bsdf)) / (wl * wl);
            const float mis = 1.0f / (1.0f + r.pdf_light_a / (g * r.pdf_bsdf));
            res += r.c_bsdf * mis;
        }
    }
    return res;
}

}  // namespace ltd
