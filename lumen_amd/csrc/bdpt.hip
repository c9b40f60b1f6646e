// bdpt.hip — the BDPT integrator on gfx950.
//
// A frame runs over chunks of 8x8 pixel tiles with two persistent kernels per
// chunk: the walk kernel traces each pixel's light and camera subpaths
// (bdpt_commons.glsl:13-270) into a structure-of-arrays vertex buffer, then
// the connection kernel evaluates every (s, t) strategy of bdpt.rgen:53-76.
// Splitting the two keeps each kernel's register footprint, and so its
// occupancy, to its own phase.  Light-tracing strategies (t = 1) add their
// splat to the target pixel with float atomics; a last kernel adds each
// pixel's splats to its camera-side estimate and applies the NaN guard and
// running mean (bdpt.rgen:77-90).
//
// MIS (calc_mis_weight, bdpt_commons.glsl:272-473) is evaluated without the
// reference's write-and-restore of the shared vertices: the ratios of the
// vertices a strategy does not modify are folded, after the walks, into one
// prefix sum per vertex (camera: P_k = r_k (c_k + P_{k-1}); light:
// Q_k = r_k (c_k + Q_{k-1})), so each strategy computes the two modified
// ratios at each end and reads one prefix: O(1) per strategy instead of the
// reference's O(s + t) walk over global memory.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bdpt.h"
#include "dev_scene.h"
#include "dev_count.h"
#include "dev_trace.h"

using namespace ltd;

namespace {

constexpr float BDPT_TMIN = 0.001f;  // bdpt_commons.glsl:23
constexpr float BDPT_TMAX = 1e6f;    // bdpt_commons.glsl:24

// vertex fields (one 32-bit plane each)
enum : int {
    F_POS = 0,    // 3
    F_NS = 3,     // 3
    F_UV = 6,     // 2
    F_THR = 8,    // 3
    F_MAT = 11,   // material index (u32)
    F_BITS = 12,  // delta | side << 1 (u32)
    F_AREA = 13,
    F_PFWD = 14,
    F_PREV = 15,
    F_PRE = 16,   // MIS prefix P_k / Q_k
    kVtxFields = 17
};

struct Vtx {
    f3 pos, n_s;
    f2 uv;
    f3 thr;
    uint32_t mat, bits;
    float area, pdf_fwd, pdf_rev, pre;
};

struct VStore {
    float* base;
    uint32_t slots;
    uint32_t nv;  // vertices per subpath (max_depth + 1)
    uint32_t slot;
    LT_DEV float* plane(int sub, int i, int f) const {
        return base + ((size_t)((sub * (int)nv + i) * kVtxFields + f)) * slots + slot;
    }
    LT_DEV void put(int sub, int i, int f, float v) const { *plane(sub, i, f) = v; }
    LT_DEV void putu(int sub, int i, int f, uint32_t v) const { *plane(sub, i, f) = __uint_as_float(v); }
    LT_DEV float get(int sub, int i, int f) const { return *plane(sub, i, f); }
    LT_DEV uint32_t getu(int sub, int i, int f) const { return __float_as_uint(*plane(sub, i, f)); }
    LT_DEV void put3(int sub, int i, int f, f3 v) const {
        put(sub, i, f, v.x);
        put(sub, i, f + 1, v.y);
        put(sub, i, f + 2, v.z);
    }
    LT_DEV f3 get3(int sub, int i, int f) const { return {get(sub, i, f), get(sub, i, f + 1), get(sub, i, f + 2)}; }
    LT_DEV void store(int sub, int i, const Vtx& v) const {
        put3(sub, i, F_POS, v.pos);
        put3(sub, i, F_NS, v.n_s);
        put(sub, i, F_UV, v.uv.x);
        put(sub, i, F_UV + 1, v.uv.y);
        put3(sub, i, F_THR, v.thr);
        putu(sub, i, F_MAT, v.mat);
        putu(sub, i, F_BITS, v.bits);
        put(sub, i, F_AREA, v.area);
        put(sub, i, F_PFWD, v.pdf_fwd);
        put(sub, i, F_PREV, v.pdf_rev);
    }
    LT_DEV Vtx load(int sub, int i) const {
        Vtx v;
        v.pos = get3(sub, i, F_POS);
        v.n_s = get3(sub, i, F_NS);
        v.uv = {get(sub, i, F_UV), get(sub, i, F_UV + 1)};
        v.thr = get3(sub, i, F_THR);
        v.mat = getu(sub, i, F_MAT);
        v.bits = getu(sub, i, F_BITS);
        v.area = get(sub, i, F_AREA);
        v.pdf_fwd = get(sub, i, F_PFWD);
        v.pdf_rev = get(sub, i, F_PREV);
        v.pre = get(sub, i, F_PRE);
        return v;
    }
};

LT_DEV uint32_t vdelta(const Vtx& v) { return v.bits & 1u; }
LT_DEV bool vside(const Vtx& v) { return (v.bits >> 1) & 1u; }
LT_DEV Vtx zero_vtx() {
    Vtx v;
    v.pos = v.n_s = v.thr = mk3(0.0f);
    v.uv = {0.0f, 0.0f};
    v.mat = v.bits = 0u;
    v.area = v.pdf_fwd = v.pdf_rev = v.pre = 0.0f;
    return v;
}
LT_DEV bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
LT_DEV float remap0(float x) { return x != 0.0f ? x : 1.0f; }
LT_DEV float mis_term(float r, bool power) { return power ? r * r : r; }

template <bool COUNT>
LT_DEV bool occluded(const SceneView& sv, f3 o, f3 d, float tmax, const TraceStack& stk, Ctr& c) {
    if (COUNT) c.shadow++;
    HitRec h;
    return trace_bvh<true, COUNT>(sv, o, d, 0.0f, tmax, h, stk, c.snodes, c.stris);
}

LT_DEV f3 eval_f(const SceneView& sv, const Vtx& v, const Mat& m, f3 wo, f3 wi) {
    float p, pr;
    return eval_bsdf(v.n_s, wo, m, vside(v), wi, p, pr, false);
}

// Walk state of one pixel's subpaths (one lane): the light subpath, then
// the camera subpath, one bounce per step.  The pixel is (seed.x, seed.y);
// `st` packs the subpath (bit 0: 0 light, mode 0; 1 camera, mode 1), a
// finite light (bit 1), the vertices written after vertex 0 (bits 8-15) and
// the light subpath's length once it has ended (bits 16-23).  The light
// vertex 0 fields MIS needs live in the pixel record, written at the start.
struct WalkState {
    RNG seed;
    f3 ray_pos, wi, thr, prev_pos, prev_n;
    float pdf_fwd;
    uint32_t slot, st;
    LT_DEV int sub() const { return (int)(st & 1u); }
    LT_DEV bool finite() const { return (st & 2u) != 0u; }
    LT_DEV int b() const { return (int)((st >> 8) & 0xFFu); }
    LT_DEV void set_b(int b) { st = (st & ~0xFF00u) | ((uint32_t)b << 8); }
    LT_DEV int n_light() const { return (int)((st >> 16) & 0xFFu); }
};

// One bounce of bdpt_commons.glsl:13-197 (random_walk) for subpath w.sub();
// false once the subpath has ended (w.b() + 1 vertices).  pdf_pos_inf is the
// light's pdf_pos if the light is infinite, else 0: the first light vertex
// then takes pdf_pos * |cos| as its forward pdf (bdpt_commons.glsl:239-241,
// applied after the walk there; nothing reads it in between).
template <bool COUNT>
LT_DEV bool walk_step(const SceneView& sv, const VStore& vs, int max_depth, WalkState& w, float pdf_pos_inf,
                      const TraceStack& stk, Ctr& c) {
    const int sub = w.sub();
    const bool eye = sub == 1;
    int b = w.b();
    if (COUNT) c.closest++;
    HitRec h;
    if (!trace_bvh<false, COUNT>(sv, w.ray_pos, w.wi, BDPT_TMIN, BDPT_TMAX, h, stk, c.nodes, c.tris)) {
        if (eye) {  // the escaped vertex keeps the zeroed fields (BDPT.cpp:76-77)
            Vtx e = zero_vtx();
            e.thr = w.thr;
            e.pdf_fwd = w.pdf_fwd;
            vs.store(sub, b + 1, e);
            w.set_b(b + 1);
        }
        return false;
    }
    Payload p;
    fill_payload(sv, h.inst, h.prim, h.u, h.v, p);
    f3 wo = w.prev_pos - p.pos;
    const float wo_len = length(wo);
    wo /= wo_len;
    f3 n_s = p.n_s;
    bool side = true;
    f3 n_g = p.n_g;
    if (dot(p.n_g, wo) < 0.0f) n_g = -n_g;
    if (dot(n_g, n_s) < 0.0f) {
        n_s = n_s * -1.0f;
        side = false;
    }
    const Mat mat = load_material(sv, p.material_idx, p.uv);
    const bool mat_specular = (mat.bsdf_props & LT_BSDF_FLAG_SPECULAR) == LT_BSDF_FLAG_SPECULAR;
    const bool mat_transmissive = (mat.bsdf_props & LT_BSDF_FLAG_TRANSMISSION) == LT_BSDF_FLAG_TRANSMISSION;
    Vtx cur;
    cur.pdf_fwd = w.pdf_fwd * fabsf(dot(wo, n_s)) / (wo_len * wo_len);
    if (!eye && b == 0 && !w.finite()) cur.pdf_fwd = pdf_pos_inf * fabsf(dot(w.wi, n_s));
    cur.n_s = n_s;
    cur.area = p.area;
    cur.pos = p.pos;
    cur.uv = p.uv;
    cur.mat = p.material_idx;
    cur.thr = w.thr;
    cur.bits = (mat_specular ? 1u : 0u) | (side ? 2u : 0u);
    cur.pdf_rev = 0.0f;
    cur.pre = 0.0f;
    vs.store(sub, b + 1, cur);
    w.set_b(++b);
    if (b >= max_depth) return false;
    float cos_theta;
    const float r0 = rnd(w.seed), r1 = rnd(w.seed), r2 = rnd(w.seed);
    const f3 f = sample_bsdf(n_s, wo, mat, eye ? 1u : 0u, side, w.wi, w.pdf_fwd, cos_theta, r0, r1, r2);
    const bool same_hem = dot(w.wi, n_s) > 0.0f && dot(wo, n_s) > 0.0f;
    if (is_zero(f) || w.pdf_fwd == 0.0f || (!same_hem && !mat_transmissive)) return false;
    w.thr *= f * fabsf(cos_theta) / w.pdf_fwd;
    float pdf_rev = w.pdf_fwd;
    if (!mat_specular) pdf_rev = bsdf_pdf(mat, n_s, w.wi, wo, side);
    const bool g_term = eye || b - 2 > -1 || w.finite();  // prev = b - 2 after the increment
    if (g_term) pdf_rev *= fabsf(dot(w.prev_n, wo)) / (wo_len * wo_len);
    vs.put(sub, b - 1, F_PREV, pdf_rev);
    w.ray_pos = offset_ray(p.pos, n_g);
    w.prev_pos = p.pos;
    w.prev_n = n_s;
    return true;
}

// The MIS prefix sums of the unmodified vertices (see the file comment).
//   camera: P_0 = 0, P_k = r_k (c_k + P_{k-1}), c_k = !delta_k && !delta_{k-1}
//   light:  Q_{-1} = 0, Q_k = r_k (c_k + Q_{k-1}), c_k = !delta_k && !dprev_k,
//           dprev_0 = the light is a delta light
LT_DEV void mis_prefixes(const VStore& vs, int n_light, int n_cam, bool light_is_delta, bool power) {
    float P = 0.0f;
    uint32_t dprev = vs.getu(1, 0, F_BITS) & 1u;
    vs.put(1, 0, F_PRE, 0.0f);
    for (int k = 1; k < n_cam; k++) {
        const uint32_t d = vs.getu(1, k, F_BITS) & 1u;
        const float r = mis_term(remap0(vs.get(1, k, F_PREV)) / remap0(vs.get(1, k, F_PFWD)), power);
        P = r * (((d == 0u && dprev == 0u) ? 1.0f : 0.0f) + P);
        vs.put(1, k, F_PRE, P);
        dprev = d;
    }
    float Q = 0.0f;
    bool dp = light_is_delta;
    for (int k = 0; k < n_light; k++) {
        const uint32_t d = vs.getu(0, k, F_BITS) & 1u;
        const float r = mis_term(remap0(vs.get(0, k, F_PREV)) / remap0(vs.get(0, k, F_PFWD)), power);
        Q = r * (((d == 0u && !dp) ? 1.0f : 0.0f) + Q);
        vs.put(0, k, F_PRE, Q);
        dp = d == 1u;
    }
}

struct LightV0 {  // light subpath vertex 0 fields kept in registers
    uint32_t flags, idx;
    float pdf_pos;  // the global light_pdf_pos of bdpt_commons.glsl:8
};

// ---- calc_mis_weight (bdpt_commons.glsl:272-473) ----
// A strategy (s, t) replaces four reverse pdfs (camera vertices t-1 and t-2,
// light vertices s-1 and s-2, :313-414) and clears the delta flags of its two
// end vertices; every other ratio comes from the prefixes P_{t-3} and Q_{s-3}.
// The helpers below compute the replaced pdfs with the reference's formulas
// and operand order; C1/C2 are camera vertices t-1/t-2, Lv1/Lv2 light
// vertices s-1/s-2 (for s = 1, Lv1 holds the NEE sample's pos, n_s and
// pdf_fwd; for t = 1, C1 is the camera vertex).

// :326-335, s >= 2
LT_DEV float rev_cam1_from_vertex(const Mat& mL, const Vtx& Lv1, const Vtx& Lv2, const Vtx& C1) {
    f3 dir = C1.pos - Lv1.pos;
    const float dl = length(dir);
    dir /= dl;
    const f3 wo = normalize(Lv2.pos - Lv1.pos);
    float r = bsdf_pdf(mL, Lv1.n_s, wo, dir, vside(Lv1));
    r *= fabsf(dot(dir, C1.n_s)) / (dl * dl);
    return r;
}
// :336-347, s == 1
LT_DEV float rev_cam1_from_light(const SceneView& sv, const LightV0& l0, const Vtx& Lv1, const Vtx& C1) {
    f3 dir = C1.pos - Lv1.pos;
    const float dl = length(dir);
    dir /= dl;
    float r;
    if (!light_finite(l0.flags)) {
        r = l0.pdf_pos;
        r *= fabsf(dot(dir, C1.n_s));
    } else {
        r = light_pdf_dir(sv, l0.idx, l0.flags, Lv1.n_s, dir);
        r *= fabsf(dot(dir, C1.n_s)) / (dl * dl);
    }
    return r;
}
// :357-377, t >= 3 (s == 0: the emitter's cosine pdf)
LT_DEV float rev_cam2(const Mat& mC, const Vtx& C1, const Vtx& C2, const Vtx& Lv1, int s) {
    f3 dir = C2.pos - C1.pos;
    const float dl = length(dir);
    dir /= dl;
    if (s > 0) {
        const f3 wo = normalize(Lv1.pos - C1.pos);
        float r = bsdf_pdf(mC, C1.n_s, wo, dir, vside(C1));
        if (r != 0.0f) r *= fabsf(dot(dir, C2.n_s)) / (dl * dl);
        return r;
    }
    const float cx = dot(C1.n_s, dir), cy = dot(C2.n_s, dir);
    return fabsf(cx * cy) / (PI * dl * dl);
}
// :379-405
LT_DEV float rev_light1_camera(const Vtx& cv0, const Vtx& Lv1, float screen_size) {  // t == 1
    f3 dir = Lv1.pos - cv0.pos;
    const float dl = length(dir);
    dir /= dl;
    const float ct = dot(cv0.n_s, dir);
    float pdf = 1.0f / (cv0.area * screen_size * ct * ct * ct);
    pdf *= fabsf(dot(dir, Lv1.n_s)) / (dl * dl);
    return pdf;
}
LT_DEV float rev_light1(const Mat& mC, const Vtx& C1, const Vtx& C2, const Vtx& Lv1, int s, bool finite) {  // t >= 2
    f3 dir = Lv1.pos - C1.pos;
    const float dl = length(dir);
    dir /= dl;
    const f3 wo = normalize(C2.pos - C1.pos);
    float r = bsdf_pdf(mC, C1.n_s, wo, dir, vside(C1));
    if ((s == 1 && finite) || s > 1) r *= fabsf(dot(dir, Lv1.n_s)) / (dl * dl);
    return r;
}
// :406-422, s >= 2
LT_DEV float rev_light2(const Mat& mL, const Vtx& Lv1, const Vtx& Lv2, const Vtx& C1, int s, bool finite) {
    f3 dir = Lv2.pos - Lv1.pos;
    const f3 wo = normalize(C1.pos - Lv1.pos);
    const float dl = length(dir);
    dir /= dl;
    float r = bsdf_pdf(mL, Lv1.n_s, wo, dir, vside(Lv1));
    if ((s == 2 && finite) || s > 2) r *= fabsf(dot(dir, Lv2.n_s)) / (dl * dl);
    return r;
}

struct MisIn {
    float C1_rev, C1_fwd, C2_rev, C2_fwd, C2pre;
    uint32_t C2_bits;  // delta | side << 1 | (delta of vertex t-3) << 2
    float L1_rev, L1_fwd, L2_rev, L2_fwd, L2pre;
    uint32_t L2_bits;  // delta | side << 1 | (delta of vertex s-3, or the light's) << 2
};

// :428-472 with the ratios of the unmodified vertices taken from the prefixes
LT_DEV float mis_combine(int s, int t, const MisIn& m, uint32_t light_flags, bool power) {
    if (s + t == 2) return 1.0f;
    float sum = 0.0f;
    // camera side: i = t-1 (delta cleared), i = t-2, then P_{t-3}
    if (t >= 2) {
        const float r1 = mis_term(remap0(m.C1_rev) / remap0(m.C1_fwd), power);
        const float c1 = (m.C2_bits & 1u) == 0u ? 1.0f : 0.0f;
        if (t == 2) {
            sum += r1 * c1;
        } else {
            const float r2 = mis_term(remap0(m.C2_rev) / remap0(m.C2_fwd), power);
            const float c2 = (m.C2_bits & 5u) == 0u ? 1.0f : 0.0f;
            sum += r1 * (c1 + r2 * (c2 + m.C2pre));
        }
    }
    // light side: i = s-1 (delta cleared), i = s-2, then Q_{s-3}
    if (s >= 1) {
        const float r1 = mis_term(remap0(m.L1_rev) / remap0(m.L1_fwd), power);
        if (s == 1) {
            sum += r1 * (light_delta(light_flags) ? 0.0f : 1.0f);
        } else {
            const float r2 = mis_term(remap0(m.L2_rev) / remap0(m.L2_fwd), power);
            const float c1 = (m.L2_bits & 1u) == 1u ? 0.0f : 1.0f;
            const float c2 = (m.L2_bits & 5u) == 0u ? 1.0f : 0.0f;
            sum += r1 * (c1 + r2 * (c2 + m.L2pre));
        }
    }
    return 1.0f / (1.0f + sum);
}

// Per-pixel record of the walk kernel for the connection kernel (one u32
// plane per field, indexed by slot like the vertices).
enum : int { H_COUNTS = 0, H_FLAGS = 1, H_IDX = 2, H_PDFPOS = 3, H_SEEDW = 4, kHdrFields = 5 };

// bdpt.rgen:41-52 and bdpt_commons.glsl:246-270: the camera vertex of pixel
// (x, y) through the pixel centre, its ray direction and solid-angle pdf.
LT_DEV Vtx camera_vertex(const BdptLaunch& L, uint32_t x, uint32_t y, f3& dir, float& pdf) {
    const float screen_size = (float)(L.width * L.height);
    const float ux = ((float)x + 0.5f) / (float)L.width, uy = ((float)y + 0.5f) / (float)L.height;
    const float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    const f4 o4 = mul4(L.inv_view, f4{0.0f, 0.0f, 0.0f, 1.0f});
    f4 ai = mul4(L.inv_projection, f4{2.0f / (float)L.width, 2.0f / (float)L.height, 0.0f, 1.0f});
    ai = f4{ai.x / ai.w, ai.y / ai.w, ai.z / ai.w, ai.w / ai.w};
    const float cam_area = fabsf(ai.x * ai.y);
    const f4 tg = mul4(L.inv_projection, f4{dx, dy, 1.0f, 1.0f});
    const f3 tn = normalize(mk3(tg.x, tg.y, tg.z));
    const f4 d4 = mul4(L.inv_view, f4{tn.x, tn.y, tn.z, 0.0f});
    const f4 fw = mul4(L.inv_view, f4{0.0f, 0.0f, 1.0f, 0.0f});
    dir = {d4.x, d4.y, d4.z};
    Vtx cv0 = zero_vtx();
    cv0.pos = {o4.x, o4.y, o4.z};
    cv0.area = cam_area;
    cv0.thr = mk3(1.0f);
    cv0.n_s = {-fw.x, -fw.y, -fw.z};
    cv0.bits = 2u;
    const float ct = dot(dir, cv0.n_s);
    pdf = 1.0f / (cam_area * screen_size * ct * ct * ct);
    return cv0;
}

// Vertex 0 and the first direction of subpath w.sub() of the pixel (light:
// bdpt_commons.glsl:199-244, camera: :246-270).  The light start writes the
// light vertex 0 fields of the pixel record (zero if the light has no
// emission sample: pdf_dir <= 0, the pixel has no light subpath, and false
// is returned).  A delta light's vertex 0 takes forward pdf 0 (:242-243,
// applied after the walk there; nothing reads it in between).
LT_DEV bool walk_begin(const SceneView& sv, const BdptLaunch& L, const VStore& vs, uint32_t* __restrict__ hdr,
                       WalkState& w) {
    w.set_b(0);
    if (w.sub() == 0) {
        f4 rp;
        rp.x = rnd(w.seed);
        rp.y = rnd(w.seed);
        rp.z = rnd(w.seed);
        rp.w = rnd(w.seed);
        const float rd0 = rnd(w.seed), rd1 = rnd(w.seed);
        float cos_theta, pdf_pos, pdf_dir;
        uint32_t flags, li;
        f3 pos, n;
        const f3 Le = sample_light_Le(sv, rp, rd0, rd1, cos_theta, flags, li, pos, w.wi, n, pdf_pos, pdf_dir);
        const bool ok = pdf_dir > 0.0f;
        const size_t S = vs.slots;
        hdr[H_FLAGS * S + vs.slot] = ok ? flags : 0u;
        hdr[H_IDX * S + vs.slot] = ok ? li : 0u;
        hdr[H_PDFPOS * S + vs.slot] = __float_as_uint(ok ? pdf_pos : 0.0f);
        if (!ok) return false;
        Vtx v0 = zero_vtx();
        v0.pos = pos;
        v0.n_s = n;
        v0.pdf_fwd = light_delta(flags) ? 0.0f : pdf_pos;
        v0.bits = 2u;  // delta 0, side 1
        v0.thr = Le;
        vs.store(0, 0, v0);
        w.ray_pos = w.prev_pos = pos;
        w.prev_n = n;
        w.thr = Le * cos_theta / (pdf_dir * pdf_pos);
        w.pdf_fwd = pdf_dir;
        if (light_finite(flags)) w.st |= 2u;
    } else {
        const Vtx cv0 = camera_vertex(L, w.seed.x, w.seed.y, w.wi, w.pdf_fwd);
        vs.store(1, 0, cv0);
        w.ray_pos = w.prev_pos = cv0.pos;
        w.prev_n = cv0.n_s;
        w.thr = mk3(1.0f);
        w.st &= ~2u;
    }
    return true;
}

// After both subpaths: the predecessor's delta flag as bit 2 of each vertex,
// the MIS prefixes, and the rest of the pixel record (subpath lengths, the
// RNG counter) for the connection kernel.
LT_DEV void walk_finish(const BdptLaunch& L, const VStore& vs, uint32_t* __restrict__ hdr, const WalkState& w) {
    const int n_light = w.n_light(), n_cam = w.b() + 1;
    const size_t S = vs.slots;
    const bool delta0 = light_delta(hdr[H_FLAGS * S + vs.slot]);
    for (int k = 1; k < n_cam; k++)
        vs.putu(1, k, F_BITS, vs.getu(1, k, F_BITS) | ((vs.getu(1, k - 1, F_BITS) & 1u) << 2));
    for (int k = 0; k < n_light; k++) {
        const uint32_t dp = k > 0 ? (vs.getu(0, k - 1, F_BITS) & 1u) : (delta0 ? 1u : 0u);
        vs.putu(0, k, F_BITS, vs.getu(0, k, F_BITS) | (dp << 2));
    }
    mis_prefixes(vs, n_light, n_cam, delta0, L.power != 0u);
    hdr[H_COUNTS * S + vs.slot] = (uint32_t)n_light | ((uint32_t)n_cam << 8);
    hdr[H_SEEDW * S + vs.slot] = w.seed.w;
}

// Connection kernel body for pixel (x, y): every (s, t) strategy of
// bdpt.rgen:53-76 in the reference's order; returns the pixel's camera-side
// estimate, t = 1 strategies add their splats to `splat`.  Each strategy is
// evaluated in full (contribution, MIS weight, splat target) before its
// occlusion query, so the loop has a single traversal call site and no vertex
// or material state is live across it.
template <bool COUNT>
LT_DEV f3 bdpt_connect(const SceneView& sv, const BdptLaunch& L, const VStore& vs, const uint32_t* __restrict__ hdr,
                       uint32_t x, uint32_t y, float* __restrict__ splat, const TraceStack& stk, Ctr& c) {
    const size_t S = vs.slots;
    const uint32_t counts = hdr[H_COUNTS * S + vs.slot];
    const int n_light = (int)(counts & 0xFFu), n_cam = (int)(counts >> 8);
    const LightV0 l0{hdr[H_FLAGS * S + vs.slot], hdr[H_IDX * S + vs.slot],
                     __uint_as_float(hdr[H_PDFPOS * S + vs.slot])};
    RNG seed = {x, y, L.frame_dev ? (*L.frame_dev ^ L.seed) : L.frame_word, hdr[H_SEEDW * S + vs.slot]};
    const bool power = L.power != 0u;
    const float screen_size = (float)(L.width * L.height);
    const int D = L.max_depth;
    f3 cam_dir;
    float cam_pdf;
    const Vtx cv0 = camera_vertex(L, x, y, cam_dir, cam_pdf);

    // ---- strategies in the reference's order (bdpt.rgen:53-76) ----
    f3 col = mk3(0.0f);
    const Vtx zero = zero_vtx();
    for (int t = 1; t <= n_cam; t++) {
        for (int s = (t == 1 ? 2 : 0); s <= n_light; s++) {
            if (s + t - 2 > D - 1) break;
            // the strategy's candidate: contribution with its MIS weight, and
            // the occlusion query that decides it
            f3 contrib = mk3(0.0f);
            bool query = false;
            f3 ro, rd;
            float rtmax = 0.0f;
            int splat_px = -1;
            {
                const bool finite = light_finite(l0.flags);
                const Vtx C1 = t == 1 ? cv0 : vs.load(1, t - 1);
                const Vtx C2 = t <= 1 ? zero : (t == 2 ? cv0 : vs.load(1, t - 2));
                Vtx Lv1 = zero, Lv2 = zero;
                MisIn m;
                m.C1_rev = m.C2_rev = m.L1_rev = m.L2_rev = 0.0f;
                m.C1_fwd = C1.pdf_fwd;
                m.C2_fwd = C2.pdf_fwd;
                m.C2_bits = C2.bits;
                m.C2pre = t >= 3 ? vs.get(1, t - 3, F_PRE) : 0.0f;
                m.L2pre = s >= 3 ? vs.get(0, s - 3, F_PRE) : 0.0f;
                if (s >= 2) {
                    Lv1 = vs.load(0, s - 1);
                    Lv2 = vs.load(0, s - 2);
                }
                m.L2_fwd = Lv2.pdf_fwd;
                m.L2_bits = Lv2.bits;
                if (t == 1) {
                    // bdpt_commons.glsl:475-533
                    f3 dir = cv0.pos - Lv1.pos;
                    const float len = length(dir);
                    dir /= len;
                    const float cos_y = dot(dir, Lv1.n_s);
                    const float cos_theta = dot(cv0.n_s, -dir);
                    if (!(cos_theta <= 0.0f)) {
                        const float c3 = cos_theta * cos_theta * cos_theta;
                        const float ratio = fabsf(cos_y) / (cv0.area * c3 * len * len);
                        f3 f;
                        {
                            const Mat mL = load_material(sv, Lv1.mat, Lv1.uv);
                            f = eval_f(sv, Lv1, mL, normalize(Lv2.pos - Lv1.pos), dir);
                            if (!is_zero(f) && ratio > 0.0f) m.L2_rev = rev_light2(mL, Lv1, Lv2, cv0, s, finite);
                        }
                        if (!is_zero(f) && ratio > 0.0f) {
                            const f3 Lc = Lv1.thr * ratio * f / screen_size;
                            const f3 nd = -dir;
                            f4 tg = mul4(L.view, f4{nd.x, nd.y, nd.z, 0.0f});
                            tg = f4{tg.x / tg.z, tg.y / tg.z, tg.z / tg.z, tg.w / tg.z};
                            const f4 pr = mul4(L.projection, tg);
                            float fx, fy;
                            {
#pragma clang fp contract(off)
                                fx = 0.5f * (1.0f - pr.x) * (float)L.width - 0.5f;
                                fy = 0.5f * (1.0f - pr.y) * (float)L.height - 0.5f;
                            }
                            // ivec2() truncates toward zero: a coordinate in (-1, W) lands in [0, W)
                            if (fx > -1.0f && fx < (float)L.width && fy > -1.0f && fy < (float)L.height &&
                                !(dot(nd, cv0.n_s) < 0.0f) && luminance(Lc) != 0.0f) {
                                m.L1_rev = rev_light1_camera(cv0, Lv1, screen_size);
                                m.L1_fwd = Lv1.pdf_fwd;
                                contrib = mis_combine(s, 1, m, l0.flags, power) * Lc;
                                if (luminance(contrib) > 0.0f) {
                                    splat_px = (int)fy * (int)L.width + (int)fx;
                                    query = true;
                                    ro = offset_ray2(Lv1.pos, Lv1.n_s);
                                    rd = dir;
                                    rtmax = len - EPS;
                                }
                            }
                        }
                    }
                } else if (s == 0) {
                    // bdpt_commons.glsl:540-548 (an escaped vertex reads material 0)
                    contrib = ld3(sv.mats[C1.mat].emissive_factor) * C1.thr;
                    if (luminance(contrib) != 0.0f) {
                        m.C1_rev = 1.0f / ((float)sv.total_light_triangle_cnt * C1.area);
                        if (t >= 3) m.C2_rev = rev_cam2(Mat{}, C1, C2, zero, 0);
                        contrib *= mis_combine(0, t, m, l0.flags, power);
                    }
                } else if (s == 1) {
                    // bdpt_commons.glsl:549-595
                    f4 r4;
                    r4.x = rnd(seed);
                    r4.y = rnd(seed);
                    r4.z = rnd(seed);
                    r4.w = rnd(seed);
                    f3 wi, n, pos;
                    float wi_len, pdf_pos_a, cos_y, pw;
                    LightRecord rec;
                    rec.triangle_idx = rec.instance_idx = 0u;
                    const f3 Le = sample_light_Li(sv, r4, C1.pos, pw, wi, wi_len, pdf_pos_a, cos_y, rec, n, pos);
                    const float cos_x = fabsf(dot(wi, C1.n_s));
                    const float ltc = (float)sv.total_light_triangle_cnt;
                    Lv1.pdf_fwd = pdf_pos_a / ltc;
                    Lv1.pos = pos;
                    Lv1.n_s = n;
                    f3 f;
                    {
                        const Mat mC = load_material(sv, C1.mat, C1.uv);
                        f = eval_f(sv, C1, mC, normalize(C2.pos - C1.pos), wi);
                        if (!is_zero(f)) {
                            if (t >= 3) m.C2_rev = rev_cam2(mC, C1, C2, Lv1, 1);
                            m.L1_rev = rev_light1(mC, C1, C2, Lv1, 1, finite);
                        }
                    }
                    if (!is_zero(f)) {
                        const float pdf_light_w = light_pdf_a_to_w(rec.flags, pdf_pos_a, wi_len * wi_len, cos_y) / ltc;
                        contrib = C1.thr * f * cos_x * Le / pdf_light_w;
                        if (luminance(contrib) != 0.0f) {
                            m.C1_rev = rev_cam1_from_light(sv, l0, Lv1, C1);
                            m.L1_fwd = Lv1.pdf_fwd;
                            contrib *= mis_combine(1, t, m, l0.flags, power);
                        }
                        query = true;
                        ro = offset_ray2(C1.pos, C1.n_s);
                        rd = wi;
                        rtmax = wi_len - EPS;
                    }
                } else {
                    // bdpt_commons.glsl:596-632
                    f3 d = Lv1.pos - C1.pos;
                    const float len = length(d);
                    d /= len;
                    const float G = dot(Lv1.n_s, -d) * dot(C1.n_s, d) / (len * len);
                    if (G > 0.0f) {
                        f3 b1 = mk3(0.0f), b2;
                        {
                            const Mat mL = load_material(sv, Lv1.mat, Lv1.uv);
                            b2 = eval_f(sv, Lv1, mL, normalize(Lv2.pos - Lv1.pos), -d);
                            if (!is_zero(b2)) {
                                m.C1_rev = rev_cam1_from_vertex(mL, Lv1, Lv2, C1);
                                m.L2_rev = rev_light2(mL, Lv1, Lv2, C1, s, finite);
                            }
                        }
                        if (!is_zero(b2)) {
                            const Mat mC = load_material(sv, C1.mat, C1.uv);
                            b1 = eval_f(sv, C1, mC, normalize(C2.pos - C1.pos), d);
                            if (!is_zero(b1)) {
                                if (t >= 3) m.C2_rev = rev_cam2(mC, C1, C2, Lv1, s);
                                m.L1_rev = rev_light1(mC, C1, C2, Lv1, s, finite);
                            }
                        }
                        if (!is_zero(b1) && !is_zero(b2)) {
                            contrib = Lv1.thr * G * b1 * b2 * C1.thr;
                            if (luminance(contrib) != 0.0f) {
                                m.L1_fwd = Lv1.pdf_fwd;
                                contrib *= mis_combine(s, t, m, l0.flags, power);
                            }
                            query = true;
                            ro = offset_ray2(C1.pos, C1.n_s);
                            rd = d;
                            rtmax = len - EPS;
                        }
                    }
                }
            }
            if (query && occluded<COUNT>(sv, ro, rd, rtmax, stk, c)) continue;
            if (splat_px >= 0) {
                float* dst = splat + 3 * (size_t)splat_px;
                unsafeAtomicAdd(dst, contrib.x);
                unsafeAtomicAdd(dst + 1, contrib.y);
                unsafeAtomicAdd(dst + 2, contrib.z);
            } else if (t >= 2) {
                col += contrib;
            }
        }
    }
    return col;
}

// Both kernels are persistent: each wave pulls 8x8 pixel tiles of the chunk
// [tile0, tile0 + n_tiles) from `work`; a pixel's slot in the vertex and
// record buffers is its position in the chunk.
// The walk kernel: persistent lanes, one bounce per iteration; a lane takes
// the next pixel of the chunk as soon as its pixel's camera subpath ends
// (refill when at least regen_min lanes of the wave are idle), so subpaths of
// uneven length do not leave lanes idle.  A pixel's slot in the vertex and
// record buffers is its position in the chunk.
#ifndef LT_BDPT_WALK_WAVES
#define LT_BDPT_WALK_WAVES 4
#endif
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_BDPT_WALK_WAVES))) void k_bdpt_walk(
    SceneView sv, BdptLaunch L, float* __restrict__ verts, uint32_t* __restrict__ hdr, uint32_t slots, uint32_t tile0,
    uint32_t n_tiles, uint32_t* __restrict__ work, uint32_t* __restrict__ spill,
    unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tiles_x = (L.width + 7) / 8;
    const uint32_t n_items = n_tiles * 64u;
    const uint32_t frame_word = L.frame_dev ? (*L.frame_dev ^ L.seed) : L.frame_word;
    const int D = L.max_depth;
    const uint32_t nv = (uint32_t)D + 1u;
    Ctr c{0, 0, 0, 0};
    WalkState w;
    bool has = false, exhausted = false;  // exhausted is wave-uniform
    for (;;) {
        if (!exhausted) {
            const unsigned long long idle = __ballot(!has);
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            if (n_idle >= L.regen_min || n_idle == 64u) {
                const int leader = __ffsll((long long)idle) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(work, n_idle);
                base = __shfl(base, leader);
                if (base + n_idle >= n_items) exhausted = true;
                if (!has) {
                    const uint32_t i = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                    if (i < n_items) {
                        const uint32_t g = tile0 + (i >> 6), q = i & 63u;
                        const uint32_t x = (g % tiles_x) * 8 + (q & 7u), y = (g / tiles_x) * 8 + (q >> 3);
                        if (x < L.width && y < L.height) {
                            w.seed = {x, y, frame_word, 0u};
                            w.slot = i;
                            w.st = 0u;
                            const VStore vs{verts, slots, nv, i};
                            if (!walk_begin(sv, L, vs, hdr, w)) {
                                w.st = 1u;  // no light subpath: n_light 0
                                walk_begin(sv, L, vs, hdr, w);
                            }
                            has = true;
                        }
                    }
                }
            }
        }
        if (__ballot(has) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (has) {
            const VStore vs{verts, slots, nv, w.slot};
            float pdf_pos_inf = 0.0f;
            if (w.st == 0u) pdf_pos_inf = __uint_as_float(hdr[H_PDFPOS * (size_t)slots + w.slot]);
            if (D == 0 || !walk_step<COUNT>(sv, vs, D, w, pdf_pos_inf, stk, c)) {
                if (w.sub() == 0) {  // the light subpath ended: the camera subpath starts
                    w.st = 1u | ((uint32_t)(w.b() + 1) << 16);
                    walk_begin(sv, L, vs, hdr, w);
                } else {
                    walk_finish(L, vs, hdr, w);
                    has = false;
                }
            }
        }
    }
    if (COUNT) flush_counts(c, counts);
}

#ifndef LT_BDPT_CONN_WAVES
#define LT_BDPT_CONN_WAVES 1
#endif
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_BDPT_CONN_WAVES))) void k_bdpt_connect(SceneView sv, BdptLaunch L, float* __restrict__ verts,
                                                              const uint32_t* __restrict__ hdr, uint32_t slots,
                                                              uint32_t tile0, uint32_t n_tiles,
                                                              float4* __restrict__ fcol, float* __restrict__ splat,
                                                              uint32_t* __restrict__ work,
                                                              uint32_t* __restrict__ spill,
                                                              unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tiles_x = (L.width + 7) / 8;
    Ctr c{0, 0, 0, 0};
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(work, 1u);
        t = __shfl(t, 0);
        if (t >= n_tiles) break;
        const uint32_t g = tile0 + t;
        const uint32_t x = (g % tiles_x) * 8 + (lane & 7), y = (g / tiles_x) * 8 + (lane >> 3);
        if (x >= L.width || y >= L.height) continue;
        const VStore vs{verts, slots, (uint32_t)L.max_depth + 1u, t * 64u + lane};
        const f3 col = bdpt_connect<COUNT>(sv, L, vs, hdr, x, y, splat, stk, c);
        fcol[(size_t)y * L.width + x] = make_float4(col.x, col.y, col.z, 0.0f);
    }
    if (COUNT) flush_counts(c, counts);
}

// bdpt.rgen:77-90: camera estimate + splats, NaN guard, running mean
__global__ void k_bdpt_resolve(const float4* __restrict__ fcol, float* __restrict__ splat, float4* __restrict__ accum,
                               uint32_t n, uint32_t frame_arg, const uint32_t* __restrict__ frame_dev) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t frame_num = frame_dev ? *frame_dev : frame_arg;
    const float4 fc = fcol[i];
    float* sp = splat + 3 * (size_t)i;
    const f3 col = mk3(fc.x, fc.y, fc.z) + mk3(sp[0], sp[1], sp[2]);
    sp[0] = sp[1] = sp[2] = 0.0f;
    if (isnan(luminance(col))) return;
    float4 a = accum[i];
    if (frame_num > 0) {
        const float w = 1.0f / (float)(frame_num + 1);
        a.x = a.x * (1.0f - w) + col.x * w;
        a.y = a.y * (1.0f - w) + col.y * w;
        a.z = a.z * (1.0f - w) + col.z * w;
    } else {
        a.x = col.x;
        a.y = col.y;
        a.z = col.z;
    }
    a.w = 1.0f;
    accum[i] = a;
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

// Grids of the walk and connection kernels: the workgroups every CU holds at
// once (persistent kernels), unless the tuning fixes a count.
static void bdpt_grids(lt_context* ctx, int count, uint32_t& walk, uint32_t& conn) {
    const uint32_t fixed = ctx->tuning.bdpt_blocks_per_cu;
    const void* kw = count ? reinterpret_cast<const void*>(&k_bdpt_walk<true>)
                           : reinterpret_cast<const void*>(&k_bdpt_walk<false>);
    const void* kc = count ? reinterpret_cast<const void*>(&k_bdpt_connect<true>)
                           : reinterpret_cast<const void*>(&k_bdpt_connect<false>);
    walk = (uint32_t)ctx->num_cus * resident_blocks(kw, fixed);
    conn = (uint32_t)ctx->num_cus * resident_blocks(kc, fixed);
}

size_t bdpt_spill_threads(lt_context* ctx) {
    size_t m = 0;
    for (int count = 0; count < 2; count++) {
        uint32_t w, c;
        bdpt_grids(ctx, count, w, c);
        m = std::max(m, (size_t)std::max(w, c));
    }
    return m * kTraceBlock;
}

lt_status bdpt_alloc(lt_context* ctx, BdptBuffers& b, uint32_t width, uint32_t height, int32_t max_depth) {
    const uint32_t tiles = ((width + 7) / 8) * ((height + 7) / 8);
    const size_t tile_bytes = (size_t)64 * 2 * (max_depth + 1) * kVtxFields * sizeof(float);
    const size_t fit = kBdptChunkBytes / tile_bytes;
    const size_t want = ctx->tuning.bdpt_chunk_tiles ? ctx->tuning.bdpt_chunk_tiles : (fit > 64 ? fit : 64);
    b.chunk_tiles = (uint32_t)(tiles < want ? tiles : want);
    b.slots = b.chunk_tiles * 64u;
    b.max_depth = max_depth;
    b.pixels = (size_t)width * height;
    const size_t vbytes = (size_t)2 * (max_depth + 1) * kVtxFields * b.slots * sizeof(float);
    LT_HIP(ctx, hipMalloc(&b.verts, vbytes));
    LT_HIP(ctx, hipMalloc(&b.hdr, (size_t)kHdrFields * b.slots * sizeof(uint32_t)));
    LT_HIP(ctx, hipMalloc(&b.fcol, b.pixels * sizeof(float4)));
    LT_HIP(ctx, hipMalloc(&b.splat, b.pixels * 3 * sizeof(float)));
    const size_t n = b.pixels * 3;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (uint32_t*)b.splat, n,
                       0u);
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}

void bdpt_free(BdptBuffers& b) {
    if (b.verts) (void)hipFree(b.verts);
    if (b.hdr) (void)hipFree(b.hdr);
    if (b.fcol) (void)hipFree(b.fcol);
    if (b.splat) (void)hipFree(b.splat);
    b = BdptBuffers{};
}

lt_status launch_bdpt(lt_context* ctx, BdptBuffers& b, const BdptLaunch& L0, uint32_t frame_num, uint32_t frames,
                      uint32_t seed, float4* accum, int count, uint64_t* d_counts) {
    if (L0.max_depth != b.max_depth || (size_t)L0.width * L0.height != b.pixels)
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "BDPT launch does not match its buffers");
    const uint32_t tiles = ((L0.width + 7) / 8) * ((L0.height + 7) / 8);
    uint32_t walk_blocks, conn_blocks;
    bdpt_grids(ctx, count, walk_blocks, conn_blocks);
    lt_status st = ensure_spill(ctx, bdpt_spill_threads(ctx));
    if (st != LT_OK) return st;
    SceneView sv = make_scene_view(ctx);
    auto* cnt = (unsigned long long*)d_counts;
    if (count)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)d_counts, (size_t)(2 * kCountSlots),
                           0u);
    const uint32_t n = (uint32_t)b.pixels;
    const dim3 wgrid(walk_blocks), cgrid(conn_blocks), block(kTraceBlock);
    for (uint32_t f = 0; f < frames; f++) {
        BdptLaunch L = L0;
        L.frame_word = (frame_num + f) ^ seed;
        L.seed = seed;
        L.regen_min = ctx->tuning.regen_min ? ctx->tuning.regen_min : 4u;
        for (uint32_t t0 = 0; t0 < tiles; t0 += b.chunk_tiles) {
            const uint32_t nt = tiles - t0 < b.chunk_tiles ? tiles - t0 : b.chunk_tiles;
            if (f == 0) phase_mark(ctx);
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->counters, (size_t)2, 0u);
            if (count)
                hipLaunchKernelGGL((k_bdpt_walk<true>), wgrid, block, 0, ctx->stream, sv, L, b.verts, b.hdr, b.slots,
                                   t0, nt, ctx->counters, ctx->spill, cnt);
            else
                hipLaunchKernelGGL((k_bdpt_walk<false>), wgrid, block, 0, ctx->stream, sv, L, b.verts, b.hdr, b.slots,
                                   t0, nt, ctx->counters, ctx->spill, cnt);
            if (f == 0) phase_mark(ctx);
            if (count)
                hipLaunchKernelGGL((k_bdpt_connect<true>), cgrid, block, 0, ctx->stream, sv, L, b.verts, b.hdr,
                                   b.slots, t0, nt, b.fcol, b.splat, ctx->counters + 1, ctx->spill, cnt);
            else
                hipLaunchKernelGGL((k_bdpt_connect<false>), cgrid, block, 0, ctx->stream, sv, L, b.verts, b.hdr,
                                   b.slots, t0, nt, b.fcol, b.splat, ctx->counters + 1, ctx->spill, cnt);
        }
        if (f == 0) phase_mark(ctx);
        hipLaunchKernelGGL(k_bdpt_resolve, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, b.fcol, b.splat, accum,
                           n, frame_num + f, L0.frame_dev);
    }
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}
