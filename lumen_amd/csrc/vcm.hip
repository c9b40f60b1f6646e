// vcm.hip — the VCM integrator on gfx950.
//
// A frame is a light pass, the photon grid, an eye pass and, for the
// corrected estimator with merging, a merge pass and a resolve:
//   k_vcm_light   one light subpath per pixel slot (vcm_light.rgen:123-166,
//                 vcm_commons.glsl:77-130, 302-474): stores its non-specular
//                 vertices as 96-B records, splats camera connections with
//                 float atomics and counts its photons into hash buckets;
//   grid          reference estimator: each bucket keeps the largest vertex
//                 reference inserted (atomicMax, a deterministic "last
//                 writer") and the insertion count; corrected estimator: an
//                 exclusive scan of the bucket counts and a scatter of every
//                 photon (k_vcm_scatter: 16-B positions, 48-B rest);
//   k_vcm_eye     one camera path per pixel (vcm_eye.rgen:42-98,
//                 vcm_commons.glsl:477-631): emission, NEE, connections to the
//                 vertices of one random light path; the reference estimator
//                 merges in place and folds in the splats and the running
//                 mean; the corrected one appends a merge query per
//                 non-specular vertex and leaves the colour in fcol;
//   k_vcm_merge   (corrected) one wave per merge query over the photons of
//                 all its cells, sums added to fcol;
//   k_vcm_resolve (corrected) fcol plus the splats into the running mean.
// The light, eye and merge kernels are persistent.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "dev_scene.h"
#include "dev_count.h"
#include "dev_trace.h"
#include "vcm.h"

using namespace ltd;

namespace {

constexpr float VCM_TMIN = 0.001f;    // vcm_light.rgen:54
constexpr float VCM_TMAX = 10000.0f;  // vcm_light.rgen:55

struct VState {  // VCMState (utils.glsl:28-39)
    f3 wi, n_s, pos, thr;
    float d_vcm, d_vc, d_vm;
};

struct LVtx {  // a stored light vertex in registers
    f3 pos, n_s, wi, wo, thr;
    f2 uv;
    float d_vcm, d_vc, d_vm;
    uint32_t mat, path_len, side;
};

LT_DEV LVtx load_lv(const VcmVertex* __restrict__ v) {
    const float4 a = v->a, b = v->b, c = v->c, d = v->d, e = v->e;
    const uint4 f = v->f;
    LVtx r;
    r.pos = {a.x, a.y, a.z};
    r.d_vcm = a.w;
    r.n_s = {b.x, b.y, b.z};
    r.d_vc = b.w;
    r.wi = {c.x, c.y, c.z};
    r.d_vm = c.w;
    r.wo = {d.x, d.y, d.z};
    r.thr = {e.x, e.y, e.z};
    r.uv = {d.w, e.w};
    r.mat = f.x;
    r.path_len = f.y;
    r.side = f.z;
    return r;
}

LT_DEV bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

// utils.glsl:62-65 (GLSL int products wrap like uint32)
LT_DEV uint32_t cell_hash(int x, int y, int z, uint32_t n) {
    const uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
    return h % (10u * n);
}
// vcm_commons.glsl:19-23 (its clamp result is discarded, so no clamp here)
LT_DEV void grid_idx(const VcmLaunch& L, f3 p, int o[3]) {
    const float qx = (p.x - L.bmin[0]) / (L.bmax[0] - L.bmin[0]);
    const float qy = (p.y - L.bmin[1]) / (L.bmax[1] - L.bmin[1]);
    const float qz = (p.z - L.bmin[2]) / (L.bmax[2] - L.bmin[2]);
    o[0] = (int)(qx * (float)L.grid_res[0]);
    o[1] = (int)(qy * (float)L.grid_res[1]);
    o[2] = (int)(qz * (float)L.grid_res[2]);
}

// The corrected estimator's grid: grid_idx with the extent's reciprocal.
// Only the set of photons within the radius enters its sum, and any grid
// whose cell index is monotone in the position finds each of them exactly
// once, so the cheaper multiply changes nothing but bucket membership.
LT_DEV void grid_idx_fast(const VcmLaunch& L, f3 p, int o[3]) {
    o[0] = (int)((p.x - L.bmin[0]) * L.inv_ext[0] * (float)L.grid_res[0]);
    o[1] = (int)((p.y - L.bmin[1]) * L.inv_ext[1] * (float)L.grid_res[1]);
    o[2] = (int)((p.z - L.bmin[2]) * L.inv_ext[2] * (float)L.grid_res[2]);
}

template <bool COUNT>
LT_DEV bool occluded(const SceneView& sv, f3 o, f3 d, float tmax, const TraceStack& stk, Ctr& c) {
    if (COUNT) c.shadow++;
    HitRec h;
    return trace_bvh<true, COUNT>(sv, o, d, 0.0f, tmax, h, stk, c.snodes, c.stris);
}

template <bool COUNT>
LT_DEV bool closest(const SceneView& sv, f3 o, f3 d, Payload& p, const TraceStack& stk, Ctr& c) {
    if (COUNT) c.closest++;
    HitRec h;
    if (!trace_bvh<false, COUNT>(sv, o, d, VCM_TMIN, VCM_TMAX, h, stk, c.nodes, c.tris)) return false;
    fill_payload(sv, h.inst, h.prim, h.u, h.v, p);
    return true;
}

struct Camera {
    f3 pos, nrm;
    float area;
};
LT_DEV Camera camera_of(const VcmLaunch& L) {
    const f4 o4 = mul4(L.inv_view, f4{0.0f, 0.0f, 0.0f, 1.0f});
    const f4 fw = mul4(L.inv_view, f4{0.0f, 0.0f, 1.0f, 0.0f});
    f4 ai = mul4(L.inv_projection, f4{2.0f / (float)L.width, 2.0f / (float)L.height, 0.0f, 1.0f});
    ai = f4{ai.x / ai.w, ai.y / ai.w, ai.z / ai.w, ai.w / ai.w};
    return Camera{{o4.x, o4.y, o4.z}, {-fw.x, -fw.y, -fw.z}, fabsf(ai.x * ai.y)};
}

// ---------------------------------------------------------------- light pass
// vcm_commons.glsl:24-75; returns the splat pixel or -1
template <bool COUNT>
LT_DEV int connect_cam(const SceneView& sv, const VcmLaunch& L, const Camera& cam, f3 n_s, f3 payload_n_s, f3 pos,
                       const VState& st, f3 wo, const Mat& mat, f3& out, const TraceStack& stk, Ctr& c) {
    const float screen_size = (float)L.n_paths;
    f3 dir = cam.pos - pos;
    const float len = length(dir);
    dir /= len;
    const float cos_y = dot(dir, n_s);
    const float cos_theta = dot(cam.nrm, -dir);
    if (cos_theta <= 0.0f) return -1;
    const float c3 = cos_theta * cos_theta * cos_theta;
    const float ratio = fabsf(cos_y) / (cam.area * c3 * len * len);
    float pdf_rev, pdf_fwd;
    const f3 f = eval_bsdf(n_s, wo, mat, dot(payload_n_s, wo) > 0.0f, dir, pdf_fwd, pdf_rev, true);
    if (is_zero(f)) return -1;
    // raster position first (no state beyond it is needed by the query)
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
    if (!(fx > -1.0f && fx < (float)L.width && fy > -1.0f && fy < (float)L.height) || dot(nd, cam.nrm) < 0.0f)
        return -1;
    if (!(ratio > 0.0f)) return -1;
    const float w_light = (ratio / screen_size) * (L.eta_vm + st.d_vcm + pdf_rev * st.d_vc);
    const float mis = 1.0f / (1.0f + w_light);
    out = mis * st.thr * ratio * f / screen_size;
    if (!(luminance(out) > 0.0f)) return -1;
    if (occluded<COUNT>(sv, offset_ray2(pos, n_s), dir, len - EPS, stk, c)) return -1;
    return (int)fy * (int)L.width + (int)fx;
}

template <bool COUNT>
LT_DEV uint32_t light_path(const SceneView& sv, const VcmLaunch& L, const Camera& cam, uint32_t p,
                           VcmVertex* __restrict__ lv, float* __restrict__ splat, const TraceStack& stk, Ctr& c) {
    const uint32_t x = p / L.height, y = p % L.height;  // pixel_idx = x * H + y (vcm_light.rgen:56)
    RNG seed = {x, y, L.light_word, 0u};
    f4 rp;
    rp.x = rnd(seed);
    rp.y = rnd(seed);
    rp.z = rnd(seed);
    rp.w = rnd(seed);
    const float rd0 = rnd(seed), rd1 = rnd(seed);
    float cos_theta, pdf_pos, pdf_dir;
    uint32_t flags, li;
    f3 pos, wi, n;
    const f3 Le = sample_light_Le(sv, rp, rd0, rd1, cos_theta, flags, li, pos, wi, n, pdf_pos, pdf_dir);
    if (pdf_dir <= 0.0f) return 0;
    VState st;
    const float pdf_emit = pdf_pos * pdf_dir, pdf_direct = pdf_pos;
    st.pos = pos;
    st.wi = wi;
    st.n_s = n;
    st.thr = Le * cos_theta / (pdf_dir * pdf_pos);
    st.d_vcm = pdf_direct / pdf_emit;
    const bool finite = light_finite(flags);
    st.d_vc = light_delta(flags) ? 0.0f : (finite ? cos_theta : 1.0f) / (pdf_dir * pdf_pos);
    st.d_vm = st.d_vc * L.eta_vc;
    uint32_t path_idx = 0;
    for (int depth = 1;; depth++) {
        Payload pl;
        if (!closest<COUNT>(sv, st.pos, st.wi, pl, stk, c)) break;
        f3 wo = st.pos - pl.pos;
        f3 n_s = pl.n_s, n_g = pl.n_g;
        bool side = true;
        if (dot(pl.n_g, wo) <= 0.0f) n_g = -n_g;
        if (dot(n_g, n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        const float dist = length(pl.pos - st.pos);
        const float dist_sqr = dist * dist;
        wo /= dist;
        const Mat mat = load_material(sv, pl.material_idx, pl.uv);
        const bool spec = (mat.bsdf_props & LT_BSDF_FLAG_SPECULAR) == LT_BSDF_FLAG_SPECULAR;
        const float cos_theta_wo = fabsf(dot(wo, n_s));
        if (depth > 1 || finite) st.d_vcm *= dist_sqr;
        st.d_vcm /= cos_theta_wo;
        st.d_vc /= cos_theta_wo;
        st.d_vm /= cos_theta_wo;
        if (!spec) {  // vertex connection is always on (VCM.cpp:7)
            VcmVertex* v = lv + path_idx;
            v->a = make_float4(pl.pos.x, pl.pos.y, pl.pos.z, st.d_vcm);
            v->b = make_float4(n_s.x, n_s.y, n_s.z, st.d_vc);
            v->c = make_float4(st.wi.x, st.wi.y, st.wi.z, st.d_vm);
            v->d = make_float4(wo.x, wo.y, wo.z, pl.uv.x);
            v->e = make_float4(st.thr.x, st.thr.y, st.thr.z, pl.uv.y);
            v->f = make_uint4(pl.material_idx, (uint32_t)depth + 1u, side ? 1u : 0u, 0u);
            path_idx++;
        }
        if (depth >= L.max_depth) break;
        if (!spec) {
            f3 sc;
            const int px = connect_cam<COUNT>(sv, L, cam, n_s, pl.n_s, pl.pos, st, wo, mat, sc, stk, c);
            if (px >= 0) {
                float* dst = splat + 3 * (size_t)px;
                unsafeAtomicAdd(dst, sc.x);
                unsafeAtomicAdd(dst + 1, sc.y);
                unsafeAtomicAdd(dst + 2, sc.z);
            }
        }
        float pdf_d, cos_t;
        const float r0 = rnd(seed), r1 = rnd(seed), r2 = rnd(seed);
        const f3 f = sample_bsdf(n_s, wo, mat, 0u, side, st.wi, pdf_d, cos_t, r0, r1, r2);
        const bool same_hem = dot(st.wi, n_s) > 0.0f && dot(wo, n_s) > 0.0f;
        const bool transm = (mat.bsdf_props & LT_BSDF_FLAG_TRANSMISSION) == LT_BSDF_FLAG_TRANSMISSION;
        if (is_zero(f) || pdf_d == 0.0f || (!same_hem && !transm)) break;
        float pdf_rev = pdf_d;
        if (!spec) pdf_rev = bsdf_pdf(mat, n_s, st.wi, wo, side);
        const float ac = fabsf(cos_t);
        st.pos = offset_ray(pl.pos, n_g);
        if (!spec) {
            st.d_vc = (ac / pdf_d) * (L.eta_vm + st.d_vcm + pdf_rev * st.d_vc);
            st.d_vm = (ac / pdf_d) * (1.0f + st.d_vcm * L.eta_vc + pdf_rev * st.d_vm);
            st.d_vcm = 1.0f / pdf_d;
        } else {
            st.d_vcm = 0.0f;
            st.d_vc *= ac;
            st.d_vm *= ac;
        }
        st.thr *= f * ac / pdf_d;
        st.n_s = n_s;
    }
    return path_idx;
}

// Three waves per SIMD (168 VGPRs, no scratch); four (128 VGPRs, 148 B of
// scratch per lane) produced wrong images in the replay tests once the BVH8
// node step picked its plane loads by octant, with no cause found in this
// source, so it is not used.
#ifndef LT_VCM_LIGHT_WAVES
#define LT_VCM_LIGHT_WAVES 3
#endif
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_VCM_LIGHT_WAVES))) void k_vcm_light(SceneView sv, VcmLaunch L, VcmVertex* __restrict__ verts,
                                                           uint32_t* __restrict__ cnt, float* __restrict__ splat,
                                                           uint32_t* __restrict__ key, uint32_t* __restrict__ count,
                                                           uint32_t* __restrict__ work, uint32_t* __restrict__ spill,
                                                           unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t chunks = (L.n_paths + 63u) / 64u;
    const uint32_t nv = (uint32_t)L.max_depth + 1u;
    const Camera cam = camera_of(L);
    Ctr c{0, 0, 0, 0};
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(work, 1u);
        t = __shfl(t, 0);
        if (t >= chunks) break;
        const uint32_t p = t * 64u + lane;
        if (p >= L.n_paths) continue;
        VcmVertex* lv = verts + (size_t)p * nv;
        const uint32_t n = light_path<COUNT>(sv, L, cam, p, lv, splat, stk, c);
        cnt[p] = n;
        if (L.use_vm) {  // vcm_commons.glsl:455-469
            for (uint32_t i = 0; i < n; i++) {
                const float4 a = lv[i].a;
                int g[3];
                if (L.corrected)
                    grid_idx_fast(L, f3{a.x, a.y, a.z}, g);
                else
                    grid_idx(L, f3{a.x, a.y, a.z}, g);
                const uint32_t h = cell_hash(g[0], g[1], g[2], L.n_paths);
                if (!L.corrected) atomicMax(&key[h], p * nv + i + 1u);
                atomicAdd(&count[h], 1u);
            }
        }
    }
    if (COUNT) flush_counts(c, counts);
}

// corrected estimator: every photon into its bucket's range
__global__ void k_vcm_scatter(VcmLaunch L, const VcmVertex* __restrict__ verts, const uint32_t* __restrict__ cnt,
                              const uint32_t* __restrict__ start, uint32_t* __restrict__ fill,
                              float4* __restrict__ ppos, VcmPhotonRest* __restrict__ prest) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L.n_paths) return;
    const uint32_t nv = (uint32_t)L.max_depth + 1u;
    for (uint32_t i = 0; i < cnt[p]; i++) {
        const VcmVertex& v = verts[(size_t)p * nv + i];
        const float4 a = v.a, b = v.b, c = v.c, e = v.e;
        const uint4 f = v.f;
        int g[3];
        grid_idx_fast(L, f3{a.x, a.y, a.z}, g);
        const uint32_t h = cell_hash(g[0], g[1], g[2], L.n_paths);
        const uint32_t k = start[h] + atomicAdd(&fill[h], 1u);
        ppos[k] = a;
        VcmPhotonRest& o = prest[k];
        o.b = make_float4(-c.x, -c.y, -c.z, c.w);
        o.c = make_float4(e.x, e.y, e.z, __uint_as_float(f.y));
        o.d = make_float4(b.x, b.y, b.z, 0.0f);
    }
}

// ------------------------------------------------------------------ eye pass
struct Grid {
    const VcmVertex* verts;
    const uint32_t *cnt, *key, *count, *start;
    const float4* ppos;
    const VcmPhotonRest* prest;
};

// vcm_commons.glsl:132-143
LT_DEV f3 light_radiance(const SceneView& sv, const VcmLaunch& L, const Mat& mat, const VState& cs, const Payload& pl,
                         int d) {
    if (d == 1) return mat.emissive;
    const float pdf_light_pos = 1.0f / (pl.area * (float)sv.total_light_triangle_cnt);
    const float pdf_light_dir = fabsf(dot(pl.n_s, -cs.wi)) / PI;
    const float w_camera = pdf_light_pos * cs.d_vcm + (pdf_light_pos * pdf_light_dir) * cs.d_vc;
    const float mis = 1.0f / (1.0f + w_camera);
    return mis * mat.emissive;
}

// vcm_commons.glsl:145-199
template <bool COUNT>
LT_DEV f3 connect_light(const SceneView& sv, const VcmLaunch& L, f3 n_s, f3 wo, const Mat& mat, bool side,
                        const VState& cs, const Payload& pl, RNG& seed, float& pdf_rev, const TraceStack& stk,
                        Ctr& c) {
    f4 r4;
    r4.x = rnd(seed);
    r4.y = rnd(seed);
    r4.z = rnd(seed);
    r4.w = rnd(seed);
    f3 wi, nn, lp;
    float wi_len, pdf_pos_w, pdf_pos_a, cos_y, pdf_pos_dir_w;
    LightRecord rec;
    rec.triangle_idx = rec.instance_idx = 0u;
    const f3 Le =
        sample_light_Li(sv, r4, pl.pos, pdf_pos_w, wi, wi_len, pdf_pos_a, cos_y, rec, nn, lp, pdf_pos_dir_w);
    const float cos_x = dot(wi, n_s);
    float pdf_fwd;
    const f3 f = eval_bsdf(n_s, wo, mat, side, wi, pdf_fwd, pdf_rev, true);
    if (is_zero(f)) return mk3(0.0f);
    if (light_delta(rec.flags)) pdf_fwd = 0.0f;
    const float ltc = (float)sv.total_light_triangle_cnt;
    const float w_light = pdf_fwd / (pdf_pos_w / ltc);
    const float w_cam = pdf_pos_dir_w * fabsf(cos_x) / (pdf_pos_w * cos_y) * (L.eta_vm + cs.d_vcm + cs.d_vc * pdf_rev);
    const float mis = 1.0f / (1.0f + w_light + w_cam);
    if (!(mis > 0.0f)) return mk3(0.0f);
    const f3 res = mis * fabsf(cos_x) * f * cs.thr * Le / (pdf_pos_w / ltc);
    if (occluded<COUNT>(sv, offset_ray2(pl.pos, n_s), wi, wi_len - EPS, stk, c)) return mk3(0.0f);
    return res;
}

// vcm_commons.glsl:201-247 over the stored vertices of light path lp
template <bool COUNT>
LT_DEV f3 connect_light_vertices(const SceneView& sv, const VcmLaunch& L, const Grid& G, uint32_t lp, int depth,
                                 f3 n_s, f3 wo, const Mat& mat, bool side, const VState& cs, float nee_pdf_rev,
                                 const Payload& pl, const TraceStack& stk, Ctr& c) {
    f3 res = mk3(0.0f);
    const uint32_t nv = (uint32_t)L.max_depth + 1u;
    const uint32_t n = G.cnt[lp];
    for (uint32_t i = 0; i < n; i++) {
        const LVtx v = load_lv(G.verts + (size_t)lp * nv + i);
        const uint32_t md = v.path_len + (uint32_t)depth - 1u;
        if (md >= (uint32_t)L.max_depth) break;
        f3 dir = v.pos - pl.pos;
        const float len = length(dir);
        const float len_sqr = len * len;
        dir /= len;
        const float cos_cam = dot(n_s, dir), cos_light = dot(v.n_s, -dir);
        const float Gt = cos_light * cos_cam / len_sqr;
        if (!(Gt > 0.0f)) continue;
        float cam_pdf_fwd, cam_pdf_rev, l_pdf_fwd, l_pdf_rev;
        const f3 f_cam = eval_bsdf(n_s, wo, mat, side, dir, cam_pdf_fwd, cam_pdf_rev, L.corrected != 0u);
        f3 f_light;
        {
            const Mat lm = load_material(sv, v.mat, v.uv);
            f_light = eval_bsdf(v.n_s, v.wo, lm, v.side == 1u, -dir, l_pdf_fwd, l_pdf_rev, true);
        }
        if (is_zero(f_light) || is_zero(f_cam)) continue;
        cam_pdf_fwd *= fabsf(cos_light) / len_sqr;
        l_pdf_fwd *= fabsf(cos_cam) / len_sqr;
        const float w_light = cam_pdf_fwd * (L.eta_vm + v.d_vcm + l_pdf_rev * v.d_vc);
        const float rev = L.corrected ? cam_pdf_rev : nee_pdf_rev;
        const float w_camera = l_pdf_fwd * (L.eta_vm + cs.d_vcm + rev * cs.d_vc);
        const float mis = 1.0f / (1.0f + w_camera + w_light);
        const f3 r = mis * Gt * cs.thr * v.thr * f_cam * f_light;
        if (!occluded<COUNT>(sv, offset_ray2(pl.pos, n_s), dir, len - EPS, stk, c)) {
            if (L.corrected)
                res += r;
            else
                res = r;
        }
    }
    return res;
}

// one merge candidate (vcm_commons.glsl:268-290) of a photon at ppos within
// radius (dist_sqr) arriving from direction pwi; false if rejected
LT_DEV bool merge_one(const VcmLaunch& L, float dist_sqr, f3 pwi, f3 pthr, f3 pnrm, float pd_vcm, float pd_vm,
                      uint32_t ppath_len, float photon_count, int depth, f3 n_s, f3 wo, const Mat& mat, bool side,
                      const VState& cs, float norm, f3& out) {
    const uint32_t dd = ppath_len + (uint32_t)depth - 1u;
    if (dd > (uint32_t)L.max_depth) return false;
    float pf, pr;
    const f3 f = eval_bsdf(n_s, wo, mat, side, pwi, pf, pr, true);
    if (is_zero(f)) return false;
    const float w_light = pd_vcm * L.eta_vc + pd_vm * pf;
    const float w_cam = cs.d_vcm * L.eta_vc + cs.d_vm * pr;
    const float mis = 1.0f / (1.0f + w_light + w_cam);
    const float cos_nrm = dot(pnrm, n_s);
    if (!(cos_nrm > EPS)) return false;
    const float w = 1.0f - sqrtf(dist_sqr) / L.radius;
    out = w * mis * photon_count * pthr * f * cs.thr * norm * 3.0f;
    return true;
}

// vcm_commons.glsl:251-296, the reference estimator: one photon per bucket
// (the deterministic "last writer") weighted by the bucket's count, the
// result overwritten per cell
LT_DEV f3 merge(const VcmLaunch& L, const Grid& G, int depth, f3 n_s, f3 wo, const Mat& mat, bool side,
                const VState& cs, const Payload& pl, float norm) {
    f3 res = mk3(0.0f);
    int lo[3], hi[3];
    grid_idx(L, pl.pos - mk3(L.radius), lo);
    grid_idx(L, pl.pos + mk3(L.radius), hi);
    for (int x = lo[0]; x <= hi[0]; x++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int z = lo[2]; z <= hi[2]; z++) {
                const uint32_t h = cell_hash(x, y, z, L.n_paths);
                const uint32_t cnt = G.count[h];
                if (cnt == 0u) continue;
                const LVtx v = load_lv(G.verts + (G.key[h] - 1u));
                const f3 pp = pl.pos - v.pos;
                const float dist_sqr = dot(pp, pp);
                if (dist_sqr > L.radius * L.radius) continue;
                f3 r;
                if (merge_one(L, dist_sqr, -v.wi, v.thr, v.n_s, v.d_vcm, v.d_vm, v.path_len, (float)(int)cnt, depth,
                              n_s, wo, mat, side, cs, norm, r))
                    res = r;
            }
    return res;
}

// A cell's coordinates modulo 1024 in one word.  Within one query every cell
// compared (those of its range, and that of a photon within the radius of
// it) lies in the query's box of cells, a few cells wide, where this packing
// is one to one.
LT_DEV uint32_t cell_pack(int x, int y, int z) {
    return (((uint32_t)x & 1023u) << 20) | (((uint32_t)y & 1023u) << 10) | ((uint32_t)z & 1023u);
}

// ---- merge pass (corrected estimator) ----
// vcm_commons.glsl:251-296 summed over every photon in range, for the merge
// queries the eye pass wrote.  A whole wave serves one query at a time: the
// query and its material are wave-uniform (scalar loads and registers, and
// the BSDF's branches on the material are uniform), the lanes load the
// photon ranges of up to 64 of its cells at once, a wave scan lays the
// cells' photons end to end and each lane takes every 64th photon.  A query
// over a dense caustic costs ceil(photons / 64) coalesced steps and no lane
// waits on another lane's serial loop.  The wave's sum is added to the
// pixel's colour with float atomics (the order of the sum is the only
// difference from a serial loop).
#ifndef LT_VCM_MERGE_WAVES
#define LT_VCM_MERGE_WAVES 3
#endif
#ifndef LT_VCM_CELL_DIV
#define LT_VCM_CELL_DIV 2
#endif
#ifndef LT_VCM_MERGE_UNROLL
#define LT_VCM_MERGE_UNROLL 4
#endif
constexpr int kMergeUnroll = LT_VCM_MERGE_UNROLL;
// pending merges per wave: at most 63 carried + 64 per unrolled photon,
// rounded up to a power of two (512 for the default of four)
constexpr uint32_t ring_size(uint32_t need) { return need <= 1u ? 1u : 2u * ring_size((need + 1u) / 2u); }
constexpr uint32_t kMergeRing = ring_size(63u + 64u * (uint32_t)kMergeUnroll);
static_assert(kMergeUnroll >= 1 && 63 + 64 * kMergeUnroll <= (int)kMergeRing, "merge ring too small");
// COUNT: the wave tallies the photons it tests (16 B position each), the
// photons it merges (48 B record each) and the algorithmic bytes of its
// queries (80-B record + 8 B of bucket range per cell) into counts[4..6].
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_VCM_MERGE_WAVES))) void k_vcm_merge(
    SceneView sv, VcmLaunch L, Grid Gr, const VcmQuery* __restrict__ queries, const uint32_t* __restrict__ n_queries,
    float4* __restrict__ fcol, unsigned long long* __restrict__ counts) {
    const uint32_t lane = threadIdx.x & 63u;
    unsigned long long n_tested = 0, n_merged = 0, n_bytes = 0;  // wave-uniform
    // photons of the wave's query that passed the distance and cell tests,
    // waiting for a full wave of merges: a ring per wave (index, squared
    // distance, the photon's d_vcm)
    __shared__ uint32_t s_k[kTraceBlock / 64u][kMergeRing];
    __shared__ float s_d2[kTraceBlock / 64u][kMergeRing];
    __shared__ float s_w[kTraceBlock / 64u][kMergeRing];
    uint32_t* const wk = s_k[threadIdx.x / 64u];
    float* const wd2 = s_d2[threadIdx.x / 64u];
    float* const ww = s_w[threadIdx.x / 64u];
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t Q = *n_queries;
    const float r2 = L.radius * L.radius;
    const float norm = 1.0f / (PI * r2 * (float)L.n_paths);
    // Queries w, w + waves, ...: neighbouring queries (neighbouring pixels)
    // cost alike, so the interleave balances.  Workgroups are dealt to the 8
    // XCDs round-robin (blockIdx mod 8), each XCD with its own L2, so the
    // wave numbering puts the waves of one XCD on consecutive queries: a
    // round of the interleave gives each XCD one contiguous run of queries
    // (the eye pass appends a tile's queries together, so a run is a compact
    // image region whose cell ranges and photons its L2 then holds).  Giving
    // each XCD one contiguous eighth of all queries instead left the XCDs
    // unevenly loaded (3.27 against 2.51 ms per caustics frame).
    const uint32_t wpb = blockDim.x / 64u, wib = threadIdx.x / 64u;
    const uint32_t q_stride = gridDim.x * wpb;
    const uint32_t q_first = (gridDim.x & 7u) == 0u
                                 ? (blockIdx.x & 7u) * (q_stride >> 3) + (blockIdx.x >> 3) * wpb + wib
                                 : blockIdx.x * wpb + wib;
    const uint32_t q_end = Q;
    for (uint32_t qi = (uint32_t)__builtin_amdgcn_readfirstlane((int)q_first); qi < q_end; qi += q_stride) {
        const VcmQuery& qr = queries[qi];
        const float4 qa = qr.a, qb = qr.b, qc = qr.c, qd = qr.d;
        const uint4 qe = qr.e;
        const f3 qpos = {qa.x, qa.y, qa.z}, qn = {qb.x, qb.y, qb.z}, qwo = {qc.x, qc.y, qc.z};
        VState cs;
        cs.thr = {qd.x, qd.y, qd.z};
        cs.d_vcm = qa.w;
        cs.d_vm = qb.w;
        const bool side = (qe.z & 1u) != 0u;
        const int depth = (int)(qe.z >> 1);
        const Mat mat = load_material(sv, qe.x, f2{qc.w, qd.w});
        int lo[3], hi[3];
        grid_idx_fast(L, qpos - mk3(L.radius), lo);
        grid_idx_fast(L, qpos + mk3(L.radius), hi);
        const int ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
        const int ncells = (hi[0] - lo[0] + 1) * ny * nz;
        if (COUNT) n_bytes += sizeof(VcmQuery) + 8ull * (unsigned long long)ncells;
        f3 part = mk3(0.0f);
        uint32_t head = 0u, pend = 0u;  // wave-uniform: ring position and fill
        auto merge_slot = [&](uint32_t slot) {
            const uint32_t k = wk[slot];
            const VcmPhotonRest& pr = Gr.prest[k];
            const float4 b = pr.b, c = pr.c, d = pr.d;
            f3 r;
            if (merge_one(L, wd2[slot], f3{b.x, b.y, b.z}, f3{c.x, c.y, c.z}, f3{d.x, d.y, d.z}, ww[slot], b.w,
                          __float_as_uint(c.w), 1.0f, depth, qn, qwo, mat, side, cs, norm, r))
                part += r;
        };
        for (int cbase = 0; cbase < ncells; cbase += 64) {
            const int ci = cbase + (int)lane;  // this lane's cell of the chunk: its photon range
            uint32_t st = 0u, cnt = 0u, cpack = 0u;
            if (ci < ncells) {
                const int cx = lo[0] + ci / (ny * nz);
                const int cy = lo[1] + (ci / nz) % ny;
                const int cz = lo[2] + ci % nz;
                cpack = cell_pack(cx, cy, cz);
                const uint32_t h = cell_hash(cx, cy, cz, L.n_paths);
                st = Gr.start[h];
                cnt = Gr.start[h + 1] - st;
            }
            uint32_t incl = cnt;  // inclusive scan over the wave
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t v = __shfl_up(incl, (unsigned)o);
                if ((int)lane >= o) incl += v;
            }
            // photon t of the chunk lies in lane j's cell at index base_j + t
            // (modulo 2^32)
            const uint32_t base = st - (incl - cnt);
            const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            if (COUNT) n_tested += total;
            // kMergeUnroll photons per lane per step: their positions are
            // loaded together, so one memory latency covers them all
            for (uint32_t t0 = 0; t0 < total; t0 += 64u * kMergeUnroll) {
                uint32_t kk[kMergeUnroll], jc[kMergeUnroll];
                bool live[kMergeUnroll];
                float4 pa[kMergeUnroll];
#pragma unroll
                for (int u = 0; u < kMergeUnroll; u++) {
                    const uint32_t t = t0 + 64u * (uint32_t)u + lane;
                    int j = 0;  // the cell holding photon t: the first lane whose scan exceeds t
                    for (int step = 32; step > 0; step >>= 1) {
                        const uint32_t v = __shfl(incl, j + step - 1);
                        if (v <= t) j += step;
                    }
                    j = j < 63 ? j : 63;
                    const uint32_t bj = __shfl(base, j);
                    jc[u] = __shfl(cpack, j);
                    live[u] = t < total;
                    kk[u] = live[u] ? bj + t : 0u;
                }
#pragma unroll
                for (int u = 0; u < kMergeUnroll; u++) pa[u] = live[u] ? Gr.ppos[kk[u]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                for (int u = 0; u < kMergeUnroll; u++) {
                    bool keep = false;
                    float dist_sqr = 0.0f;
                    if (live[u]) {
                        const f3 ppos = {pa[u].x, pa[u].y, pa[u].z};
                        const f3 pp = qpos - ppos;
                        dist_sqr = dot(pp, pp);
                        if (dist_sqr <= r2) {
                            int g[3];
                            grid_idx_fast(L, ppos, g);
                            keep = cell_pack(g[0], g[1], g[2]) == jc[u];  // else another cell of the bucket
                        }
                    }
                    const uint64_t m = __ballot(keep);
                    if (keep) {
                        const uint32_t slot = (head + pend + (uint32_t)__popcll(m & below)) & (kMergeRing - 1u);
                        wk[slot] = kk[u];
                        wd2[slot] = dist_sqr;
                        ww[slot] = pa[u].w;
                    }
                    pend += (uint32_t)__popcll(m);
                    if (COUNT) n_merged += (unsigned long long)__popcll(m);
                }
                while (pend >= 64u) {  // full waves of merges
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    merge_slot((head + lane) & (kMergeRing - 1u));
                    head += 64u;
                    pend -= 64u;
                    // later appends reuse these slots only after every lane's reads
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < pend) merge_slot((head + lane) & (kMergeRing - 1u));
        // the next query's writes must follow every lane's reads of this one
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int o = 32; o > 0; o >>= 1) {
            part.x += __shfl_xor(part.x, o);
            part.y += __shfl_xor(part.y, o);
            part.z += __shfl_xor(part.z, o);
        }
        if (lane == 0u && !is_zero(part)) {
            float* fc = reinterpret_cast<float*>(fcol + qe.y);
            atomicAdd(fc + 0, part.x);
            atomicAdd(fc + 1, part.y);
            atomicAdd(fc + 2, part.z);
        }
    }
    if (COUNT && lane == 0u) {
        atomicAdd(&counts[4], n_tested);
        atomicAdd(&counts[5], n_merged);
        atomicAdd(&counts[6], n_bytes + n_tested * 16ull + n_merged * 48ull);
    }
}

// The eye pass's colour plus the merges and the splats into the running
// mean (vcm_eye.rgen:80-97); a NaN colour drops the sample and leaves the
// pixel's splat for the next frame.
__global__ void k_vcm_resolve(const float4* __restrict__ fcol, float* __restrict__ splat, float4* __restrict__ accum,
                              uint32_t n, uint32_t frame_num) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 fc = fcol[i];
    f3 col = mk3(fc.x, fc.y, fc.z);
    if (isnan(luminance(col))) return;
    float* sp = splat + 3 * (size_t)i;
    col += mk3(sp[0], sp[1], sp[2]);
    sp[0] = sp[1] = sp[2] = 0.0f;
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

// The camera subpath of one pixel (vcm_eye.rgen:42-79 + vcm_commons.glsl:
// 477-631), one vertex per step of the wave's loop so that the whole wave
// reaches the merge of every depth together.
struct EyeState {
    RNG seed;
    VState cs;
    f3 col;
    uint32_t lp;
    bool alive;
};

LT_DEV void eye_start(const VcmLaunch& L, const Camera& cam, uint32_t x, uint32_t y, EyeState& e) {
    e.seed = {x, y, L.eye_word, 0u};
    const float screen_size = (float)L.n_paths;
    const float r0 = rnd(e.seed), r1 = rnd(e.seed);
    const float ux = ((float)x + 0.5f + (r0 - 0.5f)) / (float)L.width;
    const float uy = ((float)y + 0.5f + (r1 - 0.5f)) / (float)L.height;
    const float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    const f4 tg = mul4(L.inv_projection, f4{dx, dy, 1.0f, 1.0f});
    const f3 tn = normalize(mk3(tg.x, tg.y, tg.z));
    const f4 d4 = mul4(L.inv_view, f4{tn.x, tn.y, tn.z, 0.0f});
    e.cs.wi = {d4.x, d4.y, d4.z};
    e.cs.pos = cam.pos;
    e.cs.thr = mk3(1.0f);
    e.cs.n_s = cam.nrm;
    const float ct = fabsf(dot(e.cs.n_s, e.cs.wi));
    e.cs.d_vcm = cam.area * screen_size * ct * ct * ct;
    e.cs.d_vc = 0.0f;
    e.cs.d_vm = 0.0f;
    e.lp = (uint32_t)(rnd(e.seed) * screen_size);
    if (e.lp >= L.n_paths) e.lp = L.n_paths - 1u;  // float rounding of rand * N at very large N
    e.col = mk3(0.0f);
    e.alive = true;
}

// With the corrected estimator and merging, the merges are left to the merge
// pass: each non-specular vertex appends a query (wave-aggregated counter).
template <bool COUNT>
LT_DEV f3 eye_wave(const SceneView& sv, const VcmLaunch& L, const Grid& G, const Camera& cam, uint32_t x, uint32_t y,
                   bool valid, VcmQuery* __restrict__ queries, uint32_t* __restrict__ n_queries,
                   const TraceStack& stk, Ctr& c) {
    EyeState e;
    e.alive = false;
    e.col = mk3(0.0f);
    if (valid) eye_start(L, cam, x, y, e);
    const float norm = 1.0f / (PI * (L.radius * L.radius) * (float)L.n_paths);
    const bool coop = L.use_vm && L.corrected;
    for (int depth = 1; __ballot(e.alive) != 0ull; depth++) {
        bool want = false, side = true, spec = false;
        f3 pos = mk3(0.0f), n_s = mk3(0.0f), n_g = mk3(0.0f), wo = mk3(0.0f);
        uint32_t mat_idx = 0u;
        f2 uv = {0.0f, 0.0f};
        if (e.alive) {
            Payload pl;
            if (!closest<COUNT>(sv, e.cs.pos, e.cs.wi, pl, stk, c)) {
                e.col += e.cs.thr * ld3(L.sky_col);
                e.alive = false;
            } else {
                pos = pl.pos;
                mat_idx = pl.material_idx;
                uv = pl.uv;
                wo = e.cs.pos - pl.pos;
                const float dist = length(pl.pos - e.cs.pos);
                const float dist_sqr = dist * dist;
                wo /= dist;
                n_s = pl.n_s;
                n_g = pl.n_g;
                if (dot(pl.n_g, wo) < 0.0f) n_g = -n_g;
                if (dot(n_g, n_s) < 0.0f) {
                    n_s = -n_s;
                    side = false;
                }
                const float cos_wo = fabsf(dot(wo, n_s));
                const Mat mat = load_material(sv, pl.material_idx, pl.uv);
                spec = (mat.bsdf_props & LT_BSDF_FLAG_SPECULAR) == LT_BSDF_FLAG_SPECULAR;
                e.cs.d_vcm *= dist_sqr;
                e.cs.d_vcm /= cos_wo;
                e.cs.d_vc /= cos_wo;
                e.cs.d_vm /= cos_wo;
                if (luminance(mat.emissive) > 0.0f) e.col += e.cs.thr * light_radiance(sv, L, mat, e.cs, pl, depth);
                float pdf_rev = 0.0f;
                if (!spec && depth < L.max_depth)
                    e.col += connect_light<COUNT>(sv, L, n_s, wo, mat, side, e.cs, pl, e.seed, pdf_rev, stk, c);
                if (!spec)
                    e.col += connect_light_vertices<COUNT>(sv, L, G, e.lp, depth, n_s, wo, mat, side, e.cs, pdf_rev, pl,
                                                           stk, c);
                if (!spec && L.use_vm && !L.corrected)
                    e.col += merge(L, G, depth, n_s, wo, mat, side, e.cs, pl, norm);
                want = !spec && coop;
            }
        }
        if (coop) {
            const unsigned long long qm = __ballot(want);
            if (qm) {
                const uint32_t lane = threadIdx.x & 63u;
                const int leader = __ffsll((long long)qm) - 1;
                uint32_t qbase = 0;
                if ((int)lane == leader) qbase = atomicAdd(n_queries, (uint32_t)__popcll(qm));
                qbase = __shfl(qbase, leader);
                if (want) {
                    VcmQuery& q = queries[qbase + (uint32_t)__popcll(qm & ((1ull << lane) - 1ull))];
                    q.a = make_float4(pos.x, pos.y, pos.z, e.cs.d_vcm);
                    q.b = make_float4(n_s.x, n_s.y, n_s.z, e.cs.d_vm);
                    q.c = make_float4(wo.x, wo.y, wo.z, uv.x);
                    q.d = make_float4(e.cs.thr.x, e.cs.thr.y, e.cs.thr.z, uv.y);
                    q.e = make_uint4(mat_idx, y * L.width + x, (side ? 1u : 0u) | ((uint32_t)depth << 1), 0u);
                }
            }
        }
        if (e.alive) {
            if (depth >= L.max_depth) {
                e.alive = false;
            } else {
                const Mat mat = load_material(sv, mat_idx, uv);
                float pdf_d, cos_t;
                const float s0 = rnd(e.seed), s1 = rnd(e.seed), s2 = rnd(e.seed);
                const f3 f = sample_bsdf(n_s, wo, mat, 1u, side, e.cs.wi, pdf_d, cos_t, s0, s1, s2);
                const bool transm = (mat.bsdf_props & LT_BSDF_FLAG_TRANSMISSION) == LT_BSDF_FLAG_TRANSMISSION;
                const bool same_hem = dot(e.cs.wi, n_s) > 0.0f && dot(wo, n_s) > 0.0f;
                if (is_zero(f) || pdf_d == 0.0f || (!same_hem && !transm)) {
                    e.alive = false;
                } else {
                    float pdf_rev = pdf_d;
                    if (!spec) pdf_rev = bsdf_pdf(mat, n_s, e.cs.wi, wo, side);
                    const float ac = fabsf(cos_t);
                    e.cs.pos = offset_ray(pos, n_g);
                    if (!spec) {
                        e.cs.d_vc = (ac / pdf_d) * (L.eta_vm + e.cs.d_vcm + pdf_rev * e.cs.d_vc);
                        e.cs.d_vm = (ac / pdf_d) * (1.0f + e.cs.d_vcm * L.eta_vc + pdf_rev * e.cs.d_vm);
                        e.cs.d_vcm = 1.0f / pdf_d;
                    } else {
                        e.cs.d_vcm = 0.0f;
                        e.cs.d_vc *= ac;
                        e.cs.d_vm *= ac;
                    }
                    e.cs.thr *= f * ac / pdf_d;
                    e.cs.n_s = n_s;
                }
            }
        }
    }
    return e.col;
}

#ifndef LT_VCM_EYE_WAVES
#define LT_VCM_EYE_WAVES 2
#endif
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_VCM_EYE_WAVES, LT_VCM_EYE_WAVES))) void k_vcm_eye(SceneView sv, VcmLaunch L, Grid G,
                                                         float* __restrict__ splat, float4* __restrict__ accum,
                                                         float4* __restrict__ fcol, VcmQuery* __restrict__ queries,
                                                         uint32_t* __restrict__ n_queries,
                                                         uint32_t* __restrict__ work, uint32_t* __restrict__ spill,
                                                         unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tiles_x = (L.width + 7) / 8, tiles_y = (L.height + 7) / 8;
    const uint32_t n_tiles = tiles_x * tiles_y;
    const Camera cam = camera_of(L);
    Ctr c{0, 0, 0, 0};
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(work, 1u);
        t = __shfl(t, 0);
        if (t >= n_tiles) break;
        const uint32_t x = (t % tiles_x) * 8 + (lane & 7), y = (t / tiles_x) * 8 + (lane >> 3);
        const bool valid = x < L.width && y < L.height;
        f3 col = eye_wave<COUNT>(sv, L, G, cam, x, y, valid, queries, n_queries, stk, c);
        if (!valid) continue;
        const size_t i = (size_t)y * L.width + x;
        if (fcol) {  // merges and the resolve follow in their own passes
            fcol[i] = make_float4(col.x, col.y, col.z, 0.0f);
            continue;
        }
        if (isnan(luminance(col))) continue;  // vcm_eye.rgen:80-83: the splat stays for the next frame
        float* sp = splat + 3 * i;
        col += mk3(sp[0], sp[1], sp[2]);
        sp[0] = sp[1] = sp[2] = 0.0f;
        float4 a = accum[i];
        if (L.frame_num > 0) {
            const float w = 1.0f / (float)(L.frame_num + 1);
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
    if (COUNT) flush_counts(c, counts);
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void fill(lt_context* ctx, uint32_t* p, size_t n) {
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p, n, 0u);
}

// host restatement of pcg4d (utils.glsl:121-133) for the per-frame words
void pcg4d_host(uint32_t v[4]) {
    for (int k = 0; k < 4; k++) v[k] = v[k] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
    for (int k = 0; k < 4; k++) v[k] ^= v[k] >> 16;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
}

}  // namespace

// Per-frame RNG words: the light pass uses random_num (vcm_light.rgen:57-58)
// and the eye pass frame_num ^ random_num (vcm_eye.rgen:35-36); random_num
// (host rand(), VCM.cpp:147) is a pcg4d hash of (seed, frame) here.
void vcm_frame_words(uint32_t seed, uint32_t frame_num, uint32_t* light_word, uint32_t* eye_word) {
    uint32_t v[4] = {seed, frame_num, 0x9E3779B9u, 0x7F4A7C15u};
    pcg4d_host(v);
    *light_word = v[0];
    *eye_word = frame_num ^ v[0];
}

// VCM.cpp:140-158 (merge radius of frame frame_num, hash-grid resolution)
void vcm_radius_grid(const float smin[3], const float smax[3], float radius_factor, float alpha, uint32_t frame_num,
                     float* radius, int32_t grid_res[3]) {
#pragma clang fp contract(off)
    const float dx = smax[0] - smin[0], dy = smax[1] - smin[1], dz = smax[2] - smin[2];
    const float R = sqrtf(dx * dx + dy * dy + dz * dz) * 0.5f;
    float r = R * radius_factor / 100.0f;
    r /= (float)pow((double)frame_num + 1.0, 0.5 * (1.0 - (double)alpha));
    *radius = r;
    const float max_comp = fmaxf(dx, fmaxf(dy, dz));
    const int base = (int)(max_comp / r);
    const float g[3] = {dx * (float)base / max_comp, dy * (float)base / max_comp, dz * (float)base / max_comp};
    for (int k = 0; k < 3; k++) grid_res[k] = (int)g[k] > 1 ? (int)g[k] : 1;
}

lt_status vcm_alloc(lt_context* ctx, VcmBuffers& b, uint32_t width, uint32_t height, int32_t max_depth) {
    b.n_paths = width * height;
    b.max_depth = max_depth;
    const size_t N = b.n_paths, nv = (size_t)max_depth + 1, cells = 10 * N;
    LT_HIP(ctx, hipMalloc(&b.verts, N * nv * sizeof(VcmVertex)));
    LT_HIP(ctx, hipMalloc(&b.cnt, N * sizeof(uint32_t)));
    LT_HIP(ctx, hipMalloc(&b.key, cells * sizeof(uint32_t)));
    LT_HIP(ctx, hipMalloc(&b.count, (cells + 1) * sizeof(uint32_t)));
    LT_HIP(ctx, hipMalloc(&b.start, (cells + 1) * sizeof(uint32_t)));
    LT_HIP(ctx, hipMalloc(&b.fill, cells * sizeof(uint32_t)));
    LT_HIP(ctx, hipMalloc(&b.ppos, N * nv * sizeof(float4)));
    LT_HIP(ctx, hipMalloc(&b.prest, N * nv * sizeof(VcmPhotonRest)));
    LT_HIP(ctx, hipMalloc(&b.splat, N * 3 * sizeof(float)));
    LT_HIP(ctx, rocprim::exclusive_scan(nullptr, b.scan_bytes, b.count, b.start, 0u, cells + 1,
                                        rocprim::plus<uint32_t>(), ctx->stream));
    LT_HIP(ctx, hipMalloc(&b.scan_tmp, b.scan_bytes));
    fill(ctx, (uint32_t*)b.splat, N * 3);
    fill(ctx, b.cnt, N);
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}

void vcm_free(VcmBuffers& b) {
    void* ptrs[] = {b.verts, b.cnt, b.key, b.count, b.start, b.fill, b.ppos, b.prest, b.scan_tmp, b.splat, b.fcol,
                    b.queries, b.qctr};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    b = VcmBuffers{};
}

lt_status launch_vcm(lt_context* ctx, VcmBuffers& b, const VcmLaunch& L0, const lt_vcm_config& cfg,
                     uint32_t frame_num, uint32_t frames, uint32_t seed, float4* accum, int count,
                     uint64_t* d_counts) {
    if (L0.max_depth != b.max_depth || L0.n_paths != b.n_paths)
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "VCM launch does not match its buffers");
    // persistent grids: the workgroups every CU holds at once, unless the tuning fixes a count
    const uint32_t fixed = ctx->tuning.bdpt_blocks_per_cu;
    const uint32_t light_blocks =
        ctx->num_cus * resident_blocks(count ? reinterpret_cast<const void*>(&k_vcm_light<true>)
                                             : reinterpret_cast<const void*>(&k_vcm_light<false>),
                                       fixed);
    const uint32_t eye_blocks = ctx->num_cus * resident_blocks(count ? reinterpret_cast<const void*>(&k_vcm_eye<true>)
                                                                     : reinterpret_cast<const void*>(&k_vcm_eye<false>),
                                                               fixed);
    lt_status st = ensure_spill(ctx, (size_t)std::max(light_blocks, eye_blocks) * kTraceBlock);
    if (st != LT_OK) return st;
    // corrected estimator with merging: a merge pass over the eye pass's queries
    const bool merge_pass = L0.use_vm && L0.corrected;
    if (merge_pass && !b.queries) {
        LT_HIP(ctx, hipMalloc(&b.fcol, (size_t)b.n_paths * sizeof(float4)));
        LT_HIP(ctx, hipMalloc(&b.queries, (size_t)b.n_paths * (size_t)b.max_depth * sizeof(VcmQuery)));
        LT_HIP(ctx, hipMalloc(&b.qctr, 2 * sizeof(uint32_t)));
    }
    const uint32_t merge_blocks = ctx->num_cus * resident_blocks(reinterpret_cast<const void*>(&k_vcm_merge<false>), 0u);
    SceneView sv = make_scene_view(ctx);
    auto* cnt = (unsigned long long*)d_counts;
    if (count) fill(ctx, (uint32_t*)d_counts, 2 * kCountSlots);
    const size_t N = b.n_paths, cells = 10 * N;
    const dim3 lgrid(light_blocks), egrid(eye_blocks), block(kTraceBlock);
    const Grid G{b.verts, b.cnt, b.key, b.count, b.start, b.ppos, b.prest};
    for (uint32_t f = 0; f < frames; f++) {
        VcmLaunch L = L0;
        L.frame_num = frame_num + f;
        vcm_frame_words(seed, L.frame_num, &L.light_word, &L.eye_word);
        vcm_radius_grid(L.bmin, L.bmax, cfg.radius_factor, cfg.alpha, L.frame_num, &L.radius, L.grid_res);
        for (int k = 0; k < 3; k++) L.inv_ext[k] = 1.0f / (L.bmax[k] - L.bmin[k]);
        // The corrected estimator's cells are LT_VCM_CELL_DIV times finer per
        // axis than the reference's (about r / LT_VCM_CELL_DIV wide): a query's
        // box of +-r then covers (2 LT_VCM_CELL_DIV + 1)^3 smaller cells whose
        // photons lie closer to the radius.  Its sums do not depend on the grid.
        if (L.use_vm && L.corrected)
            for (int k = 0; k < 3; k++) L.grid_res[k] *= LT_VCM_CELL_DIV;
        const float r2 = L.radius * L.radius;
        L.eta_vcm = PI * r2 * (float)N;
        L.eta_vc = 1.0f / L.eta_vcm;
        L.eta_vm = L.use_vm ? PI * r2 * (float)N : 0.0f;
        if (f == 0) phase_mark(ctx);
        fill(ctx, ctx->counters, 2);
        if (L.use_vm) {
            if (!L.corrected) fill(ctx, b.key, cells);
            fill(ctx, b.count, cells + 1);
        }
        if (count)
            hipLaunchKernelGGL((k_vcm_light<true>), lgrid, block, 0, ctx->stream, sv, L, b.verts, b.cnt, b.splat,
                               b.key, b.count, ctx->counters, ctx->spill, cnt);
        else
            hipLaunchKernelGGL((k_vcm_light<false>), lgrid, block, 0, ctx->stream, sv, L, b.verts, b.cnt, b.splat,
                               b.key, b.count, ctx->counters, ctx->spill, cnt);
        if (L.use_vm && L.corrected) {
            LT_HIP(ctx, rocprim::exclusive_scan(b.scan_tmp, b.scan_bytes, b.count, b.start, 0u, cells + 1,
                                                rocprim::plus<uint32_t>(), ctx->stream));
            fill(ctx, b.fill, cells);
            hipLaunchKernelGGL(k_vcm_scatter, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, L,
                               b.verts, b.cnt, b.start, b.fill, b.ppos, b.prest);
        }
        float4* fcol = merge_pass ? b.fcol : nullptr;
        VcmQuery* queries = merge_pass ? b.queries : nullptr;
        if (merge_pass) fill(ctx, b.qctr, 1);
        if (f == 0) phase_mark(ctx);
        if (count)
            hipLaunchKernelGGL((k_vcm_eye<true>), egrid, block, 0, ctx->stream, sv, L, G, b.splat, accum, fcol,
                               queries, b.qctr, ctx->counters + 1, ctx->spill, cnt);
        else
            hipLaunchKernelGGL((k_vcm_eye<false>), egrid, block, 0, ctx->stream, sv, L, G, b.splat, accum, fcol,
                               queries, b.qctr, ctx->counters + 1, ctx->spill, cnt);
        if (f == 0) phase_mark(ctx);
        if (merge_pass) {
            if (count)
                hipLaunchKernelGGL(k_vcm_merge<true>, dim3(merge_blocks), block, 0, ctx->stream, sv, L, G, b.queries,
                                   b.qctr, b.fcol, cnt);
            else
                hipLaunchKernelGGL(k_vcm_merge<false>, dim3(merge_blocks), block, 0, ctx->stream, sv, L, G, b.queries,
                                   b.qctr, b.fcol, cnt);
        }
        if (f == 0) phase_mark(ctx);
        if (merge_pass) {
            hipLaunchKernelGGL(k_vcm_resolve, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, b.fcol,
                               b.splat, accum, (uint32_t)N, L.frame_num);
        }
    }
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}
