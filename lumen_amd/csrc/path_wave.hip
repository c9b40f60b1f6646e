// path_wave.hip — the Path integrator in wavefront form (gfx950).
//
// Per frame, a fixed sequence of persistent kernels over compacted queues
// (no host synchronisation; queue lengths live in a device control block):
//   generate      camera rays for every pixel -> extend queue 0
//   for depth d in [0, max_depth):
//     trace[d]    closest hits of extend queue d, and (d > 0) the shadow
//                 records (any hit: adds the NEE term if visible) and MIS
//                 records (closest hit: adds the BSDF-sampled light term if it
//                 hits the sampled light triangle) of depth d - 1, as one
//                 work range
//     shade[d]    emission, light sample (shadow record), MIS-BSDF record for
//                 area lights, BSDF sample + Russian roulette -> extend queue d+1
//   accumulate    running mean into the output (path.rgen:102-112)
// Terminated paths leave the queues (ballot + one atomic per wave), so every
// stage runs on full waves whatever the path lengths.  Each path consumes its
// pcg4d stream in exactly the order of path.rgen / pt_commons.glsl, so the
// image equals the megakernel's up to float summation order.
#include <hip/hip_runtime.h>

#include "context.h"
#include "dev_atmosphere.h"
#include "dev_scene.h"
#include "dev_trace.h"
#include "path.h"
#include "path_wave.h"

using namespace ltd;

namespace {

constexpr float PATH_TMIN = 0.001f;    // path.rgen:19
constexpr float PATH_TMAX = 10000.0f;  // path.rgen:20
constexpr int RR_MIN_DEPTH = 3;        // path.rgen:21

// frame number of frame f of the launch
LT_DEV uint32_t frame_of(const PathLaunch& L, uint32_t f) { return (L.frame_dev ? *L.frame_dev : L.frame_start) + f; }

// wave-level compaction: returns this lane's slot in `queue` (one atomic per wave)
LT_DEV uint32_t wave_append(uint32_t* len, bool pred) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t n = __popcll(mask);
    uint32_t base = 0;
    const int leader = __ffsll((long long)mask) - 1;
    if (n && (int)lane == leader) base = atomicAdd(len, n);
    base = __shfl(base, leader < 0 ? 0 : leader);
    const unsigned long long below = mask & ((1ull << lane) - 1ull);
    return base + (uint32_t)__popcll(below);
}

// Persistent wave loop helper: a wave takes 64 consecutive items.
#define WAVE_LOOP(n_expr, work_ptr)                         \
    const uint32_t lane = threadIdx.x & 63u;                \
    const uint32_t n_items = (n_expr);                      \
    for (;;) {                                              \
        uint32_t base = 0;                                  \
        if (lane == 0) base = atomicAdd((work_ptr), 64u);   \
        base = __shfl(base, 0);                             \
        if (base >= n_items) break;                         \
        const uint32_t item = base + lane;                  \
        const bool active = item < n_items;

#define WAVE_LOOP_END }

__global__ void k_wf_reset(WaveCtl* ctl) {
    uint32_t* p = (uint32_t*)ctl;
    for (uint32_t i = threadIdx.x; i < sizeof(WaveCtl) / 4; i += blockDim.x) p[i] = 0u;
}

// camera rays (path.rgen:26-38); one path per pixel
__global__ __launch_bounds__(256) void k_wf_generate(PathLaunch L, uint32_t f, WaveState S, WaveCtl* ctl) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = L.width * L.height;
    if (p >= n) return;
    const uint32_t frame_word = frame_of(L, f) ^ L.seed;
    const uint32_t x = p % L.width, y = p / L.width;
    RNG seed = {x, y, frame_word, 0u};
    const float r0 = rnd(seed), r1 = rnd(seed);
    const float ux = ((float)x + 0.5f + (r0 - 0.5f)) / (float)L.width;
    const float uy = ((float)y + 0.5f + (r1 - 0.5f)) / (float)L.height;
    const float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    const f4 o4 = mul4(L.inv_view, f4{0.0f, 0.0f, 0.0f, 1.0f});
    const f4 tg = mul4(L.inv_projection, f4{dx, dy, 1.0f, 1.0f});
    const f3 tn = normalize(mk3(tg.x, tg.y, tg.z));
    const f4 d4 = mul4(L.inv_view, f4{tn.x, tn.y, tn.z, 0.0f});
    S.ray_o[p] = make_float4(o4.x, o4.y, o4.z, 0.0f);
    S.ray_d[p] = make_float4(d4.x, d4.y, d4.z, 0.0f);
    S.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    S.col[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    S.misc[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    S.meta[p] = make_uint4(seed.w, 0u /*last_specular*/, 0u, 0u);
    S.queue[0][p] = p;
    if (p == 0) ctl->q_len[0] = n;
}

// shading of the extended paths (path.rgen:49-100, pt_commons.glsl:3-42)
#ifndef LT_WF_SHADE_WAVES
#define LT_WF_SHADE_WAVES 1
#endif
template <bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LT_WF_SHADE_WAVES))) void k_wf_shade(SceneView sv, PathLaunch L, uint32_t f, WaveState S,
                                                  WaveCtl* ctl, uint32_t depth, uint32_t sorted,
                                                  unsigned long long* __restrict__ counts) {
    const uint32_t frame_word = frame_of(L, f) ^ L.seed;
    const uint32_t* q = sorted ? S.sorted : S.queue[depth & 1];
    uint32_t* qn = S.queue[(depth + 1) & 1];
    uint32_t n_closest = 0, n_shadow = 0;
    WAVE_LOOP(ctl->q_len[depth], &ctl->work[depth][1])
    bool cont = false, want_sh = false, want_mis = false;
    uint32_t p = 0;
    ShadowRec sr;
    MisRec mr;
    float4 new_o, new_d, thr4;
    uint32_t counter = 0, last_spec = 0;
    if (active) {
        p = q[item];
        if (COUNT) n_closest++;
        const uint4 hr = S.hit[p];
        float4 col = S.col[p];
        const float4 mc = S.misc[p];
        col.x += mc.x;
        col.y += mc.y;
        col.z += mc.z;
        S.misc[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        thr4 = S.thr[p];
        const uint4 meta = S.meta[p];
        counter = meta.x;
        last_spec = meta.y;
        f3 throughput = mk3(thr4.x, thr4.y, thr4.z);
        f3 c = mk3(col.x, col.y, col.z);
        const float4 d4 = S.ray_d[p];
        const f3 direction = mk3(d4.x, d4.y, d4.z);
        const uint32_t x = p % L.width, y = p / L.width;
        RNG seed = {x, y, frame_word, counter};
        if (hr.y == 0xFFFFFFFFu) {
            if (depth > 0 || L.direct_lighting == 1) {
                const float4 o4 = S.ray_o[p];
                c += throughput * sky_radiance(sv, ld3(L.sky_col), mk3(o4.x, o4.y, o4.z), direction, PATH_TMAX);
            }
        } else {
            Payload hp;
            fill_payload(sv, hr.y, hr.x, __uint_as_float(hr.z), __uint_as_float(hr.w), hp);
            const Mat hit_mat = load_material(sv, hp.material_idx, hp.uv);
            if ((depth == 0 && L.direct_lighting == 1) || last_spec) c += throughput * hit_mat.emissive;
            if (!(depth >= (uint32_t)(L.max_depth - 1))) {
                const f3 wo = -direction;
                f3 n_s = hp.n_s;
                bool side = true;
                f3 n_g = hp.n_g;
                if (dot(hp.n_g, wo) < 0.0f) n_g = -n_g;
                if (dot(n_g, hp.n_s) < 0.0f) {
                    n_s = -n_s;
                    side = false;
                }
                const f3 origin = offset_ray(hp.pos, n_g);
                const bool spec = is_specular(hit_mat);
                last_spec = spec ? 1u : 0u;
                if (!spec && (depth > 0 || L.direct_lighting == 1)) {
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
                    const f3 Le =
                        sample_light_Li(sv, r4, hp.pos, pdf_light_w, wi, wi_len, pdf_light_a, cos_from_light, rec);
                    const float lpp = nee_pick_pdf(sv, L.nee_pick, rec.light_idx);
                    const f3 ps = offset_ray2(hp.pos, n_s);
                    float bsdf_pdf, bsdf_rev;
                    float cos_x = dot(n_s, wi);
                    f3 f = eval_bsdf(n_s, wo, hit_mat, side, wi, bsdf_pdf, bsdf_rev, false);
                    if (pdf_light_w > 0.0f) {
                        const float mis = (rec.flags & LT_LIGHT_DELTA) ? 1.0f : 1.0f / (1.0f + bsdf_pdf / pdf_light_w);
                        const f3 contrib = throughput * (mis * f * fabsf(cos_x) * Le / pdf_light_w) / lpp;
                        sr.o = make_float4(ps.x, ps.y, ps.z, wi_len - EPS);
                        sr.d = make_float4(wi.x, wi.y, wi.z, __uint_as_float(p));
                        sr.c = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
                        // a zero contribution needs no occlusion query
                        want_sh = contrib.x != 0.0f || contrib.y != 0.0f || contrib.z != 0.0f;
                    }
                    if ((rec.flags & 7u) == LT_LIGHT_AREA) {
                        const float a0 = rnd(seed), a1 = rnd(seed), a2 = rnd(seed);
                        f = sample_bsdf(n_s, wo, hit_mat, 1, side, wi, bsdf_pdf, cos_x, a0, a1, a2);
                        if (bsdf_pdf != 0.0f) {
                            const f3 w = throughput * (f * fabsf(cos_x) * Le / bsdf_pdf) / lpp;
                            mr.o = make_float4(ps.x, ps.y, ps.z, pdf_light_a);
                            mr.d = make_float4(wi.x, wi.y, wi.z, bsdf_pdf);
                            mr.w = make_float4(w.x, w.y, w.z, __uint_as_float(p));
                            mr.vpos = make_float4(hp.pos.x, hp.pos.y, hp.pos.z, 0.0f);
                            mr.sns = make_float4(hp.n_s.x, hp.n_s.y, hp.n_s.z, 0.0f);
                            mr.ids = make_uint4(rec.triangle_idx, rec.instance_idx, hp.triangle_idx, hp.instance_idx);
                            want_mis = true;
                        }
                    }
                }
                float pdf, cos_theta;
                const float s0 = rnd(seed), s1 = rnd(seed), s2 = rnd(seed);
                f3 ndir;
                const f3 f = sample_bsdf(n_s, wo, hit_mat, 1, side, ndir, pdf, cos_theta, s0, s1, s2);
                if (pdf != 0.0f) {
                    throughput *= f * fabsf(cos_theta) / pdf;
                    float rr_scale = 1.0f;
                    if (hit_mat.bsdf_props & LT_BSDF_FLAG_TRANSMISSION)
                        rr_scale *= side ? 1.0f / hit_mat.ior : hit_mat.ior;
                    cont = true;
                    if (depth > (uint32_t)RR_MIN_DEPTH) {
                        const float rr_prob = fminf(0.95f, luminance(throughput) * rr_scale);
                        if (rr_prob == 0.0f || rr_prob < rnd(seed))
                            cont = false;
                        else
                            throughput /= rr_prob;
                    }
                    new_o = make_float4(origin.x, origin.y, origin.z, 0.0f);
                    new_d = make_float4(ndir.x, ndir.y, ndir.z, 0.0f);
                    thr4 = make_float4(throughput.x, throughput.y, throughput.z, 0.0f);
                }
            }
        }
        counter = seed.w;
        S.col[p] = make_float4(c.x, c.y, c.z, 0.0f);
    }
    // queue appends (every lane of the wave takes part in the ballots)
    const uint32_t ssh = wave_append(&ctl->sh_len[depth], want_sh);
    if (want_sh) S.shadow[ssh] = sr;
    const uint32_t smis = wave_append(&ctl->mis_len[depth], want_mis);
    if (want_mis) S.mis[smis] = mr;
    const uint32_t sn = wave_append(&ctl->q_len[depth + 1], cont);
    if (cont) {
        qn[sn] = p;
        S.ray_o[p] = new_o;
        S.ray_d[p] = new_d;
        S.thr[p] = thr4;
    }
    if (active) S.meta[p] = make_uint4(counter, last_spec, 0u, 0u);
    WAVE_LOOP_END
    if (COUNT) {
        unsigned long long a = n_closest, b = n_shadow;
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o);
            b += __shfl_xor(b, o);
        }
        if ((threadIdx.x & 63u) == 0) {
            atomicAdd(&counts[0], a);
            atomicAdd(&counts[1], b);
        }
    }
}

// One trace stage: the closest hits of the extend queue of depth ext_depth
// (when with_ext), then the shadow records (any hit, pt_commons.glsl:20-27)
// and the MIS records (closest hit, pt_commons.glsl:28-40) of depth
// sec_depth (when with_sec), as one concatenated work range streamed
// through the waves.  The three queues are ready together after a shade
// stage, so one launch and one drain serve them.  On a MIS miss the
// reference compares the stale payload of the current vertex (ray.rmiss:13
// only clears material_idx), which the record carries.
// 96 VGPRs: five waves per SIMD hide more of the node-fetch latency
#ifndef LT_WF_TRACE_WAVES
#define LT_WF_TRACE_WAVES 5
#endif
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_WF_TRACE_WAVES, LT_WF_TRACE_WAVES))) void k_wf_trace(SceneView sv, WaveState S, WaveCtl* ctl, uint32_t stage,
                                                          uint32_t ext_depth, uint32_t with_ext, uint32_t sec_depth,
                                                          uint32_t with_sec, uint32_t* __restrict__ spill,
                                                          unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
    const uint32_t n_ext = with_ext ? ctl->q_len[ext_depth] : 0u;
    const uint32_t n_sh = with_sec ? ctl->sh_len[sec_depth] : 0u;
    const uint32_t n_mis = with_sec ? ctl->mis_len[sec_depth] : 0u;
    const uint32_t* q = S.queue[ext_depth & 1];
    uint32_t n_nodes = 0, n_tris = 0, n_shr = 0, n_misr = 0;
    auto load = [&](uint32_t item, f3& o, f3& d, float& tmin, float& tmax, bool& any) {
        if (item < n_ext) {
            const uint32_t p = q[item];
            const float4 ro = S.ray_o[p], rd = S.ray_d[p];
            o = mk3(ro.x, ro.y, ro.z);
            d = mk3(rd.x, rd.y, rd.z);
            tmin = PATH_TMIN;
            tmax = PATH_TMAX;
            any = false;
        } else if (item < n_ext + n_sh) {
            const ShadowRec& r = S.shadow[item - n_ext];
            const float4 ro = r.o, rd = r.d;
            o = mk3(ro.x, ro.y, ro.z);
            d = mk3(rd.x, rd.y, rd.z);
            tmin = 0.0f;
            tmax = ro.w;
            any = true;
            if (COUNT) n_shr++;
        } else {
            const MisRec& r = S.mis[item - n_ext - n_sh];
            const float4 ro = r.o, rd = r.d;
            o = mk3(ro.x, ro.y, ro.z);
            d = mk3(rd.x, rd.y, rd.z);
            tmin = PATH_TMIN;
            tmax = PATH_TMAX;
            any = false;
            if (COUNT) n_misr++;
        }
        return item;
    };
    auto finish = [&](uint32_t item, const TraceState& s) {
        const bool hit = s.best.gid != 0xFFFFFFFFu;
        if (item < n_ext) {
            const uint32_t p = q[item];
            S.hit[p] = make_uint4(hit ? s.best.prim : 0xFFFFFFFFu, hit ? s.best.inst : 0xFFFFFFFFu,
                                  __float_as_uint(s.best.u), __float_as_uint(s.best.v));
        } else if (item < n_ext + n_sh) {
            if (hit) return;  // occluded
            const ShadowRec& r = S.shadow[item - n_ext];
            const float4 rd = r.d, rc = r.c;
            const uint32_t p = __float_as_uint(rd.w);
            float4 c = S.col[p];
            c.x += rc.x;
            c.y += rc.y;
            c.z += rc.z;
            S.col[p] = c;
        } else {
            const MisRec r = S.mis[item - n_ext - n_sh];
            const f3 wi = mk3(r.d.x, r.d.y, r.d.z);
            uint32_t tri = r.ids.z, inst = r.ids.w;
            f3 ppos = mk3(r.vpos.x, r.vpos.y, r.vpos.z), pns = mk3(r.sns.x, r.sns.y, r.sns.z);
            if (hit) {
                Payload pl;
                fill_payload(sv, s.best.inst, s.best.prim, s.best.u, s.best.v, pl);
                tri = pl.triangle_idx;
                inst = pl.instance_idx;
                ppos = pl.pos;
                pns = pl.n_s;
            }
            if (tri == r.ids.x && inst == r.ids.y) {
                const f3 pos = mk3(r.vpos.x, r.vpos.y, r.vpos.z);
                const float wl = length(ppos - pos);
                const float g = fabsf(dot(pns, -wi)) / (wl * wl);
                const float mis = 1.0f / (1.0f + r.o.w / (g * r.d.w));
                // the shadow record of the same path may be finishing in
                // another lane: the MIS term goes to its own buffer
                const uint32_t p = __float_as_uint(r.w.w);
                S.misc[p] = make_float4(r.w.x * mis, r.w.y * mis, r.w.z * mis, 0.0f);
            }
        }
    };
    trace_stream_mixed<COUNT>(sv, stk, n_ext + n_sh + n_mis, &ctl->work[stage][0], load, finish, n_nodes, n_tris);
    if (COUNT) {
        unsigned long long a = n_nodes, b = n_tris, sh = n_shr, mi = n_misr;
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o);
            b += __shfl_xor(b, o);
            sh += __shfl_xor(sh, o);
            mi += __shfl_xor(mi, o);
        }
        if ((threadIdx.x & 63u) == 0) {
            atomicAdd(&counts[0], mi);
            atomicAdd(&counts[1], sh);
            atomicAdd(&counts[2], a);
            atomicAdd(&counts[3], b);
        }
    }
}

// running mean with the NaN guard (path.rgen:102-112)
// Material sort (LT_WF_SORT_BSDF): the hit paths of extend queue `depth` are
// grouped by the BSDF type of their material (bin 0: misses, bins 1-6:
// diffuse, mirror, glass, dielectric, conductor, principled, bin 7: other),
// so that the waves of the shade stage run one BSDF's code rather than
// several types' branches of bsdf_commons.glsl:68-182 one after another.
// A path's result does not depend on where its shading runs.
LT_DEV uint32_t bsdf_bin(const SceneView& sv, uint4 hit) {
    if (hit.y == 0xFFFFFFFFu) return 0u;
    const uint32_t t = sv.mats[sv.prims[hit.y].material_index].bsdf_type;
    if (t == 0u || (t & (t - 1u)) != 0u || t > LT_BSDF_PRINCIPLED) return 7u;
    return 32u - (uint32_t)__builtin_clz(t);  // 1 -> 1, 2 -> 2, ..., 32 -> 6
}

__global__ __launch_bounds__(256) void k_wf_bin_count(SceneView sv, WaveState S, WaveCtl* ctl, uint32_t depth) {
    __shared__ uint32_t h[kWfBins];
    if (threadIdx.x < kWfBins) h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t* q = S.queue[depth & 1];
    const uint32_t n = ctl->q_len[depth];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t lane = threadIdx.x & 63u;
    // wave-uniform trip count: the ballots below see the whole wave
    for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        const uint32_t i = i0 + lane;
        uint32_t b = 0xFFu;
        if (i < n) {
            b = bsdf_bin(sv, S.hit[q[i]]);
            S.bin[i] = (uint8_t)b;
        }
#pragma unroll
        for (int k = 0; k < kWfBins; k++) {
            const uint32_t c = (uint32_t)__popcll(__ballot(b == (uint32_t)k));
            if (lane == 0u && c != 0u) atomicAdd(&h[k], c);
        }
    }
    __syncthreads();
    if (threadIdx.x < kWfBins && h[threadIdx.x] != 0u) atomicAdd(&ctl->bins[depth][0][threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_wf_bin_scatter(WaveState S, WaveCtl* ctl, uint32_t depth) {
    const uint32_t* q = S.queue[depth & 1];
    const uint32_t n = ctl->q_len[depth];
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t start[kWfBins];
    uint32_t acc = 0u;
#pragma unroll
    for (int k = 0; k < kWfBins; k++) {
        start[k] = acc;
        acc += ctl->bins[depth][0][k];
    }
    const uint32_t stride = gridDim.x * blockDim.x;
    // every lane of a wave runs the same number of iterations (ballots below)
    for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        const uint32_t i = i0 + lane;
        const bool valid = i < n;
        const uint32_t b = valid ? (uint32_t)S.bin[i] : 0xFFu;
        uint32_t slot = 0u;
#pragma unroll
        for (int k = 0; k < kWfBins; k++) {
            const unsigned long long m = __ballot(b == (uint32_t)k);
            if (m == 0ull) continue;
            const int leader = __ffsll((long long)m) - 1;
            uint32_t off = 0u;
            if ((int)lane == leader) off = atomicAdd(&ctl->bins[depth][1][k], (uint32_t)__popcll(m));
            off = __shfl(off, leader);
            if (b == (uint32_t)k) slot = start[k] + off + (uint32_t)__popcll(m & below);
        }
        if (valid) S.sorted[slot] = q[i];
    }
}

__global__ __launch_bounds__(256) void k_wf_accumulate(PathLaunch L, uint32_t f, WaveState S,
                                                       float4* __restrict__ accum) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L.width * L.height) return;
    const uint32_t frame_num = frame_of(L, f);
    float4 c = S.col[p];
    const float4 mc = S.misc[p];
    c.x += mc.x;
    c.y += mc.y;
    c.z += mc.z;
    if (isnan(luminance(mk3(c.x, c.y, c.z)))) return;
    float4 acc = accum[p];
    if (frame_num > 0) {
        const float w = 1.0f / (float)(frame_num + 1);
        acc.x = acc.x * (1.0f - w) + c.x * w;
        acc.y = acc.y * (1.0f - w) + c.y * w;
        acc.z = acc.z * (1.0f - w) + c.z * w;
    } else {
        acc.x = c.x;
        acc.y = c.y;
        acc.z = c.z;
    }
    acc.w = 1.0f;
    accum[p] = acc;
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

}  // namespace

lt_status wave_alloc(lt_context* ctx, WaveBuffers& wb, uint32_t n_paths) {
    if (wb.capacity >= n_paths) return LT_OK;
    wave_free(wb);
    const size_t n = n_paths;
    auto A = [&](void** p, size_t bytes) { return hipMalloc(p, bytes); };
    hipError_t e = A((void**)&wb.S.ray_o, n * 16);
    if (e == hipSuccess) e = A((void**)&wb.S.ray_d, n * 16);
    if (e == hipSuccess) e = A((void**)&wb.S.thr, n * 16);
    if (e == hipSuccess) e = A((void**)&wb.S.col, n * 16);
    if (e == hipSuccess) e = A((void**)&wb.S.misc, n * 16);
    if (e == hipSuccess) e = A((void**)&wb.S.meta, n * 16);
    if (e == hipSuccess) e = A((void**)&wb.S.hit, n * 16);
    if (e == hipSuccess) e = A((void**)&wb.S.queue[0], n * 4);
    if (e == hipSuccess) e = A((void**)&wb.S.queue[1], n * 4);
    if (e == hipSuccess) e = A((void**)&wb.S.sorted, n * 4);
    if (e == hipSuccess) e = A((void**)&wb.S.bin, n);
    if (e == hipSuccess) e = A((void**)&wb.S.shadow, n * sizeof(ShadowRec));
    if (e == hipSuccess) e = A((void**)&wb.S.mis, n * sizeof(MisRec));
    if (e == hipSuccess) e = A((void**)&wb.ctl, sizeof(WaveCtl));
    if (e != hipSuccess) {
        wave_free(wb);
        return lt_hip_check(ctx, e, "wavefront buffers");
    }
    wb.capacity = n_paths;
    return LT_OK;
}

void wave_free(WaveBuffers& wb) {
    void* ptrs[] = {wb.S.ray_o, wb.S.ray_d, wb.S.thr, wb.S.col, wb.S.misc, wb.S.meta, wb.S.hit,
                    wb.S.queue[0], wb.S.queue[1], wb.S.sorted, wb.S.bin, wb.S.shadow, wb.S.mis, wb.ctl};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    wb = WaveBuffers{};
}

lt_status launch_path_wavefront(lt_context* ctx, WaveBuffers& wb, const PathLaunch& L, float4* accum, int count,
                                uint64_t* d_counts) {
    const uint32_t n = L.width * L.height;
    if (L.max_depth > kWaveMaxDepth) return lt_fail(ctx, LT_ERR_UNSUPPORTED, "wavefront max_depth above 32");
    lt_status st = wave_alloc(ctx, wb, n);
    if (st != LT_OK) return st;
    const size_t tblocks = (size_t)ctx->num_cus * ctx->tuning.path_blocks_per_cu;
    st = ensure_spill(ctx, tblocks * kTraceBlock);
    if (st != LT_OK) return st;
    const size_t sblocks = (size_t)ctx->num_cus * 8;
    const uint32_t sort = ctx->tuning.wf_sort == LT_WF_SORT_BSDF ? 1u : 0u;
    SceneView sv = make_scene_view(ctx);
    auto* cnt = (unsigned long long*)d_counts;
    if (count) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, cnt, (uint32_t)kCountSlots);
    phase_mark(ctx);
    const uint32_t depth_n = L.max_depth > 0 ? (uint32_t)L.max_depth : 1u;
    for (uint32_t f = 0; f < L.spp; f++) {
        hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(256), 0, ctx->stream, wb.ctl);
        hipLaunchKernelGGL(k_wf_generate, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, L, f, wb.S, wb.ctl);
        // stage 0: the camera rays; stage d + 1: the continuation rays of
        // depth d + 1 with the shadow and MIS rays of depth d.  The shade
        // stage of the last depth only adds emission (path.rgen:60-62).
        for (uint32_t d = 0; d < depth_n; d++) {
            const uint32_t with_sec = d > 0 ? 1u : 0u;
            if (count)
                hipLaunchKernelGGL((k_wf_trace<true>), dim3(tblocks), dim3(kTraceBlock), 0, ctx->stream, sv, wb.S,
                                   wb.ctl, d, d, 1u, d - with_sec, with_sec, ctx->spill, cnt);
            else
                hipLaunchKernelGGL((k_wf_trace<false>), dim3(tblocks), dim3(kTraceBlock), 0, ctx->stream, sv, wb.S,
                                   wb.ctl, d, d, 1u, d - with_sec, with_sec, ctx->spill, cnt);
            if (sort) {
                hipLaunchKernelGGL(k_wf_bin_count, dim3(sblocks), dim3(256), 0, ctx->stream, sv, wb.S, wb.ctl, d);
                hipLaunchKernelGGL(k_wf_bin_scatter, dim3(sblocks), dim3(256), 0, ctx->stream, wb.S, wb.ctl, d);
            }
            if (count)
                hipLaunchKernelGGL((k_wf_shade<true>), dim3(sblocks), dim3(256), 0, ctx->stream, sv, L, f, wb.S,
                                   wb.ctl, d, sort, cnt);
            else
                hipLaunchKernelGGL((k_wf_shade<false>), dim3(sblocks), dim3(256), 0, ctx->stream, sv, L, f, wb.S,
                                   wb.ctl, d, sort, cnt);
        }
        hipLaunchKernelGGL(k_wf_accumulate, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, L, f, wb.S, accum);
    }
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}
