// sppm.hip — the SPPM integrator on gfx950 (stochastic progressive photon
// mapping, Hachisuka & Jensen 2009, as the reference's shaders run it).
//
// Per frame (src/RayTracer/SPPM.cpp:75-176):
//   k_sppm_clear   photon table and grid bounds reset; the per-pixel state
//                  too at frame 0 (SPPM.cpp:141);
//   k_sppm_eye     one eye path per pixel through its centre: emission, the
//                  Path integrator's next-event estimate until a diffuse
//                  surface is recorded (sppm_eye.rgen:42-128);
//   k_sppm_bounds  min(p - r), max(p + r) and max(r) over every pixel
//                  (max.comp, min.comp and their reduce passes);
//   k_sppm_light   one light path per pixel slot; each vertex after the first
//                  becomes a photon of the hash bucket of its grid cell
//                  (sppm_light.rgen:43-114), the grid resolution computed from
//                  the bounds as calc_bounds.comp does;
//   k_sppm_gather  per pixel, the photons of the buckets of the cells in
//                  p +- r, then the progressive radius update and the running
//                  mean (gather.comp:42-87, composite.comp:31-62).
// The reference's photon table keeps one photon per bucket, written by
// racing threads, and counts the insertions; here the bucket's photon is the
// one with the largest (light path, vertex) reference, so the result is
// deterministic and equals the oracle's insertion in path order.
#include <hip/hip_runtime.h>

#include "context.h"
#include "dev_nee.h"
#include "dev_scene.h"
#include "dev_trace.h"
#include "sppm.h"

using namespace ltd;

namespace {

constexpr int kSppmMaxSpan = 16;        // gather cells per axis at most
constexpr float SPPM_TMIN = 0.001f;    // sppm_eye.rgen:24, sppm_light.rgen:26
constexpr float SPPM_TMAX = 10000.0f;  // sppm_eye.rgen:25, sppm_light.rgen:27

// order-preserving float <-> u32 (unsigned comparisons order the floats)
LT_DEV uint32_t ord_enc(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
LT_DEV float ord_dec(uint32_t e) { return __uint_as_float((e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e); }

// utils.glsl:62-65 over 10 * size buckets (cell coordinates as two's
// complement words)
LT_DEV uint32_t sppm_hash_u(uint32_t x, uint32_t y, uint32_t z, uint32_t size) {
    const uint32_t h = (x * 73856093u) ^ (y * 19349663u) ^ (z * 83492791u);
    return h % (10u * size);
}
LT_DEV uint32_t sppm_hash(int x, int y, int z, uint32_t size) {
    return sppm_hash_u((uint32_t)x, (uint32_t)y, (uint32_t)z, size);
}

struct SppmGrid {
    f3 mn, mx;
    int res[3];
};

// calc_bounds.comp:16-27 from the reduced bounds
LT_DEV SppmGrid sppm_grid(const SppmBounds& b) {
    SppmGrid g;
    g.mn = {ord_dec(b.mn[0]), ord_dec(b.mn[1]), ord_dec(b.mn[2])};
    g.mx = {ord_dec(b.mx[0]), ord_dec(b.mx[1]), ord_dec(b.mx[2])};
    const float max_r = ord_dec(b.max_r);
    const f3 ext = g.mx - g.mn;
    const float max_comp = fmaxf(ext.x, fmaxf(ext.y, ext.z));
    const int base = (int)(max_comp / max_r);
    if (max_comp > 0.0f && max_r > 0.0f) {
        const float fb = (float)base;
        g.res[0] = max((int)(fb * ext.x / max_comp), 1);
        g.res[1] = max((int)(fb * ext.y / max_comp), 1);
        g.res[2] = max((int)(fb * ext.z / max_comp), 1);
    } else {
        g.res[0] = g.res[1] = g.res[2] = 0;
    }
    return g;
}

// get_grid_idx (sppm_eye.rgen:32-36; the clamp's result is discarded there)
LT_DEV void sppm_cell(const SppmGrid& g, f3 p, int o[3]) {
#pragma clang fp contract(off)
    const f3 e = g.mx - g.mn;
    const f3 q = {(p.x - g.mn.x) / e.x, (p.y - g.mn.y) / e.y, (p.z - g.mn.z) / e.z};
    o[0] = (int)(q.x * (float)g.res[0]);
    o[1] = (int)(q.y * (float)g.res[1]);
    o[2] = (int)(q.z * (float)g.res[2]);
}

// |p - a.xyz|^2 without contraction, as the oracle evaluates it
LT_DEV float dist_sqr(f3 p, float4 a) {
#pragma clang fp contract(off)
    const float dx = p.x - a.x, dy = p.y - a.y, dz = p.z - a.z;
    return dx * dx + dy * dy + dz * dz;
}

LT_DEV float ld_plane(const float* st, uint32_t n, uint32_t plane, uint32_t q) { return st[(size_t)plane * n + q]; }
LT_DEV void st_plane(float* st, uint32_t n, uint32_t plane, uint32_t q, float v) { st[(size_t)plane * n + q] = v; }

__global__ void k_sppm_fill(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_sppm_clear(uint32_t* key, uint32_t* count, size_t cells, float* state, size_t state_words,
                             int clear_state, SppmBounds* bounds) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += stride) {
        key[i] = 0u;
        count[i] = 0u;
    }
    if (clear_state)
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < state_words; i += stride) state[i] = 0.0f;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int k = 0; k < 3; k++) {
            bounds->mn[k] = 0xFFFFFFFFu;  // above every encoded float
            bounds->mx[k] = 0u;
        }
        bounds->max_r = 0u;
    }
}

// sppm_eye.rgen:42-128 for pixel (x, y), state index q = y * W + x
template <bool COUNT>
LT_DEV void eye_path(const SceneView& sv, const SppmLaunch& L, uint32_t x, uint32_t y, float* st,
                     const TraceStack& stk, Ctr& c) {
    const uint32_t n = L.n_px, q = y * L.width + x;
    RNG seed = {x, y, L.word, 0u};
    const float dx = ((float)x + 0.5f) / (float)L.width * 2.0f - 1.0f;
    const float dy = ((float)y + 0.5f) / (float)L.height * 2.0f - 1.0f;
    const f4 o4 = mul4(L.inv_view, f4{0.0f, 0.0f, 0.0f, 1.0f});
    const f4 tg = mul4(L.inv_projection, f4{dx, dy, 1.0f, 1.0f});
    const f3 tn = normalize(mk3(tg.x, tg.y, tg.z));
    const f4 d4 = mul4(L.inv_view, f4{tn.x, tn.y, tn.z, 0.0f});
    f3 origin = {o4.x, o4.y, o4.z}, direction = {d4.x, d4.y, d4.z};
    f3 thr = mk3(1.0f);
    f3 col = {ld_plane(st, n, kSpColx, q), ld_plane(st, n, kSpColy, q), ld_plane(st, n, kSpColz, q)};
    bool specular = false, recorded = false;
    Payload hp;
    for (int depth = 0;; depth++) {
        trace_closest<COUNT>(sv, origin, direction, stk, hp, c);
        if (!hp.valid) {
            col += thr * ld3(L.sky_col);
            break;
        }
        if (depth >= L.max_depth - 1) break;
        const Mat mat = load_material(sv, hp.material_idx, hp.uv);
        if (depth == 0 || (specular && !recorded)) col += thr * mat.emissive;
        const f3 wo = -direction;
        f3 n_s = hp.n_s, n_g = hp.n_g;
        bool side = true;
        if (dot(hp.n_g, wo) < 0.0f) n_g = -n_g;
        if (dot(n_g, hp.n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        const f3 pos = hp.pos;
        origin = offset_ray(pos, n_g);
        specular = is_specular(mat);
        if (!recorded && !specular) {  // sppm_eye.rgen:91-96: Path's NEE / (1 / light_triangle_count)
            NeeRays nee;
            nee_sample(sv, seed, mat, hp, side, n_s, wo, thr, LT_NEE_REFERENCE, nee);
            col += nee_trace<COUNT>(sv, nee, stk, c);
        }
        if (!recorded && (mat.bsdf_type & LT_BSDF_DIFFUSE) != 0u) {  // sppm_eye.rgen:99-110
            recorded = true;
            st_plane(st, n, kSpWox, q, wo.x);
            st_plane(st, n, kSpWoy, q, wo.y);
            st_plane(st, n, kSpWoz, q, wo.z);
            st_plane(st, n, kSpPx, q, pos.x);
            st_plane(st, n, kSpPy, q, pos.y);
            st_plane(st, n, kSpPz, q, pos.z);
            st_plane(st, n, kSpMat, q, __uint_as_float(hp.material_idx));
            st_plane(st, n, kSpU, q, hp.uv.x);
            st_plane(st, n, kSpV, q, hp.uv.y);
            st_plane(st, n, kSpThrx, q, thr.x);
            st_plane(st, n, kSpThry, q, thr.y);
            st_plane(st, n, kSpThrz, q, thr.z);
            st_plane(st, n, kSpNx, q, n_s.x);
            st_plane(st, n, kSpNy, q, n_s.y);
            st_plane(st, n, kSpNz, q, n_s.z);
            st_plane(st, n, kSpPathLen, q, __int_as_float(depth));
            st_plane(st, n, kSpSide, q, __uint_as_float(side ? 1u : 0u));
        }
        float pdf, cos_theta;
        const float r0 = rnd(seed), r1 = rnd(seed), r2 = rnd(seed);
        const f3 f = sample_bsdf(n_s, wo, mat, 1u, side, direction, pdf, cos_theta, r0, r1, r2);
        if (pdf < 1e-20f || dot(f, f) < 1e-20f) break;
        thr *= f * fabsf(cos_theta) / pdf;
        if (depth > RR_MIN_DEPTH) {  // no 0.95 cap in the eye pass (sppm_eye.rgen:118-124)
            const float rr_prob = luminance(thr);
            if (rr_prob < rnd(seed)) break;
            thr /= rr_prob;
        }
    }
    st_plane(st, n, kSpColx, q, col.x);
    st_plane(st, n, kSpColy, q, col.y);
    st_plane(st, n, kSpColz, q, col.z);
    if (L.frame_num == 0u) st_plane(st, n, kSpRadius, q, L.base_radius);
}

#ifndef LT_SPPM_EYE_WAVES
#define LT_SPPM_EYE_WAVES 4
#endif
// persistent waves pulling 8x8 pixel tiles
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_SPPM_EYE_WAVES))) void k_sppm_eye(
    SceneView sv, SppmLaunch L, float* __restrict__ st, uint32_t* __restrict__ work, uint32_t* __restrict__ spill,
    unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tiles_x = (L.width + 7) / 8, n_tiles = tiles_x * ((L.height + 7) / 8);
    Ctr c{0, 0, 0, 0};
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(work, 1u);
        t = __shfl(t, 0);
        if (t >= n_tiles) break;
        const uint32_t x = (t % tiles_x) * 8 + (lane & 7u), y = (t / tiles_x) * 8 + (lane >> 3);
        if (x < L.width && y < L.height) eye_path<COUNT>(sv, L, x, y, st, stk, c);
    }
    if (COUNT) flush_counts(c, counts);
}

// max.comp / min.comp + reduce passes: a wave reduction, a workgroup
// reduction in LDS, then one atomic per workgroup and value
__global__ __launch_bounds__(256) void k_sppm_bounds(const float* __restrict__ st, uint32_t n,
                                                     SppmBounds* __restrict__ b) {
    __shared__ float red[4][7];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, mr = -INFINITY;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        const float r = ld_plane(st, n, kSpRadius, q);
        const float p[3] = {ld_plane(st, n, kSpPx, q), ld_plane(st, n, kSpPy, q), ld_plane(st, n, kSpPz, q)};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], p[k] - r);
            hi[k] = fmaxf(hi[k], p[k] + r);
        }
        mr = fmaxf(mr, r);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], o));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o));
        }
        mr = fmaxf(mr, __shfl_xor(mr, o));
    }
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u) {
        for (int k = 0; k < 3; k++) {
            red[w][k] = lo[k];
            red[w][3 + k] = hi[k];
        }
        red[w][6] = mr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t v = 1; v < blockDim.x / 64u; v++) {
            for (int k = 0; k < 3; k++) {
                lo[k] = fminf(lo[k], red[v][k]);
                hi[k] = fmaxf(hi[k], red[v][3 + k]);
            }
            mr = fmaxf(mr, red[v][6]);
        }
        for (int k = 0; k < 3; k++) {
            atomicMin(&b->mn[k], ord_enc(lo[k]));
            atomicMax(&b->mx[k], ord_enc(hi[k]));
        }
        atomicMax(&b->max_r, ord_enc(mr));
    }
}

// sppm_light.rgen:43-114 for the light path of pixel slot q (x = q % W, y = q / W)
template <bool COUNT>
LT_DEV void light_path(const SceneView& sv, const SppmLaunch& L, const SppmGrid& g, uint32_t q,
                       SppmPhoton* __restrict__ photons, uint32_t* __restrict__ key, uint32_t* __restrict__ count,
                       const TraceStack& stk, Ctr& c) {
    RNG seed = {q % L.width, q / L.width, L.word, 0u};
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
    const float abs_cos = fabsf(dot(n, wi));
    f3 thr = (abs_cos == 0.0f || pdf_dir <= 0.0f) ? mk3(0.0f) : Le * abs_cos / (pdf_dir * pdf_pos);
    for (int depth = 0; depth < L.max_depth; depth++) {
        Payload pl;
        if (COUNT) c.closest++;
        HitRec h;
        if (!trace_bvh<false, COUNT>(sv, pos, wi, SPPM_TMIN, SPPM_TMAX, h, stk, c.nodes, c.tris)) break;
        fill_payload(sv, h.inst, h.prim, h.u, h.v, pl);
        const f3 wo = -wi;
        f3 n_s = pl.n_s, n_g = pl.n_g;
        bool side = true;
        if (dot(pl.n_g, wo) < 0.0f) n_g = -n_g;
        if (dot(n_g, pl.n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        const Mat mat = load_material(sv, pl.material_idx, pl.uv);
        if (depth > 0) {  // sppm_light.rgen:81-94: the first bounce is not stored
            int gi[3];
            sppm_cell(g, pl.pos, gi);
            const uint32_t hb = sppm_hash(gi[0], gi[1], gi[2], L.n_px);
            const uint32_t ref = q * (uint32_t)L.max_depth + (uint32_t)depth;
            SppmPhoton& ph = photons[ref];
            ph.a = make_float4(pl.pos.x, pl.pos.y, pl.pos.z, __uint_as_float((uint32_t)depth + 1u));
            ph.b = make_float4(-wi.x, -wi.y, -wi.z, thr.x);
            ph.c = make_float4(thr.y, thr.z, 0.0f, 0.0f);
            atomicMax(&key[hb], ref + 1u);
            atomicAdd(&count[hb], 1u);
        }
        pos = offset_ray(pl.pos, n_g);
        float pdf, cos_t;
        const float r0 = rnd(seed), r1 = rnd(seed), r2 = rnd(seed);
        const f3 f = sample_bsdf(n_s, wo, mat, 0u, side, wi, pdf, cos_t, r0, r1, r2);
        thr *= f * fabsf(cos_t) / pdf;
        if ((f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) || pdf == 0.0f) break;
        if (depth > RR_MIN_DEPTH) {
            const float rr_prob = fminf(0.95f, luminance(thr));
            if (rr_prob < rnd(seed)) break;
            thr /= rr_prob;
        }
    }
}

#ifndef LT_SPPM_LIGHT_WAVES
#define LT_SPPM_LIGHT_WAVES 4
#endif
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_SPPM_LIGHT_WAVES))) void k_sppm_light(
    SceneView sv, SppmLaunch L, const SppmBounds* __restrict__ bounds, SppmPhoton* __restrict__ photons,
    uint32_t* __restrict__ key, uint32_t* __restrict__ count, uint32_t* __restrict__ work,
    uint32_t* __restrict__ spill, unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
    const uint32_t lane = threadIdx.x & 63u;
    const SppmGrid g = sppm_grid(*bounds);
    const uint32_t chunks = (L.n_px + 63u) / 64u;
    Ctr c{0, 0, 0, 0};
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(work, 1u);
        t = __shfl(t, 0);
        if (t >= chunks) break;
        const uint32_t q = t * 64u + lane;
        if (q < L.n_px) light_path<COUNT>(sv, L, g, q, photons, key, count, stk, c);
    }
    if (COUNT) flush_counts(c, counts);
}

// gather.comp:42-87 and composite.comp:31-62 for pixel q.  COUNT: photons
// tested (16 B position each), gathered (32 B more each) and the pass's
// algorithmic bytes (84 B of pixel state and image per pixel, 4 B per cell
// count, 4 B per occupied cell's key, 52 B of recorded point on a pixel's
// first gathered photon) into counts[4..6].
template <bool COUNT>
__global__ __launch_bounds__(256) void k_sppm_gather(SceneView sv, SppmLaunch L, const SppmBounds* __restrict__ bounds,
                                                     const SppmPhoton* __restrict__ photons,
                                                     const uint32_t* __restrict__ key,
                                                     const uint32_t* __restrict__ count, float* __restrict__ st,
                                                     float4* __restrict__ accum, unsigned long long* __restrict__ counts) {
    const uint32_t n = L.n_px;
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long n_tested = 0, n_gathered = 0, n_bytes = 0;
    if (q < n) {
    if (COUNT) n_bytes += 84;
    const SppmGrid g = sppm_grid(*bounds);
    const f3 p = {ld_plane(st, n, kSpPx, q), ld_plane(st, n, kSpPy, q), ld_plane(st, n, kSpPz, q)};
    const float radius = ld_plane(st, n, kSpRadius, q);
    const int path_len = __float_as_int(ld_plane(st, n, kSpPathLen, q));
    float r = radius;
    if (r < EPS) r = EPS;
    const float r_sqr = r * r;
    int lo[3], hi[3];
    sppm_cell(g, p - mk3(r), lo);
    sppm_cell(g, p + mk3(r), hi);
    // Cells are the largest radius wide, so p +- r spans at most 3 per axis;
    // only a gather radius far above the cell size (the EPS floor against
    // radii below 1e-3) spans more.  Such a box is cut to kSppmMaxSpan cells
    // per axis so that the loop stays bounded.
    int span[3];
    for (int k = 0; k < 3; k++) {
        const long long d = (long long)hi[k] - (long long)lo[k];
        span[k] = d < 0 ? -1 : (d >= kSppmMaxSpan ? kSppmMaxSpan - 1 : (int)d);
    }
    f3 phi = mk3(0.0f);
    int M = 0;
    bool have_mat = false;
    Mat mat;
    f3 wo, n_s, thr;
    bool side = false;
    for (int i = 0; i <= span[0]; i++)
        for (int j = 0; j <= span[1]; j++)
            for (int k = 0; k <= span[2]; k++) {
                const uint32_t h = sppm_hash_u((uint32_t)lo[0] + (uint32_t)i, (uint32_t)lo[1] + (uint32_t)j,
                                               (uint32_t)lo[2] + (uint32_t)k, n);
                const uint32_t cnt = count[h];
                if (COUNT) n_bytes += 4;
                if (cnt == 0u) continue;
                if (COUNT) {
                    n_tested++;
                    n_bytes += 4 + 16;
                }
                const SppmPhoton& ph = photons[key[h] - 1u];
                const float4 a = ph.a;
                if (dist_sqr(p, a) > r_sqr) continue;
                if ((uint32_t)L.max_depth < __float_as_uint(a.w) + (uint32_t)path_len + 1u) continue;
                if (COUNT) {
                    n_gathered++;
                    n_bytes += 32 + (have_mat ? 0 : 52);
                }
                if (!have_mat) {  // the pixel's record, loaded at its first photon
                    have_mat = true;
                    mat = load_material(sv, __float_as_uint(ld_plane(st, n, kSpMat, q)),
                                        f2{ld_plane(st, n, kSpU, q), ld_plane(st, n, kSpV, q)});
                    wo = {ld_plane(st, n, kSpWox, q), ld_plane(st, n, kSpWoy, q), ld_plane(st, n, kSpWoz, q)};
                    n_s = {ld_plane(st, n, kSpNx, q), ld_plane(st, n, kSpNy, q), ld_plane(st, n, kSpNz, q)};
                    thr = {ld_plane(st, n, kSpThrx, q), ld_plane(st, n, kSpThry, q), ld_plane(st, n, kSpThrz, q)};
                    side = __float_as_uint(ld_plane(st, n, kSpSide, q)) == 1u;
                }
                const float4 b = ph.b, cc = ph.c;
                float pf, pr;
                const f3 f = eval_bsdf(n_s, wo, mat, side, mk3(b.x, b.y, b.z), pf, pr, false);
                phi += (float)cnt * mk3(b.w, cc.x, cc.y) * f * thr;
                M += 1;
            }
    f3 col = {ld_plane(st, n, kSpColx, q), ld_plane(st, n, kSpColy, q), ld_plane(st, n, kSpColz, q)};
    if (M > 0) {  // composite.comp:40-51 (tau restarts at zero every frame)
        const float gamma = 2.0f / 3.0f;
        const float N = ld_plane(st, n, kSpN, q);
        const float n_new = N + gamma * (float)M;
        const float r_new = radius * sqrtf(n_new / (N + (float)M));
        const f3 tau = phi * r_new * r_new / (radius * radius);
        st_plane(st, n, kSpN, q, n_new);
        st_plane(st, n, kSpRadius, q, r_new);
        col += tau / ((float)n * PI * r_new * r_new);
    }
    st_plane(st, n, kSpColx, q, 0.0f);
    st_plane(st, n, kSpColy, q, 0.0f);
    st_plane(st, n, kSpColz, q, 0.0f);
    float4 acc = accum[q];
    if (L.frame_num > 0u) {
        const float w = 1.0f / (float)(L.frame_num + 1u);
        acc.x = acc.x * (1.0f - w) + col.x * w;
        acc.y = acc.y * (1.0f - w) + col.y * w;
        acc.z = acc.z * (1.0f - w) + col.z * w;
    } else {
        acc.x = col.x;
        acc.y = col.y;
        acc.z = col.z;
    }
    acc.w = 1.0f;
    accum[q] = acc;
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) {
            n_tested += __shfl_xor(n_tested, o);
            n_gathered += __shfl_xor(n_gathered, o);
            n_bytes += __shfl_xor(n_bytes, o);
        }
        if ((threadIdx.x & 63u) == 0u) {
            atomicAdd(&counts[4], n_tested);
            atomicAdd(&counts[5], n_gathered);
            atomicAdd(&counts[6], n_bytes);
        }
    }
}

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

uint32_t sppm_frame_word(uint32_t seed, uint32_t frame_num) {
    uint32_t v[4] = {seed, frame_num, 0x5A17F00Du, 0x3C6EF372u};
    pcg4d_host(v);
    return frame_num ^ v[0];
}

lt_status sppm_alloc(lt_context* ctx, SppmBuffers& b, uint32_t width, uint32_t height, int32_t max_depth) {
    b.n_px = width * height;
    b.max_depth = max_depth;
    const size_t N = b.n_px, cells = 10 * N;
    LT_HIP(ctx, hipMalloc(&b.state, N * kSpPlanes * sizeof(float)));
    LT_HIP(ctx, hipMalloc(&b.photons, N * (size_t)max_depth * sizeof(SppmPhoton)));
    LT_HIP(ctx, hipMalloc(&b.key, cells * sizeof(uint32_t)));
    LT_HIP(ctx, hipMalloc(&b.count, cells * sizeof(uint32_t)));
    LT_HIP(ctx, hipMalloc(&b.bounds, sizeof(SppmBounds)));
    return LT_OK;
}

void sppm_free(SppmBuffers& b) {
    void* ptrs[] = {b.state, b.photons, b.key, b.count, b.bounds};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    b = SppmBuffers{};
}

lt_status launch_sppm(lt_context* ctx, SppmBuffers& b, const SppmLaunch& L0, uint32_t frame_num, uint32_t frames,
                      uint32_t seed, float4* accum, int count, uint64_t* d_counts) {
    if (L0.max_depth != b.max_depth || L0.n_px != b.n_px)
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "SPPM launch does not match its buffers");
    const uint32_t fixed = ctx->tuning.bdpt_blocks_per_cu;
    const uint32_t eye_blocks = ctx->num_cus * resident_blocks(count ? reinterpret_cast<const void*>(&k_sppm_eye<true>)
                                                                     : reinterpret_cast<const void*>(&k_sppm_eye<false>),
                                                               fixed);
    const uint32_t light_blocks =
        ctx->num_cus * resident_blocks(count ? reinterpret_cast<const void*>(&k_sppm_light<true>)
                                             : reinterpret_cast<const void*>(&k_sppm_light<false>),
                                       fixed);
    lt_status st = ensure_spill(ctx, (size_t)std::max(light_blocks, eye_blocks) * kTraceBlock);
    if (st != LT_OK) return st;
    SceneView sv = make_scene_view(ctx);
    auto* cnt = (unsigned long long*)d_counts;
    if (count) hipLaunchKernelGGL(k_sppm_fill, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)d_counts, 2u * kCountSlots);
    const size_t N = b.n_px, cells = 10 * N;
    const unsigned clear_blocks = (unsigned)std::min<size_t>((cells + 255) / 256, 8192);
    const unsigned px_blocks = (unsigned)((N + 255) / 256);
    for (uint32_t f = 0; f < frames; f++) {
        SppmLaunch L = L0;
        L.frame_num = frame_num + f;
        L.word = sppm_frame_word(seed, L.frame_num);
        if (f == 0) phase_mark(ctx);
        hipLaunchKernelGGL(k_sppm_clear, dim3(clear_blocks), dim3(256), 0, ctx->stream, b.key, b.count, cells,
                           b.state, N * kSpPlanes, L.frame_num == 0u ? 1 : 0, b.bounds);
        hipLaunchKernelGGL(k_sppm_fill, dim3(1), dim3(64), 0, ctx->stream, ctx->counters, 2u);
        if (count)
            hipLaunchKernelGGL((k_sppm_eye<true>), dim3(eye_blocks), dim3(kTraceBlock), 0, ctx->stream, sv, L,
                               b.state, ctx->counters, ctx->spill, cnt);
        else
            hipLaunchKernelGGL((k_sppm_eye<false>), dim3(eye_blocks), dim3(kTraceBlock), 0, ctx->stream, sv, L,
                               b.state, ctx->counters, ctx->spill, cnt);
        if (f == 0) phase_mark(ctx);
        hipLaunchKernelGGL(k_sppm_bounds, dim3((unsigned)std::min<size_t>((N + 255) / 256, 4u * ctx->num_cus)),
                           dim3(256), 0, ctx->stream, b.state, (uint32_t)N, b.bounds);
        if (f == 0) phase_mark(ctx);
        if (count)
            hipLaunchKernelGGL((k_sppm_light<true>), dim3(light_blocks), dim3(kTraceBlock), 0, ctx->stream, sv, L,
                               b.bounds, b.photons, b.key, b.count, ctx->counters + 1, ctx->spill, cnt);
        else
            hipLaunchKernelGGL((k_sppm_light<false>), dim3(light_blocks), dim3(kTraceBlock), 0, ctx->stream, sv, L,
                               b.bounds, b.photons, b.key, b.count, ctx->counters + 1, ctx->spill, cnt);
        if (f == 0) phase_mark(ctx);
        if (count)
            hipLaunchKernelGGL(k_sppm_gather<true>, dim3(px_blocks), dim3(256), 0, ctx->stream, sv, L, b.bounds,
                               b.photons, b.key, b.count, b.state, accum, cnt);
        else
            hipLaunchKernelGGL(k_sppm_gather<false>, dim3(px_blocks), dim3(256), 0, ctx->stream, sv, L, b.bounds,
                               b.photons, b.key, b.count, b.state, accum, cnt);
    }
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}
