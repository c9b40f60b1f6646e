// path.hip — the Path integrator on gfx950 (megakernel form).
//
// One thread per pixel, persistent waves pulling 8x8 pixel tiles, all frames
// of a launch accumulated in registers with the reference's per-frame running
// mean (path.rgen:106-112) so the result equals frame-by-frame launches.
// Per-sample logic: src/shaders/integrators/path/path.rgen:26-104 and
// src/shaders/integrators/pt_commons.glsl:3-47.
#include <hip/hip_runtime.h>

#include "context.h"
#include "dev_scene.h"
#include "dev_atmosphere.h"
#include "dev_nee.h"
#include "dev_trace.h"
#include "path.h"

using namespace ltd;

namespace {

// The state of one path between bounces (path.rgen:26-100).  The hit
// payload is not part of it: a bounce's closest-hit trace either fills every
// field or misses and ends the path, so only the NEE MIS ray of the same
// bounce sees a stale payload (ray.rmiss:13), which nee_sample records.
struct PathState {
    RNG seed;
    f3 origin, direction, throughput, col;
    uint32_t pix;
    int depth;
    bool last_specular;
};

LT_DEV void path_start(const PathLaunch& L, uint32_t x, uint32_t y, uint32_t frame_word, PathState& p) {
    p.seed = {x, y, frame_word, 0u};
    const float r0 = rnd(p.seed);
    const float r1 = rnd(p.seed);
    const float jx = r0 - 0.5f, jy = r1 - 0.5f;
    const float ux = ((float)x + 0.5f + jx) / (float)L.width, uy = ((float)y + 0.5f + jy) / (float)L.height;
    const float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    const f4 o4 = mul4(L.inv_view, f4{0.0f, 0.0f, 0.0f, 1.0f});
    const f4 tg = mul4(L.inv_projection, f4{dx, dy, 1.0f, 1.0f});
    const f3 tn = normalize(mk3(tg.x, tg.y, tg.z));
    const f4 d4 = mul4(L.inv_view, f4{tn.x, tn.y, tn.z, 0.0f});
    p.origin = {o4.x, o4.y, o4.z};
    p.direction = {d4.x, d4.y, d4.z};
    p.col = mk3(0.0f);
    p.throughput = mk3(1.0f);
    p.last_specular = false;
    p.depth = 0;
}

// one iteration of path.rgen:46-100; returns false when the path has ended.
// The bounce draws its random numbers in the reference's order (NEE light
// sample, NEE BSDF sample, continuation, Russian roulette), then traces its
// NEE rays, whose terms use the throughput before the continuation.
template <bool COUNT>
LT_DEV bool path_bounce(const SceneView& sv, const PathLaunch& L, PathState& p, const TraceStack& stk, Ctr& c) {
    NeeRays nee;
    bool alive = true, has_nee = false;
    {
        Payload hp;
        trace_closest<COUNT>(sv, p.origin, p.direction, stk, hp, c);
        if (!hp.valid) {
            if (p.depth > 0 || L.direct_lighting == 1)
                p.col += p.throughput * sky_radiance(sv, ld3(L.sky_col), p.origin, p.direction, PATH_TMAX);
            return false;
        }
        const Mat hit_mat = load_material(sv, hp.material_idx, hp.uv);
        if ((p.depth == 0 && L.direct_lighting == 1) || p.last_specular) p.col += p.throughput * hit_mat.emissive;
        if (p.depth >= L.max_depth - 1) return false;
        const f3 wo = -p.direction;
        f3 n_s = hp.n_s;
        bool side = true;
        f3 n_g = hp.n_g;
        if (dot(hp.n_g, wo) < 0.0f) n_g = -n_g;
        if (dot(n_g, hp.n_s) < 0.0f) {
            n_s = -n_s;
            side = false;
        }
        p.origin = offset_ray(hp.pos, n_g);
        p.last_specular = is_specular(hit_mat);
        if (!p.last_specular && (p.depth > 0 || L.direct_lighting == 1)) {
            nee_sample(sv, p.seed, hit_mat, hp, side, n_s, wo, p.throughput, L.nee_pick, nee);
            has_nee = true;
        }
        float pdf, cos_theta;
        const float s0 = rnd(p.seed), s1 = rnd(p.seed), s2 = rnd(p.seed);
        const f3 f = sample_bsdf(n_s, wo, hit_mat, 1, side, p.direction, pdf, cos_theta, s0, s1, s2);
        if (pdf == 0.0f) {
            alive = false;
        } else {
            const f3 thr = p.throughput * (f * fabsf(cos_theta) / pdf);
            float rr_scale = 1.0f;
            if (hit_mat.bsdf_props & LT_BSDF_FLAG_TRANSMISSION) rr_scale *= side ? 1.0f / hit_mat.ior : hit_mat.ior;
            if (p.depth > RR_MIN_DEPTH) {
                const float rr_prob = fminf(0.95f, luminance(thr) * rr_scale);
                if (rr_prob == 0.0f || rr_prob < rnd(p.seed)) alive = false;
                else p.throughput = thr / rr_prob;
            } else {
                p.throughput = thr;
            }
        }
    }
    if (has_nee) p.col += nee_trace<COUNT>(sv, nee, stk, c);
    p.depth++;
    return alive;
}

// One path of pixel (x, y) (path.rgen:26-104); false when the NaN guard
// drops the sample.
template <bool COUNT>
LT_DEV bool path_sample(const SceneView& sv, const PathLaunch& L, uint32_t x, uint32_t y, uint32_t frame_word,
                        const TraceStack& stk, f3& col, Ctr& c) {
    PathState p;
    path_start(L, x, y, frame_word, p);
    while (path_bounce<COUNT>(sv, L, p, stk, c)) {
    }
    col = p.col;
    return !isnan(luminance(col));
}

template <bool COUNT>
// Four waves per SIMD (128 VGPRs; the path state across a traversal spills to
// scratch, which stays in cache): 6.5 ms per 1080p terrain frame against
// 9.0 ms at the unconstrained 203 VGPRs / 2 waves.  A cap of five waves
// (96 VGPRs, 416 B of scratch per lane) produced wrong images with no cause
// found in this source, so it is not used.
#ifndef LT_PATH_MEGA_WAVES
#define LT_PATH_MEGA_WAVES 4
#endif
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_PATH_MEGA_WAVES))) void k_path_mega(SceneView sv, PathLaunch L, float4* __restrict__ accum,
                                                           uint32_t* __restrict__ work, uint32_t* __restrict__ spill,
                                                           unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tiles_x = (L.width + 7) / 8, tiles_y = (L.height + 7) / 8;
    const uint32_t n_tiles = tiles_x * tiles_y;
    Ctr c{0, 0, 0, 0};
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(work, 1u);
        t = __shfl(t, 0);
        if (t >= n_tiles) break;
        const uint32_t x = (t % tiles_x) * 8 + (lane & 7), y = (t / tiles_x) * 8 + (lane >> 3);
        if (x >= L.width || y >= L.height) continue;
        const size_t pix = (size_t)y * L.width + x;
        float4 acc = accum[pix];
        const uint32_t frame0 = L.frame_dev ? *L.frame_dev : L.frame_start;
        for (uint32_t f = 0; f < L.spp; f++) {
            const uint32_t frame_num = frame0 + f;
            f3 col;
            if (!path_sample<COUNT>(sv, L, x, y, frame_num ^ L.seed, stk, col, c)) continue;  // path.rgen:102-104
            if (frame_num > 0) {
                const float w = 1.0f / (float)(frame_num + 1);
                acc.x = acc.x * (1.0f - w) + col.x * w;
                acc.y = acc.y * (1.0f - w) + col.y * w;
                acc.z = acc.z * (1.0f - w) + col.z * w;
            } else {
                acc.x = col.x;
                acc.y = col.y;
                acc.z = col.z;
            }
            acc.w = 1.0f;
        }
        accum[pix] = acc;
    }
    if (COUNT) flush_counts(c, counts);
}

// Path regeneration (the persistent form): each lane runs one bounce of its
// path per iteration and, once the path ends, takes the next pixel of the
// frame (a ballot over the idle lanes and one atomic per refill), so paths
// of uneven length (Russian roulette, escapes) never leave lanes idle until
// the longest path of a tile ends.  Per path the arithmetic, the RNG stream
// and the result are those of path_sample.
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(LT_PATH_MEGA_WAVES))) void k_path_regen(
    SceneView sv, PathLaunch L, float4* __restrict__ accum, uint32_t* __restrict__ work, uint32_t* __restrict__ spill,
    unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tiles_x = (L.width + 7) / 8, tiles_y = (L.height + 7) / 8;
    const uint32_t n_items = tiles_x * tiles_y * 64u;  // pixels in 8x8-tile order
    const uint32_t frame_num = L.frame_dev ? *L.frame_dev : L.frame_start;
    const uint32_t frame_word = frame_num ^ L.seed;
    Ctr c{0, 0, 0, 0};
    PathState p;
    // XCD bands: workgroups are dealt to the 8 XCDs round-robin (blockIdx mod
    // 8); with bands on, the pixels are cut into 8 contiguous bands of tiles
    // with a counter each, and a wave takes pixels from its XCD's band, so the
    // primary rays of one L2 share a compact image region; a wave whose band
    // is used up helps the next band.  Off: one counter over every pixel.
    const uint32_t n_bands = (L.xcd_bands && (gridDim.x & 7u) == 0u) ? 8u : 1u;
    const uint32_t band_len = (n_items + n_bands - 1u) / n_bands;
    uint32_t band = n_bands == 8u ? (blockIdx.x & 7u) : 0u, bands_done = 0;  // wave-uniform
    bool has = false, exhausted = false;  // exhausted is wave-uniform
    for (;;) {
        if (!exhausted) {
            const unsigned long long idle = __ballot(!has);
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            if (n_idle >= L.regen_min || n_idle == 64u) {
                const int leader = __ffsll((long long)idle) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(work + band, n_idle);
                base = __shfl(base, leader);
                const uint32_t band_lo = band * band_len, band_hi = min(n_items, band_lo + band_len);
                const uint32_t first = band_lo + base;
                if (first + n_idle >= band_hi) {  // this band is used up: the next one, until all are
                    band = band + 1u == n_bands ? 0u : band + 1u;
                    if (++bands_done == n_bands) exhausted = true;
                }
                if (!has) {
                    const uint32_t t = first + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                    if (t < band_hi) {
                        const uint32_t tile = t >> 6, w = t & 63u;
                        const uint32_t x = (tile % tiles_x) * 8 + (w & 7u), y = (tile / tiles_x) * 8 + (w >> 3);
                        if (x < L.width && y < L.height) {
                            path_start(L, x, y, frame_word, p);
                            p.pix = y * L.width + x;
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
        if (has && !path_bounce<COUNT>(sv, L, p, stk, c)) {
            has = false;
            if (!isnan(luminance(p.col))) {  // path.rgen:102-112
                float4 acc = accum[p.pix];
                if (frame_num > 0) {
                    const float w = 1.0f / (float)(frame_num + 1);
                    acc.x = acc.x * (1.0f - w) + p.col.x * w;
                    acc.y = acc.y * (1.0f - w) + p.col.y * w;
                    acc.z = acc.z * (1.0f - w) + p.col.z * w;
                } else {
                    acc.x = p.col.x;
                    acc.y = p.col.y;
                    acc.z = p.col.z;
                }
                acc.w = 1.0f;
                accum[p.pix] = acc;
            }
        }
    }
    if (COUNT) flush_counts(c, counts);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

lt_status launch_path_megakernel(lt_context* ctx, const PathLaunch& L, float4* accum, int count,
                                 uint64_t* d_counts) {
    const size_t blocks = (size_t)ctx->num_cus * ctx->tuning.path_blocks_per_cu;
    lt_status st = ensure_spill(ctx, blocks * kTraceBlock);
    if (st != LT_OK) return st;
    SceneView sv = make_scene_view(ctx);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->counters, 4u);
    if (count) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)d_counts, 2u * kCountSlots);
    auto* cnt = (unsigned long long*)d_counts;
    phase_mark(ctx);
    if (count)
        hipLaunchKernelGGL((k_path_mega<true>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, L, accum,
                           ctx->counters, ctx->spill, cnt);
    else
        hipLaunchKernelGGL((k_path_mega<false>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, L, accum,
                           ctx->counters, ctx->spill, cnt);
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}

lt_status launch_path_persistent(lt_context* ctx, const PathLaunch& L0, float4* accum, int count,
                                 uint64_t* d_counts) {
    const size_t blocks = (size_t)ctx->num_cus * ctx->tuning.path_blocks_per_cu;
    lt_status st = ensure_spill(ctx, blocks * kTraceBlock);
    if (st != LT_OK) return st;
    if (L0.frame_dev && L0.spp != 1) return lt_fail(ctx, LT_ERR_INVALID_ARG, "a captured frame renders one frame");
    SceneView sv = make_scene_view(ctx);
    auto* cnt = (unsigned long long*)d_counts;
    if (count) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)d_counts, 2u * kCountSlots);
    phase_mark(ctx);
    for (uint32_t f = 0; f < L0.spp; f++) {  // the running mean needs the frames in order
        PathLaunch L = L0;
        L.frame_start = L0.frame_start + f;
        L.spp = 1;
        L.regen_min = ctx->tuning.regen_min ? ctx->tuning.regen_min : 4u;
        L.xcd_bands = ctx->tuning.xcd_bands == LT_XCD_BANDS_ON ? 1u : 0u;
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->counters, 8u);
        if (count)
            hipLaunchKernelGGL((k_path_regen<true>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, L, accum,
                               ctx->counters, ctx->spill, cnt);
        else
            hipLaunchKernelGGL((k_path_regen<false>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, L, accum,
                               ctx->counters, ctx->spill, cnt);
    }
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}
