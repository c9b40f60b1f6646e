// trace.hip — raw closest-hit / any-hit kernels behind lt_trace_closest,
// lt_trace_any and lt_trace_count (persistent waves pulling 64 rays at a time).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "context.h"
#include "dev_trace.h"

using namespace ltd;

namespace {

// spill_stride: global stack entries per thread beyond the kLdsStack LDS ones
template <bool ANY, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_trace(SceneView sv, const lt_ray* __restrict__ rays, uint32_t n,
                                                       void* __restrict__ out, uint32_t* __restrict__ work,
                                                       uint32_t* __restrict__ spill, uint32_t spill_stride,
                                                       unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * spill_stride};
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t n_nodes = 0, n_tris = 0;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(work, 64u);
        base = __shfl(base, 0);
        if (base >= n) break;
        const uint32_t i = base + lane;
        if (i < n) {
            const lt_ray r = rays[i];
            HitRec h;
            const bool hit = trace_bvh<ANY, COUNT>(sv, ld3(r.org), ld3(r.dir), r.tmin, r.tmax, h, stk, n_nodes,
                                                   n_tris);
            if (ANY) {
                ((uint8_t*)out)[i] = hit ? 1 : 0;
            } else {
                lt_hit o;
                o.t = hit ? h.t : 0.0f;
                o.prim = h.prim;
                o.inst = h.inst;
                o.u = hit ? h.u : 0.0f;
                o.v = hit ? h.v : 0.0f;
                ((lt_hit*)out)[i] = o;
            }
        }
    }
    if (COUNT) {
        unsigned long long a = n_nodes, b = n_tris;
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o);
            b += __shfl_xor(b, o);
        }
        if (lane == 0) {
            atomicAdd(&counts[0], a);
            atomicAdd(&counts[1], b);
        }
    }
}

constexpr uint32_t kSpillStride = ltb::kStackTotal - ltb::kLdsStack;

}  // namespace

lt_status launch_trace(lt_context* ctx, const lt_ray* d_rays, size_t n, void* d_out, int any_hit, int count,
                       uint64_t* d_counts) {
    if (n == 0) return LT_OK;
    if (n > 0xFFFFFFFFull - 64) return lt_fail(ctx, LT_ERR_INVALID_ARG, "too many rays in one call");
    const size_t blocks = (size_t)ctx->num_cus * ctx->tuning.trace_blocks_per_cu;
    lt_status st = ensure_spill(ctx, blocks * kTraceBlock);
    if (st != LT_OK) return st;
    SceneView sv = make_scene_view(ctx);
    LT_HIP(ctx, hipMemsetAsync(ctx->counters, 0, 4, ctx->stream));
    if (count) LT_HIP(ctx, hipMemsetAsync(d_counts, 0, 16, ctx->stream));
    LT_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    auto* cnt = (unsigned long long*)d_counts;
    if (any_hit) {
        if (count)
            hipLaunchKernelGGL((k_trace<true, true>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, d_rays,
                               (uint32_t)n, d_out, ctx->counters, ctx->spill, kSpillStride, cnt);
        else
            hipLaunchKernelGGL((k_trace<true, false>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, d_rays,
                               (uint32_t)n, d_out, ctx->counters, ctx->spill, kSpillStride, cnt);
    } else {
        if (count)
            hipLaunchKernelGGL((k_trace<false, true>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, d_rays,
                               (uint32_t)n, d_out, ctx->counters, ctx->spill, kSpillStride, cnt);
        else
            hipLaunchKernelGGL((k_trace<false, false>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, d_rays,
                               (uint32_t)n, d_out, ctx->counters, ctx->spill, kSpillStride, cnt);
    }
    LT_HIP(ctx, hipGetLastError());
    LT_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    return LT_OK;
}

// ---- closest hits without allocation (lt_trace_closest_noalloc) -----------
//
// The persistent trace grid's work counter and its stack spill area live in a
// file-scope device array instead of the context's hipMalloc'ed buffers.  A
// BVH8 stack holds at most one entry per tree level below the root, so a BVH
// of depth D never needs more than D - 1 entries: with kNoAllocSpill global
// entries per thread after the kLdsStack LDS ones, BVHs up to depth
// kLdsStack + kNoAllocSpill + 1 are served.  Every call on a device shares that
// device's array: calls that overlap on one GPU (two streams or contexts) race.

namespace {

constexpr uint32_t kNoAllocBlocks = 1024;  // largest grid, in kTraceBlock workgroups
constexpr uint32_t kNoAllocSpill = 16;     // global stack entries per thread
constexpr uint32_t kNoAllocCounterWords = 64;
constexpr size_t kNoAllocWords = kNoAllocCounterWords + (size_t)kNoAllocBlocks * kTraceBlock * kNoAllocSpill;

// 8 MiB + 256 B per device, held until the process ends
__device__ __align__(256) uint32_t scratch[kNoAllocWords];

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Device address of `scratch` on the context's device, looked up once per
// device with that device current (a host-side use of the array's name is not
// a device address).
lt_status noalloc_scratch_init(lt_context* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return lt_fail(ctx, LT_ERR_UNSUPPORTED, "device index");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        LT_HIP(ctx, hipSetDevice(ctx->device));
        void* p = nullptr;
        LT_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[ctx->device] = p;
    }
    ctx->noalloc_scratch = (uint32_t*)g_scratch_ptr[ctx->device];
    return LT_OK;
}

extern "C" lt_status lt_trace_closest_noalloc(lt_context* ctx, const lt_ray* rays, size_t n, lt_hit* hits) {
    if (!ctx || (n && (!rays || !hits))) return LT_ERR_INVALID_ARG;
    if (!ctx->bvh.built) return lt_fail(ctx, LT_ERR_NOT_READY, "trace before lt_accel_build");
    if (!ctx->noalloc_scratch) return lt_fail(ctx, LT_ERR_NOT_READY, "no scratch address for this device");
    if (n == 0) return LT_OK;
    if (n > 0xFFFFFFFFull - 64) return lt_fail(ctx, LT_ERR_INVALID_ARG, "too many rays in one call");
    if (ctx->bvh.depth > (uint32_t)ltb::kLdsStack + kNoAllocSpill + 1u)
        return lt_fail(ctx, LT_ERR_UNSUPPORTED,
                       "lt_trace_closest_noalloc: BVH depth " + std::to_string(ctx->bvh.depth) +
                           " exceeds the " + std::to_string(ltb::kLdsStack + kNoAllocSpill + 1) +
                           " levels its fixed stack holds");
    LT_HIP(ctx, hipSetDevice(ctx->device));
    const size_t blocks = std::min<size_t>((size_t)ctx->num_cus * ctx->tuning.trace_blocks_per_cu, kNoAllocBlocks);
    uint32_t* work = ctx->noalloc_scratch;
    uint32_t* spill = work + kNoAllocCounterWords;
    SceneView sv = make_scene_view(ctx);
    // the work counter starts at zero (a fill kernel, also under graph capture)
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, work, kNoAllocCounterWords, 0u);
    hipLaunchKernelGGL((k_trace<false, false>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, rays, (uint32_t)n,
                       (void*)hits, work, spill, kNoAllocSpill, (unsigned long long*)nullptr);
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}
