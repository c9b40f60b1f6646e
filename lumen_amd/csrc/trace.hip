// trace.hip — raw closest-hit / any-hit kernels behind lt_trace_closest,
// lt_trace_any and lt_trace_count (persistent waves pulling 64 rays at a time).
#include <hip/hip_runtime.h>

#include "context.h"
#include "dev_trace.h"

using namespace ltd;

namespace {

template <bool ANY, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_trace(SceneView sv, const lt_ray* __restrict__ rays, uint32_t n,
                                                       void* __restrict__ out, uint32_t* __restrict__ work,
                                                       uint32_t* __restrict__ spill,
                                                       unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_stack[ltb::kLdsWords * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack stk{s_stack + threadIdx.x, spill + (size_t)gtid * (ltb::kStackTotal - ltb::kLdsStack)};
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
                               (uint32_t)n, d_out, ctx->counters, ctx->spill, cnt);
        else
            hipLaunchKernelGGL((k_trace<true, false>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, d_rays,
                               (uint32_t)n, d_out, ctx->counters, ctx->spill, cnt);
    } else {
        if (count)
            hipLaunchKernelGGL((k_trace<false, true>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, d_rays,
                               (uint32_t)n, d_out, ctx->counters, ctx->spill, cnt);
        else
            hipLaunchKernelGGL((k_trace<false, false>), dim3(blocks), dim3(kTraceBlock), 0, ctx->stream, sv, d_rays,
                               (uint32_t)n, d_out, ctx->counters, ctx->spill, cnt);
    }
    LT_HIP(ctx, hipGetLastError());
    LT_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    return LT_OK;
}
