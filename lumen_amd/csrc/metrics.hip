// metrics.hip — squared-error reduction behind lt_image_sq_error.
//
// The reference reduces sum_px |gt.rgb - out.rgb|^2 with subgroup adds in
// calc_rmse.comp:20-41 / reduce_rmse.comp and divides in output_rmse.comp:23.
// Here one pass: each thread accumulates a grid-stride slice in double, a
// wave64 shuffle tree reduces it, one double atomic per wave.
#include <hip/hip_runtime.h>

#include "context.h"

namespace {

__global__ void k_zero_d(double* p) { *p = 0.0; }

__global__ __launch_bounds__(256) void k_sq_error(const float4* __restrict__ a, const float4* __restrict__ b,
                                                  size_t n, double* __restrict__ out) {
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 x = a[i], y = b[i];
        const float dr = x.x - y.x, dg = x.y - y.y, db = x.z - y.z;
        acc += (double)(dr * dr) + (double)(dg * dg) + (double)(db * db);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
}

}  // namespace

lt_status launch_sq_error(lt_context* ctx, const float4* a, const float4* b, size_t n, double* d_out) {
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(1), 0, ctx->stream, d_out);
    size_t blocks = std::min<size_t>((n + 255) / 256, (size_t)ctx->num_cus * 8);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_sq_error, dim3(blocks), dim3(256), 0, ctx->stream, a, b, n, d_out);
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}

namespace {

__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

namespace {

__global__ void k_frame_word(uint32_t* p, uint32_t value, int advance) {
    if (threadIdx.x == 0) *p = advance ? *p + 1u : value;
}

}  // namespace

lt_status launch_frame_word(lt_context* ctx, uint32_t* p, uint32_t value, int advance, hipStream_t stream) {
    hipLaunchKernelGGL(k_frame_word, dim3(1), dim3(64), 0, stream, p, value, advance);
    LT_HIP(ctx, hipGetLastError());
    return LT_OK;
}

extern "C" lt_status lt_stream_copy(lt_context* ctx, void* dst, const void* src, size_t bytes, float* ms) {
    if (!ctx || !dst || !src || (bytes & 15)) return LT_ERR_INVALID_ARG;
    const size_t n = bytes / 16;
    size_t blocks = std::min<size_t>((n + 255) / 256, (size_t)ctx->num_cus * 16);
    if (blocks == 0) blocks = 1;
    LT_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4*)src, (uint4*)dst, n);
    LT_HIP(ctx, hipGetLastError());
    LT_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    LT_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float t = 0.0f;
    LT_HIP(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    if (ms) *ms = t;
    return LT_OK;
}
