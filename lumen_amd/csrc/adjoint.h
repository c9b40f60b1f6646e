// adjoint.h — launch parameters of the path-replay adjoint kernels.
#pragma once
#include "context.h"

// The differentiable Path estimator: path.rgen:26-104 without roulette,
// direct lighting on, spp samples per pixel (sample k uses the RNG frame
// word k ^ seed), over `variants` texel sets that replace the scene textures
// aid (albedo, rgb) and rid (roughness).
//
// Work items: item i covers variant i % variants of pixel p = i / variants,
// pixels ordered by 8x8 tiles (tile p / 64, row-major inside), so one wave
// holds the variants of few pixels.  Those lanes trace the same rays until a
// glossy sample on a variant surface separates them, and their gradient
// scatters go to different variants' arrays.
struct AdjLaunch {
    float inv_view[16];
    float inv_projection[16];
    uint32_t width, height;
    int32_t max_depth;
    uint32_t spp, seed, variants;
    float sky_col[3];
    int32_t aid, rid;          // replaced scene textures (-1: none)
    uint32_t aw, ah, rw, rh;   // their texel grids
    uint32_t tiles_x;
    uint32_t n_items;          // tiles * 64 * variants
    uint32_t rec_depth;        // bounces per sample in the cache's path records (0: none)
};

// Bounces per sample the sample cache records (3 flag bits per bounce in one
// 32-bit word), and the cache's size in floats: per sample the radiance (3),
// and with records two 16-byte hit records per bounce (the path's closest
// hit and the NEE MIS ray's) and the flag word.
constexpr uint32_t kAdjRecMaxDepth = 10;
inline uint32_t adj_rec_depth(int32_t max_depth) {
    return max_depth >= 1 && max_depth <= (int32_t)kAdjRecMaxDepth ? (uint32_t)max_depth : 0u;
}
inline size_t adj_cache_floats(uint32_t spp, uint32_t n_items, uint32_t rec_depth) {
    return (size_t)spp * n_items * (3u + (rec_depth ? 8u * rec_depth + 1u : 0u));
}

// Device arrays of one call (all row-major, float32):
//   albedo [V][ah][aw][3], rough [V][rh][rw]       variant texels
//   images [V][H][W][3]                            forward output
//   cache  [spp][3][n_items]                        per-sample radiance (optional)
//   d_images [V][H][W][3]                          adjoint of images
//   g_albedo, g_rough like albedo, rough           texel gradients (overwritten)
struct AdjBuffers {
    const float* albedo;
    const float* rough;
    float* images;
    float* cache;
    const float* d_images;
    float* g_albedo;
    float* g_rough;
};

// counts: NULL, or kCountSlots u64 that the counting kernel variants fill
lt_status launch_adjoint(lt_context* ctx, const AdjLaunch& L, const AdjBuffers& B, bool backward, uint32_t* work,
                         uint32_t* spill, uint32_t blocks, unsigned long long* counts, hipStream_t stream);
