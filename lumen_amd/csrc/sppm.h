// sppm.h — launch parameters and buffers of the SPPM integrator kernels.
#pragma once
#include "context.h"

constexpr int kSppmMaxDepth = 16;

// PCSPPM (src/shaders/integrators/sppm/sppm_commons.h:10-26) plus the
// SceneUBO camera matrices.
struct SppmLaunch {
    float inv_view[16];
    float inv_projection[16];
    uint32_t width, height, n_px;  // n_px = width * height pixels = light paths
    int32_t max_depth;
    uint32_t frame_num;            // running-mean index of this frame
    uint32_t word;                 // pcg4d frame word of both passes
    float base_radius;
    float sky_col[3];
};

// Per-pixel SPPM state (SPPMData, sppm_commons.h:28-42) as planes of n_px
// 32-bit words, so the lanes of a wave touch consecutive words.  phi, M and
// tau live only within one frame's gather (tau restarts at zero every
// frame, composite.comp:52) and are kept in registers there.
enum SppmPlane : uint32_t {
    kSpPx, kSpPy, kSpPz,        // recorded surface point p
    kSpWox, kSpWoy, kSpWoz,     // wo at p
    kSpColx, kSpColy, kSpColz,  // col: emission and next-event terms of this frame
    kSpThrx, kSpThry, kSpThrz,  // eye throughput at p
    kSpNx, kSpNy, kSpNz,        // shading normal at p (flipped to wo's side)
    kSpU, kSpV,                 // texture coordinates at p
    kSpMat,                     // material index (u32 bits)
    kSpN,                       // accumulated photon count N
    kSpRadius,                  // gather radius
    kSpPathLen,                 // eye bounces before p (i32 bits)
    kSpSide,                    // front face (u32 bits)
    kSpPlanes
};

// The frame's grid (AtomicData, sppm_commons.h:3-8): bounds of p -+ radius
// over every pixel and the largest radius, reduced with order-preserving
// integer atomics (min, max of floats are exact, so any order gives the
// same result).
struct SppmBounds {
    uint32_t mn[3];  // ordered encodings of the minimum of p - radius
    uint32_t mx[3];  // ... of the maximum of p + radius
    uint32_t max_r;  // ... of the largest radius
    uint32_t pad;
};

// One photon, 48 B: pos.xyz + path_len, wi.xyz (towards the light) +
// throughput.x, throughput.yz.  Every photon of light path q at vertex d is
// stored at q * max_depth + d; a bucket's photon is the one with the
// largest such reference (the deterministic "last writer").
struct SppmPhoton {
    float4 a;
    float4 b;
    float4 c;
};

struct SppmBuffers {
    float* state = nullptr;           // kSpPlanes planes of n_px words
    SppmPhoton* photons = nullptr;    // n_px * max_depth records
    uint32_t* key = nullptr;          // 10 n_px buckets: 1 + largest photon reference
    uint32_t* count = nullptr;        // 10 n_px buckets: photons inserted
    SppmBounds* bounds = nullptr;
    uint32_t n_px = 0;
    int32_t max_depth = 0;
};

lt_status sppm_alloc(lt_context* ctx, SppmBuffers& b, uint32_t width, uint32_t height, int32_t max_depth);
void sppm_free(SppmBuffers& b);
// The pcg4d word of frame frame_num: frame_num ^ random_num, random_num (a
// host rand(), SPPM.cpp:81) being a pcg4d hash of (seed, frame_num).
uint32_t sppm_frame_word(uint32_t seed, uint32_t frame_num);
// Renders `frames` frames from frame_num: per frame the eye pass, the grid
// bounds, the light pass (photon table) and the gather with the progressive
// radius update and the running mean (SPPM.cpp:75-189).
lt_status launch_sppm(lt_context* ctx, SppmBuffers& b, const SppmLaunch& L, uint32_t frame_num, uint32_t frames,
                      uint32_t seed, float4* accum, int count, uint64_t* d_counts);
