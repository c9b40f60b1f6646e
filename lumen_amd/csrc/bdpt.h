// bdpt.h — launch parameters and buffers of the BDPT integrator kernels.
#pragma once
#include "context.h"

constexpr int kBdptMaxDepth = 16;           // subpaths hold at most max_depth + 1 vertices
// Vertex-store budget of one walk/connect pass: a pass covers as many 8x8
// tiles as fit (1024x1024 at max_depth 8 is one pass of 1.3 GB).
constexpr size_t kBdptChunkBytes = (size_t)4 << 30;

// PCBDPT (src/shaders/integrators/bdpt/bdpt_commons.h:5-16) plus the
// SceneUBO camera matrices the ray generation and the splat projection read.
struct BdptLaunch {
    float view[16];
    float projection[16];
    float inv_view[16];
    float inv_projection[16];
    uint32_t width, height;
    int32_t max_depth;
    uint32_t frame_word;  // RNG frame word: frame_num ^ seed (bdpt.rgen:36-37)
    uint32_t power;       // 1: power heuristic; 0: the reference's balance heuristic
    // when set, the frame number is read here and the word is *frame_dev ^
    // seed (a frame captured in a graph and replayed frame after frame)
    const uint32_t* frame_dev;
    uint32_t seed;
    uint32_t regen_min;  // walk kernel: idle lanes that trigger taking new pixels
};

// Device state of one BDPT integrator.
//  - verts: the light and camera subpaths of the pixels of one chunk, one
//    32-bit field per plane: plane ((sub * (max_depth + 1) + i) * kVtxFields + f)
//    holds field f of vertex i of subpath sub for all `slots` pixels, so the
//    lanes of a wave touch consecutive words.
//  - hdr: per-pixel record of the walk kernel (subpath lengths, light vertex 0
//    fields, RNG counter), same slot indexing.
//  - fcol: the camera-side estimate of each pixel for the current frame.
//  - splat: light-tracing splats (float3 per pixel), cleared by the pixel that
//    reads them.
struct BdptBuffers {
    float* verts = nullptr;
    uint32_t* hdr = nullptr;
    uint32_t slots = 0;        // pixels per chunk (chunk_tiles * 64)
    uint32_t chunk_tiles = 0;
    int32_t max_depth = 0;
    float4* fcol = nullptr;
    float* splat = nullptr;
    size_t pixels = 0;
};

lt_status bdpt_alloc(lt_context* ctx, BdptBuffers& b, uint32_t width, uint32_t height, int32_t max_depth);
void bdpt_free(BdptBuffers& b);
// threads of the largest BDPT grid (the traversal spill buffer it needs)
size_t bdpt_spill_threads(lt_context* ctx);
// Renders `frames` frames starting at frame_num into the running mean `accum`.
// frames > 1 or frame_dev == nullptr; with frame_dev, one frame whose number is read there
lt_status launch_bdpt(lt_context* ctx, BdptBuffers& b, const BdptLaunch& L, uint32_t frame_num, uint32_t frames,
                      uint32_t seed, float4* accum, int count, uint64_t* d_counts);
