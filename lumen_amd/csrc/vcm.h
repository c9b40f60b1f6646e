// vcm.h — launch parameters and buffers of the VCM integrator kernels.
#pragma once
#include "context.h"

constexpr int kVcmMaxDepth = 16;

// PCVCM (src/shaders/integrators/vcm/vcm_commons.h:4-25) with the per-frame
// radius, grid and MIS constants (VCM.cpp:131-158, vcm_eye.rgen:62-67) and
// the SceneUBO camera matrices.
struct VcmLaunch {
    float view[16];
    float projection[16];
    float inv_view[16];
    float inv_projection[16];
    uint32_t width, height, n_paths;  // n_paths = width * height light paths
    int32_t max_depth;
    uint32_t light_word, eye_word;    // RNG frame words of the two passes
    uint32_t frame_num;               // running-mean index of this frame
    float radius, eta_vcm, eta_vc, eta_vm;
    int32_t grid_res[3];
    float bmin[3], bmax[3];
    float inv_ext[3];  // 1 / (bmax - bmin): the corrected estimator's grid (see k_vcm_scatter)
    float sky_col[3];
    uint32_t use_vm;
    uint32_t corrected;  // 0: the reference estimator, 1: corrected VCM
};

// A merge query of the corrected estimator, 80 B: an eye vertex and the
// camera subpath state merging needs, written by the eye pass and gathered
// by the merge pass.
struct VcmQuery {
    float4 a;  // pos.xyz, d_vcm
    float4 b;  // n_s.xyz, d_vm
    float4 c;  // wo.xyz, uv.x
    float4 d;  // throughput.xyz, uv.y
    uint4 e;   // material_idx, pixel, side | depth << 1, 0
};

// One stored light vertex (VCMVertex, vcm_commons.h:88-103), 96 B.
struct VcmVertex {
    float4 a;  // pos.xyz, d_vcm
    float4 b;  // n_s.xyz, d_vc
    float4 c;  // wi.xyz, d_vm
    float4 d;  // wo.xyz, uv.x
    float4 e;  // throughput.xyz, uv.y
    uint4 f;   // material_idx, path_len, side, 0
};

// A photon of the corrected estimator's bucketed list: its position (and
// d_vcm) in one array, 16 B, the rest in another, 48 B, so that the photons
// a gather tests and rejects cost 16 B each.
struct VcmPhotonRest {
    float4 b;  // wi.xyz (towards the light), d_vm
    float4 c;  // throughput.xyz, path_len (u32 bits)
    float4 d;  // n_s.xyz, 0
};

// Device state of one VCM integrator.
//  - verts / cnt: the stored vertices of each light path (path p holds
//    max_depth + 1 records at p * (max_depth + 1)) and their count.
//  - cells: 10 * W * H hash buckets (utils.glsl:62-65).  Reference estimator:
//    key = 1 + the largest vertex reference inserted into the bucket (the
//    deterministic "last writer") and count = insertions.  Corrected
//    estimator: count per bucket, its exclusive scan `start`, and `ppos` /
//    `prest` = the photons themselves, bucket by bucket, so a gather streams
//    each bucket's records contiguously.
//  - splat: light-tracing splats (float3 per pixel), read and cleared by the
//    eye pass (by the resolve pass when merges are gathered separately).
struct VcmBuffers {
    VcmVertex* verts = nullptr;
    uint32_t* cnt = nullptr;
    uint32_t* key = nullptr;
    uint32_t* count = nullptr;
    uint32_t* start = nullptr;
    uint32_t* fill = nullptr;
    float4* ppos = nullptr;  // pos.xyz, d_vcm
    VcmPhotonRest* prest = nullptr;
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    float* splat = nullptr;
    // corrected estimator with merging: the eye pass's colours (merges added
    // by the merge pass) and its merge queries (at most max_depth per pixel)
    float4* fcol = nullptr;
    VcmQuery* queries = nullptr;
    uint32_t* qctr = nullptr;  // queries written
    uint32_t n_paths = 0;
    int32_t max_depth = 0;
};

lt_status vcm_alloc(lt_context* ctx, VcmBuffers& b, uint32_t width, uint32_t height, int32_t max_depth);
void vcm_free(VcmBuffers& b);
// Renders `frames` frames starting at frame_num: per frame a light pass, the
// photon grid, and an eye pass that folds in the splats and the running mean.
lt_status launch_vcm(lt_context* ctx, VcmBuffers& b, const VcmLaunch& L, const lt_vcm_config& cfg,
                     uint32_t frame_num, uint32_t frames, uint32_t seed, float4* accum, int count, uint64_t* d_counts);
