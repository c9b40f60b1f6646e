// path_wave.h — state and queues of the wavefront Path integrator.
#pragma once
#include <hip/hip_runtime.h>

#include "context.h"
#include "path.h"

constexpr int kWaveMaxDepth = 32;
constexpr int kWfBins = 8;  // material sort: miss, then one bin per BSDF type, then other

// occlusion query with its pending NEE contribution (48 B)
struct ShadowRec {
    float4 o;  // xyz origin, w = tmax
    float4 d;  // xyz direction, w = path index (bits)
    float4 c;  // xyz contribution if unoccluded
};

// BSDF-sampled ray towards the sampled light triangle (pt_commons.glsl:28-40)
struct MisRec {
    float4 o;     // xyz origin, w = pdf_light_a
    float4 d;     // xyz direction, w = bsdf_pdf
    float4 w;     // xyz contribution before the MIS weight, w = path index (bits)
    float4 vpos;  // xyz vertex position (also the stale payload position)
    float4 sns;   // xyz stale payload shading normal
    uint4 ids;    // sampled light (triangle, instance), stale payload (triangle, instance)
};

// per-path state, SoA, one path per pixel
struct WaveState {
    float4* ray_o;
    float4* ray_d;
    float4* thr;
    float4* col;
    float4* misc;  // MIS-BSDF light terms, folded into col by the next shade or the accumulation
    uint4* meta;  // x = pcg4d counter word, y = last bounce specular
    uint4* hit;   // prim, inst (0xFFFFFFFF = miss), u, v
    uint32_t* queue[2];
    uint32_t* sorted;  // extend queue of a depth grouped by BSDF type (LT_WF_SORT_BSDF)
    uint8_t* bin;      // each queue item's BSDF bin
    ShadowRec* shadow;
    MisRec* mis;
};

// queue lengths and work counters of one frame (zeroed by a kernel, so a
// frame's launch sequence can be captured in a graph)
struct WaveCtl {
    uint32_t q_len[kWaveMaxDepth + 1];
    uint32_t sh_len[kWaveMaxDepth];
    uint32_t mis_len[kWaveMaxDepth];
    uint32_t work[kWaveMaxDepth][4];
    uint32_t bins[kWaveMaxDepth][2][kWfBins];  // material sort: items per bin, then scatter fill
};

struct WaveBuffers {
    WaveState S{};
    WaveCtl* ctl = nullptr;
    uint32_t capacity = 0;
};

lt_status wave_alloc(lt_context* ctx, WaveBuffers& wb, uint32_t n_paths);
void wave_free(WaveBuffers& wb);
lt_status launch_path_wavefront(lt_context* ctx, WaveBuffers& wb, const PathLaunch& L, float4* accum, int count,
                                uint64_t* d_counts);
