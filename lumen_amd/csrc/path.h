// path.h — launch parameters of the Path integrator kernels.
#pragma once
#include "context.h"
#include "dev_math.h"

// Push-constant equivalent (PCPath, src/shaders/integrators/path/path_commons.h:3-15)
// plus the camera matrices of the SceneUBO.
struct PathLaunch {
    float inv_view[16];
    float inv_projection[16];
    uint32_t width, height;
    int32_t max_depth;
    float sky_col[3];
    uint32_t direct_lighting;
    uint32_t frame_start;  // frame_num of the first frame of this launch
    uint32_t spp;          // frames accumulated by this launch
    uint32_t seed;         // XOR'ed into the RNG frame word
    // when set, frame_start is read here instead (a frame captured in a
    // graph and replayed frame after frame; the graph advances the word)
    const uint32_t* frame_dev;
    uint32_t regen_min;  // persistent kernel: idle lanes that trigger a refill
    uint32_t nee_pick;   // LT_NEE_REFERENCE or LT_NEE_CORRECTED
    uint32_t xcd_bands;  // persistent kernel: pixels drawn per XCD from 8 bands
};

// Probability of the NEE light pick that sample_light_Li made (light `li`):
// the reference's 1 / light_triangle_count (path.rgen:76-78), or
// 1 / (lights x the picked light's triangles).
LT_DEV float nee_pick_pdf(const SceneView& sv, uint32_t nee_pick, uint32_t li) {
    if (nee_pick != LT_NEE_CORRECTED) return 1.0f / (float)sv.total_light_triangle_cnt;
    const lt_light& l = sv.lights[li];
    return 1.0f / ((float)sv.num_lights * ((l.light_flags & 7u) == LT_LIGHT_AREA ? (float)l.num_triangles : 1.0f));
}

// persistent lanes with path regeneration, one launch per frame
lt_status launch_path_persistent(lt_context* ctx, const PathLaunch& L, float4* accum, int count, uint64_t* d_counts);
lt_status launch_path_megakernel(lt_context* ctx, const PathLaunch& L, float4* accum, int count,
                                 uint64_t* d_counts);
