// context.h — host-side state behind the C ABI (one HIP device, one stream).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/lt.h"
#include "bvh_layout.h"

// texture descriptor in HBM (texels are float4, row-major)
struct TexDesc {
    uint32_t width, height;
    uint64_t offset;  // in float4 units into the texel pool
};

// Everything a kernel needs to know about the scene, passed by value.
struct SceneView {
    const lt_vertex* vtx;
    const uint32_t* idx;
    const lt_prim_mesh_info* prims;
    const lt_material* mats;
    const lt_light* lights;
    const float* cones;      // 2 per light
    const float4* texels;
    const TexDesc* texs;
    const int32_t* rough;    // per material
    // per mesh, kXfFloats each (NULL: every mesh has the identity transform):
    // object-to-world rows (3 x 4), then the normal matrix rows (3 x 3,
    // inverse transpose of the linear part)
    const float* xf;
    // per light, kXfFloats each (NULL: every light's world_matrix is the
    // identity): the same two matrices from lt_light::world_matrix
    const float* lxf;
    const ltb::Node8* nodes;
    const ltb::TriRec* tris;
    const uint32_t* tri_start;  // num_prims + 1: first flattened triangle of each mesh
    uint32_t num_prims;
    uint32_t num_lights;
    uint32_t num_nodes;
    uint32_t refill_min;  // wavefront trace stages (lt_tuning)
    uint32_t total_light_triangle_cnt;
    int32_t dir_light_idx;
    float total_light_area;
    float scene_min[3], scene_max[3];
};

// floats per mesh / light transform record: object-to-world 3 x 4, normal
// matrix 3 x 3, 3 unused
constexpr int kXfFloats = 24;

struct DeviceScene {
    lt_vertex* vtx = nullptr;
    uint32_t* idx = nullptr;
    lt_prim_mesh_info* prims = nullptr;
    lt_material* mats = nullptr;
    lt_light* lights = nullptr;
    float* cones = nullptr;
    float4* texels = nullptr;
    TexDesc* texs = nullptr;
    int32_t* rough = nullptr;
    float* xf = nullptr;           // per mesh transforms (SceneView::xf), or NULL
    float* lxf = nullptr;          // per light transforms (SceneView::lxf), or NULL
    uint32_t* tri_start = nullptr; // nprims + 1: first flattened triangle of each mesh
    uint32_t nv = 0, ni = 0, nprims = 0, nmats = 0, nlights = 0, ntex = 0;
    uint32_t ntris = 0;            // triangles over every mesh (instances counted apart)
    uint32_t total_light_triangle_cnt = 0;
    int32_t dir_light_idx = -1;
    float total_light_area = 0.0f;
    float smin[3] = {0, 0, 0}, smax[3] = {0, 0, 0};
    // host copies for validation of later calls (materials, roughness
    // texture ids per material, texture sizes)
    std::vector<lt_material> h_mats;
    std::vector<int32_t> h_rough;
    std::vector<TexDesc> h_texs;
    bool uploaded = false;
};

struct DeviceBVH {
    ltb::Node8* nodes = nullptr;
    ltb::TriRec* tris = nullptr;
    uint32_t n_nodes = 0, n_tris = 0, depth = 0;
    bool built = false;
};

struct lt_context {
    int device = 0;
    int num_cus = 0;
    std::string arch;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    DeviceScene scene;
    DeviceBVH bvh;
    // traversal stack spill (kStackTotal - kLdsStack entries per resident thread)
    uint32_t* spill = nullptr;
    size_t spill_threads = 0;
    // persistent-kernel work counters
    uint32_t* counters = nullptr;  // 64 words
    // scratch for raw trace calls with host pointers
    void* ray_buf = nullptr;
    size_t ray_buf_bytes = 0;
    void* hit_buf = nullptr;
    size_t hit_buf_bytes = 0;
    uint64_t* count_buf = nullptr;  // kCountSlots counters of a counting render
    // device timing of the phases of a render (lt_render_set_timing): the
    // launch code marks each phase's start and the frame's end
    int phase_on = 0;
    int phase_n = 0;
    hipEvent_t phase_ev[9] = {};
    // device address of the no-allocation entry point's scratch array (shared
    // by every context on this device; not owned)
    uint32_t* noalloc_scratch = nullptr;
    float last_ms = 0.0f;
    lt_tuning tuning{10u, 8u, 32u, 0u, 4u, 0u, LT_WF_SORT_NONE, LT_XCD_BANDS_OFF, {}};
    // bumped whenever device state that launches capture by pointer or value
    // changes (scene upload, BVH build, spill reallocation, launch shapes):
    // a captured frame graph older than this is stale
    uint64_t generation = 1;
};

// helpers shared by the .cpp / .hip translation units
lt_status lt_fail(lt_context* ctx, lt_status st, const std::string& msg);
lt_status lt_hip_check(lt_context* ctx, hipError_t e, const char* what);
#define LT_HIP(ctx, call)                                              \
    do {                                                               \
        hipError_t _e = (call);                                        \
        if (_e != hipSuccess) return lt_hip_check((ctx), _e, #call);   \
    } while (0)

SceneView make_scene_view(const lt_context* ctx);
// persistent grid: number of resident threads the spill buffer is sized for
lt_status ensure_spill(lt_context* ctx, size_t threads);
constexpr int kTraceBlock = 128;
// counters of a counting render, u64 each: closest rays, shadow rays, BVH8
// nodes popped, triangles tested, photons tested against a gather radius,
// photons gathered, algorithmic bytes of the gather pass, the nodes and the
// triangles of occlusion queries (included in slots 2 and 3), spare
constexpr int kCountSlots = 10;
constexpr int kMaxPhases = 8;
// records the next phase boundary of the first frame of a render when phase
// timing is on (at most kMaxPhases phases)
void phase_mark(lt_context* ctx);
// workgroups of kTraceBlock threads of `kernel` that one CU holds at once
// (its register and LDS use), at least 1; `fixed` (a tuning value) wins if set
uint32_t resident_blocks(const void* kernel, uint32_t fixed);

// builder / kernels (implemented in the .hip files)
lt_status bvh_build(lt_context* ctx, const lt_build_opts* opts, lt_build_stats* stats);
// writes `value` into the word at p (advance = 0), or increments it (advance = 1)
lt_status launch_frame_word(lt_context* ctx, uint32_t* p, uint32_t value, int advance, hipStream_t stream);
lt_status launch_trace(lt_context* ctx, const lt_ray* d_rays, size_t n, void* d_out, int any_hit, int count,
                       uint64_t* d_counts);
// resolves ctx->noalloc_scratch (once per device)
lt_status noalloc_scratch_init(lt_context* ctx);
