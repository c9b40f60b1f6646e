/*
 * lt.h — C ABI of the MI355X light-transport core.
 *
 * Every entry point returns an lt_status; no C++ exception crosses this
 * boundary.  All structures are plain data with the byte layouts of the
 * reference's host/device wire format (src/shaders/commons.h) so that a
 * LumenScene's CPU vectors can be handed over unchanged.
 *
 * Threading: calls on one lt_context are not thread-safe; each context owns one
 * HIP stream.  lt_render / lt_trace_* with device pointers are asynchronous on
 * that stream; every entry point that returns host data synchronises it.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository root).
 */
#ifndef LT_H
#define LT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LT_ABI_VERSION 1

typedef enum lt_status {
    LT_OK = 0,
    LT_ERR_INVALID_ARG = 1,   /* null pointer, bad size, out-of-range index   */
    LT_ERR_HIP = 2,           /* a HIP runtime call failed                    */
    LT_ERR_OUT_OF_MEMORY = 3, /* device or host allocation failed             */
    LT_ERR_NOT_READY = 4,     /* e.g. render before scene upload / accel build */
    LT_ERR_PARSE = 5,         /* scene JSON / OBJ parse error                 */
    LT_ERR_UNSUPPORTED = 6,   /* feature not built                            */
    LT_ERR_NO_DEVICE = 7      /* no HIP device visible                        */
} lt_status;

/* ---- reference wire format (src/shaders/commons.h) ------------------------ */

/* BSDF type flags, commons.h:13-18 */
#define LT_BSDF_DIFFUSE 1u
#define LT_BSDF_MIRROR 2u
#define LT_BSDF_GLASS 4u
#define LT_BSDF_DIELECTRIC 8u
#define LT_BSDF_CONDUCTOR 16u
#define LT_BSDF_PRINCIPLED 32u
/* BSDF property flags, commons.h:22-26 */
#define LT_BSDF_FLAG_DIFFUSE 1u
#define LT_BSDF_FLAG_SPECULAR 2u
#define LT_BSDF_FLAG_GLOSSY 4u
#define LT_BSDF_FLAG_REFLECTION 8u
#define LT_BSDF_FLAG_TRANSMISSION 16u
/* Light type = flags & 7, commons.h:36-38; bit 4 finite, bit 5 delta
 * (commons.glsl:42-46).  LT_LIGHT_POINT is an addition of this core: an omni
 * point light (finite + delta), which the reference data model lacks. */
#define LT_LIGHT_SPOT 1u
#define LT_LIGHT_AREA 2u
#define LT_LIGHT_DIRECTIONAL 3u
#define LT_LIGHT_POINT 4u
#define LT_LIGHT_FINITE (1u << 4)
#define LT_LIGHT_DELTA (1u << 5)

/* Vertex, commons.h:194-198 (32 B) */
typedef struct lt_vertex {
    float pos[3];
    float normal[3];
    float uv0[2];
} lt_vertex;

/* PrimMeshInfo, commons.h:333-340 (48 B) */
typedef struct lt_prim_mesh_info {
    uint32_t index_offset;
    uint32_t vertex_offset;
    uint32_t material_index;
    uint32_t pad;
    float min_pos[4];
    float max_pos[4];
} lt_prim_mesh_info;

/* Material, commons.h:212-234 (104 B).  texture_id indexes lt_scene_desc
 * textures and multiplies albedo (bsdf_commons.glsl:16-22).  A roughness
 * texture per material, which the reference lacks, is given by
 * lt_scene_desc.roughness_texture_ids. */
typedef struct lt_material {
    float albedo[3];
    float ior;
    float emissive_factor[3];
    uint32_t bsdf_type;
    uint32_t bsdf_props;
    float k[3];
    int32_t texture_id;
    float roughness;
    float diffuse_trans;
    float spec_trans;
    float metallic;
    float specular_tint;
    float sheen_tint;
    float clearcoat;
    float clearcoat_gloss;
    float sheen;
    float subsurface;
    float flatness;
    float anisotropy;
    uint32_t thin;
} lt_material;

/* Light, commons.h:200-210 (128 B).  world_matrix is column-major (glm). */
typedef struct lt_light {
    float world_matrix[16];
    float pos[3];
    uint32_t prim_mesh_idx;
    float to[3];
    uint32_t num_triangles;
    float L[3];
    uint32_t light_flags;
    float world_center[3];
    float world_radius;
} lt_light;

/* A float RGBA texture (texel (x,y) at data[4*(y*width+x)]), sampled
 * bilinearly with repeat addressing (the reference sampler,
 * src/RayTracer/LumenScene.cpp:190-195).  Texels are linear floats. */
typedef struct lt_texture {
    uint32_t width;
    uint32_t height;
    const float* data;
} lt_texture;

/* Flattened scene: the CPU vectors of LumenScene (src/RayTracer/LumenScene.h:47-99)
 * after load_scene (src/RayTracer/LumenScene.cpp:52-229).  Copied at upload; the
 * caller keeps ownership.  Vertices of meshes whose world matrix is not the
 * identity must already be in world space (the reference's JSON path always
 * uses identity world matrices, LumenScene.cpp:328). */
typedef struct lt_scene_desc {
    const lt_vertex* vertices;
    uint32_t num_vertices;
    const uint32_t* indices;
    uint32_t num_indices;
    const lt_prim_mesh_info* prim_infos;
    uint32_t num_prim_infos;
    const lt_material* materials;
    uint32_t num_materials;
    const lt_light* lights;
    uint32_t num_lights;
    const lt_texture* textures;
    uint32_t num_textures;
    /* per material: texture whose .x replaces roughness, or -1; NULL = none */
    const int32_t* roughness_texture_ids;
    /* per light: spot cone (cos_width, cos_falloff_start); NULL = the reference's
     * fixed 30 / 25 degrees (commons.glsl:264-265).  An addition of this core. */
    const float* spot_cones;
    uint32_t total_light_triangle_cnt; /* LumenScene.cpp:91,106 */
    float total_light_area;            /* LumenScene.cpp:111-129 */
    int32_t dir_light_idx;             /* -1 if none, LumenScene.cpp:107-109 */
    float scene_min[3];                /* m_dimensions, LumenScene.cpp:732-751 */
    float scene_max[3];
    /* Per prim mesh, optional (NULL = identity for every mesh, as JSON scenes
     * load, LumenScene.cpp:328): LumenPrimMesh::world_matrix (LumenScene.h:32),
     * 16 floats column-major (glm), the object-to-world transform of the mesh's
     * TLAS instance (Integrator.cpp:147-158).  Hits report world-space
     * positions and normals as ray.rchit:62-80 computes them; area lights on
     * the mesh sample world-space points (commons.glsl:112-149). */
    const float* world_matrices;
    /* Per prim mesh, optional: its index count.  Meshes that are instances of
     * one geometry share index_offset / vertex_offset and differ in their world
     * matrix; NULL = each mesh's indices end where the next mesh's begin. */
    const uint32_t* index_counts;
} lt_scene_desc;

/* SceneUBO subset used by the kernels (commons.h:180-192): column-major
 * matrices as produced by Integrator::update_uniform_buffers
 * (src/RayTracer/Integrator.cpp:60-72). */
typedef struct lt_camera {
    float view[16];
    float projection[16];
    float inv_view[16];
    float inv_projection[16];
} lt_camera;

/* A ray for the parity / microbenchmark entry points (32 B). */
typedef struct lt_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} lt_ray;

/* Closest-hit result (20 B).  prim = BLAS-local triangle index
 * (gl_PrimitiveID), inst = mesh index (gl_InstanceCustomIndexEXT); both are
 * 0xFFFFFFFF on a miss.  u, v are the barycentric weights of v1 and v2. */
typedef struct lt_hit {
    float t;
    uint32_t prim;
    uint32_t inst;
    float u;
    float v;
} lt_hit;

/* ---- context -------------------------------------------------------------- */

typedef struct lt_context lt_context;
typedef struct lt_integrator lt_integrator;

/* Replaces vk::init (src/Framework/VulkanBase.cpp:619): selects HIP device
 * `device`, creates one stream. */
lt_status lt_context_create(int device, lt_context** out);
/* Replaces Integrator::destroy / LumenScene::destroy
 * (src/RayTracer/Integrator.cpp:128-135, LumenScene.cpp:753-766).  Idempotent
 * on NULL. */
lt_status lt_context_destroy(lt_context* ctx);
/* Human-readable text of the last error on this context (never NULL). */
const char* lt_last_error(const lt_context* ctx);
/* Number of compute units and the gfx arch name of the context's device. */
lt_status lt_device_info(lt_context* ctx, int* num_cus, char* arch, size_t arch_len);

/* Replaces LumenScene::load_scene's buffer uploads (LumenScene.cpp:146-216). */
lt_status lt_scene_upload(lt_context* ctx, const lt_scene_desc* scene);
/* The host-side checks of lt_scene_upload without a device (index ranges,
 * references, transforms, finite vertex positions); err receives the reason
 * of a refusal.  No reference counterpart. */
lt_status lt_scene_validate(const lt_scene_desc* scene, char* err, size_t err_len);

/* Build options for lt_accel_build. */
typedef struct lt_build_opts {
    uint32_t max_leaf_tris; /* triangles per BVH8 leaf child, 1..4 (0: the default, 1) */
    uint32_t collapse;      /* LT_COLLAPSE_SAH (0: the default) or LT_COLLAPSE_GREEDY */
    uint32_t binary_tree;   /* LT_TREE_PLOC (0: the default) or LT_TREE_LBVH */
    uint32_t reserved[5];
} lt_build_opts;
/* The binary tree the collapse starts from: PLOC, parallel locally-ordered
 * clustering of the Morton-sorted triangles by the surface area of merged
 * boxes (Meister and Bittner 2018, radius 16), or the LBVH of the Morton
 * codes' radix tree (Karras 2012). */
#define LT_TREE_PLOC 1u
#define LT_TREE_LBVH 2u
/* Collapse of the binary LBVH into BVH8 nodes: the SAH-optimal choice of each
 * node's children by dynamic programming over the binary tree (Ylitie, Karras
 * and Laine 2017, node cost 1, triangle cost 0.5), or the greedy expansion of
 * the child with the largest surface area. */
#define LT_COLLAPSE_SAH 1u
#define LT_COLLAPSE_GREEDY 2u

/* Build statistics. */
typedef struct lt_build_stats {
    uint32_t num_nodes;     /* BVH8 nodes (128 B each)          */
    uint32_t num_tris;      /* triangle records (48 B each)     */
    uint32_t max_depth;     /* BVH8 depth (root = 1)            */
    uint32_t binary_depth;  /* depth of the LBVH binary tree    */
    float build_ms;         /* device time of the whole build   */
} lt_build_stats;

/* Replaces Integrator::create_accel -> vk::build_blas / vk::build_tlas
 * (src/RayTracer/Integrator.cpp:137-160, src/Framework/AccelerationStructure.cpp:171-315):
 * LBVH (30-bit Morton, radix sort, Karras radix tree) collapsed to a BVH8 of
 * 128-byte quantised nodes.  opts may be NULL. */
lt_status lt_accel_build(lt_context* ctx, const lt_build_opts* opts, lt_build_stats* stats);
/* Copies the built BVH8 (128-B nodes) and its 48-B triangle records to host
 * memory (for inspection and for host traversals over the identical tree).
 * No reference counterpart: the reference's BVH is opaque to the host. */
lt_status lt_accel_export(lt_context* ctx, void* nodes, size_t nodes_bytes, void* tris, size_t tris_bytes);

/* ---- raw queries (parity and microbenchmarks) ---------------------------- */

#define LT_PTR_DEVICE 1u /* rays/hits are device pointers; call is asynchronous */

/* Closest hit: the traceRayEXT(tlas, gl_RayFlagsOpaqueEXT, ...) query plus the
 * ray.rchit / ray.rmiss identity fields (src/shaders/ray.rchit:24-84,
 * ray.rmiss:13).  Ties in t resolve to the smaller (inst, prim). */
lt_status lt_trace_closest(lt_context* ctx, const lt_ray* rays, size_t n, lt_hit* hits, uint32_t flags);
/* Closest hits like lt_trace_closest with LT_PTR_DEVICE (rays and hits are
 * device pointers; asynchronous on the context stream), for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  No counterpart in the reference.
 * It calls no hipMalloc / hipFree: the persistent grid's work counter and its
 * traversal-stack spill area are a fixed device array of 8 MiB + 256 B per
 * device (at most 1024 workgroups of 128 threads, 16 spill entries each),
 * held until the process ends and shared by every call on that device, so
 * calls that overlap on one GPU race.  A BVH deeper than 33 levels (more than
 * the 16 LDS + 16 spill stack entries) is refused with LT_ERR_UNSUPPORTED. */
lt_status lt_trace_closest_noalloc(lt_context* ctx, const lt_ray* rays, size_t n, lt_hit* hits);
/* Occlusion: traceRayEXT(TerminateOnFirstHit|SkipClosestHit) + ray.rahit /
 * ray_shadow.rmiss (src/shaders/integrators/pt_commons.glsl:20-23,
 * ray.rahit:18-20, ray_shadow.rmiss:12).  occluded[i] = 1 iff any triangle is
 * hit with tmin <= t <= tmax. */
lt_status lt_trace_any(lt_context* ctx, const lt_ray* rays, size_t n, uint8_t* occluded, uint32_t flags);
/* Per-ray traversal work of the last lt_trace_closest / lt_trace_any call in a
 * counting pass: total BVH8 nodes visited and triangles tested. */
lt_status lt_trace_count(lt_context* ctx, const lt_ray* rays, size_t n, int any_hit, uint32_t flags,
                         uint64_t* nodes, uint64_t* tris);

/* Launch shapes of the persistent kernels (workgroups per compute unit; 0
 * keeps the current value).  Set by the autotuner (lumen_amd/autotune.py).
 * No reference counterpart: the reference's launch shapes are the driver's. */
typedef struct lt_tuning {
    uint32_t path_blocks_per_cu;   /* Path megakernel and wavefront trace stages, 128-thread
                                      workgroups (default 10)                           */
    uint32_t trace_blocks_per_cu;  /* raw queries, 128-thread workgroups (default 8)     */
    uint32_t refill_min;           /* wavefront trace stages: idle lanes that trigger a
                                      refill from the queue, 1..64 (default 32)         */
    uint32_t bdpt_blocks_per_cu;   /* BDPT and VCM persistent kernels, 128-thread
                                      workgroups; default 0 = each kernel's resident
                                      workgroups per CU (its register and LDS use).
                                      lt_set_tuning leaves a 0 field unchanged.         */
    uint32_t regen_min;            /* persistent Path kernel and BDPT walk: idle lanes
                                      that trigger taking new pixels, 1..64 (default 4) */
    uint32_t bdpt_chunk_tiles;     /* BDPT: 8x8 pixel tiles per walk/connection pass,
                                      read when an integrator is created; default 0 =
                                      as many as fit a 4 GiB vertex store               */
    uint32_t wf_sort;              /* wavefront Path shade stages: LT_WF_SORT_NONE
                                      (default) or LT_WF_SORT_BSDF (the hit paths of a
                                      depth grouped by the BSDF type of their material
                                      before shading); 0 = unchanged                     */
    uint32_t xcd_bands;            /* persistent Path kernel: LT_XCD_BANDS_OFF (default) or
                                      LT_XCD_BANDS_ON (the pixels cut into 8 bands, one
                                      per XCD's workgroups, each with its own counter);
                                      0 = unchanged                                      */
    uint32_t reserved[8];
} lt_tuning;
#define LT_WF_SORT_NONE 1u
#define LT_WF_SORT_BSDF 2u
#define LT_XCD_BANDS_OFF 1u
#define LT_XCD_BANDS_ON 2u
lt_status lt_set_tuning(lt_context* ctx, const lt_tuning* t);
lt_status lt_get_tuning(lt_context* ctx, lt_tuning* t);

/* Device-time of the last kernel launch issued by lt_trace_* / lt_render, in ms
 * (HIP events on the context stream). */
lt_status lt_last_kernel_ms(lt_context* ctx, float* ms);

/* Device buffers (for callers that keep rays/hits resident in HBM). */
lt_status lt_device_alloc(lt_context* ctx, size_t bytes, void** dptr);
lt_status lt_device_free(lt_context* ctx, void* dptr);
lt_status lt_memcpy_h2d(lt_context* ctx, void* dst, const void* src, size_t bytes);
lt_status lt_memcpy_d2h(lt_context* ctx, void* dst, const void* src, size_t bytes);
lt_status lt_synchronize(lt_context* ctx);
/* Plain device-to-device copy kernel (16-B vector loads and stores, grid
 * stride).  No reference counterpart: it measures the attainable HBM bandwidth
 * and calibrates the rocprof byte counters used for the roofline. */
lt_status lt_stream_copy(lt_context* ctx, void* dst, const void* src, size_t bytes, float* ms);

/* ---- integrators ------------------------------------------------------------ */

/* IntegratorType names of src/RayTracer/SceneConfig.h:13 that this core builds. */
typedef enum lt_integrator_kind {
    LT_INTEGRATOR_PATH = 0,
    LT_INTEGRATOR_BDPT = 1,
    LT_INTEGRATOR_VCM = 2,
    LT_INTEGRATOR_SPPM = 3
} lt_integrator_kind;

/* Path execution strategies. */
#define LT_PATH_MEGAKERNEL 0u /* one thread per pixel, whole path in one kernel */
#define LT_PATH_WAVEFRONT 1u  /* persistent stages with compacted ray queues   */
#define LT_PATH_PERSISTENT 2u /* persistent lanes with path regeneration: a lane
                                 takes the next pixel as soon as its path ends */

/* Weight of the Path integrator's next-event sample.  The reference divides
 * by 1 / light_triangle_count whichever light was picked (path.rgen:76-78,
 * each analytic light counting one; LumenScene.cpp:106), while it picks a
 * light uniformly and then one of its triangles (commons.glsl:227-238): exact
 * for a single emitter mesh, biased otherwise.  The corrected weight is the
 * pick's own probability, 1 / (lights x the picked light's triangles). */
#define LT_NEE_REFERENCE 0u
#define LT_NEE_CORRECTED 1u

/* Path integrator config: SceneConfig/PathConfig (src/RayTracer/SceneConfig.h:15-25)
 * plus the PCPath fields set by Path::render (src/RayTracer/Path.cpp:27-38). */
typedef struct lt_path_config {
    uint32_t width;
    uint32_t height;
    int32_t max_depth;        /* path_length (default 6)                 */
    float sky_col[3];
    uint32_t direct_lighting; /* Path.h: default 1                       */
    uint32_t strategy;        /* LT_PATH_MEGAKERNEL, LT_PATH_WAVEFRONT or
                                 LT_PATH_PERSISTENT                        */
    uint32_t nee_pick;        /* LT_NEE_REFERENCE (default) or LT_NEE_CORRECTED */
    uint32_t reserved[5];
} lt_path_config;

/* MIS heuristics of the BDPT integrator. */
#define LT_MIS_BALANCE 0u /* calc_mis_weight of the reference (bdpt_commons.glsl:428-472) */
#define LT_MIS_POWER 1u   /* power heuristic, beta = 2                                     */

/* BDPT integrator config: SceneConfig/BDPTConfig (src/RayTracer/SceneConfig.h:27-29)
 * plus the PCBDPT fields set by BDPT::render (src/RayTracer/BDPT.cpp:55-64).
 * Camera rays go through pixel centres (bdpt.rgen:41-42); light and camera
 * subpaths hold up to max_depth + 1 vertices each. */
typedef struct lt_bdpt_config {
    uint32_t width;
    uint32_t height;
    int32_t max_depth;   /* path_length, 1..16 (default 6) */
    uint32_t mis;        /* LT_MIS_BALANCE or LT_MIS_POWER */
    uint32_t reserved[8];
} lt_bdpt_config;

/* VCM estimators. */
#define LT_VCM_REFERENCE 0u /* the reference's shader: last-writer photon table, overwritten
                               vertex-connection and merge results (vcm_commons.glsl:240,287,455) */
#define LT_VCM_CORRECTED 1u /* summed connections and merges over every stored photon      */

/* VCM integrator config: SceneConfig/VCMConfig (src/RayTracer/SceneConfig.h:38-42)
 * plus the PCVCM fields set by VCM::render (src/RayTracer/VCM.cpp:131-158).
 * One light subpath per pixel per frame; vertex connection is always on
 * (VCM.cpp:7). */
typedef struct lt_vcm_config {
    uint32_t width;
    uint32_t height;
    int32_t max_depth;    /* path_length, 1..16 (default 6)                              */
    float sky_col[3];
    float radius_factor;  /* merge radius in percent of the scene's bounding radius
                             (half diagonal), default 0.025                              */
    float alpha;          /* radius reduction: r_n = r_0 / (n+1)^((1-alpha)/2); the
                             reference hard-codes 2/3 (VCM.cpp:141)                      */
    uint32_t enable_vm;   /* vertex merging (default 0)                                  */
    uint32_t estimator;   /* LT_VCM_REFERENCE or LT_VCM_CORRECTED                         */
    uint32_t reserved[6];
} lt_vcm_config;

/* SPPM integrator config: SceneConfig/SPPMConfig (src/RayTracer/SceneConfig.h:33-36)
 * plus the PCSPPM fields set by SPPM::render (src/RayTracer/SPPM.cpp:75-104).
 * Eye rays go through pixel centres (sppm_eye.rgen:38-40); one light path per
 * pixel slot per frame; the per-pixel state (recorded point, radius, photon
 * count N) persists from frame 0 on.  The reference's photon table keeps one
 * photon per hash bucket written by racing threads (sppm_light.rgen:85-93);
 * here the bucket keeps the photon of the largest (light path, vertex), so
 * images are deterministic. */
typedef struct lt_sppm_config {
    uint32_t width;
    uint32_t height;
    int32_t max_depth;   /* path_length, 1..16 (default 6)                     */
    float sky_col[3];
    float base_radius;   /* initial gather radius (default 0.03; SPPM.cpp:86-88
                            raises it to at least 1e-7)                          */
    uint32_t reserved[8];
} lt_sppm_config;

/* Per-frame parameters (SceneUBO + push-constant frame_num). */
typedef struct lt_frame_params {
    lt_camera camera;
    uint32_t frame_num; /* running-mean index: 0 overwrites */
    uint32_t seed;      /* XOR'ed into the RNG frame word (0 = reference Path) */
    uint32_t spp;       /* frames rendered by this call (>=1); frame_num advances */
    uint32_t reserved[5];
} lt_frame_params;

/* Render statistics of the last lt_render call. */
typedef struct lt_render_stats {
    uint64_t closest_rays;   /* closest-hit queries traced            */
    uint64_t shadow_rays;    /* occlusion queries traced              */
    uint64_t nodes_visited;  /* BVH8 nodes popped (128 B each)        */
    uint64_t tris_tested;    /* triangle records tested (48 B each)   */
    float kernel_ms;         /* device time of the render             */
    uint32_t frames;
    /* Photon gathers (VCM corrected estimator's merge pass, SPPM gather), with
     * counting on: photons whose position was tested against the radius,
     * photons gathered, and the gather pass's algorithmic bytes (records of
     * the queries, cell ranges, photon positions and photon records read). */
    uint64_t photons_tested;
    uint64_t photons_gathered;
    uint64_t gather_bytes;
    /* The part of nodes_visited / tris_tested that occlusion queries took. */
    uint64_t shadow_nodes_visited;
    uint64_t shadow_tris_tested;
    /* With lt_render_set_timing on: device time of each phase of the first
     * frame of the last lt_render.  Phases: Path persistent / megakernel: 1
     * (the frame); Path wavefront: 1; BDPT: walk and connect of each pixel
     * chunk (one chunk at 1024x1024, depth 8), then the splat resolve; VCM: 4 (light pass + photon grid, eye pass, merge pass,
     * resolve; the last two 0 without the corrected estimator's merging);
     * SPPM: 4 (eye pass, grid bounds, light pass, gather). */
    uint32_t num_phases;
    float phase_ms[8];
} lt_render_stats;

/* Replaces RayTracer::create_integrator + Integrator::init
 * (src/RayTracer/RayTracer.cpp:244-282, Integrator.cpp:8-50).  cfg points to
 * the config struct of `kind` (lt_path_config for LT_INTEGRATOR_PATH,
 * lt_bdpt_config for LT_INTEGRATOR_BDPT, lt_vcm_config for LT_INTEGRATOR_VCM,
 * lt_sppm_config for LT_INTEGRATOR_SPPM) and cfg_size must equal its size. */
lt_status lt_integrator_create(lt_context* ctx, lt_integrator_kind kind, const void* cfg, size_t cfg_size,
                               lt_integrator** out);
lt_status lt_integrator_destroy(lt_integrator* it);
/* Replaces Integrator::render + RenderGraph::run (src/RayTracer/Path.cpp:27-59):
 * renders params->spp frames starting at params->frame_num into the running
 * mean.  Asynchronous. */
lt_status lt_render(lt_integrator* it, const lt_frame_params* params);
/* Ray counts and device time of the last lt_render (synchronises; with
 * counting enabled the kernels tally rays). */
lt_status lt_render_stats_get(lt_integrator* it, lt_render_stats* stats);
/* Enables per-ray counting in later lt_render calls (slower; for Mrays/s). */
lt_status lt_render_set_counting(lt_integrator* it, int enable);
/* Enables per-phase device timing of later lt_render calls (HIP events on the
 * context stream between the phases of the first frame; see lt_render_stats).
 * No reference counterpart.  Not with frame graphs. */
lt_status lt_render_set_timing(lt_integrator* it, int enable);
/* Frames from a captured hipGraph (Path and BDPT; VCM and SPPM: LT_ERR_UNSUPPORTED).
 * No reference counterpart.  The first lt_render after enabling, or after a
 * change of camera, seed, counting, scene, BVH or launch shapes, captures one
 * frame's launch sequence; each frame is then one graph launch whose kernels
 * read the frame number from device memory and whose last kernel advances
 * it.  Images equal those of the plain launches. */
lt_status lt_render_set_graph(lt_integrator* it, int enable);
/* Reset of the accumulation (Path::update with updated=true, Path.cpp:61-68). */
lt_status lt_accum_reset(lt_integrator* it);
/* Copies output_tex (RGBA32F, row-major: texel (x,y) at 4*(y*width+x)) to host. */
lt_status lt_readback_rgba32f(lt_integrator* it, float* host, size_t n_floats);

/* ---- differentiable rendering ---------------------------------------------------- */

/* No reference counterpart: the reference renders but does not
 * differentiate.  The estimator is the Path integrator (path.rgen:26-104,
 * pt_commons.glsl:3-42) with direct lighting on and Russian roulette removed,
 * averaging spp samples per pixel; sample k of pixel (x, y) uses the RNG frame
 * word k ^ seed, so with max_depth <= 5 one sample equals one Path frame.
 * Per call, `variants` sets of texels replace the scene textures
 * albedo_texture (rgb) and roughness_texture (first channel).  Gradients are
 * path-replay backpropagation with detached sampling, for Lambertian albedo
 * and for the principled BSDF with metallic, spec_trans, specular_tint,
 * clearcoat, flatness and anisotropy zero; lt_adjoint_create refuses
 * (LT_ERR_UNSUPPORTED) a scene where another material reads either texture. */
typedef struct lt_adjoint lt_adjoint;
typedef struct lt_adjoint_config {
    uint32_t width;
    uint32_t height;
    int32_t max_depth;          /* path_length (1..64)                                   */
    uint32_t spp;               /* samples per pixel per call                            */
    uint32_t variants;          /* texel sets per call (>= 1)                            */
    int32_t albedo_texture;     /* scene texture id replaced by the variants (-1: none)  */
    int32_t roughness_texture;  /* scene texture id replaced by the variants (-1: none)  */
    float sky_col[3];
    uint32_t reserved[8];
} lt_adjoint_config;
/* Requires lt_scene_upload and lt_accel_build on ctx. */
lt_status lt_adjoint_create(lt_context* ctx, const lt_adjoint_config* cfg, lt_adjoint** out);
lt_status lt_adjoint_destroy(lt_adjoint* adj);
/* Texel grids of the two replaced textures (0 x 0 for a texture set to -1). */
lt_status lt_adjoint_texture_size(lt_adjoint* adj, uint32_t* albedo_w, uint32_t* albedo_h, uint32_t* rough_w,
                                  uint32_t* rough_h);
/* Floats of the optional per-sample cache: spp * items * 3 (radiance), plus,
 * for max_depth <= 10, spp * items * (8 * max_depth + 1) for the path records
 * (each bounce's closest hit and NEE MIS hit, 16 B each, and a word of flags
 * for the occlusion tests). */
lt_status lt_adjoint_cache_floats(lt_adjoint* adj, size_t* n);
/* All pointers are device pointers, float32, row-major:
 *   albedo [V][ah][aw][3], roughness [V][rh][rw], images [V][H][W][3].
 * sample_cache (may be NULL): filled by the forward, read by a backward with
 * the same camera, seed and texels, which then traces each path once instead
 * of twice (with path records: only the occlusion rays the forward skipped
 * because their light's contribution was zero).  Another cache gives wrong
 * gradients but no access out of bounds.  stream: the hipStream_t the kernels run on, in order with the
 * caller's other work on it (NULL: HIP's null stream, e.g. torch's default
 * stream); the calls are asynchronous. */
lt_status lt_adjoint_forward(lt_adjoint* adj, const lt_camera* camera, uint32_t seed, const float* albedo,
                             const float* roughness, float* images, float* sample_cache, void* stream);
/* Gradients of sum(d_images * images) with respect to the texels; grad_albedo
 * and grad_roughness (shaped like albedo and roughness) are overwritten. */
lt_status lt_adjoint_backward(lt_adjoint* adj, const lt_camera* camera, uint32_t seed, const float* albedo,
                              const float* roughness, const float* d_images, const float* sample_cache,
                              float* grad_albedo, float* grad_roughness, void* stream);
/* Counting mode (no counterpart in the reference; for measurement): while on,
 * forward and backward calls run the counting variants of their kernels and
 * record the work of the call; lt_adjoint_counts waits for the device and
 * returns that of the last call: closest-hit rays, occlusion rays, BVH8 nodes
 * and triangles visited by all queries, and those of the occlusion queries. */
lt_status lt_adjoint_set_counting(lt_adjoint* adj, int32_t on);
lt_status lt_adjoint_counts(lt_adjoint* adj, uint64_t counts[6]);

/* ---- metrics ------------------------------------------------------------------- */

/* Squared-error sum between two device or host RGBA32F images over rgb:
 * S = sum_px |a.rgb - b.rgb|^2.  The reference's RMSE is sqrt(S)/(3*W*H)
 * (src/shaders/rmse/output_rmse.comp:23); standard RMSE is sqrt(S/(3*W*H)). */
lt_status lt_image_sq_error(lt_context* ctx, const float* a, const float* b, size_t n_pixels, uint32_t flags,
                            double* sum_sq);

/* ---- scenes ---------------------------------------------------------------------- */

/* A host-side scene owned by the library: the flattened LumenScene vectors and
 * the camera of its JSON, as produced by a loader or a procedural generator. */
typedef struct lt_host_scene lt_host_scene;

/* Procedural workloads. */
typedef enum lt_procedural_kind {
    LT_SCENE_CORNELL = 0,   /* Cornell box: 5 wall quads, 2 rotated boxes, 0.25-area light, 36 triangles */
    LT_SCENE_TERRAIN = 1,   /* fBm heightfield + instanced boxes, Disney materials, point + area lights  */
    LT_SCENE_CAUSTICS = 2,  /* glass sphere on a diffuse plane, one spot light                           */
    LT_SCENE_DIFF_WALL = 3  /* Cornell box with a tessellated, textured back wall                        */
} lt_procedural_kind;

/* Procedural generator parameters; zero fields take the documented default. */
typedef struct lt_procedural_params {
    uint32_t seed;
    uint32_t terrain_tris;   /* TERRAIN: heightfield triangles (default 1,000,000) */
    uint32_t num_instances;  /* TERRAIN: instanced boxes (default 2,000)           */
    uint32_t num_point_lights; /* TERRAIN: point lights (default 64)              */
    uint32_t sphere_tris;    /* CAUSTICS: sphere triangles (default 20,000)        */
    uint32_t wall_tris;      /* DIFF_WALL: back-wall triangles (default 16,384)    */
    uint32_t tex_res;        /* DIFF_WALL: texture resolution (default 256)        */
    uint32_t reserved[9];
} lt_procedural_params;

lt_status lt_scene_procedural(lt_procedural_kind kind, const lt_procedural_params* params, lt_host_scene** out);
/* Lumen JSON + OBJ loader (src/RayTracer/LumenScene.cpp:231-513). */
lt_status lt_scene_load_json(const char* path, lt_host_scene** out, char* err, size_t err_len);
lt_status lt_scene_free(lt_host_scene* s);
/* Borrowed view of the scene arrays (valid until lt_scene_free). */
lt_status lt_scene_view(const lt_host_scene* s, lt_scene_desc* out);
/* Camera of the scene for a given image size: PerspectiveCamera(fov, 0.01, 1000,
 * W/H, dir, pos) then update_view_matrix (src/Framework/Camera.h:28-38,67-85,
 * 109-121; src/RayTracer/Integrator.cpp:60-72). */
lt_status lt_scene_camera(const lt_host_scene* s, uint32_t width, uint32_t height, lt_camera* out);
/* Replaces the JSON "camera" block (fov in degrees, position, dir;
 * LumenScene.cpp:487-491). */
lt_status lt_scene_set_camera(lt_host_scene* s, float fov, const float pos[3], const float dir[3]);
/* Integrator block of the scene JSON (LumenScene.cpp:237-269): type
 * (lt_integrator_kind), path_length, sky_col, VCM enable_vm / radius_factor,
 * SPPM base_radius. */
typedef struct lt_scene_integrator_cfg {
    int32_t kind;
    int32_t path_length;
    float sky_col[3];
    int32_t enable_vm;
    float radius_factor;
    float base_radius;
} lt_scene_integrator_cfg;
lt_status lt_scene_integrator(const lt_host_scene* s, lt_scene_integrator_cfg* out);

/* ---- OpenEXR image files (host only, no device needed) ------------------- */

/* Stored pixel type of lt_exr_save (the OpenEXR pixel type codes). */
#define LT_EXR_HALF 1
#define LT_EXR_FLOAT 2
/* Writes the RGB of a width x height RGBA32F image (rows top to bottom) as a
 * single-part scanline OpenEXR file: channels B, G, R, no compression, HALF
 * (the reference's choice) or FLOAT.  Replaces ImageUtils::save_exr
 * (src/Framework/ImageUtils.cpp:23-89). */
lt_status lt_exr_save(const char* path, const float* rgba, uint32_t width, uint32_t height, int32_t pixel_type,
                      char* err, size_t err_len);
/* Reads a single-part scanline OpenEXR file (channels R, G, B, A in any subset,
 * HALF / FLOAT / UINT, compression NONE, ZIPS or ZIP) into RGBA32F, rows top to
 * bottom; missing colour channels read 0 and a missing alpha 1.  With
 * rgba == NULL only the size is returned; otherwise capacity (floats) must be
 * at least 4 * width * height.  Replaces ImageUtils::load_exr
 * (src/Framework/ImageUtils.cpp:9-21), the ground-truth loader of the RMSE pass
 * (src/RayTracer/RayTracer.cpp:199-242). */
lt_status lt_exr_load(const char* path, float* rgba, size_t capacity, uint32_t* width, uint32_t* height, char* err,
                      size_t err_len);
/* Decodes a baseline (sequential Huffman, 8-bit) JPEG file into 8-bit RGBA,
 * alpha 255, rows top to bottom: the texture decode of the scene loader,
 * stbi_load(path, &x, &y, &n, 4) in src/RayTracer/LumenScene.cpp:205-212.
 * With rgba == NULL only the size is returned; otherwise capacity (bytes)
 * must be at least 4 * width * height.  Progressive, arithmetic-coded and
 * 12-bit files return LT_ERR_UNSUPPORTED. */
lt_status lt_image_load_rgba8(const char* path, uint8_t* rgba, size_t capacity, uint32_t* width, uint32_t* height,
                              char* err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* LT_H */
