// context.cpp — the C ABI (include/lt.h): context, scene upload, queries,
// integrators and readback.  No exception crosses the ABI.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <set>
#include <tuple>
#include <new>
#include <string>
#include <vector>

#include "context.h"
#include "path.h"
#include "scene.h"

// ------------------------------------------------------------- errors ----
lt_status lt_fail(lt_context* ctx, lt_status st, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return st;
}
lt_status lt_hip_check(lt_context* ctx, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return lt_fail(ctx, e == hipErrorOutOfMemory ? LT_ERR_OUT_OF_MEMORY : LT_ERR_HIP, m);
}

SceneView make_scene_view(const lt_context* ctx) {
    SceneView v{};
    const DeviceScene& s = ctx->scene;
    v.vtx = s.vtx;
    v.idx = s.idx;
    v.prims = s.prims;
    v.mats = s.mats;
    v.lights = s.lights;
    v.cones = s.cones;
    v.texels = s.texels;
    v.texs = s.texs;
    v.rough = s.rough;
    v.xf = s.xf;
    v.lxf = s.lxf;
    v.nodes = ctx->bvh.nodes;
    v.tris = ctx->bvh.tris;
    v.tri_start = s.tri_start;
    v.num_prims = s.nprims;
    v.num_lights = s.nlights;
    v.num_nodes = ctx->bvh.built ? ctx->bvh.n_nodes : 0;
    v.refill_min = ctx->tuning.refill_min;
    v.total_light_triangle_cnt = s.total_light_triangle_cnt;
    v.dir_light_idx = s.dir_light_idx;
    v.total_light_area = s.total_light_area;
    for (int k = 0; k < 3; k++) {
        v.scene_min[k] = s.smin[k];
        v.scene_max[k] = s.smax[k];
    }
    return v;
}

lt_status ensure_spill(lt_context* ctx, size_t threads) {
    if (threads <= ctx->spill_threads) return LT_OK;
    if (ctx->spill) (void)hipFree(ctx->spill);
    ctx->spill = nullptr;
    ctx->spill_threads = 0;
    ctx->generation++;
    LT_HIP(ctx, hipMalloc(&ctx->spill, threads * (ltb::kStackTotal - ltb::kLdsStack) * sizeof(uint32_t)));
    ctx->spill_threads = threads;
    return LT_OK;
}

uint32_t resident_blocks(const void* kernel, uint32_t fixed) {
    if (fixed) return fixed;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kTraceBlock, 0) != hipSuccess || n < 1) return 1u;
    return (uint32_t)n;
}

static void free_scene(DeviceScene& s) {
    void* ptrs[] = {s.vtx, s.idx, s.prims, s.mats, s.lights, s.cones, s.texels, s.texs, s.rough, s.xf, s.lxf, s.tri_start};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    s = DeviceScene{};
}
static void free_bvh(DeviceBVH& b) {
    if (b.nodes) (void)hipFree(b.nodes);
    if (b.tris) (void)hipFree(b.tris);
    b = DeviceBVH{};
}

// Index range of mesh m: its index_offset and either index_counts[m] or up
// to the next mesh's offset.  false if the range is malformed.
static bool mesh_range(const lt_scene_desc* s, uint32_t m, uint64_t& off, uint64_t& cnt) {
    off = s->prim_infos[m].index_offset;
    if (s->index_counts) {
        cnt = s->index_counts[m];
    } else {
        const uint64_t end = (m + 1 < s->num_prim_infos) ? s->prim_infos[m + 1].index_offset : s->num_indices;
        if (end < off) return false;
        cnt = end - off;
    }
    return cnt % 3 == 0 && off + cnt <= s->num_indices;
}

// The transform record of a column-major 4x4 matrix (kXfFloats floats): the
// object-to-world rows (3 x 4) and the normal matrix, the inverse transpose
// of the linear part, from its cofactors in double precision.  false if the
// matrix is not finite or its linear part is singular.
static bool make_xf(const float* m, float* out, bool& identity) {
    identity = true;
    for (int k = 0; k < 16; k++) {
        if (!std::isfinite(m[k])) return false;
        if (m[k] != ((k % 5 == 0) ? 1.0f : 0.0f)) identity = false;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = m[j * 4 + i];
    double a[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = m[j * 4 + i];
    double c[3][3];
    c[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    c[0][1] = -(a[1][0] * a[2][2] - a[1][2] * a[2][0]);
    c[0][2] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    c[1][0] = -(a[0][1] * a[2][2] - a[0][2] * a[2][1]);
    c[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    c[1][2] = -(a[0][0] * a[2][1] - a[0][1] * a[2][0]);
    c[2][0] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    c[2][1] = -(a[0][0] * a[1][2] - a[0][2] * a[1][0]);
    c[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    const double det = a[0][0] * c[0][0] + a[0][1] * c[0][1] + a[0][2] * c[0][2];
    if (!(std::fabs(det) > 1e-30)) return false;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[12 + 3 * i + j] = (float)(c[i][j] / det);
    out[21] = out[22] = out[23] = 0.0f;
    return true;
}

template <class T>
static lt_status upload(lt_context* ctx, T** dst, const T* src, size_t n) {
    *dst = nullptr;
    if (n == 0) return LT_OK;
    LT_HIP(ctx, hipMalloc((void**)dst, n * sizeof(T)));
    LT_HIP(ctx, hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return LT_OK;
}

extern "C" {

lt_status lt_context_create(int device, lt_context** out) {
    if (!out) return LT_ERR_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LT_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return LT_ERR_INVALID_ARG;
    lt_context* ctx = new (std::nothrow) lt_context;
    if (!ctx) return LT_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipMalloc(&ctx->counters, 64 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&ctx->count_buf, kCountSlots * sizeof(uint64_t)) != hipSuccess) {
        lt_context_destroy(ctx);
        return LT_ERR_HIP;
    }
    ctx->num_cus = prop.multiProcessorCount;
    ctx->arch = prop.gcnArchName;
    if (noalloc_scratch_init(ctx) != LT_OK) {
        lt_context_destroy(ctx);
        return LT_ERR_HIP;
    }
    *out = ctx;
    return LT_OK;
}

lt_status lt_context_destroy(lt_context* ctx) {
    if (!ctx) return LT_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_scene(ctx->scene);
    free_bvh(ctx->bvh);
    void* ptrs[] = {ctx->spill, ctx->counters, ctx->ray_buf, ctx->hit_buf, ctx->count_buf};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t& e : ctx->phase_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LT_OK;
}

const char* lt_last_error(const lt_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

lt_status lt_device_info(lt_context* ctx, int* num_cus, char* arch, size_t arch_len) {
    if (!ctx) return LT_ERR_INVALID_ARG;
    if (num_cus) *num_cus = ctx->num_cus;
    if (arch && arch_len) {
        strncpy(arch, ctx->arch.c_str(), arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return LT_OK;
}

// What lt_scene_upload derives from a valid scene on the host.
struct SceneHost {
    std::vector<uint32_t> tri_start;  // first flattened triangle of each mesh (+ the total)
    std::vector<float> xf, lxf;       // per mesh / per light transform records
    bool any_xf = false, any_lxf = false;
};

// Host-side validation: kernels assume every index and reference in range.
static lt_status validate_scene(const lt_scene_desc* s, SceneHost& h, std::string& err) {
    auto bad = [&](lt_status st, const std::string& m) {
        err = m;
        return st;
    };
    if (s->num_indices % 3 != 0) return bad(LT_ERR_INVALID_ARG, "num_indices is not a multiple of 3");
    if ((s->num_indices && (!s->indices || !s->vertices || !s->prim_infos)) || (s->num_materials && !s->materials) ||
        (s->num_lights && !s->lights) || (s->num_textures && !s->textures))
        return bad(LT_ERR_INVALID_ARG, "scene array pointer is NULL");
    if (s->num_indices && s->num_prim_infos == 0) return bad(LT_ERR_INVALID_ARG, "indices without meshes");
    h.tri_start.assign(s->num_prim_infos + 1, 0u);
    std::set<std::tuple<uint64_t, uint64_t, uint32_t>> checked;  // geometry shared by instances: once
    for (uint32_t m = 0; m < s->num_prim_infos; m++) {
        const lt_prim_mesh_info& pi = s->prim_infos[m];
        uint64_t off, cnt;
        if (!mesh_range(s, m, off, cnt) || pi.material_index >= s->num_materials)
            return bad(LT_ERR_INVALID_ARG, "prim_infos[" + std::to_string(m) + "] out of range");
        if ((uint64_t)h.tri_start[m] + cnt / 3 >= (1ull << 31))
            return bad(LT_ERR_UNSUPPORTED, "more than 2^31 triangles over all meshes");
        h.tri_start[m + 1] = h.tri_start[m] + (uint32_t)(cnt / 3);
        if (!checked.insert({off, cnt, pi.vertex_offset}).second) continue;
        for (uint64_t i = off; i < off + cnt; i++)
            if ((uint64_t)s->indices[i] + pi.vertex_offset >= s->num_vertices)
                return bad(LT_ERR_INVALID_ARG, "index " + std::to_string(i) + " out of range");
    }
    if (!s->index_counts && s->num_prim_infos && s->prim_infos[0].index_offset != 0)
        return bad(LT_ERR_INVALID_ARG, "prim_infos[0].index_offset must be 0");
    // a NaN or infinite position would poison the BVH build's bounds and the
    // Morton codes of every primitive
    for (uint32_t v = 0; v < s->num_vertices; v++) {
        const float* p = s->vertices[v].pos;
        if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]))
            return bad(LT_ERR_INVALID_ARG, "vertices[" + std::to_string(v) + "].pos is not finite");
    }
    // transforms: uploaded only where some mesh or area light is not the identity
    if (s->world_matrices) {
        h.xf.resize((size_t)kXfFloats * s->num_prim_infos);
        for (uint32_t m = 0; m < s->num_prim_infos; m++) {
            bool id;
            if (!make_xf(s->world_matrices + 16 * (size_t)m, h.xf.data() + (size_t)kXfFloats * m, id))
                return bad(LT_ERR_INVALID_ARG, "world_matrices[" + std::to_string(m) + "] is singular or not finite");
            h.any_xf |= !id;
        }
    }
    static const float kIdentity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    h.lxf.resize((size_t)kXfFloats * s->num_lights);
    for (uint32_t l = 0; l < s->num_lights; l++) {
        const lt_light& L = s->lights[l];
        const bool area = (L.light_flags & 7u) == LT_LIGHT_AREA;
        bool id;
        // only area lights read their world matrix (commons.glsl:112-149)
        if (!make_xf(area ? L.world_matrix : kIdentity, h.lxf.data() + (size_t)kXfFloats * l, id))
            return bad(LT_ERR_INVALID_ARG, "lights[" + std::to_string(l) + "].world_matrix is singular or not finite");
        h.any_lxf |= !id;
        if (area) {
            if (L.prim_mesh_idx >= s->num_prim_infos)
                return bad(LT_ERR_INVALID_ARG, "light " + std::to_string(l) + ": mesh out of range");
            uint64_t off, cnt;
            mesh_range(s, L.prim_mesh_idx, off, cnt);
            if (L.num_triangles == 0 || 3ull * L.num_triangles > cnt)
                return bad(LT_ERR_INVALID_ARG, "light " + std::to_string(l) + ": triangle count");
        }
    }
    for (uint32_t m = 0; m < s->num_materials; m++) {
        int t = s->materials[m].texture_id;
        int r = s->roughness_texture_ids ? s->roughness_texture_ids[m] : -1;
        if (t >= (int)s->num_textures || r >= (int)s->num_textures)
            return bad(LT_ERR_INVALID_ARG, "material " + std::to_string(m) + ": texture out of range");
    }
    for (uint32_t t = 0; t < s->num_textures; t++)
        if (!s->textures[t].data || !s->textures[t].width || !s->textures[t].height)
            return bad(LT_ERR_INVALID_ARG, "texture " + std::to_string(t) + " is empty");
    return LT_OK;
}

lt_status lt_scene_validate(const lt_scene_desc* s, char* err, size_t err_len) {
    if (!s) return LT_ERR_INVALID_ARG;
    SceneHost h;
    std::string e;
    const lt_status st = validate_scene(s, h, e);
    if (err && err_len) {
        strncpy(err, e.c_str(), err_len - 1);
        err[err_len - 1] = 0;
    }
    return st;
}

lt_status lt_scene_upload(lt_context* ctx, const lt_scene_desc* s) {
    if (!ctx || !s) return LT_ERR_INVALID_ARG;
    SceneHost h;
    std::string e;
    const lt_status vs = validate_scene(s, h, e);
    if (vs != LT_OK) return lt_fail(ctx, vs, e);
    const std::vector<uint32_t>& tri_start = h.tri_start;
    const std::vector<float>&xf = h.xf, &lxf = h.lxf;
    const bool any_xf = h.any_xf, any_lxf = h.any_lxf;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->generation++;
    free_scene(ctx->scene);
    free_bvh(ctx->bvh);
    DeviceScene& d = ctx->scene;
    lt_status st;
    if ((st = upload(ctx, &d.vtx, s->vertices, s->num_vertices)) != LT_OK) return st;
    if ((st = upload(ctx, &d.idx, s->indices, s->num_indices)) != LT_OK) return st;
    if ((st = upload(ctx, &d.prims, s->prim_infos, s->num_prim_infos)) != LT_OK) return st;
    if ((st = upload(ctx, &d.mats, s->materials, s->num_materials)) != LT_OK) return st;
    if ((st = upload(ctx, &d.lights, s->lights, s->num_lights)) != LT_OK) return st;
    std::vector<float> cones(2 * (size_t)s->num_lights);
    for (uint32_t l = 0; l < s->num_lights; l++) {
        cones[2 * l] = s->spot_cones ? s->spot_cones[2 * l] : cosf(3.14159265359f / 6.0f);
        cones[2 * l + 1] = s->spot_cones ? s->spot_cones[2 * l + 1] : cosf(25.0f * 3.14159265359f / 180.0f);
    }
    if ((st = upload(ctx, &d.cones, cones.data(), cones.size())) != LT_OK) return st;
    std::vector<TexDesc> descs(s->num_textures);
    std::vector<float4> pool;
    for (uint32_t t = 0; t < s->num_textures; t++) {
        descs[t] = TexDesc{s->textures[t].width, s->textures[t].height, pool.size()};
        size_t n = (size_t)s->textures[t].width * s->textures[t].height;
        const float4* src = (const float4*)s->textures[t].data;
        pool.insert(pool.end(), src, src + n);
    }
    if ((st = upload(ctx, &d.texs, descs.data(), descs.size())) != LT_OK) return st;
    if ((st = upload(ctx, &d.texels, pool.data(), pool.size())) != LT_OK) return st;
    if (s->roughness_texture_ids)
        if ((st = upload(ctx, &d.rough, s->roughness_texture_ids, s->num_materials)) != LT_OK) return st;
    if (any_xf && (st = upload(ctx, &d.xf, xf.data(), xf.size())) != LT_OK) return st;
    if (any_lxf && (st = upload(ctx, &d.lxf, lxf.data(), lxf.size())) != LT_OK) return st;
    if ((st = upload(ctx, &d.tri_start, tri_start.data(), tri_start.size())) != LT_OK) return st;
    d.ntris = tri_start.back();
    d.nv = s->num_vertices;
    d.ni = s->num_indices;
    d.nprims = s->num_prim_infos;
    d.nmats = s->num_materials;
    d.nlights = s->num_lights;
    d.ntex = s->num_textures;
    d.total_light_triangle_cnt = s->total_light_triangle_cnt;
    d.total_light_area = s->total_light_area;
    d.dir_light_idx = s->dir_light_idx;
    memcpy(d.smin, s->scene_min, 12);
    memcpy(d.smax, s->scene_max, 12);
    d.h_mats.assign(s->materials, s->materials + s->num_materials);
    d.h_rough.assign(s->num_materials, -1);
    if (s->roughness_texture_ids) d.h_rough.assign(s->roughness_texture_ids, s->roughness_texture_ids + s->num_materials);
    d.h_texs = descs;
    LT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    d.uploaded = true;
    return LT_OK;
}

lt_status lt_accel_build(lt_context* ctx, const lt_build_opts* opts, lt_build_stats* stats) {
    if (!ctx) return LT_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    ctx->generation++;
    return bvh_build(ctx, opts, stats);
}

lt_status lt_accel_export(lt_context* ctx, void* nodes, size_t nodes_bytes, void* tris, size_t tris_bytes) {
    if (!ctx) return LT_ERR_INVALID_ARG;
    if (!ctx->bvh.built) return lt_fail(ctx, LT_ERR_NOT_READY, "lt_accel_export before lt_accel_build");
    const size_t nb = (size_t)ctx->bvh.n_nodes * sizeof(ltb::Node8), tb = (size_t)ctx->bvh.n_tris * sizeof(ltb::TriRec);
    if ((nodes && nodes_bytes < nb) || (tris && tris_bytes < tb))
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "lt_accel_export: buffer too small");
    if (nodes && nb) LT_HIP(ctx, hipMemcpyAsync(nodes, ctx->bvh.nodes, nb, hipMemcpyDeviceToHost, ctx->stream));
    if (tris && tb) LT_HIP(ctx, hipMemcpyAsync(tris, ctx->bvh.tris, tb, hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LT_OK;
}

static lt_status staged_trace(lt_context* ctx, const lt_ray* rays, size_t n, void* out, size_t out_elem,
                              uint32_t flags, int any_hit, int count) {
    if (!ctx || (!rays && n) || (!out && n && !count)) return LT_ERR_INVALID_ARG;
    if (!ctx->bvh.built) return lt_fail(ctx, LT_ERR_NOT_READY, "trace before lt_accel_build");
    (void)hipSetDevice(ctx->device);
    if (n == 0) return LT_OK;
    if (flags & LT_PTR_DEVICE) return launch_trace(ctx, rays, n, out, any_hit, count, ctx->count_buf);
    if (ctx->ray_buf_bytes < n * sizeof(lt_ray)) {
        if (ctx->ray_buf) (void)hipFree(ctx->ray_buf);
        ctx->ray_buf = nullptr;
        ctx->ray_buf_bytes = 0;
        LT_HIP(ctx, hipMalloc(&ctx->ray_buf, n * sizeof(lt_ray)));
        ctx->ray_buf_bytes = n * sizeof(lt_ray);
    }
    size_t ob = n * std::max<size_t>(out_elem, 1);
    if (ctx->hit_buf_bytes < ob) {
        if (ctx->hit_buf) (void)hipFree(ctx->hit_buf);
        ctx->hit_buf = nullptr;
        ctx->hit_buf_bytes = 0;
        LT_HIP(ctx, hipMalloc(&ctx->hit_buf, ob));
        ctx->hit_buf_bytes = ob;
    }
    LT_HIP(ctx, hipMemcpyAsync(ctx->ray_buf, rays, n * sizeof(lt_ray), hipMemcpyHostToDevice, ctx->stream));
    lt_status st = launch_trace(ctx, (const lt_ray*)ctx->ray_buf, n, ctx->hit_buf, any_hit, count, ctx->count_buf);
    if (st != LT_OK) return st;
    if (out && !count) LT_HIP(ctx, hipMemcpyAsync(out, ctx->hit_buf, n * out_elem, hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LT_OK;
}

lt_status lt_trace_closest(lt_context* ctx, const lt_ray* rays, size_t n, lt_hit* hits, uint32_t flags) {
    return staged_trace(ctx, rays, n, hits, sizeof(lt_hit), flags, 0, 0);
}

lt_status lt_trace_any(lt_context* ctx, const lt_ray* rays, size_t n, uint8_t* occluded, uint32_t flags) {
    return staged_trace(ctx, rays, n, occluded, 1, flags, 1, 0);
}

lt_status lt_trace_count(lt_context* ctx, const lt_ray* rays, size_t n, int any_hit, uint32_t flags, uint64_t* nodes,
                         uint64_t* tris) {
    if (!nodes || !tris) return LT_ERR_INVALID_ARG;
    lt_status st = staged_trace(ctx, rays, n, nullptr, any_hit ? 1 : sizeof(lt_hit), flags, any_hit, 1);
    if (st != LT_OK) return st;
    uint64_t h[2] = {0, 0};
    if (n) {
        LT_HIP(ctx, hipMemcpyAsync(h, ctx->count_buf, 16, hipMemcpyDeviceToHost, ctx->stream));
        LT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    *nodes = h[0];
    *tris = h[1];
    return LT_OK;
}

lt_status lt_set_tuning(lt_context* ctx, const lt_tuning* t) {
    if (!ctx || !t) return LT_ERR_INVALID_ARG;
    // validate everything before applying anything
    if (t->path_blocks_per_cu > 64 || t->trace_blocks_per_cu > 64 || t->bdpt_blocks_per_cu > 64)
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "blocks per CU must be <= 64");
    if (t->refill_min > 64) return lt_fail(ctx, LT_ERR_INVALID_ARG, "refill_min must be in 1..64");
    if (t->regen_min > 64) return lt_fail(ctx, LT_ERR_INVALID_ARG, "regen_min must be in 1..64");
    if (t->wf_sort > LT_WF_SORT_BSDF) return lt_fail(ctx, LT_ERR_INVALID_ARG, "unknown wf_sort");
    if (t->xcd_bands > LT_XCD_BANDS_ON) return lt_fail(ctx, LT_ERR_INVALID_ARG, "unknown xcd_bands");
    if (t->path_blocks_per_cu) ctx->tuning.path_blocks_per_cu = t->path_blocks_per_cu;
    if (t->bdpt_blocks_per_cu) ctx->tuning.bdpt_blocks_per_cu = t->bdpt_blocks_per_cu;
    if (t->trace_blocks_per_cu) ctx->tuning.trace_blocks_per_cu = t->trace_blocks_per_cu;
    if (t->refill_min) ctx->tuning.refill_min = t->refill_min;
    if (t->regen_min) ctx->tuning.regen_min = t->regen_min;
    if (t->bdpt_chunk_tiles) ctx->tuning.bdpt_chunk_tiles = t->bdpt_chunk_tiles;
    if (t->wf_sort) ctx->tuning.wf_sort = t->wf_sort;
    if (t->xcd_bands) ctx->tuning.xcd_bands = t->xcd_bands;
    ctx->generation++;
    return LT_OK;
}

lt_status lt_get_tuning(lt_context* ctx, lt_tuning* t) {
    if (!ctx || !t) return LT_ERR_INVALID_ARG;
    *t = ctx->tuning;
    return LT_OK;
}

lt_status lt_last_kernel_ms(lt_context* ctx, float* ms) {
    if (!ctx || !ms) return LT_ERR_INVALID_ARG;
    LT_HIP(ctx, hipEventSynchronize(ctx->ev1));
    LT_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return LT_OK;
}

lt_status lt_device_alloc(lt_context* ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return LT_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    LT_HIP(ctx, hipMalloc(dptr, bytes ? bytes : 1));
    return LT_OK;
}
lt_status lt_device_free(lt_context* ctx, void* dptr) {
    if (!ctx) return LT_ERR_INVALID_ARG;
    if (dptr) {
        (void)hipStreamSynchronize(ctx->stream);
        LT_HIP(ctx, hipFree(dptr));
    }
    return LT_OK;
}
lt_status lt_memcpy_h2d(lt_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst || !src) && bytes) return LT_ERR_INVALID_ARG;
    LT_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    LT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LT_OK;
}
lt_status lt_memcpy_d2h(lt_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst || !src) && bytes) return LT_ERR_INVALID_ARG;
    LT_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LT_OK;
}
lt_status lt_synchronize(lt_context* ctx) {
    if (!ctx) return LT_ERR_INVALID_ARG;
    LT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LT_OK;
}

}  // extern "C"

void phase_mark(lt_context* ctx) {
    if (!ctx->phase_on || ctx->phase_n > kMaxPhases) return;
    if (hipEventRecord(ctx->phase_ev[ctx->phase_n], ctx->stream) == hipSuccess) ctx->phase_n++;
}
