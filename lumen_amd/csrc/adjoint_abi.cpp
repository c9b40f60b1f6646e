// adjoint_abi.cpp — lt_adjoint_* of the C ABI: validation, private work
// state and launches of the path-replay adjoint kernels (adjoint.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>

#include "adjoint.h"
#include "context.h"

struct lt_adjoint {
    lt_context* ctx = nullptr;
    lt_adjoint_config cfg{};
    uint32_t aw = 0, ah = 0, rw = 0, rh = 0;
    uint32_t tiles_x = 0, n_items = 0;
    uint32_t blocks = 0;
    // private to this object so that it may run on a caller's stream next
    // to the context's own work
    uint32_t* work = nullptr;   // persistent-grid work counter
    uint32_t* spill = nullptr;  // traversal stack spill of the resident threads
    // counting mode: kCountSlots counters of the last forward / backward call
    unsigned long long* counts = nullptr;
    bool counting = false;
};

namespace {

// the texel-driven materials this adjoint differentiates correctly
bool differentiable(const lt_material& m) {
    if (m.bsdf_type == LT_BSDF_DIFFUSE) return true;
    return m.bsdf_type == LT_BSDF_PRINCIPLED && m.metallic == 0.0f && m.spec_trans == 0.0f &&
           m.specular_tint == 0.0f && m.clearcoat == 0.0f && m.flatness == 0.0f && m.anisotropy == 0.0f;
}

lt_status check_scene(lt_context* ctx, const lt_adjoint_config& c, uint32_t& aw, uint32_t& ah, uint32_t& rw,
                      uint32_t& rh) {
    const DeviceScene& s = ctx->scene;
    if (!s.uploaded || !ctx->bvh.built) return lt_fail(ctx, LT_ERR_NOT_READY, "adjoint before scene upload / accel build");
    if (s.total_light_triangle_cnt == 0) return lt_fail(ctx, LT_ERR_INVALID_ARG, "adjoint: the scene has no lights");
    const int nt = (int)s.h_texs.size();
    if (c.albedo_texture < -1 || c.albedo_texture >= nt || c.roughness_texture < -1 || c.roughness_texture >= nt)
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "adjoint: texture id out of range");
    aw = ah = rw = rh = 0;
    if (c.albedo_texture >= 0) {
        aw = s.h_texs[c.albedo_texture].width;
        ah = s.h_texs[c.albedo_texture].height;
    }
    if (c.roughness_texture >= 0) {
        rw = s.h_texs[c.roughness_texture].width;
        rh = s.h_texs[c.roughness_texture].height;
    }
    for (size_t m = 0; m < s.h_mats.size(); m++) {
        const lt_material& mat = s.h_mats[m];
        const bool reads = (c.albedo_texture >= 0 && mat.texture_id == c.albedo_texture) ||
                           (c.roughness_texture >= 0 && s.h_rough[m] == c.roughness_texture);
        if (reads && !differentiable(mat))
            return lt_fail(ctx, LT_ERR_UNSUPPORTED,
                           "adjoint: material " + std::to_string(m) +
                               " reads a differentiated texture but is neither diffuse nor a principled BSDF with "
                               "metallic, spec_trans, specular_tint, clearcoat, flatness and anisotropy all zero");
    }
    return LT_OK;
}

void fill_launch(const lt_adjoint* a, const lt_camera* cam, uint32_t seed, AdjLaunch& L) {
    memcpy(L.inv_view, cam->inv_view, sizeof(L.inv_view));
    memcpy(L.inv_projection, cam->inv_projection, sizeof(L.inv_projection));
    L.width = a->cfg.width;
    L.height = a->cfg.height;
    L.max_depth = a->cfg.max_depth;
    L.spp = a->cfg.spp;
    L.seed = seed;
    L.variants = a->cfg.variants;
    memcpy(L.sky_col, a->cfg.sky_col, sizeof(L.sky_col));
    L.aid = a->cfg.albedo_texture;
    L.rid = a->cfg.roughness_texture;
    L.aw = a->aw;
    L.ah = a->ah;
    L.rw = a->rw;
    L.rh = a->rh;
    L.tiles_x = a->tiles_x;
    L.n_items = a->n_items;
    L.rec_depth = adj_rec_depth(a->cfg.max_depth);
}

// the scene may have been replaced since create: the texture grids the
// kernels index must still be the ones the caller's arrays were sized for
lt_status revalidate(lt_adjoint* a) {
    uint32_t aw, ah, rw, rh;
    lt_status st = check_scene(a->ctx, a->cfg, aw, ah, rw, rh);
    if (st != LT_OK) return st;
    if (aw != a->aw || ah != a->ah || rw != a->rw || rh != a->rh)
        return lt_fail(a->ctx, LT_ERR_INVALID_ARG, "adjoint: the scene's textures changed size since create");
    return LT_OK;
}

}  // namespace

extern "C" {

lt_status lt_adjoint_create(lt_context* ctx, const lt_adjoint_config* cfg, lt_adjoint** out) {
    if (!ctx || !cfg || !out) return LT_ERR_INVALID_ARG;
    *out = nullptr;
    const lt_adjoint_config& c = *cfg;
    if (c.width == 0 || c.height == 0 || c.width > 16384 || c.height > 16384)
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "adjoint: image size out of range");
    if (c.max_depth < 1 || c.max_depth > 64) return lt_fail(ctx, LT_ERR_INVALID_ARG, "adjoint: max_depth must be in 1..64");
    if (c.spp == 0 || c.variants == 0) return lt_fail(ctx, LT_ERR_INVALID_ARG, "adjoint: spp and variants must be >= 1");
    if (c.albedo_texture < 0 && c.roughness_texture < 0)
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "adjoint: no texture to differentiate");
    uint32_t aw, ah, rw, rh;
    lt_status st = check_scene(ctx, c, aw, ah, rw, rh);
    if (st != LT_OK) return st;
    const uint64_t tiles_x = (c.width + 7) / 8, tiles_y = (c.height + 7) / 8;
    const uint64_t items = tiles_x * tiles_y * 64 * c.variants;
    if (items >= (1ull << 32) - 64) return lt_fail(ctx, LT_ERR_UNSUPPORTED, "adjoint: variants x pixels exceed 2^32");
    if ((uint64_t)aw * ah * 3 >= (1ull << 31) || (uint64_t)rw * rh >= (1ull << 31))
        return lt_fail(ctx, LT_ERR_UNSUPPORTED, "adjoint: texture too large");
    lt_adjoint* a = new (std::nothrow) lt_adjoint;
    if (!a) return LT_ERR_OUT_OF_MEMORY;
    a->ctx = ctx;
    a->cfg = c;
    a->aw = aw;
    a->ah = ah;
    a->rw = rw;
    a->rh = rh;
    a->tiles_x = (uint32_t)tiles_x;
    a->n_items = (uint32_t)items;
    a->blocks = (uint32_t)ctx->num_cus * std::max(1u, ctx->tuning.path_blocks_per_cu);
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMalloc(&a->work, 256);
    if (e == hipSuccess)
        e = hipMalloc(&a->spill, (size_t)a->blocks * kTraceBlock * (ltb::kStackTotal - ltb::kLdsStack) * sizeof(uint32_t));
    if (e != hipSuccess) {
        lt_adjoint_destroy(a);
        return lt_hip_check(ctx, e, "lt_adjoint_create: hipMalloc");
    }
    *out = a;
    return LT_OK;
}

lt_status lt_adjoint_destroy(lt_adjoint* a) {
    if (!a) return LT_ERR_INVALID_ARG;
    (void)hipSetDevice(a->ctx->device);
    (void)hipDeviceSynchronize();
    if (a->work) (void)hipFree(a->work);
    if (a->spill) (void)hipFree(a->spill);
    if (a->counts) (void)hipFree(a->counts);
    delete a;
    return LT_OK;
}

lt_status lt_adjoint_texture_size(lt_adjoint* a, uint32_t* aw, uint32_t* ah, uint32_t* rw, uint32_t* rh) {
    if (!a) return LT_ERR_INVALID_ARG;
    if (aw) *aw = a->aw;
    if (ah) *ah = a->ah;
    if (rw) *rw = a->rw;
    if (rh) *rh = a->rh;
    return LT_OK;
}

lt_status lt_adjoint_cache_floats(lt_adjoint* a, size_t* n) {
    if (!a || !n) return LT_ERR_INVALID_ARG;
    *n = adj_cache_floats(a->cfg.spp, a->n_items, adj_rec_depth(a->cfg.max_depth));
    return LT_OK;
}

lt_status lt_adjoint_forward(lt_adjoint* a, const lt_camera* cam, uint32_t seed, const float* albedo,
                             const float* roughness, float* images, float* sample_cache, void* stream) {
    if (!a || !cam || !images) return LT_ERR_INVALID_ARG;
    lt_context* ctx = a->ctx;
    if ((a->cfg.albedo_texture >= 0 && !albedo) || (a->cfg.roughness_texture >= 0 && !roughness))
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "adjoint: texel array is NULL");
    lt_status st = revalidate(a);
    if (st != LT_OK) return st;
    (void)hipSetDevice(ctx->device);
    AdjLaunch L{};
    fill_launch(a, cam, seed, L);
    AdjBuffers B{albedo, roughness, images, sample_cache, nullptr, nullptr, nullptr};
    return launch_adjoint(ctx, L, B, false, a->work, a->spill, a->blocks, a->counting ? a->counts : nullptr,
                          (hipStream_t)stream);
}

lt_status lt_adjoint_backward(lt_adjoint* a, const lt_camera* cam, uint32_t seed, const float* albedo,
                              const float* roughness, const float* d_images, const float* sample_cache,
                              float* grad_albedo, float* grad_roughness, void* stream) {
    if (!a || !cam || !d_images) return LT_ERR_INVALID_ARG;
    lt_context* ctx = a->ctx;
    if ((a->cfg.albedo_texture >= 0 && (!albedo || !grad_albedo)) ||
        (a->cfg.roughness_texture >= 0 && (!roughness || !grad_roughness)))
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "adjoint: texel or gradient array is NULL");
    lt_status st = revalidate(a);
    if (st != LT_OK) return st;
    (void)hipSetDevice(ctx->device);
    AdjLaunch L{};
    fill_launch(a, cam, seed, L);
    AdjBuffers B{albedo, roughness, nullptr, (float*)sample_cache, d_images, grad_albedo, grad_roughness};
    return launch_adjoint(ctx, L, B, true, a->work, a->spill, a->blocks, a->counting ? a->counts : nullptr,
                          (hipStream_t)stream);
}

lt_status lt_adjoint_set_counting(lt_adjoint* a, int32_t on) {
    if (!a) return LT_ERR_INVALID_ARG;
    lt_context* ctx = a->ctx;
    (void)hipSetDevice(ctx->device);
    if (on && !a->counts) {
        hipError_t e = hipMalloc((void**)&a->counts, kCountSlots * sizeof(unsigned long long));
        if (e != hipSuccess) {
            a->counts = nullptr;
            return lt_hip_check(ctx, e, "lt_adjoint_set_counting: hipMalloc");
        }
        LT_HIP(ctx, hipMemset(a->counts, 0, kCountSlots * sizeof(unsigned long long)));
    }
    a->counting = on != 0;
    return LT_OK;
}

lt_status lt_adjoint_counts(lt_adjoint* a, uint64_t counts[6]) {
    if (!a || !counts) return LT_ERR_INVALID_ARG;
    lt_context* ctx = a->ctx;
    if (!a->counts) return lt_fail(ctx, LT_ERR_NOT_READY, "lt_adjoint_counts: counting was never on");
    (void)hipSetDevice(ctx->device);
    unsigned long long h[kCountSlots];
    LT_HIP(ctx, hipDeviceSynchronize());
    LT_HIP(ctx, hipMemcpy(h, a->counts, sizeof(h), hipMemcpyDeviceToHost));
    // closest rays, occlusion rays, nodes and triangles of all queries, those of the occlusion queries
    counts[0] = h[0];
    counts[1] = h[1];
    counts[2] = h[2];
    counts[3] = h[3];
    counts[4] = h[7];
    counts[5] = h[8];
    return LT_OK;
}

}  // extern "C"
