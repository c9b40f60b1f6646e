// integrator_abi.cpp — integrator objects of the C ABI (lt_integrator_*,
// lt_render*, lt_readback_rgba32f, lt_image_sq_error) and the scene entry
// points (lt_scene_*).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "bdpt.h"
#include "context.h"
#include "path.h"
#include "path_wave.h"
#include "scene.h"
#include "sppm.h"
#include "vcm.h"

lt_status launch_sq_error(lt_context* ctx, const float4* a, const float4* b, size_t n, double* d_out);

struct lt_integrator {
    lt_context* ctx = nullptr;
    lt_integrator_kind kind = LT_INTEGRATOR_PATH;
    uint32_t width = 0, height = 0;
    lt_path_config path{};
    lt_bdpt_config bdpt{};
    lt_vcm_config vcm{};
    lt_sppm_config sppm{};
    float4* accum = nullptr;  // output_tex, RGBA32F, row-major
    WaveBuffers wave;         // wavefront state (LT_PATH_WAVEFRONT)
    BdptBuffers bb;           // BDPT subpaths, camera estimates and splats
    VcmBuffers vb;            // VCM light vertices, photon grid and splats
    SppmBuffers sb;           // SPPM per-pixel state, photons and photon table
    int counting = 0;
    int timing = 0;   // per-phase events (lt_render_set_timing)
    int phases = 0;   // phase marks recorded by the last render
    lt_render_stats last{};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool rendered = false;
    // frame graph (lt_render_set_graph): one frame's launch sequence whose
    // kernels read the frame number at d_frame and whose last kernel
    // advances it, replayed once per frame
    int graph_on = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint32_t* d_frame = nullptr;
    bool graph_word_known = false;
    uint32_t graph_word = 0;  // frame number the device word holds
    lt_camera graph_cam{};
    uint32_t graph_seed = 0;
    int graph_counting = 0;
    uint64_t graph_gen = 0;
};

static void drop_graph(lt_integrator* it) {
    if (it->gexec) (void)hipGraphExecDestroy(it->gexec);
    if (it->graph) (void)hipGraphDestroy(it->graph);
    it->gexec = nullptr;
    it->graph = nullptr;
    it->graph_word_known = false;
}

extern "C" {

lt_status lt_integrator_create(lt_context* ctx, lt_integrator_kind kind, const void* cfg, size_t cfg_size,
                               lt_integrator** out) {
    if (!ctx || !cfg || !out) return LT_ERR_INVALID_ARG;
    *out = nullptr;
    uint32_t W = 0, H = 0;
    if (kind == LT_INTEGRATOR_PATH) {
        if (cfg_size != sizeof(lt_path_config)) return lt_fail(ctx, LT_ERR_INVALID_ARG, "cfg_size mismatch");
        const lt_path_config* pc = (const lt_path_config*)cfg;
        if (pc->strategy != LT_PATH_MEGAKERNEL && pc->strategy != LT_PATH_WAVEFRONT &&
            pc->strategy != LT_PATH_PERSISTENT)
            return lt_fail(ctx, LT_ERR_INVALID_ARG, "unknown path strategy");
        if (pc->nee_pick != LT_NEE_REFERENCE && pc->nee_pick != LT_NEE_CORRECTED)
            return lt_fail(ctx, LT_ERR_INVALID_ARG, "unknown nee_pick");
        if (pc->strategy == LT_PATH_WAVEFRONT && pc->max_depth > kWaveMaxDepth)
            return lt_fail(ctx, LT_ERR_UNSUPPORTED, "wavefront path length is limited to 32");
        W = pc->width;
        H = pc->height;
    } else if (kind == LT_INTEGRATOR_BDPT) {
        if (cfg_size != sizeof(lt_bdpt_config)) return lt_fail(ctx, LT_ERR_INVALID_ARG, "cfg_size mismatch");
        const lt_bdpt_config* bc = (const lt_bdpt_config*)cfg;
        if (bc->max_depth < 1 || bc->max_depth > kBdptMaxDepth)
            return lt_fail(ctx, LT_ERR_UNSUPPORTED, "BDPT path length must be in 1..16");
        if (bc->mis != LT_MIS_BALANCE && bc->mis != LT_MIS_POWER)
            return lt_fail(ctx, LT_ERR_INVALID_ARG, "unknown MIS heuristic");
        W = bc->width;
        H = bc->height;
    } else if (kind == LT_INTEGRATOR_VCM) {
        if (cfg_size != sizeof(lt_vcm_config)) return lt_fail(ctx, LT_ERR_INVALID_ARG, "cfg_size mismatch");
        const lt_vcm_config* vc = (const lt_vcm_config*)cfg;
        if (vc->max_depth < 1 || vc->max_depth > kVcmMaxDepth)
            return lt_fail(ctx, LT_ERR_UNSUPPORTED, "VCM path length must be in 1..16");
        if (vc->estimator != LT_VCM_REFERENCE && vc->estimator != LT_VCM_CORRECTED)
            return lt_fail(ctx, LT_ERR_INVALID_ARG, "unknown VCM estimator");
        if (!(vc->radius_factor > 0.0f) || !(vc->alpha > 0.0f && vc->alpha <= 1.0f))
            return lt_fail(ctx, LT_ERR_INVALID_ARG, "VCM radius_factor must be > 0 and alpha in (0, 1]");
        if ((uint64_t)vc->width * vc->height * 10u >= (1ull << 31) ||
            (uint64_t)vc->width * vc->height * (vc->max_depth + 1u) >= (1ull << 32) - 1u)
            return lt_fail(ctx, LT_ERR_UNSUPPORTED, "VCM image too large for 32-bit photon references");
        W = vc->width;
        H = vc->height;
    } else if (kind == LT_INTEGRATOR_SPPM) {
        if (cfg_size != sizeof(lt_sppm_config)) return lt_fail(ctx, LT_ERR_INVALID_ARG, "cfg_size mismatch");
        const lt_sppm_config* sc = (const lt_sppm_config*)cfg;
        if (sc->max_depth < 1 || sc->max_depth > kSppmMaxDepth)
            return lt_fail(ctx, LT_ERR_UNSUPPORTED, "SPPM path length must be in 1..16");
        if (!(sc->base_radius >= 0.0f) || !std::isfinite(sc->base_radius))
            return lt_fail(ctx, LT_ERR_INVALID_ARG, "SPPM base_radius must be finite and >= 0");
        if ((uint64_t)sc->width * sc->height * 10u >= (1ull << 31) ||
            (uint64_t)sc->width * sc->height * (uint64_t)sc->max_depth >= (1ull << 32) - 1u)
            return lt_fail(ctx, LT_ERR_UNSUPPORTED, "SPPM image too large for 32-bit photon references");
        W = sc->width;
        H = sc->height;
    } else {
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "unknown integrator kind");
    }
    if (W == 0 || H == 0 || W > 16384 || H > 16384) return lt_fail(ctx, LT_ERR_INVALID_ARG, "image size out of range");
    lt_integrator* it = new (std::nothrow) lt_integrator;
    if (!it) return LT_ERR_OUT_OF_MEMORY;
    it->ctx = ctx;
    it->kind = kind;
    it->width = W;
    it->height = H;
    if (kind == LT_INTEGRATOR_PATH)
        it->path = *(const lt_path_config*)cfg;
    else if (kind == LT_INTEGRATOR_BDPT)
        it->bdpt = *(const lt_bdpt_config*)cfg;
    else if (kind == LT_INTEGRATOR_VCM)
        it->vcm = *(const lt_vcm_config*)cfg;
    else
        it->sppm = *(const lt_sppm_config*)cfg;
    (void)hipSetDevice(ctx->device);
    size_t bytes = (size_t)W * H * sizeof(float4);
    hipError_t e = hipMalloc(&it->accum, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(it->accum, 0, bytes, ctx->stream);
    if (e == hipSuccess) e = hipEventCreate(&it->e0);
    if (e == hipSuccess) e = hipEventCreate(&it->e1);
    if (e != hipSuccess) {
        lt_status st = lt_hip_check(ctx, e, "lt_integrator_create");
        lt_integrator_destroy(it);
        return st;
    }
    if (kind == LT_INTEGRATOR_BDPT || kind == LT_INTEGRATOR_VCM || kind == LT_INTEGRATOR_SPPM) {
        lt_status st = kind == LT_INTEGRATOR_BDPT  ? bdpt_alloc(ctx, it->bb, W, H, it->bdpt.max_depth)
                       : kind == LT_INTEGRATOR_VCM ? vcm_alloc(ctx, it->vb, W, H, it->vcm.max_depth)
                                                   : sppm_alloc(ctx, it->sb, W, H, it->sppm.max_depth);
        if (st != LT_OK) {
            lt_integrator_destroy(it);
            return st;
        }
    }
    *out = it;
    return LT_OK;
}

lt_status lt_integrator_destroy(lt_integrator* it) {
    if (!it) return LT_OK;
    if (it->ctx) (void)hipStreamSynchronize(it->ctx->stream);
    if (it->accum) (void)hipFree(it->accum);
    wave_free(it->wave);
    bdpt_free(it->bb);
    vcm_free(it->vb);
    sppm_free(it->sb);
    if (it->e0) (void)hipEventDestroy(it->e0);
    if (it->e1) (void)hipEventDestroy(it->e1);
    drop_graph(it);
    if (it->d_frame) (void)hipFree(it->d_frame);
    delete it;
    return LT_OK;
}

// Launches `spp` frames from frame_num (frame_dev == nullptr), or one frame
// whose number is read at frame_dev.
static lt_status launch_frames(lt_integrator* it, const lt_frame_params* p, uint32_t frame_num, uint32_t spp,
                               const uint32_t* frame_dev) {
    lt_context* ctx = it->ctx;
    if (it->kind == LT_INTEGRATOR_BDPT) {
        BdptLaunch L{};
        memcpy(L.view, p->camera.view, sizeof(L.view));
        memcpy(L.projection, p->camera.projection, sizeof(L.projection));
        memcpy(L.inv_view, p->camera.inv_view, sizeof(L.inv_view));
        memcpy(L.inv_projection, p->camera.inv_projection, sizeof(L.inv_projection));
        L.width = it->width;
        L.height = it->height;
        L.max_depth = it->bdpt.max_depth;
        L.power = it->bdpt.mis == LT_MIS_POWER ? 1u : 0u;
        L.frame_dev = frame_dev;
        return launch_bdpt(ctx, it->bb, L, frame_num, spp, p->seed, it->accum, it->counting, ctx->count_buf);
    }
    if (it->kind == LT_INTEGRATOR_VCM) {
        VcmLaunch L{};
        memcpy(L.view, p->camera.view, sizeof(L.view));
        memcpy(L.projection, p->camera.projection, sizeof(L.projection));
        memcpy(L.inv_view, p->camera.inv_view, sizeof(L.inv_view));
        memcpy(L.inv_projection, p->camera.inv_projection, sizeof(L.inv_projection));
        L.width = it->width;
        L.height = it->height;
        L.n_paths = it->width * it->height;
        L.max_depth = it->vcm.max_depth;
        memcpy(L.bmin, ctx->scene.smin, 12);
        memcpy(L.bmax, ctx->scene.smax, 12);
        memcpy(L.sky_col, it->vcm.sky_col, 12);
        L.use_vm = it->vcm.enable_vm ? 1u : 0u;
        L.corrected = it->vcm.estimator == LT_VCM_CORRECTED ? 1u : 0u;
        return launch_vcm(ctx, it->vb, L, it->vcm, frame_num, spp, p->seed, it->accum, it->counting, ctx->count_buf);
    }
    if (it->kind == LT_INTEGRATOR_SPPM) {
        SppmLaunch L{};
        memcpy(L.inv_view, p->camera.inv_view, sizeof(L.inv_view));
        memcpy(L.inv_projection, p->camera.inv_projection, sizeof(L.inv_projection));
        L.width = it->width;
        L.height = it->height;
        L.n_px = it->width * it->height;
        L.max_depth = it->sppm.max_depth;
        L.base_radius = it->sppm.base_radius < 1e-7f ? 1e-7f : it->sppm.base_radius;  // SPPM.cpp:86-88
        memcpy(L.sky_col, it->sppm.sky_col, 12);
        return launch_sppm(ctx, it->sb, L, frame_num, spp, p->seed, it->accum, it->counting, ctx->count_buf);
    }
    PathLaunch L{};
    memcpy(L.inv_view, p->camera.inv_view, sizeof(L.inv_view));
    memcpy(L.inv_projection, p->camera.inv_projection, sizeof(L.inv_projection));
    L.width = it->path.width;
    L.height = it->path.height;
    L.max_depth = it->path.max_depth;
    memcpy(L.sky_col, it->path.sky_col, 12);
    L.direct_lighting = it->path.direct_lighting;
    L.nee_pick = it->path.nee_pick;
    L.frame_start = frame_num;
    L.spp = spp;
    L.seed = p->seed;
    L.frame_dev = frame_dev;
    if (it->path.strategy == LT_PATH_WAVEFRONT)
        return launch_path_wavefront(ctx, it->wave, L, it->accum, it->counting, ctx->count_buf);
    if (it->path.strategy == LT_PATH_PERSISTENT)
        return launch_path_persistent(ctx, L, it->accum, it->counting, ctx->count_buf);
    return launch_path_megakernel(ctx, L, it->accum, it->counting, ctx->count_buf);
}

// Captures one frame into it->gexec.  Buffers the launch functions would
// (re)allocate are sized first: nothing may allocate while capturing.
static lt_status capture_frame(lt_integrator* it, const lt_frame_params* p) {
    lt_context* ctx = it->ctx;
    drop_graph(it);
    if (!it->d_frame) LT_HIP(ctx, hipMalloc(&it->d_frame, 256));
    const size_t pblocks = (size_t)ctx->num_cus * ctx->tuning.path_blocks_per_cu;
    lt_status st = ensure_spill(ctx, it->kind == LT_INTEGRATOR_BDPT ? bdpt_spill_threads(ctx) : pblocks * kTraceBlock);
    if (st == LT_OK && it->kind == LT_INTEGRATOR_PATH && it->path.strategy == LT_PATH_WAVEFRONT)
        st = wave_alloc(ctx, it->wave, it->width * it->height);
    if (st != LT_OK) return st;
    LT_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    st = launch_frames(it, p, 0u, 1u, it->d_frame);
    if (st == LT_OK) st = launch_frame_word(ctx, it->d_frame, 0u, 1, ctx->stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (st != LT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return lt_hip_check(ctx, e, "hipStreamEndCapture");
    it->graph = g;
    LT_HIP(ctx, hipGraphInstantiate(&it->gexec, g, nullptr, nullptr, 0));
    it->graph_cam = p->camera;
    it->graph_seed = p->seed;
    it->graph_counting = it->counting;
    it->graph_gen = ctx->generation;
    return LT_OK;
}

lt_status lt_render(lt_integrator* it, const lt_frame_params* p) {
    if (!it || !p) return LT_ERR_INVALID_ARG;
    lt_context* ctx = it->ctx;
    if (!ctx->bvh.built) return lt_fail(ctx, LT_ERR_NOT_READY, "lt_render before lt_accel_build");
    if (ctx->scene.nlights == 0 || ctx->scene.total_light_triangle_cnt == 0)
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "scene has no lights");
    (void)hipSetDevice(ctx->device);
    const uint32_t spp = p->spp ? p->spp : 1u;
    lt_status st;
    if (it->graph_on) {
        ctx->phase_on = 0;
        const bool stale = !it->gexec || memcmp(&it->graph_cam, &p->camera, sizeof(lt_camera)) != 0 ||
                           it->graph_seed != p->seed || it->graph_counting != it->counting ||
                           it->graph_gen != ctx->generation;
        if (stale && (st = capture_frame(it, p)) != LT_OK) return st;
        LT_HIP(ctx, hipEventRecord(it->e0, ctx->stream));
        if (!it->graph_word_known || it->graph_word != p->frame_num)
            if ((st = launch_frame_word(ctx, it->d_frame, p->frame_num, 0, ctx->stream)) != LT_OK) return st;
        for (uint32_t f = 0; f < spp; f++) LT_HIP(ctx, hipGraphLaunch(it->gexec, ctx->stream));
        it->graph_word_known = true;
        it->graph_word = p->frame_num + spp;
    } else {
        ctx->phase_on = it->timing;
        ctx->phase_n = 0;
        if (it->timing)
            for (hipEvent_t& e : ctx->phase_ev)
                if (!e) LT_HIP(ctx, hipEventCreate(&e));
        LT_HIP(ctx, hipEventRecord(it->e0, ctx->stream));
        st = launch_frames(it, p, p->frame_num, spp, nullptr);
        phase_mark(ctx);  // the end of the last phase (of a one-frame render)
        it->phases = ctx->phase_n;
        ctx->phase_on = 0;
        if (st != LT_OK) return st;
    }
    LT_HIP(ctx, hipEventRecord(it->e1, ctx->stream));
    it->last.frames = spp;
    it->rendered = true;
    return LT_OK;
}

lt_status lt_render_set_graph(lt_integrator* it, int enable) {
    if (!it) return LT_ERR_INVALID_ARG;
    if (enable && it->kind == LT_INTEGRATOR_VCM)
        return lt_fail(it->ctx, LT_ERR_UNSUPPORTED, "VCM frames set their radius and grid on the host");
    if (enable && it->kind == LT_INTEGRATOR_SPPM)
        return lt_fail(it->ctx, LT_ERR_UNSUPPORTED, "SPPM frames set their RNG word and state reset on the host");
    it->graph_on = enable ? 1 : 0;
    if (!enable) drop_graph(it);
    return LT_OK;
}

lt_status lt_render_stats_get(lt_integrator* it, lt_render_stats* s) {
    if (!it || !s) return LT_ERR_INVALID_ARG;
    lt_context* ctx = it->ctx;
    if (!it->rendered) return lt_fail(ctx, LT_ERR_NOT_READY, "no render yet");
    LT_HIP(ctx, hipEventSynchronize(it->e1));
    float ms = 0.0f;
    LT_HIP(ctx, hipEventElapsedTime(&ms, it->e0, it->e1));
    it->last.kernel_ms = ms;
    it->last.closest_rays = it->last.shadow_rays = it->last.nodes_visited = it->last.tris_tested = 0;
    it->last.photons_tested = it->last.photons_gathered = it->last.gather_bytes = 0;
    it->last.shadow_nodes_visited = it->last.shadow_tris_tested = 0;
    if (it->counting) {
        uint64_t h[kCountSlots];
        LT_HIP(ctx, hipMemcpyAsync(h, ctx->count_buf, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        LT_HIP(ctx, hipStreamSynchronize(ctx->stream));
        it->last.closest_rays = h[0];
        it->last.shadow_rays = h[1];
        it->last.nodes_visited = h[2];
        it->last.tris_tested = h[3];
        it->last.photons_tested = h[4];
        it->last.photons_gathered = h[5];
        it->last.gather_bytes = h[6];
        it->last.shadow_nodes_visited = h[7];
        it->last.shadow_tris_tested = h[8];
    }
    it->last.num_phases = 0;
    for (float& v : it->last.phase_ms) v = 0.0f;
    if (it->timing && it->phases > 1 && it->last.frames == 1) {
        it->last.num_phases = (uint32_t)(it->phases - 1);
        for (int i = 0; i + 1 < it->phases; i++)
            LT_HIP(ctx, hipEventElapsedTime(&it->last.phase_ms[i], ctx->phase_ev[i], ctx->phase_ev[i + 1]));
    }
    *s = it->last;
    return LT_OK;
}

lt_status lt_render_set_timing(lt_integrator* it, int enable) {
    if (!it) return LT_ERR_INVALID_ARG;
    it->timing = enable ? 1 : 0;
    return LT_OK;
}

lt_status lt_render_set_counting(lt_integrator* it, int enable) {
    if (!it) return LT_ERR_INVALID_ARG;
    it->counting = enable ? 1 : 0;
    return LT_OK;
}

lt_status lt_accum_reset(lt_integrator* it) {
    if (!it) return LT_ERR_INVALID_ARG;
    lt_context* ctx = it->ctx;
    LT_HIP(ctx, hipMemsetAsync(it->accum, 0, (size_t)it->width * it->height * sizeof(float4), ctx->stream));
    return LT_OK;
}

lt_status lt_readback_rgba32f(lt_integrator* it, float* host, size_t n_floats) {
    if (!it || !host) return LT_ERR_INVALID_ARG;
    lt_context* ctx = it->ctx;
    size_t need = (size_t)it->width * it->height * 4;
    if (n_floats < need) return lt_fail(ctx, LT_ERR_INVALID_ARG, "readback buffer too small");
    LT_HIP(ctx, hipMemcpyAsync(host, it->accum, need * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LT_OK;
}

lt_status lt_image_sq_error(lt_context* ctx, const float* a, const float* b, size_t n_pixels, uint32_t flags,
                            double* sum_sq) {
    if (!ctx || !sum_sq || (n_pixels && (!a || !b))) return LT_ERR_INVALID_ARG;
    *sum_sq = 0.0;
    if (n_pixels == 0) return LT_OK;
    (void)hipSetDevice(ctx->device);
    const float4 *da = (const float4*)a, *db = (const float4*)b;
    float4 *ta = nullptr, *tb = nullptr;
    double* dout = nullptr;
    lt_status st = LT_OK;
    auto cleanup = [&]() {
        if (ta) (void)hipFree(ta);
        if (tb) (void)hipFree(tb);
        if (dout) (void)hipFree(dout);
    };
    hipError_t e = hipMalloc(&dout, sizeof(double));
    if (e == hipSuccess && !(flags & LT_PTR_DEVICE)) {
        e = hipMalloc(&ta, n_pixels * sizeof(float4));
        if (e == hipSuccess) e = hipMalloc(&tb, n_pixels * sizeof(float4));
        if (e == hipSuccess) e = hipMemcpyAsync(ta, a, n_pixels * sizeof(float4), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(tb, b, n_pixels * sizeof(float4), hipMemcpyHostToDevice, ctx->stream);
        da = ta;
        db = tb;
    }
    if (e != hipSuccess) {
        st = lt_hip_check(ctx, e, "lt_image_sq_error");
        cleanup();
        return st;
    }
    st = launch_sq_error(ctx, da, db, n_pixels, dout);
    if (st == LT_OK) {
        e = hipMemcpyAsync(sum_sq, dout, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = lt_hip_check(ctx, e, "lt_image_sq_error");
    }
    cleanup();
    return st;
}

// ------------------------------------------------------------- scenes ----
lt_status lt_scene_procedural(lt_procedural_kind kind, const lt_procedural_params* params, lt_host_scene** out) {
    if (!out) return LT_ERR_INVALID_ARG;
    try {
        return lt::make_procedural(kind, params, out);
    } catch (const std::bad_alloc&) {
        return LT_ERR_OUT_OF_MEMORY;
    } catch (...) {
        return LT_ERR_INVALID_ARG;
    }
}

lt_status lt_scene_load_json(const char* path, lt_host_scene** out, char* err, size_t err_len) {
    if (!path || !out) return LT_ERR_INVALID_ARG;
    std::string e;
    lt_status st;
    try {
        st = lt::load_json_scene(path, out, e);
    } catch (const std::bad_alloc&) {
        st = LT_ERR_OUT_OF_MEMORY;
        e = "out of memory";
    } catch (const std::exception& ex) {
        st = LT_ERR_PARSE;
        e = ex.what();
    }
    if (err && err_len) {
        strncpy(err, e.c_str(), err_len - 1);
        err[err_len - 1] = 0;
    }
    return st;
}

lt_status lt_scene_free(lt_host_scene* s) {
    delete s;
    return LT_OK;
}

lt_status lt_scene_view(const lt_host_scene* s, lt_scene_desc* d) {
    if (!s || !d) return LT_ERR_INVALID_ARG;
    memset(d, 0, sizeof(*d));
    d->vertices = s->vertices.data();
    d->num_vertices = (uint32_t)s->vertices.size();
    d->indices = s->indices.data();
    d->num_indices = (uint32_t)s->indices.size();
    d->prim_infos = s->prims.data();
    d->num_prim_infos = (uint32_t)s->prims.size();
    d->materials = s->materials.data();
    d->num_materials = (uint32_t)s->materials.size();
    d->lights = s->lights.data();
    d->num_lights = (uint32_t)s->lights.size();
    d->textures = s->textures.data();
    d->num_textures = (uint32_t)s->textures.size();
    d->roughness_texture_ids = s->rough_tex.empty() ? nullptr : s->rough_tex.data();
    d->spot_cones = s->spot_cones.empty() ? nullptr : s->spot_cones.data();
    d->total_light_triangle_cnt = s->total_light_triangle_cnt;
    d->total_light_area = s->total_light_area;
    d->dir_light_idx = s->dir_light_idx;
    memcpy(d->scene_min, s->smin, 12);
    memcpy(d->scene_max, s->smax, 12);
    d->world_matrices = s->world.empty() ? nullptr : s->world.data();
    d->index_counts = s->index_counts.empty() ? nullptr : s->index_counts.data();
    return LT_OK;
}

lt_status lt_scene_camera(const lt_host_scene* s, uint32_t width, uint32_t height, lt_camera* out) {
    if (!s || !out || !width || !height) return LT_ERR_INVALID_ARG;
    lt::make_camera(s->cam_fov, s->cam_pos, s->cam_dir, (float)width / (float)height, out);
    return LT_OK;
}

lt_status lt_scene_set_camera(lt_host_scene* s, float fov, const float pos[3], const float dir[3]) {
    if (!s || !pos || !dir || !(fov > 0.0f && fov < 180.0f)) return LT_ERR_INVALID_ARG;
    if (dir[0] == 0.0f && dir[1] == 0.0f && dir[2] == 0.0f) return LT_ERR_INVALID_ARG;
    s->cam_fov = fov;
    memcpy(s->cam_pos, pos, 12);
    memcpy(s->cam_dir, dir, 12);
    return LT_OK;
}

lt_status lt_scene_integrator(const lt_host_scene* s, lt_scene_integrator_cfg* out) {
    if (!s || !out) return LT_ERR_INVALID_ARG;
    *out = s->icfg;
    return LT_OK;
}

}  // extern "C"
