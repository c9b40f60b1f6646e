// hip_integrator.hpp — C++ host adapter: Lumen's Integrator contract over the
// C ABI of include/lt.h (header-only, no HIP headers needed by the caller).
//
// Mirrors src/RayTracer/Integrator.h:14-33 (init / render / gui / update /
// destroy / create_accel, public output_tex / updated / frame_num) and the
// factory RayTracer::create_integrator (src/RayTracer/RayTracer.cpp:244-282).
// Failures are turned into std::runtime_error, as LUMEN_ERROR does
// (src/Framework/Logger.h:22-26); nothing throws across the C ABI itself.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lt.h"

namespace lumen_amd {

inline void check(lt_status st, lt_context* ctx, const char* what) {
    if (st != LT_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string((int)st) +
                                 "): " + (ctx ? lt_last_error(ctx) : ""));
}

// The host scene (LumenScene's flattened vectors) owned by the library.
class LumenScene {
   public:
    explicit LumenScene(lt_host_scene* s) : s_(s) { check(lt_scene_view(s_, &desc_), nullptr, "lt_scene_view"); }
    static std::unique_ptr<LumenScene> load(const std::string& json_path) {
        lt_host_scene* s = nullptr;
        char err[512] = {0};
        lt_status st = lt_scene_load_json(json_path.c_str(), &s, err, sizeof(err));
        if (st != LT_OK) throw std::runtime_error("lt_scene_load_json(" + json_path + "): " + err);
        return std::unique_ptr<LumenScene>(new LumenScene(s));
    }
    static std::unique_ptr<LumenScene> procedural(lt_procedural_kind kind, uint32_t seed = 0) {
        lt_procedural_params p{};
        p.seed = seed;
        lt_host_scene* s = nullptr;
        check(lt_scene_procedural(kind, &p, &s), nullptr, "lt_scene_procedural");
        return std::unique_ptr<LumenScene>(new LumenScene(s));
    }
    ~LumenScene() { lt_scene_free(s_); }
    LumenScene(const LumenScene&) = delete;
    LumenScene& operator=(const LumenScene&) = delete;
    const lt_scene_desc& desc() const { return desc_; }
    lt_camera camera(uint32_t w, uint32_t h) const {
        lt_camera c;
        check(lt_scene_camera(s_, w, h, &c), nullptr, "lt_scene_camera");
        return c;
    }
    lt_scene_integrator_cfg config() const {
        lt_scene_integrator_cfg c;
        check(lt_scene_integrator(s_, &c), nullptr, "lt_scene_integrator");
        return c;
    }

   private:
    lt_host_scene* s_;
    lt_scene_desc desc_{};
};

// Integrator::output_tex: an RGBA32F image readable on the host.
struct OutputTex {
    uint32_t width = 0, height = 0;
    std::vector<float> rgba;  // row-major, 4 floats per texel
};

// The abstract plugin interface (src/RayTracer/Integrator.h:14-33).
class Integrator {
   public:
    Integrator(LumenScene* scene, lt_context* ctx, uint32_t w, uint32_t h) : lumen_scene(scene), ctx_(ctx), w_(w), h_(h) {}
    virtual ~Integrator() { destroy(); }
    virtual void init() = 0;
    virtual void render() = 0;
    virtual bool gui() { return false; }
    // Integrator::update + Path::update (Path.cpp:61-68): frame_num advances,
    // and restarts from 0 when the camera / settings changed (`updated`).
    virtual bool update() {
        frame_num++;
        bool was = updated;
        if (updated) {
            frame_num = 0;
            updated = false;
        }
        return was;
    }
    virtual void destroy() {
        if (it_) lt_integrator_destroy(it_);
        it_ = nullptr;
    }
    virtual lt_build_stats create_accel() {
        lt_build_stats s{};
        check(lt_accel_build(ctx_, nullptr, &s), ctx_, "lt_accel_build");
        return s;
    }
    // Reads output_tex back (synchronises).
    const OutputTex& output_tex() {
        tex_.width = w_;
        tex_.height = h_;
        tex_.rgba.resize((size_t)w_ * h_ * 4);
        check(lt_readback_rgba32f(it_, tex_.rgba.data(), tex_.rgba.size()), ctx_, "lt_readback_rgba32f");
        return tex_;
    }
    bool updated = false;
    uint32_t frame_num = 0;
    uint32_t seed = 0;

   protected:
    LumenScene* lumen_scene;
    lt_context* ctx_;
    uint32_t w_, h_;
    lt_integrator* it_ = nullptr;
    OutputTex tex_;
};

// The unidirectional path tracer (src/RayTracer/Path.{h,cpp}).
class Path final : public Integrator {
   public:
    Path(LumenScene* scene, lt_context* ctx, uint32_t w, uint32_t h, uint32_t strategy = LT_PATH_MEGAKERNEL)
        : Integrator(scene, ctx, w, h), strategy_(strategy) {}
    void init() override {
        check(lt_scene_upload(ctx_, &lumen_scene->desc()), ctx_, "lt_scene_upload");
        lt_scene_integrator_cfg sc = lumen_scene->config();
        lt_path_config cfg{};
        cfg.width = w_;
        cfg.height = h_;
        cfg.max_depth = sc.path_length > 0 ? sc.path_length : 6;  // SceneConfig.h:16
        for (int k = 0; k < 3; k++) cfg.sky_col[k] = sc.sky_col[k];
        cfg.direct_lighting = 1;  // Path.h: direct_lighting = true
        cfg.strategy = strategy_;
        check(lt_integrator_create(ctx_, LT_INTEGRATOR_PATH, &cfg, sizeof(cfg), &it_), ctx_, "lt_integrator_create");
        camera_ = lumen_scene->camera(w_, h_);
        frame_num = 0;
    }
    void render() override {
        lt_frame_params p{};
        p.camera = camera_;
        p.frame_num = frame_num;
        p.seed = seed;
        p.spp = 1;
        check(lt_render(it_, &p), ctx_, "lt_render");
    }

   private:
    uint32_t strategy_;
    lt_camera camera_{};
};

// Bidirectional path tracing (src/RayTracer/BDPT.{h,cpp}): subpaths of up to
// path_length + 1 vertices (BDPT.cpp:8-18), MIS with the reference's balance
// heuristic (calc_mis_weight, bdpt_commons.glsl:272-473) or the power heuristic.
class BDPT final : public Integrator {
   public:
    BDPT(LumenScene* scene, lt_context* ctx, uint32_t w, uint32_t h, uint32_t mis = LT_MIS_BALANCE)
        : Integrator(scene, ctx, w, h), mis_(mis) {}
    void init() override {
        check(lt_scene_upload(ctx_, &lumen_scene->desc()), ctx_, "lt_scene_upload");
        lt_scene_integrator_cfg sc = lumen_scene->config();
        lt_bdpt_config cfg{};
        cfg.width = w_;
        cfg.height = h_;
        cfg.max_depth = sc.path_length > 0 ? sc.path_length : 6;  // SceneConfig.h:16
        cfg.mis = mis_;
        check(lt_integrator_create(ctx_, LT_INTEGRATOR_BDPT, &cfg, sizeof(cfg), &it_), ctx_, "lt_integrator_create");
        camera_ = lumen_scene->camera(w_, h_);
        frame_num = 0;  // BDPT.cpp:42
    }
    // BDPT::render (BDPT.cpp:55-91): the reference reseeds each frame with
    // rand() (pc_ray.time); here the RNG frame word is frame_num ^ seed, so a
    // run is reproducible.
    void render() override {
        lt_frame_params p{};
        p.camera = camera_;
        p.frame_num = frame_num;
        p.seed = seed;
        p.spp = 1;
        check(lt_render(it_, &p), ctx_, "lt_render");
    }

   private:
    uint32_t mis_;
    lt_camera camera_{};
};

// Vertex connection and merging (src/RayTracer/VCM.{h,cpp}): the scene's
// VCMConfig (enable_vm, radius_factor; SceneConfig.h:38-42), the merge radius
// r_0 (n + 1)^(-(1 - alpha) / 2) with the reference's alpha = 2/3 by default
// (VCM.cpp:140-141), and VCM::update's frame counter (VCM.cpp:273-280).
class VCM final : public Integrator {
   public:
    VCM(LumenScene* scene, lt_context* ctx, uint32_t w, uint32_t h, uint32_t estimator = LT_VCM_REFERENCE,
        float alpha = 2.0f / 3.0f)
        : Integrator(scene, ctx, w, h), estimator_(estimator), alpha_(alpha) {}
    void init() override {
        check(lt_scene_upload(ctx_, &lumen_scene->desc()), ctx_, "lt_scene_upload");
        lt_scene_integrator_cfg sc = lumen_scene->config();
        lt_vcm_config cfg{};
        cfg.width = w_;
        cfg.height = h_;
        cfg.max_depth = sc.path_length > 0 ? sc.path_length : 6;
        for (int k = 0; k < 3; k++) cfg.sky_col[k] = sc.sky_col[k];
        cfg.radius_factor = sc.radius_factor > 0.0f ? sc.radius_factor : 0.025f;  // SceneConfig.h:39
        cfg.alpha = alpha_;
        cfg.enable_vm = sc.enable_vm ? 1u : 0u;
        cfg.estimator = estimator_;
        check(lt_integrator_create(ctx_, LT_INTEGRATOR_VCM, &cfg, sizeof(cfg), &it_), ctx_, "lt_integrator_create");
        camera_ = lumen_scene->camera(w_, h_);
        frame_num = 0;  // VCM.cpp:105
    }
    void render() override {
        lt_frame_params p{};
        p.camera = camera_;
        p.frame_num = frame_num;
        p.seed = seed;
        p.spp = 1;
        check(lt_render(it_, &p), ctx_, "lt_render");
    }

   private:
    uint32_t estimator_;
    float alpha_;
    lt_camera camera_{};
};

// Stochastic progressive photon mapping (src/RayTracer/SPPM.{h,cpp}): the
// scene's SPPMConfig base_radius (SceneConfig.h:33-36) and SPPM::update's
// frame counter (SPPM.cpp:178-185); the per-pixel state restarts at frame 0.
class SPPM final : public Integrator {
   public:
    SPPM(LumenScene* scene, lt_context* ctx, uint32_t w, uint32_t h) : Integrator(scene, ctx, w, h) {}
    void init() override {
        check(lt_scene_upload(ctx_, &lumen_scene->desc()), ctx_, "lt_scene_upload");
        lt_scene_integrator_cfg sc = lumen_scene->config();
        lt_sppm_config cfg{};
        cfg.width = w_;
        cfg.height = h_;
        cfg.max_depth = sc.path_length > 0 ? sc.path_length : 6;
        for (int k = 0; k < 3; k++) cfg.sky_col[k] = sc.sky_col[k];
        cfg.base_radius = sc.base_radius > 0.0f ? sc.base_radius : 0.03f;  // SceneConfig.h:34
        check(lt_integrator_create(ctx_, LT_INTEGRATOR_SPPM, &cfg, sizeof(cfg), &it_), ctx_, "lt_integrator_create");
        camera_ = lumen_scene->camera(w_, h_);
        frame_num = 0;  // SPPM.cpp:60
    }
    void render() override {
        lt_frame_params p{};
        p.camera = camera_;
        p.frame_num = frame_num;
        p.seed = seed;
        p.spp = 1;
        check(lt_render(it_, &p), ctx_, "lt_render");
    }

   private:
    lt_camera camera_{};
};

// Options of the factory beyond the scene's own integrator block.
struct IntegratorOptions {
    uint32_t path_strategy = LT_PATH_MEGAKERNEL;  // Path execution form
    uint32_t mis = LT_MIS_BALANCE;                // BDPT heuristic
    uint32_t vcm_estimator = LT_VCM_REFERENCE;    // VCM estimator
    float vcm_alpha = 2.0f / 3.0f;                // VCM radius reduction (VCM.cpp:141)
};

// RayTracer::create_integrator, keyed by the IntegratorType names
// (src/RayTracer/SceneConfig.h:13) that this core builds.
inline std::unique_ptr<Integrator> create_integrator(const std::string& name, LumenScene* scene, lt_context* ctx,
                                                     uint32_t w, uint32_t h,
                                                     const IntegratorOptions& o = IntegratorOptions()) {
    if (name == "path" || name == "Path")
        return std::unique_ptr<Integrator>(new Path(scene, ctx, w, h, o.path_strategy));
    if (name == "bdpt" || name == "BDPT") return std::unique_ptr<Integrator>(new BDPT(scene, ctx, w, h, o.mis));
    if (name == "vcm" || name == "VCM")
        return std::unique_ptr<Integrator>(new VCM(scene, ctx, w, h, o.vcm_estimator, o.vcm_alpha));
    if (name == "sppm" || name == "SPPM") return std::unique_ptr<Integrator>(new SPPM(scene, ctx, w, h));
    throw std::runtime_error("integrator '" + name + "' is not built");
}

// The IntegratorType name of the scene's JSON integrator block (lt_integrator_kind).
inline std::string integrator_name(lt_integrator_kind k) {
    switch (k) {
        case LT_INTEGRATOR_BDPT: return "bdpt";
        case LT_INTEGRATOR_VCM: return "vcm";
        case LT_INTEGRATOR_SPPM: return "sppm";
        default: return "path";
    }
}

}  // namespace lumen_amd
