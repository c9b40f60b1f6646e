// lt_render_cli.cpp — headless driver: the RayTracer frame loop
// (src/RayTracer/RayTracer.cpp:383-466) without window or GUI.  Loads a scene,
// creates the integrator through the C++ adapter, renders `spp` frames
// (render + update per frame, as RayTracer::update does) and writes the
// running-mean image: an OpenEXR file when the name ends in .exr (HALF as
// ImageUtils::save_exr writes it, src/Framework/ImageUtils.cpp:23-89, or FLOAT
// with --exr-float), otherwise a PFM (little-endian RGB float, rows bottom to top).
//
//   lt_render_cli --scene cornell|terrain|caustics|diff_wall|<scene.json>
//                 [--integrator path|bdpt|vcm|sppm] [--width W] [--height H] [--spp N]
//                 [--seed S] [--strategy megakernel|persistent|wavefront]
//                 [--mis balance|power] [--estimator reference|corrected] [--alpha A]
//                 [--out image.exr|image.pfm] [--exr-float]
//
// Without --integrator the scene's own integrator block chooses (the JSON
// "integrator" type, LumenScene.cpp:237-269; path for procedural scenes).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "hip_integrator.hpp"

using namespace lumen_amd;

static bool write_pfm(const std::string& path, const OutputTex& t) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    fprintf(f, "PF\n%u %u\n-1.0\n", t.width, t.height);
    for (int y = (int)t.height - 1; y >= 0; y--)
        for (uint32_t x = 0; x < t.width; x++) fwrite(&t.rgba[4 * ((size_t)y * t.width + x)], 4, 3, f);
    fclose(f);
    return true;
}

int main(int argc, char** argv) {
    std::string scene = "cornell", integrator, out;
    bool exr_float = false;
    uint32_t w = 1920, h = 1080, spp = 16;  // 1920x1080: src/main.cpp:14-15
    uint32_t seed = 0;
    IntegratorOptions opts;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() { return i + 1 < argc ? std::string(argv[++i]) : std::string(); };
        if (a == "--scene") scene = next();
        else if (a == "--integrator") integrator = next();
        else if (a == "--width") w = (uint32_t)std::stoul(next());
        else if (a == "--height") h = (uint32_t)std::stoul(next());
        else if (a == "--spp") spp = (uint32_t)std::stoul(next());
        else if (a == "--out") out = next();
        else if (a == "--exr-float") exr_float = true;
        else if (a == "--seed") seed = (uint32_t)std::stoul(next());
        else if (a == "--alpha") opts.vcm_alpha = std::stof(next());
        else if (a == "--strategy") {
            std::string v = next();
            opts.path_strategy = v == "wavefront" ? LT_PATH_WAVEFRONT : v == "persistent" ? LT_PATH_PERSISTENT
                                                                                         : LT_PATH_MEGAKERNEL;
        } else if (a == "--mis") opts.mis = next() == "power" ? LT_MIS_POWER : LT_MIS_BALANCE;
        else if (a == "--estimator") opts.vcm_estimator = next() == "corrected" ? LT_VCM_CORRECTED : LT_VCM_REFERENCE;
        else {
            fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    try {
        std::unique_ptr<LumenScene> sc;
        if (scene == "cornell") sc = LumenScene::procedural(LT_SCENE_CORNELL);
        else if (scene == "terrain") sc = LumenScene::procedural(LT_SCENE_TERRAIN);
        else if (scene == "caustics") sc = LumenScene::procedural(LT_SCENE_CAUSTICS);
        else if (scene == "diff_wall") sc = LumenScene::procedural(LT_SCENE_DIFF_WALL);
        else sc = LumenScene::load(scene);
        lt_context* ctx = nullptr;
        check(lt_context_create(0, &ctx), nullptr, "lt_context_create");
        if (integrator.empty()) integrator = integrator_name((lt_integrator_kind)sc->config().kind);
        auto it = create_integrator(integrator, sc.get(), ctx, w, h, opts);
        it->seed = seed;
        it->init();
        lt_build_stats bs = it->create_accel();
        auto t0 = std::chrono::steady_clock::now();
        for (uint32_t f = 0; f < spp; f++) {
            it->render();
            it->update();
        }
        check(lt_synchronize(ctx), ctx, "lt_synchronize");
        double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const OutputTex& tex = it->output_tex();
        double sum = 0.0;
        for (size_t i = 0; i < tex.rgba.size(); i += 4) sum += tex.rgba[i] + tex.rgba[i + 1] + tex.rgba[i + 2];
        printf("{\"scene\": \"%s\", \"integrator\": \"%s\", \"width\": %u, \"height\": %u, \"spp\": %u, "
               "\"seconds\": %.6f, \"bvh_nodes\": %u, \"bvh_build_ms\": %.3f, \"mean_rgb\": %.6f}\n",
               scene.c_str(), integrator.c_str(), w, h, spp, secs, bs.num_nodes, bs.build_ms, sum / (3.0 * w * h));
        if (out.size() > 4 && out.compare(out.size() - 4, 4, ".exr") == 0) {
            char err[512] = {0};
            if (lt_exr_save(out.c_str(), tex.rgba.data(), tex.width, tex.height,
                            exr_float ? LT_EXR_FLOAT : LT_EXR_HALF, err, sizeof(err)) != LT_OK)
                fprintf(stderr, "%s\n", err);
        } else if (!out.empty() && !write_pfm(out, tex)) {
            fprintf(stderr, "cannot write %s\n", out.c_str());
        }
        it->destroy();
        lt_context_destroy(ctx);
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
