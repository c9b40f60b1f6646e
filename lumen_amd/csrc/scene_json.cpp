// scene_json.cpp — Lumen JSON + OBJ scene loader (LumenScene.cpp:231-513).
#include "scene.h"

namespace lt {

lt_status load_json_scene(const char* path, lt_host_scene** out, std::string& err) {
    (void)path;
    *out = nullptr;
    err = "JSON scene loading is not built yet";
    return LT_ERR_UNSUPPORTED;
}

}  // namespace lt
