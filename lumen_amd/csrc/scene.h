// scene.h — host-side scene model of the core: the flattened LumenScene
// vectors (src/RayTracer/LumenScene.h:47-99) plus camera and integrator
// settings, built by the JSON/OBJ loader or by the procedural generators.
#pragma once
#include <string>
#include <vector>

#include "../../include/lt.h"

struct lt_host_scene {
    std::vector<lt_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<lt_prim_mesh_info> prims;
    std::vector<std::string> prim_names;
    std::vector<lt_material> materials;
    std::vector<lt_light> lights;
    std::vector<float> spot_cones;  // 2 per light
    std::vector<std::vector<float>> tex_data;
    std::vector<uint32_t> tex_dims;  // width, height per texture
    std::vector<lt_texture> textures;
    std::vector<int32_t> rough_tex;  // per material
    // per prim mesh, empty unless some mesh is an instance: world matrix (16
    // floats, column-major) and index count (instances share index ranges)
    std::vector<float> world;
    std::vector<uint32_t> index_counts;
    uint32_t total_light_triangle_cnt = 0;
    float total_light_area = 0.0f;
    int32_t dir_light_idx = -1;
    float smin[3] = {0, 0, 0}, smax[3] = {0, 0, 0};
    // camera settings of the JSON (LumenScene.cpp:487-491)
    float cam_fov = 45.0f;
    float cam_pos[3] = {0, 0, 0};
    float cam_dir[3] = {0, 0, 1};
    lt_scene_integrator_cfg icfg{};
};

namespace lt {

// Analytic light as parsed from the JSON "lights" array (LumenScene.cpp:493-512).
struct AnalyticLight {
    float pos[3], to[3], L[3];
    uint32_t flags;
    float cos_width, cos_falloff;  // spot cone
};

// Appends a de-indexed triangle mesh (3 vertices per face, local indices
// 0..3T-1, as LumenScene.cpp:302-325 lays out OBJ shapes).
uint32_t add_mesh(lt_host_scene& s, const std::string& name, const std::vector<float>& pos,
                  const std::vector<float>& nrm, const std::vector<float>& uv, uint32_t material);

// Appends an instance of prim mesh `proto`'s geometry (the same index and
// vertex ranges) with its own material and object-to-world matrix (16 floats,
// column-major; LumenPrimMesh::world_matrix, LumenScene.h:32).
uint32_t add_instance(lt_host_scene& s, const std::string& name, uint32_t proto, uint32_t material,
                      const float world[16]);
// index count of prim mesh m
uint32_t mesh_index_count(const lt_host_scene& s, uint32_t m);
// sets prim mesh m's world matrix (16 floats, column-major)
void set_world(lt_host_scene& s, uint32_t m, const float world[16]);

// Lights, light triangle count, light area and scene bounds exactly as
// LumenScene::load_scene:70-129 and compute_scene_dimensions:732-751.
void finalize_scene(lt_host_scene& s, const std::vector<AnalyticLight>& analytic);

// Camera: PerspectiveCamera(fov, 0.01, 1000, aspect, dir, pos) followed by
// Integrator::update_uniform_buffers (Camera.h:28-38,67-85,109-121;
// Integrator.cpp:60-72).
void make_camera(float fov_deg, const float pos[3], const float dir[3], float aspect, lt_camera* out);

lt_status load_json_scene(const char* path, lt_host_scene** out, std::string& err);
// (re)builds s.textures, the lt_texture views of s.tex_data
void texture_views(lt_host_scene& s);
// Decodes a baseline JPEG byte stream into 8-bit RGBA rows, top to bottom
// (image_jpeg.cpp).
lt_status decode_jpeg(const uint8_t* data, size_t n, std::vector<uint8_t>& rgba, uint32_t& w, uint32_t& h,
                      std::string& err);
lt_status make_procedural(lt_procedural_kind kind, const lt_procedural_params* p, lt_host_scene** out);

}  // namespace lt
