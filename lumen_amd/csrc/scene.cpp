// scene.cpp — scene finalisation, camera matrices and the procedural
// workloads of the core (host C++).
#include "scene.h"

#include <cmath>
#include <cstring>
#include <random>

namespace lt {

// ------------------------------------------------------------------ mat4 ----
// glm conventions: column-major, m[c][r].
struct M4 {
    float m[4][4];
};
static M4 identity() {
    M4 r{};
    for (int i = 0; i < 4; i++) r.m[i][i] = 1.0f;
    return r;
}
static float radians(float d) { return d * 0.01745329251994329576923690768489f; }

// glm::inverse for 4x4, evaluated in double and rounded (ulp-level deviation
// from glm's float cofactor expansion; the kernels only see the rounded result).
static M4 inverse(const M4& a) {
    double m[16], inv[16];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) m[c * 4 + r] = a.m[c][r];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    M4 r{};
    for (int c = 0; c < 4; c++)
        for (int rr = 0; rr < 4; rr++) r.m[c][rr] = (float)(inv[c * 4 + rr] / det);
    return r;
}

// glm::translate
static M4 translate(const M4& m, const float v[3]) {
    M4 r = m;
    for (int k = 0; k < 4; k++) r.m[3][k] = m.m[0][k] * v[0] + m.m[1][k] * v[1] + m.m[2][k] * v[2] + m.m[3][k];
    return r;
}

// glm::rotate(m, angle, axis)
static M4 rotate(const M4& m, float angle, const float v[3]) {
    const float c = cosf(angle), s = sinf(angle);
    float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    float axis[3] = {v[0] / len, v[1] / len, v[2] / len};
    float temp[3] = {(1.0f - c) * axis[0], (1.0f - c) * axis[1], (1.0f - c) * axis[2]};
    float R[3][3];
    R[0][0] = c + temp[0] * axis[0];
    R[0][1] = temp[0] * axis[1] + s * axis[2];
    R[0][2] = temp[0] * axis[2] - s * axis[1];
    R[1][0] = temp[1] * axis[0] - s * axis[2];
    R[1][1] = c + temp[1] * axis[1];
    R[1][2] = temp[1] * axis[2] + s * axis[0];
    R[2][0] = temp[2] * axis[0] + s * axis[1];
    R[2][1] = temp[2] * axis[1] - s * axis[0];
    R[2][2] = c + temp[2] * axis[2];
    M4 r;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 4; k++) r.m[i][k] = m.m[0][k] * R[i][0] + m.m[1][k] * R[i][1] + m.m[2][k] * R[i][2];
    for (int k = 0; k < 4; k++) r.m[3][k] = m.m[3][k];
    return r;
}

// glm::lookAtLH
static M4 look_at_lh(const float eye[3], const float center[3], const float up[3]) {
    float f[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    float fl = sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (float& x : f) x /= fl;
    float s[3] = {up[1] * f[2] - up[2] * f[1], up[2] * f[0] - up[0] * f[2], up[0] * f[1] - up[1] * f[0]};
    float sl = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (float& x : s) x /= sl;
    float u[3] = {f[1] * s[2] - f[2] * s[1], f[2] * s[0] - f[0] * s[2], f[0] * s[1] - f[1] * s[0]};
    M4 r = identity();
    r.m[0][0] = s[0]; r.m[1][0] = s[1]; r.m[2][0] = s[2];
    r.m[0][1] = u[0]; r.m[1][1] = u[1]; r.m[2][1] = u[2];
    r.m[0][2] = f[0]; r.m[1][2] = f[1]; r.m[2][2] = f[2];
    r.m[3][0] = -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]);
    r.m[3][1] = -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]);
    r.m[3][2] = -(f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]);
    return r;
}

// glm::decompose (glm 0.9.9.8) restricted to the rotation: the matrices here
// are rigid (no perspective, unit scale), so the quaternion of the
// orthonormalised upper 3x3 (Graphics Gems II "unmatrix").
static void decompose_quat(const M4& M, float q[4] /* x,y,z,w */) {
    float Row[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Row[i][j] = M.m[i][j] / M.m[3][3];
    auto len = [](const float* v) { return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
    auto dot3 = [](const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto scale1 = [&](float* v) {
        float l = len(v);
        for (int k = 0; k < 3; k++) v[k] = v[k] * (1.0f / l);
    };
    scale1(Row[0]);
    float skz = dot3(Row[0], Row[1]);
    for (int k = 0; k < 3; k++) Row[1][k] = Row[1][k] * 1.0f + Row[0][k] * -skz;
    scale1(Row[1]);
    float sky = dot3(Row[0], Row[2]);
    for (int k = 0; k < 3; k++) Row[2][k] = Row[2][k] * 1.0f + Row[0][k] * -sky;
    float skx = dot3(Row[1], Row[2]);
    for (int k = 0; k < 3; k++) Row[2][k] = Row[2][k] * 1.0f + Row[1][k] * -skx;
    scale1(Row[2]);
    float P[3] = {Row[1][1] * Row[2][2] - Row[1][2] * Row[2][1], Row[1][2] * Row[2][0] - Row[1][0] * Row[2][2],
                  Row[1][0] * Row[2][1] - Row[1][1] * Row[2][0]};
    if (dot3(Row[0], P) < 0.0f)
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) Row[i][k] *= -1.0f;
    float trace = Row[0][0] + Row[1][1] + Row[2][2];
    if (trace > 0.0f) {
        float root = sqrtf(trace + 1.0f);
        q[3] = 0.5f * root;
        root = 0.5f / root;
        q[0] = root * (Row[1][2] - Row[2][1]);
        q[1] = root * (Row[2][0] - Row[0][2]);
        q[2] = root * (Row[0][1] - Row[1][0]);
    } else {
        static const int Next[3] = {1, 2, 0};
        int i = 0;
        if (Row[1][1] > Row[0][0]) i = 1;
        if (Row[2][2] > Row[i][i]) i = 2;
        int j = Next[i], k = Next[j];
        float root = sqrtf(Row[i][i] - Row[j][j] - Row[k][k] + 1.0f);
        q[i] = 0.5f * root;
        root = 0.5f / root;
        q[j] = root * (Row[i][j] + Row[j][i]);
        q[k] = root * (Row[i][k] + Row[k][i]);
        q[3] = root * (Row[j][k] - Row[k][j]);
    }
}

// glm::mat4_cast
static M4 quat_to_mat(const float q[4]) {
    float x = q[0], y = q[1], z = q[2], w = q[3];
    float qxx = x * x, qyy = y * y, qzz = z * z, qxz = x * z, qxy = x * y, qyz = y * z, qwx = w * x, qwy = w * y,
          qwz = w * z;
    M4 r = identity();
    r.m[0][0] = 1.0f - 2.0f * (qyy + qzz);
    r.m[0][1] = 2.0f * (qxy + qwz);
    r.m[0][2] = 2.0f * (qxz - qwy);
    r.m[1][0] = 2.0f * (qxy - qwz);
    r.m[1][1] = 1.0f - 2.0f * (qxx + qzz);
    r.m[1][2] = 2.0f * (qyz + qwx);
    r.m[2][0] = 2.0f * (qxz + qwy);
    r.m[2][1] = 2.0f * (qyz - qwx);
    r.m[2][2] = 1.0f - 2.0f * (qxx + qyy);
    return r;
}

// glm::extractEulerAngleXYZ
static void extract_euler_xyz(const M4& M, float& t1, float& t2, float& t3) {
    float T1 = atan2f(M.m[2][1], M.m[2][2]);
    float C2 = sqrtf(M.m[0][0] * M.m[0][0] + M.m[1][0] * M.m[1][0]);
    float T2 = atan2f(-M.m[2][0], C2);
    float S1 = sinf(T1), C1 = cosf(T1);
    float T3 = atan2f(S1 * M.m[0][2] - C1 * M.m[0][1], C1 * M.m[1][1] - S1 * M.m[1][2]);
    t1 = -T1;
    t2 = -T2;
    t3 = -T3;
}

static void store(const M4& a, float out[16]) {
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) out[c * 4 + r] = a.m[c][r];
}

void make_camera(float fov, const float pos[3], const float dir[3], float aspect, lt_camera* out) {
    const float cam_near = 0.01f, cam_far = 1000.0f;
    // PerspectiveCamera::make_projection_matrix(true), Camera.h:109-116; the
    // matrix starts as the identity, so [3][3] stays 1.
    M4 proj = identity();
    proj.m[0][0] = 1.0f / (aspect * tanf(radians(fov / 2.0f)));
    proj.m[1][1] = -1.0f / (tanf(radians(fov / 2.0f)));
    proj.m[2][2] = cam_far / (cam_near - cam_far);
    proj.m[2][3] = -1.0f;
    proj.m[3][2] = cam_near * cam_far / (cam_near - cam_far);
    // Camera.h:74-84: rotation from lookAtLH via decompose + Euler XYZ (degrees)
    float center[3] = {pos[0] + dir[0], pos[1] + dir[1], pos[2] + dir[2]};
    float up[3] = {0.0f, 1.0f, 0.0f};
    M4 la = look_at_lh(pos, center, up);
    float q[4];
    decompose_quat(la, q);
    float rx, ry, rz;
    extract_euler_xyz(quat_to_mat(q), rx, ry, rz);
    const float to_deg = (float)(180.0 / (double)3.14159265358979323846f);
    rx *= to_deg;
    ry *= to_deg;
    rz *= to_deg;
    // Camera::update_view_matrix, Camera.h:28-38
    const float UP[3] = {0, 1, 0}, RIGHT[3] = {1, 0, 0}, FORWARD[3] = {0, 0, 1};
    M4 res = identity();
    res = translate(res, pos);
    res = rotate(res, radians(ry), UP);
    res = rotate(res, radians(rx), RIGHT);
    res = rotate(res, radians(rz), FORWARD);
    M4 view = inverse(res);
    // Integrator::update_uniform_buffers, Integrator.cpp:60-72
    store(view, out->view);
    store(proj, out->projection);
    store(inverse(view), out->inv_view);
    store(inverse(proj), out->inv_projection);
}

// ------------------------------------------------------------- meshes ----
uint32_t add_mesh(lt_host_scene& s, const std::string& name, const std::vector<float>& pos,
                  const std::vector<float>& nrm, const std::vector<float>& uv, uint32_t material) {
    lt_prim_mesh_info pi{};
    pi.index_offset = (uint32_t)s.indices.size();
    pi.vertex_offset = (uint32_t)s.vertices.size();
    pi.material_index = material;
    size_t nv = pos.size() / 3;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < nv; i++) {
        lt_vertex v{};
        for (int k = 0; k < 3; k++) {
            v.pos[k] = pos[3 * i + k];
            v.normal[k] = nrm.empty() ? 0.0f : nrm[3 * i + k];
            lo[k] = fminf(lo[k], v.pos[k]);
            hi[k] = fmaxf(hi[k], v.pos[k]);
        }
        v.uv0[0] = uv.empty() ? 0.0f : uv[2 * i];
        v.uv0[1] = uv.empty() ? 0.0f : uv[2 * i + 1];
        s.vertices.push_back(v);
        s.indices.push_back((uint32_t)i);
    }
    for (int k = 0; k < 3; k++) {
        pi.min_pos[k] = lo[k];
        pi.max_pos[k] = hi[k];
    }
    s.prims.push_back(pi);
    s.prim_names.push_back(name);
    if (!s.index_counts.empty()) {  // the scene has instances: this mesh's own range, identity transform
        s.index_counts.push_back((uint32_t)nv);
        for (int k = 0; k < 16; k++) s.world.push_back(k % 5 == 0 ? 1.0f : 0.0f);
    }
    return (uint32_t)s.prims.size() - 1;
}

uint32_t mesh_index_count(const lt_host_scene& s, uint32_t m) {
    if (!s.index_counts.empty()) return s.index_counts[m];
    const uint32_t end = (m + 1 < s.prims.size()) ? s.prims[m + 1].index_offset : (uint32_t)s.indices.size();
    return end - s.prims[m].index_offset;
}

// the per-mesh transform and index-count arrays, filled for the meshes so
// far (their own index ranges, identity transforms)
static void ensure_instancing(lt_host_scene& s) {
    if (!s.index_counts.empty()) return;
    std::vector<uint32_t> counts;
    for (uint32_t m = 0; m < s.prims.size(); m++) counts.push_back(mesh_index_count(s, m));
    s.index_counts = counts;
    s.world.assign(16 * s.prims.size(), 0.0f);
    for (size_t m = 0; m < s.prims.size(); m++)
        for (int i = 0; i < 4; i++) s.world[16 * m + 5 * i] = 1.0f;
}

void set_world(lt_host_scene& s, uint32_t m, const float world[16]) {
    ensure_instancing(s);
    memcpy(&s.world[16 * (size_t)m], world, 64);
}

uint32_t add_instance(lt_host_scene& s, const std::string& name, uint32_t proto, uint32_t material,
                      const float world[16]) {
    ensure_instancing(s);
    lt_prim_mesh_info pi = s.prims[proto];
    pi.material_index = material;
    s.index_counts.push_back(s.index_counts[proto]);
    s.world.insert(s.world.end(), world, world + 16);
    s.prims.push_back(pi);
    s.prim_names.push_back(name);
    return (uint32_t)s.prims.size() - 1;
}

void finalize_scene(lt_host_scene& s, const std::vector<AnalyticLight>& analytic) {
    // scene bounds (compute_scene_dimensions, LumenScene.cpp:732-751): the
    // meshes' object-space boxes, since the transformed box Bbox::transform
    // returns (BBox.h:48-61) is discarded there (LumenScene.cpp:735)
    float lo[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f};
    float hi[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
    for (auto& p : s.prims)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], std::min(p.min_pos[k], p.max_pos[k]));
            hi[k] = std::max(hi[k], std::max(p.min_pos[k], p.max_pos[k]));
        }
    int rank = (lo[0] < hi[0]) + (lo[1] < hi[1]) + (lo[2] < hi[2]);
    if (s.prims.empty() || rank != 3) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], -1.0f);
            hi[k] = std::max(hi[k], 1.0f);
        }
    }
    float radius = 0.0f;
    {
        float e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        radius = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) * 0.5f;
    }
    memcpy(s.smin, lo, 12);
    memcpy(s.smax, hi, 12);
    // emissive meshes become area lights (LumenScene.cpp:73-96)
    s.lights.clear();
    s.spot_cones.clear();
    s.total_light_triangle_cnt = 0;
    uint32_t n_mesh = (uint32_t)s.prims.size();
    for (uint32_t m = 0; m < n_mesh; m++) {
        const lt_material& mat = s.materials[s.prims[m].material_index];
        if (mat.emissive_factor[0] > 0 || mat.emissive_factor[1] > 0 || mat.emissive_factor[2] > 0) {
            lt_light L{};
            if (s.world.empty())
                for (int i = 0; i < 4; i++) L.world_matrix[i * 5] = 1.0f;
            else
                memcpy(L.world_matrix, &s.world[16 * (size_t)m], 64);  // LumenScene.cpp:86
            L.num_triangles = mesh_index_count(s, m) / 3;
            L.prim_mesh_idx = m;
            L.light_flags = LT_LIGHT_AREA | LT_LIGHT_FINITE;
            memcpy(L.L, mat.emissive_factor, 12);
            s.lights.push_back(L);
            s.spot_cones.push_back(cosf(3.14159265359f / 6.0f));
            s.spot_cones.push_back(cosf(25.0f * 3.14159265359f / 180.0f));
            s.total_light_triangle_cnt += L.num_triangles;
        }
    }
    // analytic lights (LumenScene.cpp:98-110)
    s.dir_light_idx = -1;
    for (size_t i = 0; i < analytic.size(); i++) {
        const AnalyticLight& a = analytic[i];
        lt_light L{};
        memcpy(L.L, a.L, 12);
        L.light_flags = a.flags;
        memcpy(L.pos, a.pos, 12);
        memcpy(L.to, a.to, 12);
        s.total_light_triangle_cnt++;
        L.world_radius = radius;
        for (int k = 0; k < 3; k++) L.world_center[k] = 0.5f * (hi[k] + lo[k]);
        // LumenScene.cpp:107 tests the directional bits with a mask; the index
        // stored is the index into the analytic list (as the reference does)
        if ((a.flags & LT_LIGHT_DIRECTIONAL) == LT_LIGHT_DIRECTIONAL && (a.flags & 7u) == LT_LIGHT_DIRECTIONAL)
            s.dir_light_idx = (int32_t)i;
        s.lights.push_back(L);
        s.spot_cones.push_back(a.cos_width);
        s.spot_cones.push_back(a.cos_falloff);
    }
    // light area (LumenScene.cpp:111-129)
    float area = 0.0f;
    for (auto& L : s.lights) {
        if ((L.light_flags & 7u) != LT_LIGHT_AREA) continue;
        const lt_prim_mesh_info& pm = s.prims[L.prim_mesh_idx];
        const float* W = L.world_matrix;
        auto world_pos = [&](const float* p, float* o) {  // world_matrix * vec4(p, 1) (LumenScene.cpp:126-128)
            for (int i = 0; i < 3; i++) o[i] = W[i] * p[0] + W[4 + i] * p[1] + W[8 + i] * p[2] + W[12 + i];
        };
        for (uint32_t t = 0; t < L.num_triangles; t++) {
            uint32_t io = pm.index_offset + 3 * t;
            float a[3], b[3], c[3];
            world_pos(s.vertices[s.indices[io] + pm.vertex_offset].pos, a);
            world_pos(s.vertices[s.indices[io + 1] + pm.vertex_offset].pos, b);
            world_pos(s.vertices[s.indices[io + 2] + pm.vertex_offset].pos, c);
            float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            area += 0.5f * sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        }
    }
    s.total_light_area = area;
    if (s.rough_tex.size() < s.materials.size()) s.rough_tex.resize(s.materials.size(), -1);
    s.textures.clear();  // views are built by texture_views once tex_data is complete
}

// ------------------------------------------------------------ helpers ----
static lt_material diffuse(float r, float g, float b) {
    lt_material m{};
    m.albedo[0] = r;
    m.albedo[1] = g;
    m.albedo[2] = b;
    m.bsdf_type = LT_BSDF_DIFFUSE;
    m.bsdf_props = LT_BSDF_FLAG_DIFFUSE | LT_BSDF_FLAG_REFLECTION;
    m.texture_id = -1;
    return m;
}

// principled defaults of LumenScene.cpp:444-475
static lt_material principled(float r, float g, float b, float roughness, float metallic) {
    lt_material m{};
    m.albedo[0] = r;
    m.albedo[1] = g;
    m.albedo[2] = b;
    m.bsdf_type = LT_BSDF_PRINCIPLED;
    m.ior = 1.5f;
    m.roughness = roughness;
    m.metallic = metallic;
    m.sheen_tint = 0.5f;
    m.clearcoat_gloss = 1.0f;
    m.texture_id = -1;
    if (m.roughness < 1.0f) m.bsdf_props |= LT_BSDF_FLAG_REFLECTION;
    if (m.spec_trans > 0.0f) m.bsdf_props |= LT_BSDF_FLAG_TRANSMISSION;
    m.bsdf_props |= m.roughness > 0.08f ? LT_BSDF_FLAG_GLOSSY : LT_BSDF_FLAG_SPECULAR;
    return m;
}

struct MeshBuf {
    std::vector<float> p, n, uv;
    void tri(const float a[3], const float b[3], const float c[3], const float nn[3]) {
        const float* vs[3] = {a, b, c};
        for (auto* v : vs) {
            p.insert(p.end(), v, v + 3);
            n.insert(n.end(), nn, nn + 3);
            uv.push_back(0.0f);
            uv.push_back(0.0f);
        }
    }
    // quad a,b,c,d (counter-clockwise seen from the normal side)
    void quad(const float a[3], const float b[3], const float c[3], const float d[3], const float nn[3]) {
        tri(a, b, c, nn);
        tri(a, c, d, nn);
    }
};

// axis-aligned box of half-extents h, rotated about +y by `deg`, centred at c
static void box(MeshBuf& mb, const float c[3], const float h[3], float deg, const float* quat = nullptr) {
    float R[3][3];
    if (quat) {
        float x = quat[0], y = quat[1], z = quat[2], w = quat[3];
        R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - w * z); R[0][2] = 2 * (x * z + w * y);
        R[1][0] = 2 * (x * y + w * z); R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - w * x);
        R[2][0] = 2 * (x * z - w * y); R[2][1] = 2 * (y * z + w * x); R[2][2] = 1 - 2 * (x * x + y * y);
    } else {
        float a = deg * 3.14159265359f / 180.0f, cs = cosf(a), sn = sinf(a);
        R[0][0] = cs; R[0][1] = 0; R[0][2] = sn;
        R[1][0] = 0; R[1][1] = 1; R[1][2] = 0;
        R[2][0] = -sn; R[2][1] = 0; R[2][2] = cs;
    }
    auto xf = [&](float lx, float ly, float lz, float* o) {
        float l[3] = {lx * h[0], ly * h[1], lz * h[2]};
        for (int i = 0; i < 3; i++) o[i] = c[i] + R[i][0] * l[0] + R[i][1] * l[1] + R[i][2] * l[2];
    };
    auto xn = [&](float lx, float ly, float lz, float* o) {
        for (int i = 0; i < 3; i++) o[i] = R[i][0] * lx + R[i][1] * ly + R[i][2] * lz;
    };
    // 6 faces: (normal, 4 corners ccw)
    const float F[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int f = 0; f < 6; f++) {
        // build a local basis u, v with u x v = n
        float n[3] = {F[f][0], F[f][1], F[f][2]};
        float u[3], v[3];
        if (fabsf(n[0]) > 0.5f) { u[0] = 0; u[1] = 1; u[2] = 0; }
        else if (fabsf(n[1]) > 0.5f) { u[0] = 0; u[1] = 0; u[2] = 1; }
        else { u[0] = 1; u[1] = 0; u[2] = 0; }
        v[0] = n[1] * u[2] - n[2] * u[1];
        v[1] = n[2] * u[0] - n[0] * u[2];
        v[2] = n[0] * u[1] - n[1] * u[0];
        float cs[4][3];
        const float su[4] = {-1, 1, 1, -1}, sv[4] = {-1, -1, 1, 1};
        for (int k = 0; k < 4; k++)
            xf(n[0] + su[k] * u[0] + sv[k] * v[0], n[1] + su[k] * u[1] + sv[k] * v[1],
               n[2] + su[k] * u[2] + sv[k] * v[2], cs[k]);
        float wn[3];
        xn(n[0], n[1], n[2], wn);
        mb.quad(cs[0], cs[1], cs[2], cs[3], wn);
    }
}

static void add_buf(lt_host_scene& s, const std::string& name, const MeshBuf& mb, uint32_t mat) {
    add_mesh(s, name, mb.p, mb.n, mb.uv, mat);
}

// ------------------------------------------------------------ Cornell ----
// Walls of [-1,1]^3 open towards +z; light 0.5 x 0.5 (area 0.25) at y = 0.99.
static void cornell_walls(lt_host_scene& s, uint32_t white, uint32_t red, uint32_t green, bool back_wall) {
    const float A[3] = {-1, -1, -1}, B[3] = {1, -1, -1}, C[3] = {1, -1, 1}, D[3] = {-1, -1, 1};  // floor
    const float E[3] = {-1, 1, -1}, F[3] = {1, 1, -1}, G[3] = {1, 1, 1}, H[3] = {-1, 1, 1};      // ceiling
    {
        MeshBuf m;
        const float n[3] = {0, 1, 0};
        m.quad(A, D, C, B, n);
        add_buf(s, "floor", m, white);
    }
    {
        MeshBuf m;
        const float n[3] = {0, -1, 0};
        m.quad(E, F, G, H, n);
        add_buf(s, "ceiling", m, white);
    }
    if (back_wall) {
        MeshBuf m;
        const float n[3] = {0, 0, 1};
        m.quad(A, B, F, E, n);
        add_buf(s, "back_wall", m, white);
    }
    {
        MeshBuf m;
        const float n[3] = {1, 0, 0};
        m.quad(A, E, H, D, n);
        add_buf(s, "left_wall", m, red);
    }
    {
        MeshBuf m;
        const float n[3] = {-1, 0, 0};
        m.quad(B, C, G, F, n);
        add_buf(s, "right_wall", m, green);
    }
}

static void cornell_boxes_and_light(lt_host_scene& s, uint32_t white, uint32_t light) {
    {
        MeshBuf m;
        const float c[3] = {0.35f, -0.7f, 0.3f}, h[3] = {0.3f, 0.3f, 0.3f};
        box(m, c, h, -17.0f);
        add_buf(s, "short_box", m, white);
    }
    {
        MeshBuf m;
        const float c[3] = {-0.35f, -0.4f, -0.3f}, h[3] = {0.3f, 0.6f, 0.3f};
        box(m, c, h, 17.0f);
        add_buf(s, "tall_box", m, white);
    }
    {
        MeshBuf m;
        const float a[3] = {-0.25f, 0.99f, -0.25f}, b[3] = {0.25f, 0.99f, -0.25f}, c[3] = {0.25f, 0.99f, 0.25f},
                    d[3] = {-0.25f, 0.99f, 0.25f};
        const float n[3] = {0, -1, 0};
        m.quad(a, b, c, d, n);
        add_buf(s, "light", m, light);
    }
}

static lt_host_scene* make_cornell(bool tess_wall, const lt_procedural_params& p) {
    lt_host_scene* s = new lt_host_scene;
    s->materials.push_back(diffuse(0.73f, 0.73f, 0.73f));  // 0 white
    s->materials.push_back(diffuse(0.65f, 0.05f, 0.05f));  // 1 red
    s->materials.push_back(diffuse(0.12f, 0.45f, 0.15f));  // 2 green
    lt_material lm = diffuse(1.0f, 1.0f, 1.0f);           // 3 light (Lumen's JSON light: diffuse, albedo 1)
    lm.emissive_factor[0] = 17.0f;
    lm.emissive_factor[1] = 12.0f;
    lm.emissive_factor[2] = 4.0f;
    s->materials.push_back(lm);
    cornell_walls(*s, 0, 1, 2, !tess_wall);
    if (tess_wall) {
        // back wall as an nx x ny grid of quads (2 triangles each) carrying
        // albedo and roughness textures (seeded smooth value noise)
        uint32_t target = p.wall_tris ? p.wall_tris : 16384;
        uint32_t nx = (uint32_t)floorf(sqrtf((float)target / 2.0f));
        uint32_t ny = (target / 2 + nx - 1) / nx;
        uint32_t res = p.tex_res ? p.tex_res : 256;
        lt_material wm = principled(1.0f, 1.0f, 1.0f, 0.5f, 0.0f);
        wm.texture_id = 0;
        s->materials.push_back(wm);  // 4
        uint32_t wall_mat = 4;
        MeshBuf m;
        const float n[3] = {0, 0, 1};
        for (uint32_t j = 0; j < ny; j++)
            for (uint32_t i = 0; i < nx; i++) {
                float x0 = -1.0f + 2.0f * i / nx, x1 = -1.0f + 2.0f * (i + 1) / nx;
                float y0 = -1.0f + 2.0f * j / ny, y1 = -1.0f + 2.0f * (j + 1) / ny;
                float a[3] = {x0, y0, -1}, b[3] = {x1, y0, -1}, c[3] = {x1, y1, -1}, d[3] = {x0, y1, -1};
                size_t base = m.uv.size();
                m.quad(a, b, c, d, n);
                const float us[6] = {x0, x1, x1, x0, x1, x1}, vs[6] = {y0, y0, y1, y0, y1, y1};
                const float* P[6] = {a, b, c, a, c, d};
                for (int k = 0; k < 6; k++) {
                    m.uv[base + 2 * k] = (P[k][0] + 1.0f) * 0.5f;
                    m.uv[base + 2 * k + 1] = (P[k][1] + 1.0f) * 0.5f;
                }
                (void)us;
                (void)vs;
            }
        add_buf(*s, "back_wall", m, wall_mat);
        // textures: 0 = albedo (rgb), 1 = roughness (x)
        std::mt19937 rng(p.seed ? p.seed : 7u);
        std::uniform_real_distribution<float> U(0.0f, 1.0f);
        const int G = 9;
        float lat[3][G][G];
        for (int c = 0; c < 3; c++)
            for (int a = 0; a < G; a++)
                for (int b = 0; b < G; b++) lat[c][a][b] = U(rng);
        auto smooth = [&](int c, float x, float y) {
            float fx = x * (G - 1), fy = y * (G - 1);
            int ix = std::min((int)fx, G - 2), iy = std::min((int)fy, G - 2);
            float ax = fx - ix, ay = fy - iy;
            ax = ax * ax * (3 - 2 * ax);
            ay = ay * ay * (3 - 2 * ay);
            float t = lat[c][iy][ix] * (1 - ax) + lat[c][iy][ix + 1] * ax;
            float b = lat[c][iy + 1][ix] * (1 - ax) + lat[c][iy + 1][ix + 1] * ax;
            return t * (1 - ay) + b * ay;
        };
        std::vector<float> alb(4 * res * res), rough(4 * res * res);
        for (uint32_t y = 0; y < res; y++)
            for (uint32_t x = 0; x < res; x++) {
                float u = (x + 0.5f) / res, v = (y + 0.5f) / res;
                size_t o = 4 * ((size_t)y * res + x);
                alb[o] = 0.2f + 0.7f * smooth(0, u, v);
                alb[o + 1] = 0.2f + 0.7f * smooth(1, u, v);
                alb[o + 2] = 0.2f + 0.7f * smooth(2, u, v);
                alb[o + 3] = 1.0f;
                float r = 0.1f + 0.85f * smooth((0 + 1) % 3, v, u);
                rough[o] = rough[o + 1] = rough[o + 2] = r;
                rough[o + 3] = 1.0f;
            }
        s->tex_data.push_back(std::move(alb));
        s->tex_data.push_back(std::move(rough));
        s->rough_tex.assign(s->materials.size(), -1);
        s->rough_tex[wall_mat] = 1;
    }
    cornell_boxes_and_light(*s, 0, 3);
    finalize_scene(*s, {});
    s->cam_fov = 45.0f;
    s->cam_pos[0] = 0.0f;
    s->cam_pos[1] = 0.0f;
    s->cam_pos[2] = 3.9f;
    s->cam_dir[0] = 0.0f;
    s->cam_dir[1] = 0.0f;
    s->cam_dir[2] = 1.0f;  // Lumen's camera looks along -dir (see make_camera)
    s->icfg.kind = LT_INTEGRATOR_PATH;
    s->icfg.path_length = 6;
    return s;
}

// ------------------------------------------------------------ terrain ----
static uint32_t hash3(uint32_t x, uint32_t y, uint32_t seed) {
    uint32_t h = x * 374761393u + y * 668265263u + seed * 2246822519u;
    h = (h ^ (h >> 13)) * 1274126177u;
    return h ^ (h >> 16);
}
static float value_noise(float x, float y, uint32_t seed) {
    float fx = floorf(x), fy = floorf(y);
    int ix = (int)fx, iy = (int)fy;
    float ax = x - fx, ay = y - fy;
    ax = ax * ax * (3 - 2 * ax);
    ay = ay * ay * (3 - 2 * ay);
    auto v = [&](int a, int b) { return (hash3((uint32_t)a, (uint32_t)b, seed) & 0xFFFFFF) / 16777215.0f; };
    float t = v(ix, iy) * (1 - ax) + v(ix + 1, iy) * ax;
    float b = v(ix, iy + 1) * (1 - ax) + v(ix + 1, iy + 1) * ax;
    return t * (1 - ay) + b * ay;
}
static float fbm(float x, float y, uint32_t seed) {
    float s = 0.0f, amp = 0.5f, f = 1.0f;
    for (int o = 0; o < 6; o++) {
        s += amp * value_noise(x * f, y * f, seed + o);
        amp *= 0.5f;
        f *= 2.0f;
    }
    return s;
}

static lt_host_scene* make_terrain(const lt_procedural_params& p) {
    lt_host_scene* s = new lt_host_scene;
    const uint32_t seed = p.seed ? p.seed : 1u;
    const uint32_t ntri = p.terrain_tris ? p.terrain_tris : 1000000u;
    const uint32_t ninst = p.num_instances ? p.num_instances : 2000u;
    const uint32_t npl = p.num_point_lights ? p.num_point_lights : 64u;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    std::bernoulli_distribution metal(0.3);
    auto rand_mat = [&]() {
        float r = 0.05f + 0.95f * U(rng);
        float m = metal(rng) ? 1.0f : 0.0f;
        return principled(0.2f + 0.7f * U(rng), 0.2f + 0.7f * U(rng), 0.2f + 0.7f * U(rng), r, m);
    };
    // heightfield: n x n quads over [-50,50]^2, split into 4 x 4 tiles (meshes)
    const uint32_t n = (uint32_t)floorf(sqrtf((float)ntri / 2.0f));
    const float ext = 50.0f, hscale = 12.0f, freq = 1.0f / 12.0f;
    auto height = [&](float x, float z) { return hscale * fbm(x * freq + 100.0f, z * freq + 100.0f, seed); };
    auto normal = [&](float x, float z, float* o) {
        const float e = 2.0f * ext / n;
        float dx = height(x + e, z) - height(x - e, z), dz = height(x, z + e) - height(x, z - e);
        float v[3] = {-dx, 2.0f * e, -dz};
        float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; k++) o[k] = v[k] / l;
    };
    std::vector<float> H((size_t)(n + 1) * (n + 1)), N(3 * (size_t)(n + 1) * (n + 1));
    for (uint32_t j = 0; j <= n; j++)
        for (uint32_t i = 0; i <= n; i++) {
            float x = -ext + 2.0f * ext * i / n, z = -ext + 2.0f * ext * j / n;
            H[(size_t)j * (n + 1) + i] = height(x, z);
            normal(x, z, &N[3 * ((size_t)j * (n + 1) + i)]);
        }
    const uint32_t T = 4;
    for (uint32_t tj = 0; tj < T; tj++)
        for (uint32_t ti = 0; ti < T; ti++) {
            s->materials.push_back(rand_mat());
            uint32_t mat = (uint32_t)s->materials.size() - 1;
            MeshBuf m;
            uint32_t i0 = ti * n / T, i1 = (ti + 1) * n / T, j0 = tj * n / T, j1 = (tj + 1) * n / T;
            for (uint32_t j = j0; j < j1; j++)
                for (uint32_t i = i0; i < i1; i++) {
                    auto P = [&](uint32_t a, uint32_t b, float* o) {
                        o[0] = -ext + 2.0f * ext * a / n;
                        o[1] = H[(size_t)b * (n + 1) + a];
                        o[2] = -ext + 2.0f * ext * b / n;
                    };
                    float a[3], b[3], c[3], d[3];
                    P(i, j, a);
                    P(i + 1, j, b);
                    P(i + 1, j + 1, c);
                    P(i, j + 1, d);
                    const float* na = &N[3 * ((size_t)j * (n + 1) + i)];
                    const float* nb = &N[3 * ((size_t)j * (n + 1) + i + 1)];
                    const float* nc = &N[3 * ((size_t)(j + 1) * (n + 1) + i + 1)];
                    const float* nd = &N[3 * ((size_t)(j + 1) * (n + 1) + i)];
                    const float* vs[6] = {a, c, b, a, d, c};
                    const float* ns[6] = {na, nc, nb, na, nd, nc};
                    for (int k = 0; k < 6; k++) {
                        m.p.insert(m.p.end(), vs[k], vs[k] + 3);
                        m.n.insert(m.n.end(), ns[k], ns[k] + 3);
                        m.uv.push_back((vs[k][0] + ext) / (2 * ext));
                        m.uv.push_back((vs[k][2] + ext) / (2 * ext));
                    }
                }
            add_buf(*s, "terrain", m, mat);
        }
    // instanced boxes: one unit cube geometry, one instance (a mesh with its
    // own material and world matrix T R S: uniform random position on the
    // terrain, uniform random rotation, random size) per box
    uint32_t proto = 0;
    for (uint32_t b = 0; b < ninst; b++) {
        s->materials.push_back(rand_mat());
        uint32_t mat = (uint32_t)s->materials.size() - 1;
        float x = -ext * 0.95f + 1.9f * ext * U(rng), z = -ext * 0.95f + 1.9f * ext * U(rng);
        float sz = 0.3f + 1.2f * U(rng);
        float c[3] = {x, height(x, z) + sz * 0.5f, z};
        // uniform random rotation (Shoemake)
        float u1 = U(rng), u2 = U(rng), u3 = U(rng);
        float q[4] = {sqrtf(1 - u1) * sinf(6.2831853f * u2), sqrtf(1 - u1) * cosf(6.2831853f * u2),
                      sqrtf(u1) * sinf(6.2831853f * u3), sqrtf(u1) * cosf(6.2831853f * u3)};
        const float qx = q[0], qy = q[1], qz = q[2], qw = q[3];
        const float R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy)},
                               {2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx)},
                               {2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)}};
        float W[16] = {0};
        for (int col = 0; col < 3; col++)
            for (int row = 0; row < 3; row++) W[4 * col + row] = R[row][col] * sz;
        W[12] = c[0];
        W[13] = c[1];
        W[14] = c[2];
        W[15] = 1.0f;
        if (b == 0) {  // the cube geometry, as the first box with its own pose
            MeshBuf m;
            const float c0[3] = {0.0f, 0.0f, 0.0f}, h[3] = {0.5f, 0.5f, 0.5f};
            box(m, c0, h, 0.0f);
            proto = add_mesh(*s, "box", m.p, m.n, m.uv, mat);
            set_world(*s, proto, W);
        } else {
            add_instance(*s, "box", proto, mat, W);
        }
    }
    // one area light: 16 x 16 quad facing down at y = 45
    lt_material lm = diffuse(1.0f, 1.0f, 1.0f);
    lm.emissive_factor[0] = 6.0f;
    lm.emissive_factor[1] = 6.0f;
    lm.emissive_factor[2] = 5.0f;
    s->materials.push_back(lm);
    {
        MeshBuf m;
        const float a[3] = {-8, 45, -8}, b[3] = {8, 45, -8}, c[3] = {8, 45, 8}, d[3] = {-8, 45, 8};
        const float nn[3] = {0, -1, 0};
        m.quad(a, b, c, d, nn);
        add_buf(*s, "area_light", m, (uint32_t)s->materials.size() - 1);
    }
    // point lights above the terrain
    std::vector<AnalyticLight> al;
    for (uint32_t i = 0; i < npl; i++) {
        AnalyticLight L{};
        float x = -ext + 2 * ext * U(rng), z = -ext + 2 * ext * U(rng);
        L.pos[0] = x;
        L.pos[1] = height(x, z) + 2.0f + 8.0f * U(rng);
        L.pos[2] = z;
        L.to[0] = x;
        L.to[1] = 0;
        L.to[2] = z;
        for (int k = 0; k < 3; k++) L.L[k] = 20.0f + 60.0f * U(rng);
        L.flags = LT_LIGHT_POINT | LT_LIGHT_FINITE | LT_LIGHT_DELTA;
        L.cos_width = cosf(30.0f * 3.14159265359f / 180.0f);
        L.cos_falloff = cosf(25.0f * 3.14159265359f / 180.0f);
        al.push_back(L);
    }
    finalize_scene(*s, al);
    // Lumen's camera round trip (lookAtLH -> Euler XYZ -> T*Ry*Rx*Rz) looks
    // along (-dir.x, dir.y, -dir.z): this views the terrain from above its +z edge
    s->cam_fov = 55.0f;
    s->cam_pos[0] = 0.0f;
    s->cam_pos[1] = 40.0f;
    s->cam_pos[2] = 58.0f;
    s->cam_dir[0] = 0.0f;
    s->cam_dir[1] = -0.75f;
    s->cam_dir[2] = 1.0f;
    s->icfg.kind = LT_INTEGRATOR_PATH;
    s->icfg.path_length = 12;
    s->icfg.sky_col[0] = 0.25f;
    s->icfg.sky_col[1] = 0.3f;
    s->icfg.sky_col[2] = 0.4f;
    return s;
}

// ----------------------------------------------------------- caustics ----
static lt_host_scene* make_caustics(const lt_procedural_params& p) {
    lt_host_scene* s = new lt_host_scene;
    s->materials.push_back(diffuse(0.8f, 0.8f, 0.8f));  // 0 plane
    lt_material g{};                                     // 1 glass: dielectric, ior 1.5, roughness 0
    g.albedo[0] = g.albedo[1] = g.albedo[2] = 1.0f;
    g.bsdf_type = LT_BSDF_DIELECTRIC;
    g.ior = 1.5f;
    g.bsdf_props = LT_BSDF_FLAG_TRANSMISSION | LT_BSDF_FLAG_REFLECTION | LT_BSDF_FLAG_SPECULAR;
    g.texture_id = -1;
    s->materials.push_back(g);
    {
        MeshBuf m;
        const float a[3] = {-10, 0, -10}, b[3] = {10, 0, -10}, c[3] = {10, 0, 10}, d[3] = {-10, 0, 10};
        const float n[3] = {0, 1, 0};
        m.quad(a, d, c, b, n);
        add_buf(*s, "plane", m, 0);
    }
    {
        // UV sphere, radius 1, centre (0,1,0): slices x stacks quads, poles as triangles
        uint32_t target = p.sphere_tris ? p.sphere_tris : 20000u;
        uint32_t st = (uint32_t)floorf(sqrtf((float)target / 2.0f));
        uint32_t sl = (target + 2 * st) / (2 * st) + 0;
        MeshBuf m;
        auto P = [&](uint32_t i, uint32_t j, float* o, float* nn) {
            float th = 3.14159265359f * j / st, ph = 6.28318530718f * i / sl;
            nn[0] = sinf(th) * cosf(ph);
            nn[1] = cosf(th);
            nn[2] = sinf(th) * sinf(ph);
            o[0] = nn[0];
            o[1] = 1.0f + nn[1];
            o[2] = nn[2];
        };
        for (uint32_t j = 0; j < st; j++)
            for (uint32_t i = 0; i < sl; i++) {
                float a[3], b[3], c[3], d[3], na[3], nb[3], nc[3], nd[3];
                P(i, j, a, na);
                P(i + 1, j, b, nb);
                P(i + 1, j + 1, c, nc);
                P(i, j + 1, d, nd);
                auto put = [&](const float* v, const float* nn) {
                    m.p.insert(m.p.end(), v, v + 3);
                    m.n.insert(m.n.end(), nn, nn + 3);
                    m.uv.push_back(0);
                    m.uv.push_back(0);
                };
                if (j != 0) { put(a, na); put(c, nc); put(b, nb); }
                if (j != st - 1) { put(a, na); put(d, nd); put(c, nc); }
            }
        add_buf(*s, "glass_sphere", m, 1);
    }
    std::vector<AnalyticLight> al(1);
    AnalyticLight& L = al[0];
    L.pos[0] = 2.0f;
    L.pos[1] = 6.0f;
    L.pos[2] = 1.0f;
    L.to[0] = 0.0f;
    L.to[1] = 1.0f;
    L.to[2] = 0.0f;
    L.L[0] = L.L[1] = L.L[2] = 50.0f;
    L.flags = LT_LIGHT_SPOT | LT_LIGHT_FINITE | LT_LIGHT_DELTA;
    L.cos_width = cosf(15.0f * 3.14159265359f / 180.0f);
    L.cos_falloff = cosf(12.5f * 3.14159265359f / 180.0f);
    finalize_scene(*s, al);
    s->cam_fov = 40.0f;
    s->cam_pos[0] = 0.0f;
    s->cam_pos[1] = 3.0f;
    s->cam_pos[2] = 7.0f;
    s->cam_dir[0] = 0.0f;
    s->cam_dir[1] = 0.0f;
    s->cam_dir[2] = 1.0f;
    s->icfg.kind = LT_INTEGRATOR_BDPT;
    s->icfg.path_length = 8;
    return s;
}

lt_status make_procedural(lt_procedural_kind kind, const lt_procedural_params* pp, lt_host_scene** out) {
    lt_procedural_params p{};
    if (pp) p = *pp;
    lt_host_scene* s = nullptr;
    switch (kind) {
        case LT_SCENE_CORNELL: s = make_cornell(false, p); break;
        case LT_SCENE_DIFF_WALL: s = make_cornell(true, p); break;
        case LT_SCENE_TERRAIN: s = make_terrain(p); break;
        case LT_SCENE_CAUSTICS: s = make_caustics(p); break;
        default: return LT_ERR_INVALID_ARG;
    }
    // the procedural textures are square
    s->tex_dims.clear();
    for (auto& td : s->tex_data) {
        const uint32_t r = (uint32_t)lroundf(sqrtf((float)(td.size() / 4)));
        s->tex_dims.push_back(r);
        s->tex_dims.push_back(r);
    }
    texture_views(*s);
    *out = s;
    return LT_OK;
}

void texture_views(lt_host_scene& s) {
    s.textures.clear();
    for (size_t i = 0; i < s.tex_data.size(); i++) {
        lt_texture t;
        t.width = s.tex_dims[2 * i];
        t.height = s.tex_dims[2 * i + 1];
        t.data = s.tex_data[i].data();
        s.textures.push_back(t);
    }
}

}  // namespace lt
