// oracle.h — CPU restatement of the reference light-transport path (TEST INFRASTRUCTURE).
//
// This directory is the parity oracle: a scalar, host-only restatement of the
// reference's GLSL (yuphin/Lumen src/shaders/*) used by tests/, the smoke check
// and bench.py's cpu_baseline leg ONLY.  Nothing in the product (lumen_amd/)
// includes, links or calls it.
//
// Arithmetic rules: everything is float32 like the GLSL (literals carry the f
// suffix, GLSL's unsuffixed literals are float), compiled with
// -ffp-contract=off so that no multiply-add is fused behind our back.  The
// traversal restatement (watertight test) is bit-exact with the device; the
// shading restatement uses libm transcendentals, so images agree with the
// device within float tolerance, not bitwise.
//
// Pinning status per function (how each is checked besides by the device):
//   pcg4d / rand          : closed-form known answers + uniformity (tests/test_oracle_kat.py)
//   offset_ray(2)         : bit-level known answers of the RT Gems ch.6 formula
//   concentric disk, ONB  : closed-form points, orthonormality invariants
//   watertight hit        : brute force = host BVH traversal; closed-form hits on axis quads
//   BSDFs                 : Lambert albedo/pi, Fresnel normal incidence, GGX normalisation,
//                           sample/eval/pdf consistency, white furnace <= 1
//   lights                : triangle sampling pdf = 1/A, spot falloff closed form
//   Path integrator       : furnace scene closed form; analytic direct lighting of a plane
//   BDPT                  : furnace closed form (balance and power MIS), spot-lit plane closed
//                           form, agreement with Path on a sealed scene (tests/test_oracle_bdpt.py)
//   VCM                   : furnace closed form (both estimators, merging on/off), corrected
//                           estimator agrees with Path, photon grid = brute-force radius search,
//                           radius schedule closed form (tests/test_oracle_vcm.py)
//   adjoint               : BSDF derivatives = central differences of eval_bsdf, forward =
//                           Path frames, texel gradients = central differences of the forward
//                           (tests/test_oracle_adjoint.py)
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../include/lt.h"

namespace ora {

// A genuine 32-bit float, also in the float64 build of oracle_f64.cpp (which
// widens the keyword `float`; a float literal keeps its type there): the
// bit-level steps (uint_to_float, the integer ulp push of offset_ray) use it.
typedef __typeof__(1.0f) f32;

constexpr float PI = 3.14159265359f;       // utils.glsl:4
constexpr float TWO_PI = 6.28318530718f;   // utils.glsl:5
constexpr float INV_PI = 1.0f / PI;        // utils.glsl:6 (float division)
constexpr float EPS = 0.001f;              // utils.glsl:8

struct vec2 {
    float x, y;
};
struct vec3 {
    float x, y, z;
    float& operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
};
struct vec4 {
    float x, y, z, w;
};
struct uvec4 {
    uint32_t x, y, z, w;
};

inline vec2 v2(float a, float b) { return {a, b}; }
inline vec3 v3(float a) { return {a, a, a}; }
inline vec3 v3(float a, float b, float c) { return {a, b, c}; }
inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator*(float s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3 operator/(vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline vec3& operator+=(vec3& a, vec3 b) { a = a + b; return a; }
inline vec3& operator*=(vec3& a, vec3 b) { a = a * b; return a; }
inline vec3& operator*=(vec3& a, float s) { a = a * s; return a; }
inline vec3& operator/=(vec3& a, float s) { a = a / s; return a; }
// GLSL dot/length/normalize evaluated left to right without fusion.
inline float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(vec3 a, vec3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float length(vec3 a) { return sqrtf(dot(a, a)); }
inline vec3 normalize(vec3 a) { return a * (1.0f / length(a)); }
inline vec3 reflect(vec3 i, vec3 n) { return i - 2.0f * dot(n, i) * n; }
inline vec3 vabs(vec3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
inline vec3 vsqrt(vec3 a) { return {sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)}; }
inline vec3 mix(vec3 a, vec3 b, float t) { return a * (1.0f - t) + b * t; }
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
inline float luminance(vec3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }  // utils.glsl:99
inline vec3 load3(const float* p) { return {p[0], p[1], p[2]}; }
#ifdef ORA_F64
inline vec3 load3(const f32* p) { return {p[0], p[1], p[2]}; }
#endif

// Column-major 4x4 (glm): m[c*4 + r].
template <class R>  // R = the matrix's element type (float32 in the wire format)
inline vec4 mul(const R* m, vec4 v) {
    vec4 r;
    r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
    r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
    r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
    r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
    return r;
}

// ---- RNG (utils.glsl:121-154) ----
uvec4 pcg4d(uvec4 v);
float uint_to_float(uint32_t x);
float rand(uvec4& s);

// ---- geometry helpers (utils.glsl) ----
vec3 offset_ray(vec3 p, vec3 n);
vec3 offset_ray2(vec3 p, vec3 n);
void branchless_onb(vec3 n, vec3& b1, vec3& b2);
vec2 concentric_sample_disk(vec2 r);
vec3 sample_hemisphere(vec2 xi);                 // local frame, SAMPLING_MODE 1
vec3 sample_hemisphere_n(vec2 xi, vec3 n, float& phi);

// ---- scene view ----
struct Scene {
    lt_scene_desc d;
};

// ---- traversal (watertight, Woop/Benthin/Wald 2013) ----
struct Hit {
    float t, u, v;
    uint32_t prim, inst;  // 0xFFFFFFFF on miss
};
struct Tri {
    vec3 v0, v1, v2;
    uint32_t prim, inst, gid;
};
// Instance transforms (lt_scene_desc::world_matrices, LumenScene.h:32).  The
// record of a column-major 4x4 matrix: object-to-world rows (3 x 4), then the
// rows of the normal matrix, the inverse transpose of the linear part, from
// its cofactors in double precision.  xf_record returns false for the
// identity (and for a singular matrix, which the upload refuses).
bool xf_record(const f32* m, f32 out[24]);
// mesh m's transform record, false if it has none (identity)
bool mesh_xf(const lt_scene_desc& s, uint32_t m, f32 out[24]);
// mesh m's index range (index_counts, or up to the next mesh's offset)
void mesh_range(const lt_scene_desc& s, uint32_t m, uint32_t& off, uint32_t& cnt);
struct Accel;  // host BVH2 over Tri
Accel* accel_build(const lt_scene_desc& s);
void accel_free(Accel* a);
Hit accel_closest(const Accel* a, vec3 o, vec3 d, float tmin, float tmax);
bool accel_any(const Accel* a, vec3 o, vec3 d, float tmin, float tmax);
Hit brute_closest(const Accel* a, vec3 o, vec3 d, float tmin, float tmax);
bool brute_any(const Accel* a, vec3 o, vec3 d, float tmin, float tmax);
// single triangle test; returns true and t/u/v if hit in [tmin, tmax]
bool intersect_tri(const Tri& tr, vec3 o, vec3 d, float tmin, float tmax, float& t, float& u, float& v);

// ---- hit payload (ray.rchit:24-84) ----
struct Payload {
    vec3 n_g, n_s, pos;
    vec2 uv;
    uint32_t material_idx, triangle_idx, instance_idx;
    float area, dist;
};
void fill_payload(const lt_scene_desc& s, const Hit& h, vec3 o, vec3 d, float tmin, Payload& p);

// ---- materials / BSDF (bsdf_commons.glsl, bsdf/*.glsl) ----
struct Material {
    vec3 albedo;
    float ior;
    vec3 emissive_factor;
    uint32_t bsdf_type, bsdf_props;
    vec3 k;
    int32_t texture_id;
    float roughness, diffuse_trans, spec_trans, metallic, specular_tint, sheen_tint, clearcoat, clearcoat_gloss,
        sheen, subsurface, flatness, anisotropy;
    uint32_t thin;
};
// Field-wise copy of the wire-format material (so that the float64 build,
// whose Material holds doubles, converts instead of reinterpreting bytes).
inline Material material_from(const lt_material& w) {
    Material m;
    m.albedo = load3(w.albedo);
    m.ior = w.ior;
    m.emissive_factor = load3(w.emissive_factor);
    m.bsdf_type = w.bsdf_type;
    m.bsdf_props = w.bsdf_props;
    m.k = load3(w.k);
    m.texture_id = w.texture_id;
    m.roughness = w.roughness;
    m.diffuse_trans = w.diffuse_trans;
    m.spec_trans = w.spec_trans;
    m.metallic = w.metallic;
    m.specular_tint = w.specular_tint;
    m.sheen_tint = w.sheen_tint;
    m.clearcoat = w.clearcoat;
    m.clearcoat_gloss = w.clearcoat_gloss;
    m.sheen = w.sheen;
    m.subsurface = w.subsurface;
    m.flatness = w.flatness;
    m.anisotropy = w.anisotropy;
    m.thin = w.thin;
    return m;
}
Material load_material(const lt_scene_desc& s, uint32_t idx, vec2 uv);
vec3 sample_bsdf(vec3 n_s, vec3 wo, const Material& m, uint32_t mode, bool side, vec3& wi, float& pdf_w,
                 float& cos_theta, vec3 rands);
vec3 eval_bsdf(vec3 n_s, vec3 wo, const Material& m, uint32_t mode, bool side, vec3 wi, float& pdf_w,
               float& pdf_rev_w, bool eval_reverse);
float bsdf_pdf(const Material& m, vec3 n_s, vec3 wo, vec3 wi, bool side);
inline bool is_specular(const Material& m) { return (m.bsdf_props & LT_BSDF_FLAG_SPECULAR) != 0; }
float fresnel_dielectric(float cos_i, float eta, bool forward_facing);
float fresnel_conductor(float cos_i, float eta, float k);
float D_GGX_isotropic(float alpha_sqr, float cos_theta);
float D_GGX_anisotropic(vec2 alpha, vec3 h);

// ---- lights (commons.glsl:112-406) ----
struct LightRecord {
    uint32_t light_idx, flags;
    vec2 bary;
    uint32_t triangle_idx, instance_idx;
};
struct TriangleRecord {
    vec3 pos, n_s;
    float triangle_pdf;
};
TriangleRecord sample_triangle(const lt_scene_desc& s, const lt_prim_mesh_info& pi, vec2 rands, uint32_t tri,
                               const f32* world, vec2& uv);
vec3 sample_light_Li(const lt_scene_desc& s, vec4 rands_pos, vec3 p, int num_lights, float& pdf_pos_w, vec3& wi,
                     float& wi_len, float& pdf_pos_a, float& cos_from_light, LightRecord& rec, vec3& n, vec3& pos,
                     float& pdf_pos_dir_w);

// commons.glsl:335-406 (+ point lights): one emission sample; pdf_pos_a is
// divided by total_light (oracle_bdpt.cpp)
vec3 sample_light_Le(const lt_scene_desc& s, vec4 rp, vec2 rd, int num_lights, int total_light, float& cos_from_light,
                     uint32_t& flags, vec3& pos, vec3& wi, vec3& n, float& pdf_pos_a, float& pdf_dir_w);

// ---- Path integrator (integrators/path/path.rgen, pt_commons.glsl) ----
struct PathParams {
    uint32_t width, height;
    int32_t max_depth;
    vec3 sky_col;
    uint32_t direct_lighting;
    uint32_t frame_num;  // RNG frame word
    uint32_t nee_pick;   // 0: the reference's 1 / light_triangle_count, 1: the sampled light's own pick pdf
};
struct Counters {
    uint64_t closest = 0, shadow = 0;
};
// pt_commons.glsl:3-42 at a vertex with closest-hit payload `hit`
// sky of scenes with a directional light (oracle_atmosphere.cpp)
float phase_rayleigh(float costh);
float phase_mie(float costh, float g);
vec3 integrate_scattering(vec3 start, vec3 dir, float len, vec3 ldir, vec3 Lc);
vec3 shade_atmosphere(const lt_scene_desc& s, vec3 sky_col, vec3 o, vec3 d, float len);
vec3 nee_uniform_sample_light(const lt_scene_desc& s, const Accel* a, uvec4& seed, const Material& mat, vec3 pos,
                              bool side, vec3 n_s, vec3 wo, const Payload& hit, Counters* cnt);
// One sample of pixel (x, y).  `nan_dropped` reports the reference's NaN guard.
vec3 path_sample(const lt_scene_desc& s, const Accel* a, const lt_camera& cam, const PathParams& pp, uint32_t x,
                 uint32_t y, bool& nan_dropped, Counters* cnt);

}  // namespace ora
