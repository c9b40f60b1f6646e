// dev_scene.h — closest-hit payload, material fetch and light sampling on
// the device (src/shaders/ray.rchit, bsdf_commons.glsl:16-22,
// commons.glsl:112-300).
#pragma once
#include "context.h"
#include "dev_bsdf.h"
#include "dev_math.h"

namespace ltd {

// HitPayload (utils.glsl:15-26); persists across traces like the GLSL
// rayPayloadEXT: a miss only clears `valid` (material_idx = -1, ray.rmiss:13).
struct Payload {
    f3 n_g, n_s, pos;
    f2 uv;
    uint32_t material_idx, triangle_idx, instance_idx;
    float area;  // world-space triangle area
    bool valid;
};

// A mesh's transform record (kXfFloats floats, SceneView::xf): rows of the
// object-to-world 3 x 4 matrix, then rows of the normal matrix (the inverse
// transpose of its linear part).
LT_DEV f3 xf_point(const float* m, f3 p) {
    return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
LT_DEV f3 xf_vector(const float* m, f3 d) {
    return {m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
            m[8] * d.x + m[9] * d.y + m[10] * d.z};
}
LT_DEV f3 xf_normal(const float* m, f3 n) {
    return {m[12] * n.x + m[13] * n.y + m[14] * n.z, m[15] * n.x + m[16] * n.y + m[17] * n.z,
            m[18] * n.x + m[19] * n.y + m[20] * n.z};
}

// ray.rchit:24-84: barycentric interpolation in object space, then the
// instance's object-to-world transform for the position and its inverse
// transpose for the normals (identity when the scene has no transforms)
LT_DEV void fill_payload(const SceneView& sv, uint32_t inst, uint32_t prim, float u, float v, Payload& p) {
    const lt_prim_mesh_info pi = sv.prims[inst];
    const uint32_t io = pi.index_offset + 3 * prim;
    const lt_vertex a = sv.vtx[sv.idx[io] + pi.vertex_offset];
    const lt_vertex b = sv.vtx[sv.idx[io + 1] + pi.vertex_offset];
    const lt_vertex c = sv.vtx[sv.idx[io + 2] + pi.vertex_offset];
    const f3 v0 = ld3(a.pos), v1 = ld3(b.pos), v2 = ld3(c.pos);
    const float b0 = 1.0f - u - v;
    p.pos = v0 * b0 + v1 * u + v2 * v;
    const f3 nrm = normalize(ld3(a.normal) * b0 + ld3(b.normal) * u + ld3(c.normal) * v);
    p.uv = {a.uv0[0] * b0 + b.uv0[0] * u + c.uv0[0] * v, a.uv0[1] * b0 + b.uv0[1] * u + c.uv0[1] * v};
    const f3 e0 = v2 - v0, e1 = v1 - v0;
    const f3 cg = cross(e0, e1);
    if (sv.xf) {
        const float* m = sv.xf + (size_t)kXfFloats * inst;
        p.pos = xf_point(m, p.pos);
        p.n_s = normalize(xf_normal(m, nrm));
        p.n_g = normalize(xf_normal(m, cg));
        p.area = 0.5f * length(cross(xf_vector(m, e0), xf_vector(m, e1)));
    } else {
        p.n_s = normalize(nrm);
        p.n_g = normalize(cg);
        p.area = 0.5f * length(cg);
    }
    p.material_idx = pi.material_index;
    p.triangle_idx = prim;
    p.instance_idx = inst;
    p.valid = true;
}

// bilinear, repeat addressing, texel centres at (i + 0.5) / size
LT_DEV float4 tex_bilinear(const SceneView& sv, int tid, f2 uv) {
    const TexDesc t = sv.texs[tid];
    float fx = uv.x * (float)t.width - 0.5f, fy = uv.y * (float)t.height - 0.5f;
    float x0f = floorf(fx), y0f = floorf(fy);
    float ax = fx - x0f, ay = fy - y0f;
    int w = (int)t.width, h = (int)t.height;
    int x0 = ((int)x0f % w + w) % w, y0 = ((int)y0f % h + h) % h;
    int x1 = (x0 + 1) % w, y1 = (y0 + 1) % h;
    const float4* base = sv.texels + t.offset;
    const float4 A = base[(size_t)y0 * w + x0], B = base[(size_t)y0 * w + x1];
    const float4 C = base[(size_t)y1 * w + x0], D = base[(size_t)y1 * w + x1];
    auto lerp2 = [&](float a, float b, float c, float d) {
        float top = a * (1.0f - ax) + b * ax;
        float bot = c * (1.0f - ax) + d * ax;
        return top * (1.0f - ay) + bot * ay;
    };
    return make_float4(lerp2(A.x, B.x, C.x, D.x), lerp2(A.y, B.y, C.y, D.y), lerp2(A.z, B.z, C.z, D.z),
                       lerp2(A.w, B.w, C.w, D.w));
}

// bsdf_commons.glsl:16-22 (+ per-material roughness texture)
LT_DEV Mat load_material(const SceneView& sv, uint32_t idx, f2 uv) {
    Mat m;
    const uint4* src = (const uint4*)(sv.mats + idx);  // 104 B = 6.5 x 16 B
    uint32_t* dst = (uint32_t*)&m;
    const uint32_t* s32 = (const uint32_t*)(sv.mats + idx);
#pragma unroll
    for (int i = 0; i < 26; i++) dst[i] = s32[i];
    (void)src;
    if (m.texture_id > -1) {
        float4 t = tex_bilinear(sv, m.texture_id, uv);
        m.albedo = m.albedo * mk3(t.x, t.y, t.z);
    }
    if (sv.rough) {
        int rt = sv.rough[idx];
        if (rt > -1) m.roughness = tex_bilinear(sv, rt, uv).x;
    }
    return m;
}

struct LightRecord {
    uint32_t light_idx, flags;
    f2 bary;
    uint32_t triangle_idx, instance_idx;
};

struct TriangleRecord {
    f3 pos, n_s;
    float triangle_pdf;
};

// commons.glsl:112-149: world = the light's world matrix (column-major),
// lxf = its transform record or NULL for the identity.  The reference
// transforms the edges as points (w = 1, :136-137), which only matters for a
// translated light; here they are vectors (w = 0), so the area pdf is the
// world triangle's.
LT_DEV TriangleRecord sample_triangle(const SceneView& sv, const lt_prim_mesh_info& pi, float rx, float ry,
                                      uint32_t tri, const float* world, const float* lxf, f2& uv) {
    TriangleRecord r;
    const uint32_t io = pi.index_offset + 3 * tri;
    const lt_vertex a = sv.vtx[sv.idx[io] + pi.vertex_offset];
    const lt_vertex b = sv.vtx[sv.idx[io + 1] + pi.vertex_offset];
    const lt_vertex c = sv.vtx[sv.idx[io + 2] + pi.vertex_offset];
    const f3 v0 = ld3(a.pos), v1 = ld3(b.pos), v2 = ld3(c.pos);
    const float sr = sqrtf(rx);
    uv = {1.0f - sr, ry * sr};
    const f3 bc = {1.0f - uv.x - uv.y, uv.x, uv.y};
    const f3 d1 = v1 - v0, d2 = v2 - v0;
    const f4 e0 = mul4(world, f4{d1.x, d1.y, d1.z, 0.0f});
    const f4 e1 = mul4(world, f4{d2.x, d2.y, d2.z, 0.0f});
    const f3 pos = v0 * bc.x + v1 * bc.y + v2 * bc.z;
    const f3 nrm = normalize(ld3(a.normal) * bc.x + ld3(b.normal) * bc.y + ld3(c.normal) * bc.z);
    const f4 wp = mul4(world, f4{pos.x, pos.y, pos.z, 1.0f});
    r.n_s = lxf ? normalize(xf_normal(lxf, nrm)) : normalize(nrm);
    r.triangle_pdf = 2.0f / length(cross(mk3(e0.x, e0.y, e0.z), mk3(e1.x, e1.y, e1.z)));
    r.pos = {wp.x, wp.y, wp.z};
    return r;
}

LT_DEV float spot_falloff(float cos_from_light, float cos_width, float cos_falloff) {
    if (cos_from_light < cos_width) return 0.0f;
    if (cos_from_light >= cos_falloff) return 1.0f;
    float d = (cos_from_light - cos_width) / (cos_falloff - cos_width);
    return (d * d) * (d * d);
}

// commons.glsl:224-300 (+ LT_LIGHT_POINT).  rp = rand4; n and pos are the
// sampled light point's normal and position (commons.glsl:302-307) and
// pdf_pos_dir_w the pdf of emitting along -wi from it (the light's cone for
// spot lights).
LT_DEV f3 sample_light_Li(const SceneView& sv, f4 rp, f3 p, float& pdf_pos_w, f3& wi, float& wi_len,
                          float& pdf_pos_a, float& cos_from_light, LightRecord& rec, f3& n, f3& pos,
                          float& pdf_pos_dir_w) {
    pdf_pos_dir_w = 0.0f;
    rec.light_idx = (uint32_t)(rp.x * (float)sv.num_lights);
    const lt_light& L = sv.lights[rec.light_idx];
    const uint32_t flags = L.light_flags;
    const uint32_t lt = flags & 7u;
    f3 Le = mk3(0.0f);
    rec.flags = flags;
    n = pos = mk3(0.0f);
    pdf_pos_w = 0.0f;
    wi = mk3(0.0f);
    wi_len = 0.0f;
    pdf_pos_a = 0.0f;
    cos_from_light = 0.0f;
    if (lt == LT_LIGHT_AREA) {
        const lt_prim_mesh_info pi = sv.prims[L.prim_mesh_idx];
        rec.triangle_idx = (uint32_t)(rp.y * (float)L.num_triangles);
        TriangleRecord r = sample_triangle(sv, pi, rp.z, rp.w, rec.triangle_idx, L.world_matrix,
                                           sv.lxf ? sv.lxf + (size_t)kXfFloats * rec.light_idx : nullptr, rec.bary);
        Mat lm = load_material(sv, pi.material_index, rec.bary);
        wi = r.pos - p;
        const float d2 = dot(wi, wi);
        wi_len = sqrtf(d2);
        wi /= wi_len;
        cos_from_light = fabsf(dot(r.n_s, -wi));
        Le = lm.emissive;
        pdf_pos_a = r.triangle_pdf;
        pdf_pos_w = pdf_pos_a * d2 / cos_from_light;
        pdf_pos_dir_w = cos_from_light * INV_PI * r.triangle_pdf;
        rec.instance_idx = L.prim_mesh_idx;
        n = r.n_s;
        pos = r.pos;
    } else if (lt == LT_LIGHT_SPOT) {
        wi = ld3(L.pos) - p;
        const float d2 = dot(wi, wi);
        wi_len = sqrtf(d2);
        wi /= wi_len;
        const f3 ld = normalize(ld3(L.to) - ld3(L.pos));
        cos_from_light = dot(-wi, ld);
        const float fo = spot_falloff(cos_from_light, sv.cones[2 * rec.light_idx], sv.cones[2 * rec.light_idx + 1]);
        pdf_pos_a = 1.0f;
        pdf_pos_w = d2;
        pdf_pos_dir_w = 1.0f / (TWO_PI * (1.0f - sv.cones[2 * rec.light_idx]));
        Le = ld3(L.L) * fo;
        n = -wi;
        pos = ld3(L.pos);
    } else if (lt == LT_LIGHT_DIRECTIONAL) {
        const f3 dir = normalize(ld3(L.pos) - ld3(L.to));
        const f3 lp = p + dir * (2.0f * L.world_radius);
        wi = lp - p;
        wi_len = length(wi);
        wi /= wi_len;
        pdf_pos_a = 1.0f;
        pdf_pos_w = 1.0f;
        pdf_pos_dir_w = INV_PI / (L.world_radius * L.world_radius);
        Le = ld3(L.L);
        cos_from_light = 1.0f;
        n = -wi;
        pos = lp;
    } else if (lt == LT_LIGHT_POINT) {
        wi = ld3(L.pos) - p;
        const float d2 = dot(wi, wi);
        wi_len = sqrtf(d2);
        wi /= wi_len;
        cos_from_light = 1.0f;
        pdf_pos_a = 1.0f;
        pdf_pos_w = d2;
        pdf_pos_dir_w = 1.0f / (4.0f * PI);
        Le = ld3(L.L);
        n = -wi;
        pos = ld3(L.pos);
    }
    return Le;
}

LT_DEV f3 sample_light_Li(const SceneView& sv, f4 rp, f3 p, float& pdf_pos_w, f3& wi, float& wi_len,
                          float& pdf_pos_a, float& cos_from_light, LightRecord& rec, f3& n, f3& pos) {
    float pdw;
    return sample_light_Li(sv, rp, p, pdf_pos_w, wi, wi_len, pdf_pos_a, cos_from_light, rec, n, pos, pdw);
}

LT_DEV f3 sample_light_Li(const SceneView& sv, f4 rp, f3 p, float& pdf_pos_w, f3& wi, float& wi_len,
                          float& pdf_pos_a, float& cos_from_light, LightRecord& rec) {
    f3 n, pos;
    return sample_light_Li(sv, rp, p, pdf_pos_w, wi, wi_len, pdf_pos_a, cos_from_light, rec, n, pos);
}

// ---- emission sampling and light pdfs (BDPT / VCM) ----

LT_DEV bool light_finite(uint32_t f) { return ((f >> 4) & 1u) != 0; }  // commons.glsl:42
LT_DEV bool light_delta(uint32_t f) { return ((f >> 5) & 1u) != 0; }   // commons.glsl:44
LT_DEV float uniform_cone_pdf(float cos_max) { return 1.0f / (TWO_PI * (1.0f - cos_max)); }  // commons.glsl:40

// commons.glsl:81-95, the spot cone of light `li` (+ LT_LIGHT_POINT)
LT_DEV float light_pdf_dir(const SceneView& sv, uint32_t li, uint32_t flags, f3 n_s, f3 wi) {
    switch (flags & 7u) {
        case LT_LIGHT_AREA: return fmaxf(dot(n_s, wi) / PI, 0.0f);
        case LT_LIGHT_SPOT: return uniform_cone_pdf(sv.cones[2 * li]);
        case LT_LIGHT_POINT: return 1.0f / (4.0f * PI);
        default: return 0.0f;
    }
}

// commons.glsl:64-79 (+ LT_LIGHT_POINT)
LT_DEV float light_pdf_a_to_w(uint32_t flags, float pdf_a, float wi_len_sqr, float cos_from_light) {
    switch (flags & 7u) {
        case LT_LIGHT_AREA: return pdf_a * wi_len_sqr / cos_from_light;
        case LT_LIGHT_SPOT: return wi_len_sqr / cos_from_light;
        case LT_LIGHT_DIRECTIONAL: return 1.0f;
        case LT_LIGHT_POINT: return wi_len_sqr;
        default: return 0.0f;
    }
}

// utils.glsl:157-173: rotation taking v to +z, and its application
LT_DEV f4 to_local_quat(f3 v) {
    if (v.z < -0.99999f) return f4{1.0f, 0.0f, 0.0f, 0.0f};
    const float x = v.y, y = -v.x, z = 0.0f, w = 1.0f + v.z;
    const float il = 1.0f / sqrtf(x * x + y * y + z * z + w * w);
    return f4{x * il, y * il, z * il, w * il};
}
LT_DEV f3 rot_quat(f4 q, f3 v) {
    const f3 qa = {q.x, q.y, q.z};
    return 2.0f * dot(qa, v) * qa + (q.w * q.w - dot(qa, qa)) * v + 2.0f * q.w * cross(qa, v);
}
// commons.glsl:217-222
LT_DEV f3 uniform_sample_cone(float u0, float u1, float cos_max) {
    const float ct = (1.0f - u0) + u0 * cos_max;
    const float st = sqrtf(1.0f - ct * ct);
    const float phi = u1 * TWO_PI;
    return {cosf(phi) * st, sinf(phi) * st, ct};
}

// commons.glsl:335-406 (+ LT_LIGHT_POINT: uniform sphere; spot cones per
// light).  rp = rand4, rd = rand2.  pdf_pos_a is divided by total_light.
LT_DEV f3 sample_light_Le(const SceneView& sv, f4 rp, float rd0, float rd1, float& cos_from_light, uint32_t& flags,
                          uint32_t& li, f3& pos, f3& wi, f3& n, float& pdf_pos_a, float& pdf_dir_w) {
    li = (uint32_t)(rp.x * (float)sv.num_lights);
    const lt_light& L = sv.lights[li];
    f3 Le = mk3(0.0f);
    flags = L.light_flags;
    cos_from_light = 0.0f;
    pdf_pos_a = 0.0f;
    pdf_dir_w = 0.0f;
    pos = wi = n = mk3(0.0f);
    const uint32_t lt = flags & 7u;
    if (lt == LT_LIGHT_AREA) {
        const lt_prim_mesh_info pi = sv.prims[L.prim_mesh_idx];
        const uint32_t tri = (uint32_t)(rp.y * (float)L.num_triangles);
        f2 bary;
        const TriangleRecord r = sample_triangle(sv, pi, rp.z, rp.w, tri, L.world_matrix,
                                                 sv.lxf ? sv.lxf + (size_t)kXfFloats * li : nullptr, bary);
        const Mat lm = load_material(sv, pi.material_index, bary);
        pos = r.pos;
        f3 T, B;
        branchless_onb(r.n_s, T, B);
        wi = to_world(sample_hemisphere(rd0, rd1), T, B, r.n_s);
        Le = lm.emissive;
        cos_from_light = fmaxf(dot(r.n_s, wi), 0.0f);
        pdf_pos_a = r.triangle_pdf;
        pdf_dir_w = dot(wi, r.n_s) / PI;
        n = r.n_s;
    } else if (lt == LT_LIGHT_SPOT) {
        const float cw = sv.cones[2 * li], cf = sv.cones[2 * li + 1];
        const f3 ld = normalize(ld3(L.to) - ld3(L.pos));
        const f4 q = to_local_quat(ld);
        wi = rot_quat(f4{-q.x, -q.y, -q.z, q.w}, uniform_sample_cone(rd0, rd1, cw));
        pos = ld3(L.pos);
        cos_from_light = dot(wi, ld);
        Le = ld3(L.L) * spot_falloff(cos_from_light, cw, cf);
        pdf_pos_a = 1.0f;
        pdf_dir_w = uniform_cone_pdf(cw);
        n = wi;
    } else if (lt == LT_LIGHT_POINT) {
        const float z = 1.0f - 2.0f * rd0, r = sqrtf(fmaxf(0.0f, 1.0f - z * z)), phi = TWO_PI * rd1;
        wi = {r * cosf(phi), r * sinf(phi), z};
        pos = ld3(L.pos);
        cos_from_light = 1.0f;
        Le = ld3(L.L);
        pdf_pos_a = 1.0f;
        pdf_dir_w = 1.0f / (4.0f * PI);
        n = wi;
    }
    pdf_pos_a /= (float)sv.total_light_triangle_cnt;
    return Le;
}

}  // namespace ltd
