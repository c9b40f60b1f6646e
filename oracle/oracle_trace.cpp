// oracle_trace.cpp — watertight closest/any hit, brute force and a host BVH2,
// and the closest-hit payload (TEST INFRASTRUCTURE, see oracle.h).
//
// The reference delegates traversal to the Vulkan driver and RT cores
// (src/shaders/integrators/path/path.rgen:48, ray.rchit, ray.rahit), so this is
// not a restatement of reference source but of the query semantics:
//   closest hit = nearest triangle with tmin <= t <= tmax, ties broken by the
//                 smaller global triangle index (= smaller (inst, prim));
//   any hit     = some triangle with tmin <= t <= tmax;
// with the triangle test of Woop, Benthin & Wald, "Watertight Ray/Triangle
// Intersection", JCGT 2(1), 2013 (including its double-precision edge fallback).
#include <algorithm>
#include <cmath>
#include <vector>

#include "oracle.h"

namespace ora {

struct Node2 {
    float lo[3], hi[3];
    int left;   // internal: index of left child (right = left + 1); leaf: -1
    int first;  // leaf: first tri
    int count;  // leaf: count
};
struct Accel {
    std::vector<Tri> tris;
    std::vector<Node2> nodes;
};

// Woop et al. 2013, listing 1 (no back-face culling)
bool intersect_tri(const Tri& tr, vec3 org, vec3 dir, float tmin, float tmax, float& t_out, float& u_out,
                   float& v_out) {
    float ax = fabsf(dir.x), ay = fabsf(dir.y), az = fabsf(dir.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    if (dir[kz] < 0.0f) std::swap(kx, ky);
    const float Sx = dir[kx] / dir[kz];
    const float Sy = dir[ky] / dir[kz];
    const float Sz = 1.0f / dir[kz];
    const vec3 A = tr.v0 - org, B = tr.v1 - org, C = tr.v2 - org;
    const float Ax = A[kx] - Sx * A[kz], Ay = A[ky] - Sy * A[kz];
    const float Bx = B[kx] - Sx * B[kz], By = B[ky] - Sy * B[kz];
    const float Cx = C[kx] - Sx * C[kz], Cy = C[ky] - Sy * C[kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = Sz * A[kz], Bz = Sz * B[kz], Cz = Sz * C[kz];
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    const float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    u_out = V * rcp;
    v_out = W * rcp;
    return true;
}

bool xf_record(const f32* m, f32 out[24]) {
    bool identity = true;
    for (int k = 0; k < 16; k++)
        if (m[k] != ((k % 5 == 0) ? 1.0f : 0.0f)) identity = false;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = m[j * 4 + i];
    double a[3][3], c[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = m[j * 4 + i];
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
        for (int j = 0; j < 3; j++) out[12 + 3 * i + j] = (f32)(c[i][j] / det);
    out[21] = out[22] = out[23] = 0.0f;
    return !identity;
}

bool mesh_xf(const lt_scene_desc& s, uint32_t m, f32 out[24]) {
    return s.world_matrices && xf_record(s.world_matrices + 16 * (size_t)m, out);
}

void mesh_range(const lt_scene_desc& s, uint32_t m, uint32_t& off, uint32_t& cnt) {
    off = s.prim_infos[m].index_offset;
    cnt = s.index_counts ? s.index_counts[m]
                         : ((m + 1 < s.num_prim_infos) ? s.prim_infos[m + 1].index_offset : s.num_indices) - off;
}

// world-space point through the rows of a transform record, in the order
// ((m0 x + m1 y) + m2 z) + m3 (the device builder's order)
static vec3 xf_point_rows(const f32* m, vec3 p) {
    return {((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3], ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7],
            ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11]};
}

// The flattened triangle list: mesh by mesh (instances of one geometry each
// with their own triangles), vertices in world space; gid = position in the
// list, the tie-break key of equal hit distances.
static void collect_tris(const lt_scene_desc& s, std::vector<Tri>& out) {
    uint32_t gid = 0;
    for (uint32_t m = 0; m < s.num_prim_infos; m++) {
        const lt_prim_mesh_info& pi = s.prim_infos[m];
        uint32_t off, cnt;
        mesh_range(s, m, off, cnt);
        f32 xf[24];
        const bool has_xf = mesh_xf(s, m, xf);
        for (uint32_t t = 0; t < cnt / 3; t++) {
            uint32_t io = off + 3 * t;
            Tri tr;
            tr.v0 = load3(s.vertices[s.indices[io] + pi.vertex_offset].pos);
            tr.v1 = load3(s.vertices[s.indices[io + 1] + pi.vertex_offset].pos);
            tr.v2 = load3(s.vertices[s.indices[io + 2] + pi.vertex_offset].pos);
            if (has_xf) {
                tr.v0 = xf_point_rows(xf, tr.v0);
                tr.v1 = xf_point_rows(xf, tr.v1);
                tr.v2 = xf_point_rows(xf, tr.v2);
            }
            tr.prim = t;
            tr.inst = m;
            tr.gid = gid++;
            out.push_back(tr);
        }
    }
}

// top-down binned-SAH-free median split; quality is irrelevant for the oracle
static int build_rec(Accel* a, std::vector<int>& idx, std::vector<vec3>& cent, int begin, int end) {
    int ni = (int)a->nodes.size();
    a->nodes.push_back(Node2{});
    Node2 n;
    for (int k = 0; k < 3; k++) {
        n.lo[k] = INFINITY;
        n.hi[k] = -INFINITY;
    }
    vec3 clo = v3(INFINITY), chi = v3(-INFINITY);
    for (int i = begin; i < end; i++) {
        const Tri& t = a->tris[idx[i]];
        for (int k = 0; k < 3; k++) {
            n.lo[k] = fminf(n.lo[k], fminf(t.v0[k], fminf(t.v1[k], t.v2[k])));
            n.hi[k] = fmaxf(n.hi[k], fmaxf(t.v0[k], fmaxf(t.v1[k], t.v2[k])));
            clo[k] = fminf(clo[k], cent[idx[i]][k]);
            chi[k] = fmaxf(chi[k], cent[idx[i]][k]);
        }
    }
    if (end - begin <= 4) {
        n.left = -1;
        n.first = begin;
        n.count = end - begin;
        a->nodes[ni] = n;
        return ni;
    }
    int axis = 0;
    float ext = chi.x - clo.x;
    if (chi.y - clo.y > ext) { axis = 1; ext = chi.y - clo.y; }
    if (chi.z - clo.z > ext) { axis = 2; }
    int mid = (begin + end) / 2;
    std::nth_element(idx.begin() + begin, idx.begin() + mid, idx.begin() + end,
                     [&](int p, int q) { return cent[p][axis] < cent[q][axis]; });
    n.left = (int)a->nodes.size();
    n.first = 0;
    n.count = 0;
    a->nodes[ni] = n;
    // reserve both child slots contiguously
    a->nodes.push_back(Node2{});
    a->nodes.push_back(Node2{});
    int l = build_rec(a, idx, cent, begin, mid);
    int r = build_rec(a, idx, cent, mid, end);
    a->nodes[n.left] = a->nodes[l];
    a->nodes[n.left + 1] = a->nodes[r];
    return ni;
}

Accel* accel_build(const lt_scene_desc& s) {
    Accel* a = new Accel;
    collect_tris(s, a->tris);
    if (a->tris.empty()) return a;
    std::vector<int> idx(a->tris.size());
    std::vector<vec3> cent(a->tris.size());
    for (size_t i = 0; i < a->tris.size(); i++) {
        idx[i] = (int)i;
        const Tri& t = a->tris[i];
        cent[i] = (t.v0 + t.v1 + t.v2) * (1.0f / 3.0f);
    }
    // simple iterative rebuild: recursion writes children into reserved slots
    a->nodes.reserve(a->tris.size() * 3);
    build_rec(a, idx, cent, 0, (int)idx.size());
    std::vector<Tri> sorted(a->tris.size());
    for (size_t i = 0; i < idx.size(); i++) sorted[i] = a->tris[idx[i]];
    a->tris.swap(sorted);
    return a;
}

void accel_free(Accel* a) { delete a; }

// Conservative slab test (PBRT's 1 + 2*gamma(3) widening on both ends).  The
// interval [tmin, tmax] is widened by 2^-16 of the box's far t along the
// dominant axis kz: a contained triangle's watertight t may undercut its true
// t by a few ulp of its vertices' sheared z, which that far t bounds.
static inline bool box_hit(const Node2& n, vec3 o, vec3 d, vec3 inv, float tmin, float tmax, float& tnear) {
    const float g = 1.0f + 2.0f * (3.0f * 5.9604645e-8f) / (1.0f - 3.0f * 5.9604645e-8f);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    float margin = fmaxf(fabsf((n.lo[kz] - o[kz]) * inv[kz]), fabsf((n.hi[kz] - o[kz]) * inv[kz])) *
                   (1.0f / 65536.0f);
    float t0 = tmin - margin, t1 = tmax + margin;
    for (int k = 0; k < 3; k++) {
        float a = (n.lo[k] - o[k]) * inv[k];
        float b = (n.hi[k] - o[k]) * inv[k];
        if (a != a) a = -INFINITY;  // 0 * inf
        if (b != b) b = INFINITY;
        float lo = fminf(a, b), hi = fmaxf(a, b);
        hi = hi > 0.0f ? hi * g : hi / g;
        lo = lo > 0.0f ? lo / g : lo * g;
        t0 = fmaxf(t0, lo);
        t1 = fminf(t1, hi);
    }
    tnear = t0;
    return t0 <= t1;
}

static inline bool better(float t, uint32_t gid, const Hit& h, uint32_t hgid) {
    return t < h.t || (t == h.t && gid < hgid);
}

Hit accel_closest(const Accel* a, vec3 o, vec3 d, float tmin, float tmax) {
    Hit h{INFINITY, 0, 0, 0xFFFFFFFFu, 0xFFFFFFFFu};
    uint32_t hgid = 0xFFFFFFFFu;
    if (a->nodes.empty()) return h;
    vec3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    float tcur = tmax;
    while (sp) {
        const Node2& n = a->nodes[stack[--sp]];
        float tn;
        if (!box_hit(n, o, d, inv,tmin, tcur, tn)) continue;
        if (n.left < 0) {
            for (int i = n.first; i < n.first + n.count; i++) {
                const Tri& tr = a->tris[i];
                float t, u, v;
                if (intersect_tri(tr, o, d, tmin, tcur, t, u, v) && better(t, tr.gid, h, hgid)) {
                    h = {t, u, v, tr.prim, tr.inst};
                    hgid = tr.gid;
                    tcur = t;
                }
            }
        } else {
            stack[sp++] = n.left + 1;
            stack[sp++] = n.left;
        }
    }
    return h;
}

bool accel_any(const Accel* a, vec3 o, vec3 d, float tmin, float tmax) {
    if (a->nodes.empty()) return false;
    vec3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node2& n = a->nodes[stack[--sp]];
        float tn;
        if (!box_hit(n, o, d, inv,tmin, tmax, tn)) continue;
        if (n.left < 0) {
            for (int i = n.first; i < n.first + n.count; i++) {
                float t, u, v;
                if (intersect_tri(a->tris[i], o, d, tmin, tmax, t, u, v)) return true;
            }
        } else {
            stack[sp++] = n.left + 1;
            stack[sp++] = n.left;
        }
    }
    return false;
}

Hit brute_closest(const Accel* a, vec3 o, vec3 d, float tmin, float tmax) {
    Hit h{INFINITY, 0, 0, 0xFFFFFFFFu, 0xFFFFFFFFu};
    uint32_t hgid = 0xFFFFFFFFu;
    for (const Tri& tr : a->tris) {
        float t, u, v;
        if (intersect_tri(tr, o, d, tmin, tmax, t, u, v) && better(t, tr.gid, h, hgid)) {
            h = {t, u, v, tr.prim, tr.inst};
            hgid = tr.gid;
        }
    }
    return h;
}

bool brute_any(const Accel* a, vec3 o, vec3 d, float tmin, float tmax) {
    for (const Tri& tr : a->tris) {
        float t, u, v;
        if (intersect_tri(tr, o, d, tmin, tmax, t, u, v)) return true;
    }
    return false;
}

// ray.rchit:24-84: object-space interpolation, then the instance's
// object-to-world transform (gl_ObjectToWorldEXT) for the position and the
// inverse transpose (nrm * gl_WorldToObjectEXT) for the normals; identity for
// meshes without a transform (the JSON path, LumenScene.cpp:328).
// dist = gl_RayTminEXT + gl_HitTEXT (ray.rchit:81).
void fill_payload(const lt_scene_desc& s, const Hit& h, vec3 o, vec3 d, float tmin, Payload& p) {
    (void)o;
    (void)d;
    const lt_prim_mesh_info& pi = s.prim_infos[h.inst];
    uint32_t io = pi.index_offset + 3 * h.prim;
    uint32_t vo = pi.vertex_offset;
    const lt_vertex& a = s.vertices[s.indices[io + 0] + vo];
    const lt_vertex& b = s.vertices[s.indices[io + 1] + vo];
    const lt_vertex& c = s.vertices[s.indices[io + 2] + vo];
    vec3 v0 = load3(a.pos), v1 = load3(b.pos), v2p = load3(c.pos);
    vec3 n0 = load3(a.normal), n1 = load3(b.normal), n2 = load3(c.normal);
    vec3 bc = {1.0f - h.u - h.v, h.u, h.v};
    vec3 pos = v0 * bc.x + v1 * bc.y + v2p * bc.z;
    vec3 nrm = normalize(n0 * bc.x + n1 * bc.y + n2 * bc.z);
    vec2 uv = {a.uv0[0] * bc.x + b.uv0[0] * bc.y + c.uv0[0] * bc.z, a.uv0[1] * bc.x + b.uv0[1] * bc.y + c.uv0[1] * bc.z};
    vec3 e0 = v2p - v0, e1 = v1 - v0;
    f32 xf[24];
    if (mesh_xf(s, h.inst, xf)) {
        auto nrm_xf = [&](vec3 n) {
            return vec3{xf[12] * n.x + xf[13] * n.y + xf[14] * n.z, xf[15] * n.x + xf[16] * n.y + xf[17] * n.z,
                        xf[18] * n.x + xf[19] * n.y + xf[20] * n.z};
        };
        auto vec_xf = [&](vec3 d) {
            return vec3{xf[0] * d.x + xf[1] * d.y + xf[2] * d.z, xf[4] * d.x + xf[5] * d.y + xf[6] * d.z,
                        xf[8] * d.x + xf[9] * d.y + xf[10] * d.z};
        };
        p.n_g = normalize(nrm_xf(cross(e0, e1)));
        p.n_s = normalize(nrm_xf(nrm));
        p.pos = vec_xf(pos) + v3(xf[3], xf[7], xf[11]);
        p.area = 0.5f * length(cross(vec_xf(e0), vec_xf(e1)));
    } else {
        p.n_g = normalize(cross(e0, e1));
        p.n_s = normalize(nrm);
        p.pos = pos;
        p.area = 0.5f * length(cross(e0, e1));
    }
    p.uv = uv;
    p.material_idx = pi.material_index;
    p.triangle_idx = h.prim;
    p.instance_idx = h.inst;
    p.dist = tmin + h.t;
}

}  // namespace ora
