// oracle_bvh8.cpp — host traversal of the device-built BVH8 (TEST INFRASTRUCTURE).
//
// Walks the exported 128-byte nodes with the same child decode, slab test and
// child order as the device kernel (nearest child first for closest-hit
// queries, slot order for occlusion queries), testing a node's leaf
// triangles at once (the device queues them and tests one per step, so its
// closest-hit bound shrinks later): its per-ray node and triangle counts are
// the device's lower bound, and its hits must equal the device's and brute
// force (oracle_trace.cpp), which isolates build defects from traversal
// defects.
#include <vector>

#include "oracle.h"

namespace ora {

struct Node8h {
    float p[3];
    int8_t e[3];
    uint8_t imask;
    uint32_t child_base, tri_base;
    uint8_t meta[8];
    uint16_t qlo[3][8], qhi[3][8];
};
static_assert(sizeof(Node8h) == 128, "node layout");
struct TriRech {
    float v0[3], v1[3], v2[3];
    uint32_t gid, prim, inst;
};
static_assert(sizeof(TriRech) == 48, "tri layout");

static const float kSlabErr = 4.76837158e-7f;  // 2^-21

// value of an fp16 bit pattern (the node planes); exact
static float f16v(uint16_t h) {
    const int ex = (h >> 10) & 0x1F, man = h & 0x3FF;
    const float mag = ex == 0 ? ldexpf((float)man, -24) : ldexpf((float)(man | 0x400), ex - 25);
    return (h & 0x8000) ? -mag : mag;
}

}  // namespace ora

using namespace ora;

extern "C" void ora_bvh8_trace(const void* nodes_v, uint32_t n_nodes, const void* tris_v, const lt_ray* rays,
                               size_t n, lt_hit* hits, uint8_t* occ, int any_hit, uint64_t* node_counts,
                               uint64_t* tri_counts) {
    const Node8h* nodes = (const Node8h*)nodes_v;
    const TriRech* tris = (const TriRech*)tris_v;
    for (size_t r = 0; r < n; r++) {
        const vec3 o = load3(rays[r].org), d = load3(rays[r].dir);
        const float tmin = rays[r].tmin, tmax = rays[r].tmax;
        const float tiny = 1e-30f;
        const vec3 inv = {1.0f / (fabsf(d.x) < tiny ? copysignf(tiny, d.x) : d.x),
                          1.0f / (fabsf(d.y) < tiny ? copysignf(tiny, d.y) : d.y),
                          1.0f / (fabsf(d.z) < tiny ? copysignf(tiny, d.z) : d.z)};
        float best_t = tmax, best_u = 0, best_v = 0, tcur = tmax;
        uint32_t best_gid = 0xFFFFFFFFu, best_prim = 0xFFFFFFFFu, best_inst = 0xFFFFFFFFu;
        uint64_t nn = 0, nt = 0;
        bool done = false;
        std::vector<uint32_t> stack;
        if (n_nodes) stack.push_back(0);
        while (!stack.empty() && !done) {
            const Node8h& nd = nodes[stack.back()];
            stack.pop_back();
            nn++;
            // child slab distances t = q * a + b, a = 2^e / d, b = (p - o) / d,
            // widened by the rounding bound E = 2^-21 (65535 |a| + |b|)
            float a[3], b[3], bn[3], bf[3];
            for (int k = 0; k < 3; k++) {
                uint32_t bits = (uint32_t)(nd.e[k] + 127) << 23;
                float s;
                memcpy(&s, &bits, 4);
                a[k] = s * inv[k];
                b[k] = (nd.p[k] - o[k]) * inv[k];
                float E = kSlabErr * fmaf(65535.0f, fabsf(a[k]), fabsf(b[k]));
                bn[k] = b[k] - E;
                bf[k] = b[k] + E;
            }
            // cull margin: 2^-16 of the node's far t along the dominant axis kz
            // (bounds the error of a contained triangle's watertight t)
            float adx = fabsf(d.x), ady = fabsf(d.y), adz = fabsf(d.z);
            int kz = (adx > ady) ? (adx > adz ? 0 : 2) : (ady > adz ? 1 : 2);
            float far_kz = fmaxf(fabsf(b[kz]), fabsf(fmaf(65535.0f, a[kz], b[kz])));
            float margin = far_kz * (1.0f / 65536.0f);
            uint32_t hit_int = 0, hit_leaf = 0;
            float tn[8];
            for (int c = 0; c < 8; c++) {
                float qn[3], qf[3];
                for (int k = 0; k < 3; k++) {
                    bool neg = inv[k] < 0.0f;  // ray octant picks the near plane
                    qn[k] = f16v(neg ? nd.qhi[k][c] : nd.qlo[k][c]);
                    qf[k] = f16v(neg ? nd.qlo[k][c] : nd.qhi[k][c]);
                }
                float t0 = fmaxf(fmaxf(fmaf(qn[0], a[0], bn[0]), fmaf(qn[1], a[1], bn[1])), fmaf(qn[2], a[2], bn[2]));
                float t1 = fminf(fminf(fmaf(qf[0], a[0], bf[0]), fmaf(qf[1], a[1], bf[1])), fmaf(qf[2], a[2], bf[2]));
                t0 = fmaxf(t0, tmin - margin);
                t1 = fminf(t1, tcur + margin);
                bool h = t0 <= t1;  // an empty slot carries no internal bit and no count
                tn[c] = t0;
                bool internal = (nd.imask >> c) & 1u;
                if (h && internal) hit_int |= 1u << c;
                if (h && !internal && (nd.meta[c] & 7u)) hit_leaf |= 1u << c;
            }
            for (int c = 0; c < 8 && !done; c++) {
                if (!((hit_leaf >> c) & 1u)) continue;
                uint32_t first = nd.tri_base + (nd.meta[c] >> 3), cnt = nd.meta[c] & 7u;
                for (uint32_t k = 0; k < cnt; k++) {
                    const TriRech& tr = tris[first + k];
                    nt++;
                    Tri t3{load3(tr.v0), load3(tr.v1), load3(tr.v2), tr.prim, tr.inst, tr.gid};
                    float t, u, v;
                    if (!intersect_tri(t3, o, d, tmin, tmax, t, u, v)) continue;
                    if (any_hit) {
                        best_t = t; best_u = u; best_v = v; best_gid = tr.gid; best_prim = tr.prim; best_inst = tr.inst;
                        done = true;
                        break;
                    }
                    if (t < best_t || (t == best_t && tr.gid < best_gid)) {
                        best_t = t; best_u = u; best_v = v; best_gid = tr.gid; best_prim = tr.prim; best_inst = tr.inst;
                        tcur = t;
                    }
                }
            }
            if (done || !hit_int) continue;
            int nearest = __builtin_ctz(hit_int);  // occlusion queries: slot order, as the device
            float tb = INFINITY;
            if (!any_hit)
                for (int c = 0; c < 8; c++)
                    if (((hit_int >> c) & 1u) && tn[c] < tb) {
                        tb = tn[c];
                        nearest = c;
                    }
            for (int c = 0; c < 8; c++)
                if (((hit_int >> c) & 1u) && c != nearest) stack.push_back(nd.child_base + nd.meta[c]);
            stack.push_back(nd.child_base + nd.meta[nearest]);
        }
        bool hit = best_gid != 0xFFFFFFFFu;
        if (hits) {
            hits[r].t = hit ? best_t : 0.0f;
            hits[r].prim = best_prim;
            hits[r].inst = best_inst;
            hits[r].u = hit ? best_u : 0.0f;
            hits[r].v = hit ? best_v : 0.0f;
        }
        if (occ) occ[r] = hit ? 1 : 0;
        if (node_counts) node_counts[r] = nn;
        if (tri_counts) tri_counts[r] = nt;
    }
}
