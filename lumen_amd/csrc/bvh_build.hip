// bvh_build.hip — GPU LBVH build collapsed to a quantised BVH8 (gfx950).
//
// Replaces the driver build behind Integrator::create_accel ->
// vk::build_blas / vk::build_tlas (src/RayTracer/Integrator.cpp:137-160,
// src/Framework/AccelerationStructure.cpp:171-315).  Stages:
//   1. k_make_tris     : 48-B triangle records + centroid bounds (atomics)
//   2. k_morton        : 30-bit Morton codes of the centroids
//   3. radix sort      : (morton, triangle) pairs, rocPRIM radix sort (bits 0..29)
//   4. k_karras        : binary radix tree (Karras 2012), one thread per node
//   5. k_refit         : bottom-up AABBs with arrival counters
//   5'. (in k_refit)    : SAH dynamic programme over the binary tree (Ylitie,
//                        Karras & Laine 2017): per node the cheapest cover of
//                        its subtree by at most i = 1..8 BVH8 subtree roots;
//   6. k_collapse      : top-down BVH8 collapse, one thread per BVH8 node and
//                        level; a node's children are the roots the programme
//                        chose (or, greedy, the expansion of the child with the
//                        largest surface area), subtrees of <= max_leaf
//                        triangles become leaves; 16-bit quantised, 128-B nodes.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <cmath>
#include <vector>

#include "context.h"

using ltb::Node8;
using ltb::TriRec;

namespace {

struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// object-to-world point of the transform rows xf (3 x 4), evaluated without
// contraction in the order the host replay uses
__device__ __forceinline__ void xf_point(const float* __restrict__ xf, const float* p, float* o) {
#pragma clang fp contract(off)
    for (int i = 0; i < 3; i++) o[i] = ((xf[4 * i] * p[0] + xf[4 * i + 1] * p[1]) + xf[4 * i + 2] * p[2]) + xf[4 * i + 3];
}

// One thread per triangle g of the flattened triangle list (mesh by mesh,
// instances of one geometry each with their own triangles): the mesh by
// binary search over the meshes' first triangles; vertices in world space
// through the mesh's transform, the flattening that gives one BVH the
// instance semantics of the reference's TLAS (Integrator.cpp:147-158).
__global__ void k_make_tris(const lt_vertex* __restrict__ vtx, const uint32_t* __restrict__ idx,
                            const lt_prim_mesh_info* __restrict__ prims, const uint32_t* __restrict__ tri_start,
                            const float* __restrict__ xf, uint32_t nprims, uint32_t ntri, TriRec* __restrict__ out,
                            uint32_t* __restrict__ cbounds /* 6 ordered uints */) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (g < ntri) {
        uint32_t lo = 0, hi = nprims;  // last mesh whose first triangle is <= g
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (tri_start[mid] <= g)
                lo = mid;
            else
                hi = mid;
        }
        const lt_prim_mesh_info pi = prims[lo];
        const uint32_t prim = g - tri_start[lo];
        const uint32_t io = pi.index_offset + 3 * prim;
        TriRec r;
        const float* a = vtx[idx[io] + pi.vertex_offset].pos;
        const float* b = vtx[idx[io + 1] + pi.vertex_offset].pos;
        const float* c = vtx[idx[io + 2] + pi.vertex_offset].pos;
        if (xf) {
            const float* m = xf + (size_t)kXfFloats * lo;
            xf_point(m, a, r.v0);
            xf_point(m, b, r.v1);
            xf_point(m, c, r.v2);
        } else {
            for (int k = 0; k < 3; k++) {
                r.v0[k] = a[k];
                r.v1[k] = b[k];
                r.v2[k] = c[k];
            }
        }
        for (int k = 0; k < 3; k++) {
            float cen = (r.v0[k] + r.v1[k] + r.v2[k]) * (1.0f / 3.0f);
            clo[k] = cen;
            chi[k] = cen;
        }
        r.gid = g;
        r.prim = prim;
        r.inst = lo;
        out[g] = r;
    }
    // block reduction of the centroid bounds, then one atomic per block
    for (int k = 0; k < 3; k++) {
        float l = clo[k], h = chi[k];
        for (int o = 32; o > 0; o >>= 1) {
            l = fminf(l, __shfl_xor(l, o));
            h = fmaxf(h, __shfl_xor(h, o));
        }
        if ((threadIdx.x & 63) == 0 && l <= h) {
            atomicMin(&cbounds[k], f2ord(l));
            atomicMax(&cbounds[3 + k], f2ord(h));
        }
    }
}

// fp16 bit patterns of non-negative plane values, rounded outward (the
// largest fp16 <= x / the smallest fp16 >= x, both within [0, 65504]): the
// double -> float -> fp16 conversion lands on one of the two fp16 values
// around x, and one step along the (monotone) bit pattern fixes the side.
__device__ __forceinline__ float f16_value(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }
__device__ __forceinline__ uint16_t f16_floor(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 65504.0) return 0x7BFF;
    uint16_t h = __builtin_bit_cast(uint16_t, (_Float16)(float)x);
    if ((double)f16_value(h) > x) h -= 1;
    return h;
}
__device__ __forceinline__ uint16_t f16_ceil(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 65504.0) return 0x7BFF;
    uint16_t h = __builtin_bit_cast(uint16_t, (_Float16)(float)x);
    if ((double)f16_value(h) < x) h += 1;
    return h;
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void k_morton(const TriRec* __restrict__ tris, uint32_t n, const uint32_t* __restrict__ cb,
                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const TriRec r = tris[g];
    uint32_t code = 0;
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        float lo = ord2f(cb[k]), hi = ord2f(cb[3 + k]);
        float cen = (r.v0[k] + r.v1[k] + r.v2[k]) * (1.0f / 3.0f);
        float ext = hi - lo;
        float t = ext > 0.0f ? (cen - lo) / ext : 0.5f;
        q[k] = (uint32_t)fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f);
    }
    code = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
    keys[g] = code;
    vals[g] = g;
}

__device__ __forceinline__ int delta(const uint32_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((uint32_t)(i ^ j));
    return __clz(a ^ b);
}

// Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees
// and k-d trees".  Children: bit 31 set = leaf (index into sorted order).
__global__ void k_karras(const uint32_t* __restrict__ keys, int n, uint32_t* __restrict__ left,
                         uint32_t* __restrict__ right, uint32_t* __restrict__ parent_int,
                         uint32_t* __restrict__ parent_leaf, uint32_t* __restrict__ range_lo,
                         uint32_t* __restrict__ range_hi) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? -1 : 0);
    int lo = min(i, j), hi = max(i, j);
    uint32_t lc = (lo == gamma) ? (0x80000000u | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t rc = (hi == gamma + 1) ? (0x80000000u | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    range_lo[i] = (uint32_t)lo;
    range_hi[i] = (uint32_t)hi;
    if (lc & 0x80000000u)
        parent_leaf[lc & 0x7fffffffu] = (uint32_t)i;
    else
        parent_int[lc] = (uint32_t)i;
    if (rc & 0x80000000u)
        parent_leaf[rc & 0x7fffffffu] = (uint32_t)i;
    else
        parent_int[rc] = (uint32_t)i;
}

__device__ __forceinline__ Box tri_box(const TriRec& r) {
    Box b;
    for (int k = 0; k < 3; k++) {
        b.lo[k] = fminf(r.v0[k], fminf(r.v1[k], r.v2[k]));
        b.hi[k] = fmaxf(r.v0[k], fmaxf(r.v1[k], r.v2[k]));
    }
    return b;
}

__device__ __forceinline__ float half_area(const Box& b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// ---- PLOC binary tree (Meister and Bittner 2018, "Parallel locally-ordered
// clustering for bounding volume hierarchy construction") ----
// The clusters start as the Morton-sorted triangles.  Each round, every
// cluster finds the neighbour within kPlocRadius positions whose union box has
// the smallest surface area; mutual nearest neighbours merge into a new
// internal node (allocated from n - 2 down, so the last merge is the root 0
// and a parent's index is below its children's), and the survivors are
// compacted in order.  Ties are broken by the pair's indices, a strict order
// under which the globally closest pair is mutual, so every round merges.
#ifndef LT_PLOC_RADIUS
#define LT_PLOC_RADIUS 16
#endif
constexpr int kPlocRadius = LT_PLOC_RADIUS;

__device__ __forceinline__ Box box_union(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; k++) {
        r.lo[k] = fminf(a.lo[k], b.lo[k]);
        r.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
    return r;
}

__global__ void k_ploc_init(const TriRec* __restrict__ tris, const uint32_t* __restrict__ sorted, uint32_t n,
                            uint32_t* __restrict__ cl_node, Box* __restrict__ cl_box) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cl_node[i] = 0x80000000u | i;
    cl_box[i] = tri_box(tris[sorted[i]]);
}

// nearest neighbour of each of the m clusters within the radius: (area, lower
// index, higher index) lexicographically smallest
__global__ void k_ploc_nn(const Box* __restrict__ cl_box, uint32_t m, uint32_t* __restrict__ nn) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const Box bi = cl_box[i];
    const uint32_t lo = i > (uint32_t)kPlocRadius ? i - kPlocRadius : 0u;
    const uint32_t hi = min(m - 1u, i + (uint32_t)kPlocRadius);
    float best = INFINITY;
    uint32_t bj = i, bmin = 0xFFFFFFFFu, bmax = 0xFFFFFFFFu;
    for (uint32_t j = lo; j <= hi; j++) {
        if (j == i) continue;
        const float a = half_area(box_union(bi, cl_box[j]));
        const uint32_t pmin = min(i, j), pmax = max(i, j);
        if (a < best || (a == best && (pmin < bmin || (pmin == bmin && pmax < bmax)))) {
            best = a;
            bj = j;
            bmin = pmin;
            bmax = pmax;
        }
    }
    nn[i] = bj;
}

// mutual pairs merge at the lower index; keep[i] = 1 for the survivors
__global__ void k_ploc_merge(const uint32_t* __restrict__ cl_node, const Box* __restrict__ cl_box, uint32_t m,
                             const uint32_t* __restrict__ nn, uint32_t n, uint32_t* __restrict__ ctr,
                             uint32_t* __restrict__ left, uint32_t* __restrict__ right,
                             uint32_t* __restrict__ parent_int, uint32_t* __restrict__ parent_leaf,
                             uint32_t* __restrict__ out_node, Box* __restrict__ out_box, uint32_t* __restrict__ keep) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t j = nn[i];
    const bool mutual = j != i && nn[j] == i;
    if (!mutual) {
        out_node[i] = cl_node[i];
        out_box[i] = cl_box[i];
        keep[i] = 1u;
        return;
    }
    if (i > j) {  // merged by its partner
        keep[i] = 0u;
        return;
    }
    const uint32_t idx = n - 2u - atomicAdd(ctr, 1u);
    const uint32_t a = cl_node[i], b = cl_node[j];
    left[idx] = a;
    right[idx] = b;
    if (a & 0x80000000u)
        parent_leaf[a & 0x7fffffffu] = idx;
    else
        parent_int[a] = idx;
    if (b & 0x80000000u)
        parent_leaf[b & 0x7fffffffu] = idx;
    else
        parent_int[b] = idx;
    out_node[i] = idx;
    out_box[i] = box_union(cl_box[i], cl_box[j]);
    keep[i] = 1u;
}

__global__ void k_ploc_compact(const uint32_t* __restrict__ in_node, const Box* __restrict__ in_box, uint32_t m,
                               const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos,
                               uint32_t* __restrict__ out_node, Box* __restrict__ out_box) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || !keep[i]) return;
    out_node[pos[i]] = in_node[i];
    out_box[pos[i]] = in_box[i];
}

// PLOC's subtrees are not contiguous runs of the Morton order; the collapse
// wants every subtree's triangles contiguous.  Leaf counts bottom-up (second
// arrival sums), then each node's first position is the sum, over the
// ancestors it lies to the right of, of their left subtrees' counts.
__global__ void k_tree_count(const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                             const uint32_t* __restrict__ parent_int, const uint32_t* __restrict__ parent_leaf,
                             uint32_t n, uint32_t* __restrict__ flags, uint32_t* __restrict__ cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = parent_leaf[i];
    for (;;) {
        if (atomicAdd(&flags[p], 1u) == 0u) return;
        __threadfence();
        const uint32_t a = left[p], b = right[p];
        const uint32_t ca = (a & 0x80000000u) ? 1u : ((const volatile uint32_t*)cnt)[a];
        const uint32_t cb = (b & 0x80000000u) ? 1u : ((const volatile uint32_t*)cnt)[b];
        cnt[p] = ca + cb;
        __threadfence();
        if (p == 0u) return;
        p = parent_int[p];
    }
}

__device__ __forceinline__ uint32_t first_pos(uint32_t code, const uint32_t* __restrict__ left,
                                              const uint32_t* __restrict__ right,
                                              const uint32_t* __restrict__ parent_int,
                                              const uint32_t* __restrict__ parent_leaf,
                                              const uint32_t* __restrict__ cnt) {
    uint32_t pos = 0u;
    uint32_t c = code;
    uint32_t p = (c & 0x80000000u) ? parent_leaf[c & 0x7fffffffu] : (c == 0u ? 0xFFFFFFFFu : parent_int[c]);
    while (p != 0xFFFFFFFFu) {
        if (right[p] == c) {
            const uint32_t a = left[p];
            pos += (a & 0x80000000u) ? 1u : cnt[a];
        }
        c = p;
        p = p == 0u ? 0xFFFFFFFFu : parent_int[p];
    }
    return pos;
}

// internal nodes: their leaf range; leaves: their position in the new order
__global__ void k_tree_ranges(const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                              const uint32_t* __restrict__ parent_int, const uint32_t* __restrict__ parent_leaf,
                              const uint32_t* __restrict__ cnt, uint32_t n, uint32_t* __restrict__ range_lo,
                              uint32_t* __restrict__ range_hi, uint32_t* __restrict__ leaf_pos) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n - 1u) {
        const uint32_t lo = first_pos(i, left, right, parent_int, parent_leaf, cnt);
        range_lo[i] = lo;
        range_hi[i] = lo + cnt[i] - 1u;
    }
    if (i < n) leaf_pos[i] = first_pos(0x80000000u | i, left, right, parent_int, parent_leaf, cnt);
}

// leaves renumbered to their new positions (children, triangle order, parents)
__global__ void k_tree_relabel(uint32_t* __restrict__ left, uint32_t* __restrict__ right,
                               const uint32_t* __restrict__ leaf_pos, const uint32_t* __restrict__ sorted_in,
                               const uint32_t* __restrict__ parent_leaf_in, uint32_t n,
                               uint32_t* __restrict__ sorted_out, uint32_t* __restrict__ parent_leaf_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n - 1u) {
        const uint32_t a = left[i], b = right[i];
        if (a & 0x80000000u) left[i] = 0x80000000u | leaf_pos[a & 0x7fffffffu];
        if (b & 0x80000000u) right[i] = 0x80000000u | leaf_pos[b & 0x7fffffffu];
    }
    if (i < n) {
        sorted_out[leaf_pos[i]] = sorted_in[i];
        parent_leaf_out[leaf_pos[i]] = parent_leaf_in[i];
    }
}

// SAH weights of a BVH8 node visit and of a triangle test (a watertight
// test costs about half a node step)
#ifndef LT_SAH_TRI_COST
#define LT_SAH_TRI_COST 0.5f
#endif
constexpr float kSahNode = 1.0f;
constexpr float kSahTri = LT_SAH_TRI_COST;
constexpr uint8_t kDecLeaf = 0xFF;

// The dynamic programme at internal node p (children a, b):
//   R(p, i) = cheapest cover of p's subtree by at most i subtree roots,
//   R(p, 1) = min(leaf cost (<= max_leaf triangles), A(p) kSahNode + min_k R(a, k) + R(b, 8 - k)),
//   R(p, i) = min(R(p, i - 1), min_k R(a, k) + R(b, i - k)),  i = 2..8,
// a binary leaf costing A kSahTri for any i.  dec[p][0] = kDecLeaf or the k of
// the node's own 8 roots; dec[p][i - 1] = 0 (fewer roots) or the k of the split.
__device__ void sah_node(uint32_t p, uint32_t a, uint32_t b, const Box& pb, uint32_t cnt, uint32_t max_leaf,
                         const volatile Box* lbox, const volatile float* R_all, float* __restrict__ R,
                         uint8_t* __restrict__ dec) {
    float ra[8], rb[8];
    auto load = [&](uint32_t c, float* r) {
        if (c & 0x80000000u) {
            const volatile Box& bx = lbox[c & 0x7fffffffu];
            Box q;
            for (int k = 0; k < 3; k++) {
                q.lo[k] = bx.lo[k];
                q.hi[k] = bx.hi[k];
            }
            const float v = half_area(q) * kSahTri;
            for (int i = 0; i < 8; i++) r[i] = v;
        } else {
            for (int i = 0; i < 8; i++) r[i] = R_all[(size_t)c * 8 + i];
        }
    };
    load(a, ra);
    load(b, rb);
    float out[8];
    uint8_t d[8];
    float node = INFINITY;
    uint8_t node_k = 1;
    for (int k = 1; k < 8; k++) {
        const float v = ra[k - 1] + rb[8 - k - 1];
        if (v < node) {
            node = v;
            node_k = (uint8_t)k;
        }
    }
    node += half_area(pb) * kSahNode;
    const float leaf = cnt <= max_leaf ? half_area(pb) * kSahTri * (float)cnt : INFINITY;
    out[0] = leaf <= node ? leaf : node;
    d[0] = leaf <= node ? kDecLeaf : node_k;
    for (int i = 2; i <= 8; i++) {
        float best = out[i - 2];
        uint8_t bk = 0;
        for (int k = 1; k < i; k++) {
            const float v = ra[k - 1] + rb[i - k - 1];
            if (v < best) {
                best = v;
                bk = (uint8_t)k;
            }
        }
        out[i - 1] = best;
        d[i - 1] = bk;
    }
    for (int i = 0; i < 8; i++) {
        R[(size_t)p * 8 + i] = out[i];
        dec[(size_t)p * 8 + i] = d[i];
    }
}

// bottom-up refit: second arrival at a node computes its box (and, when R is
// given, the SAH programme of the node)
__global__ void k_refit(const TriRec* __restrict__ tris, const uint32_t* __restrict__ sorted, int n,
                        const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                        const uint32_t* __restrict__ parent_int, const uint32_t* __restrict__ parent_leaf,
                        uint32_t* __restrict__ flags, Box* __restrict__ ibox, Box* __restrict__ lbox,
                        const uint32_t* __restrict__ rlo, const uint32_t* __restrict__ rhi, uint32_t max_leaf,
                        float* __restrict__ R, uint8_t* __restrict__ dec) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lbox[i] = tri_box(tris[sorted[i]]);
    if (n == 1) return;
    __threadfence();
    uint32_t p = parent_leaf[i];
    for (;;) {
        uint32_t old = atomicAdd(&flags[p], 1u);
        if (old == 0) return;
        __threadfence();
        uint32_t a = left[p], b = right[p];
        // volatile reads: the sibling's box was written by another wave
        const volatile Box* ba = (a & 0x80000000u) ? (const volatile Box*)&lbox[a & 0x7fffffffu]
                                                   : (const volatile Box*)&ibox[a];
        const volatile Box* bb = (b & 0x80000000u) ? (const volatile Box*)&lbox[b & 0x7fffffffu]
                                                   : (const volatile Box*)&ibox[b];
        Box r;
        for (int k = 0; k < 3; k++) {
            r.lo[k] = fminf(ba->lo[k], bb->lo[k]);
            r.hi[k] = fmaxf(ba->hi[k], bb->hi[k]);
        }
        ibox[p] = r;
        if (R) sah_node(p, a, b, r, rhi[p] - rlo[p] + 1, max_leaf, lbox, R, R, dec);
        __threadfence();
        if (p == 0) return;
        p = parent_int[p];
    }
}

struct Task {
    uint32_t bnode;  // binary node (bit 31: leaf)
    uint32_t slot;   // BVH8 node index
};

__device__ __forceinline__ uint32_t count_of(uint32_t c, const uint32_t* rlo, const uint32_t* rhi) {
    if (c & 0x80000000u) return 1;
    return rhi[c] - rlo[c] + 1;
}

__device__ __forceinline__ Box box_of(uint32_t c, const Box* ibox, const Box* lbox) {
    return (c & 0x80000000u) ? lbox[c & 0x7fffffffu] : ibox[c];
}

__global__ void k_collapse(const Task* __restrict__ in, uint32_t n_in, Task* __restrict__ out,
                           uint32_t* __restrict__ ctr /* [0] nodes, [1] tris, [2] out tasks */,
                           const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                           const uint32_t* __restrict__ rlo, const uint32_t* __restrict__ rhi,
                           const Box* __restrict__ ibox, const Box* __restrict__ lbox,
                           const uint32_t* __restrict__ sorted, const TriRec* __restrict__ tris_in,
                           TriRec* __restrict__ tris_out, Node8* __restrict__ nodes, uint32_t max_leaf,
                           const uint8_t* __restrict__ dec) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_in) return;
    const Task task = in[g];
    uint32_t ch[8];
    int nc = 0;
    if (task.bnode & 0x80000000u) {
        ch[nc++] = task.bnode;  // a single-triangle scene
    } else if (count_of(task.bnode, rlo, rhi) <= max_leaf && (!dec || dec[(size_t)task.bnode * 8] == kDecLeaf)) {
        ch[nc++] = task.bnode;  // root small enough to be one leaf
    } else if (dec) {
        // the roots the SAH programme chose: the node's own split into 8,
        // then each (subtree, i roots) pair resolved by its decision
        uint32_t st_m[8], st_i[8];
        int sp = 0;
        const uint32_t k0 = dec[(size_t)task.bnode * 8] == kDecLeaf ? 4u : dec[(size_t)task.bnode * 8];
        st_m[sp] = left[task.bnode];
        st_i[sp++] = k0;
        st_m[sp] = right[task.bnode];
        st_i[sp++] = 8u - k0;
        while (sp > 0) {
            sp--;
            const uint32_t m = st_m[sp], i = st_i[sp];
            if ((m & 0x80000000u) || i == 1u) {
                ch[nc++] = m;
                continue;
            }
            const uint32_t dk = dec[(size_t)m * 8 + (i - 1)];
            if (dk == 0u) {
                st_m[sp] = m;
                st_i[sp++] = i - 1u;
            } else {
                st_m[sp] = left[m];
                st_i[sp++] = dk;
                st_m[sp] = right[m];
                st_i[sp++] = i - dk;
            }
        }
    } else {
        ch[nc++] = left[task.bnode];
        ch[nc++] = right[task.bnode];
        while (nc < 8) {
            int best = -1;
            float best_a = -1.0f;
            for (int c = 0; c < nc; c++) {
                if (ch[c] & 0x80000000u) continue;
                if (count_of(ch[c], rlo, rhi) <= max_leaf) continue;
                float a = half_area(ibox[ch[c]]);
                if (a > best_a) {
                    best_a = a;
                    best = c;
                }
            }
            if (best < 0) break;
            uint32_t b = ch[best];
            ch[best] = left[b];
            ch[nc++] = right[b];
        }
    }
    // classify; internal children take the first slots (in their order), so
    // the traversal finds an internal child's node at child_base + slot
    uint32_t n_int = 0, n_leaf_tris = 0;
    bool is_leaf[8];
    {
        // a child is a leaf if it is a binary leaf, or holds at most max_leaf
        // triangles (and the SAH programme made it a leaf)
        auto leaf_child = [&](uint32_t c) {
            if (c & 0x80000000u) return true;
            if (count_of(c, rlo, rhi) > max_leaf) return false;
            return !dec || dec[(size_t)c * 8] == kDecLeaf || nc == 1;
        };
        uint32_t ord[8];
        int ni = 0, nl = 0;
        for (int c = 0; c < nc; c++) {
            if (leaf_child(ch[c])) {
                n_leaf_tris += count_of(ch[c], rlo, rhi);
            } else {
                ord[ni++] = ch[c];
                n_int++;
            }
        }
        for (int c = 0; c < nc; c++)
            if (leaf_child(ch[c])) ord[ni + nl++] = ch[c];
        for (int c = 0; c < nc; c++) {
            ch[c] = ord[c];
            is_leaf[c] = c >= (int)n_int;
        }
    }
    uint32_t child_base = n_int ? atomicAdd(&ctr[0], n_int) : 0u;
    uint32_t tri_base = n_leaf_tris ? atomicAdd(&ctr[1], n_leaf_tris) : 0u;
    uint32_t task_base = n_int ? atomicAdd(&ctr[2], n_int) : 0u;
    // node box and quantisation
    Box cb[8];
    Box nb;
    for (int k = 0; k < 3; k++) {
        nb.lo[k] = INFINITY;
        nb.hi[k] = -INFINITY;
    }
    for (int c = 0; c < nc; c++) {
        cb[c] = box_of(ch[c], ibox, lbox);
        for (int k = 0; k < 3; k++) {
            nb.lo[k] = fminf(nb.lo[k], cb[c].lo[k]);
            nb.hi[k] = fmaxf(nb.hi[k], cb[c].hi[k]);
        }
    }
    Node8 nd;
    int ex[3];
    for (int k = 0; k < 3; k++) {
        nd.p[k] = nb.lo[k];
        double extent = (double)nb.hi[k] - (double)nb.lo[k];
        double s = extent / 65504.0;
        int e;
        if (s <= 0.0) {
            e = -100;
        } else {
            double m = frexp(s, &e);  // s = m * 2^e, m in [0.5, 1)
            if (m == 0.5) e -= 1;
            if (ldexp(65504.0, e) < extent) e += 1;
            if (e < -100) e = -100;
        }
        ex[k] = e;
        nd.e[k] = (int8_t)e;
    }
    nd.imask = 0;
    nd.child_base = child_base;
    nd.tri_base = tri_base;
    uint32_t int_slot = 0, tri_off = 0;
    for (int c = 0; c < 8; c++) {
        if (c >= nc) {
            nd.meta[c] = 0;
            for (int k = 0; k < 3; k++) {
                nd.qlo[k][c] = 0x7BFF;  // empty: lo (65504) > hi (0)
                nd.qhi[k][c] = 0;
            }
            continue;
        }
        for (int k = 0; k < 3; k++) {
            double inv = ldexp(1.0, -ex[k]);
            nd.qlo[k][c] = f16_floor(((double)cb[c].lo[k] - (double)nd.p[k]) * inv);
            nd.qhi[k][c] = f16_ceil(((double)cb[c].hi[k] - (double)nd.p[k]) * inv);
        }
        if (is_leaf[c]) {
            uint32_t first, cnt;
            if (ch[c] & 0x80000000u) {
                first = ch[c] & 0x7fffffffu;
                cnt = 1;
            } else {
                first = rlo[ch[c]];
                cnt = rhi[ch[c]] - rlo[ch[c]] + 1;
            }
            nd.meta[c] = (uint8_t)((tri_off << 3) | cnt);
            for (uint32_t t = 0; t < cnt; t++) tris_out[tri_base + tri_off + t] = tris_in[sorted[first + t]];
            tri_off += cnt;
        } else {
            nd.imask |= (uint8_t)(1u << c);
            nd.meta[c] = (uint8_t)int_slot;
            out[task_base + int_slot] = Task{ch[c], child_base + int_slot};
            int_slot++;
        }
    }
    nodes[task.slot] = nd;
}

}  // namespace

lt_status bvh_build(lt_context* ctx, const lt_build_opts* opts, lt_build_stats* stats) {
    if (!ctx->scene.uploaded) return lt_fail(ctx, LT_ERR_NOT_READY, "lt_accel_build: no scene uploaded");
    uint32_t max_leaf = opts && opts->max_leaf_tris ? opts->max_leaf_tris : 1u;
    if (max_leaf < 1 || max_leaf > 4) return lt_fail(ctx, LT_ERR_INVALID_ARG, "max_leaf_tris must be in 1..4");
    const uint32_t collapse = opts && opts->collapse ? opts->collapse : LT_COLLAPSE_SAH;
    if (collapse != LT_COLLAPSE_SAH && collapse != LT_COLLAPSE_GREEDY)
        return lt_fail(ctx, LT_ERR_INVALID_ARG, "unknown collapse");
    const bool sah = collapse == LT_COLLAPSE_SAH;
    const uint32_t tree = opts && opts->binary_tree ? opts->binary_tree : LT_TREE_PLOC;
    if (tree != LT_TREE_PLOC && tree != LT_TREE_LBVH) return lt_fail(ctx, LT_ERR_INVALID_ARG, "unknown binary tree");
    hipStream_t st = ctx->stream;
    DeviceBVH& bvh = ctx->bvh;
    if (bvh.nodes) (void)hipFree(bvh.nodes);
    if (bvh.tris) (void)hipFree(bvh.tris);
    bvh = DeviceBVH{};
    const uint32_t n = ctx->scene.ntris;
    LT_HIP(ctx, hipEventRecord(ctx->ev0, st));
    if (n == 0) {
        bvh.built = true;
        if (stats) *stats = lt_build_stats{};
        return LT_OK;
    }
    // temporaries
    TriRec* tris_unsorted = nullptr;
    uint32_t *cb = nullptr, *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr;
    uint32_t *left = nullptr, *right = nullptr, *pint = nullptr, *pleaf = nullptr, *rlo = nullptr, *rhi = nullptr,
             *flags = nullptr, *ctr = nullptr;
    Box *ibox = nullptr, *lbox = nullptr;
    Task *q0 = nullptr, *q1 = nullptr;
    void* sort_tmp = nullptr;
    size_t sort_bytes = 0;
    float* sah_R = nullptr;
    uint8_t* sah_dec = nullptr;
    // PLOC temporaries
    uint32_t *cl_node0 = nullptr, *cl_node1 = nullptr, *nn = nullptr, *keep = nullptr, *kpos = nullptr,
             *cnt = nullptr, *leaf_pos = nullptr, *vals3 = nullptr, *pleaf3 = nullptr, *mctr = nullptr;
    Box *cl_box0 = nullptr, *cl_box1 = nullptr;
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    lt_status rc = LT_OK;
    auto fail = [&](hipError_t e, const char* what) {
        rc = lt_hip_check(ctx, e, what);
        return rc;
    };
#define BCHK(x)                                       \
    do {                                              \
        hipError_t _e = (x);                          \
        if (_e != hipSuccess) {                       \
            fail(_e, #x);                             \
            goto cleanup;                             \
        }                                             \
    } while (0)
    {
        const uint32_t ni = n > 1 ? n - 1 : 1;
        BCHK(hipMalloc(&tris_unsorted, sizeof(TriRec) * n));
        BCHK(hipMalloc(&bvh.tris, sizeof(TriRec) * n));
        BCHK(hipMalloc(&bvh.nodes, sizeof(Node8) * (size_t)std::max<uint32_t>(n, 1)));
        BCHK(hipMalloc(&cb, 6 * sizeof(uint32_t)));
        BCHK(hipMalloc(&keys, 4ull * n));
        BCHK(hipMalloc(&vals, 4ull * n));
        BCHK(hipMalloc(&keys2, 4ull * n));
        BCHK(hipMalloc(&vals2, 4ull * n));
        BCHK(hipMalloc(&left, 4ull * ni));
        BCHK(hipMalloc(&right, 4ull * ni));
        BCHK(hipMalloc(&pint, 4ull * ni));
        BCHK(hipMalloc(&pleaf, 4ull * n));
        BCHK(hipMalloc(&rlo, 4ull * ni));
        BCHK(hipMalloc(&rhi, 4ull * ni));
        BCHK(hipMalloc(&flags, 4ull * ni));
        BCHK(hipMalloc(&ibox, sizeof(Box) * ni));
        BCHK(hipMalloc(&lbox, sizeof(Box) * n));
        BCHK(hipMalloc(&q0, sizeof(Task) * (size_t)n));
        BCHK(hipMalloc(&q1, sizeof(Task) * (size_t)n));
        BCHK(hipMalloc(&ctr, 4 * sizeof(uint32_t)));
        uint32_t init_cb[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
        BCHK(hipMemcpyAsync(cb, init_cb, sizeof(init_cb), hipMemcpyHostToDevice, st));
        BCHK(hipMemsetAsync(flags, 0, 4ull * ni, st));
        const int B = 256;
        const uint32_t G = (n + B - 1) / B;
        hipLaunchKernelGGL(k_make_tris, dim3(G), dim3(B), 0, st, ctx->scene.vtx, ctx->scene.idx, ctx->scene.prims,
                           ctx->scene.tri_start, ctx->scene.xf, ctx->scene.nprims, n, tris_unsorted, cb);
        BCHK(hipGetLastError());
        hipLaunchKernelGGL(k_morton, dim3(G), dim3(B), 0, st, tris_unsorted, n, cb, keys, vals);
        BCHK(hipGetLastError());
        BCHK(rocprim::radix_sort_pairs(nullptr, sort_bytes, keys, keys2, vals, vals2, n, 0u, 30u, st));
        BCHK(hipMalloc(&sort_tmp, sort_bytes));
        BCHK(rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys, keys2, vals, vals2, n, 0u, 30u, st));
        if (n > 1 && tree == LT_TREE_LBVH) {
            hipLaunchKernelGGL(k_karras, dim3((n - 1 + B - 1) / B), dim3(B), 0, st, keys2, (int)n, left, right, pint,
                               pleaf, rlo, rhi);
            BCHK(hipGetLastError());
        }
        if (n > 1 && tree == LT_TREE_PLOC) {
            BCHK(hipMalloc(&cl_node0, 4ull * n));
            BCHK(hipMalloc(&cl_node1, 4ull * n));
            BCHK(hipMalloc(&cl_box0, sizeof(Box) * n));
            BCHK(hipMalloc(&cl_box1, sizeof(Box) * n));
            BCHK(hipMalloc(&nn, 4ull * n));
            BCHK(hipMalloc(&keep, 4ull * n));
            BCHK(hipMalloc(&kpos, 4ull * n));
            BCHK(hipMalloc(&cnt, 4ull * ni));
            BCHK(hipMalloc(&leaf_pos, 4ull * n));
            BCHK(hipMalloc(&vals3, 4ull * n));
            BCHK(hipMalloc(&pleaf3, 4ull * n));
            BCHK(hipMalloc(&mctr, 4));
            BCHK(hipMemsetAsync(mctr, 0, 4, st));
            BCHK(rocprim::exclusive_scan(nullptr, scan_bytes, keep, kpos, 0u, n, rocprim::plus<uint32_t>(), st));
            BCHK(hipMalloc(&scan_tmp, scan_bytes));
            hipLaunchKernelGGL(k_ploc_init, dim3(G), dim3(B), 0, st, tris_unsorted, vals2, n, cl_node0, cl_box0);
            BCHK(hipGetLastError());
            uint32_t m = n;
            while (m > 1) {
                const uint32_t gm = (m + B - 1) / B;
                hipLaunchKernelGGL(k_ploc_nn, dim3(gm), dim3(B), 0, st, cl_box0, m, nn);
                hipLaunchKernelGGL(k_ploc_merge, dim3(gm), dim3(B), 0, st, cl_node0, cl_box0, m, nn, n, mctr, left,
                                   right, pint, pleaf, cl_node1, cl_box1, keep);
                BCHK(hipGetLastError());
                BCHK(rocprim::exclusive_scan(scan_tmp, scan_bytes, keep, kpos, 0u, m, rocprim::plus<uint32_t>(), st));
                hipLaunchKernelGGL(k_ploc_compact, dim3(gm), dim3(B), 0, st, cl_node1, cl_box1, m, keep, kpos,
                                   cl_node0, cl_box0);
                BCHK(hipGetLastError());
                uint32_t tail[2];
                BCHK(hipMemcpyAsync(&tail[0], kpos + (m - 1), 4, hipMemcpyDeviceToHost, st));
                BCHK(hipMemcpyAsync(&tail[1], keep + (m - 1), 4, hipMemcpyDeviceToHost, st));
                BCHK(hipStreamSynchronize(st));
                const uint32_t m_next = tail[0] + tail[1];
                if (m_next >= m) {
                    rc = lt_fail(ctx, LT_ERR_HIP, "PLOC round merged no clusters");
                    goto cleanup;
                }
                m = m_next;
            }
            // subtrees as contiguous runs: counts, ranges, new leaf order
            BCHK(hipMemsetAsync(flags, 0, 4ull * ni, st));
            hipLaunchKernelGGL(k_tree_count, dim3(G), dim3(B), 0, st, left, right, pint, pleaf, n, flags, cnt);
            hipLaunchKernelGGL(k_tree_ranges, dim3(G), dim3(B), 0, st, left, right, pint, pleaf, cnt, n, rlo, rhi,
                               leaf_pos);
            hipLaunchKernelGGL(k_tree_relabel, dim3(G), dim3(B), 0, st, left, right, leaf_pos, vals2, pleaf, n, vals3,
                               pleaf3);
            BCHK(hipGetLastError());
            std::swap(vals2, vals3);
            std::swap(pleaf, pleaf3);
            BCHK(hipMemsetAsync(flags, 0, 4ull * ni, st));
        }
        if (sah && n > 1) {
            BCHK(hipMalloc(&sah_R, sizeof(float) * 8 * (size_t)ni));
            BCHK(hipMalloc(&sah_dec, 8 * (size_t)ni));
        }
        hipLaunchKernelGGL(k_refit, dim3(G), dim3(B), 0, st, tris_unsorted, vals2, (int)n, left, right, pint, pleaf,
                           flags, ibox, lbox, rlo, rhi, max_leaf, sah_R, sah_dec);
        BCHK(hipGetLastError());
        // top-down collapse, one launch per BVH8 level
        uint32_t h_ctr[4] = {1u, 0u, 0u, 0u};  // node 0 is the root
        BCHK(hipMemcpyAsync(ctr, h_ctr, sizeof(h_ctr), hipMemcpyHostToDevice, st));
        Task root{n > 1 ? 0u : 0x80000000u, 0u};
        BCHK(hipMemcpyAsync(q0, &root, sizeof(Task), hipMemcpyHostToDevice, st));
        uint32_t n_in = 1, depth = 0;
        while (n_in > 0) {
            depth++;
            BCHK(hipMemsetAsync(ctr + 2, 0, 4, st));
            hipLaunchKernelGGL(k_collapse, dim3((n_in + 63) / 64), dim3(64), 0, st, q0, n_in, q1, ctr, left, right,
                               rlo, rhi, ibox, lbox, vals2, tris_unsorted, bvh.tris, bvh.nodes, max_leaf, sah_dec);
            BCHK(hipGetLastError());
            BCHK(hipMemcpyAsync(h_ctr, ctr, sizeof(h_ctr), hipMemcpyDeviceToHost, st));
            BCHK(hipStreamSynchronize(st));
            n_in = h_ctr[2];
            std::swap(q0, q1);
        }
        bvh.n_nodes = h_ctr[0];
        bvh.n_tris = h_ctr[1];
        bvh.depth = depth;
        BCHK(hipEventRecord(ctx->ev1, st));
        BCHK(hipEventSynchronize(ctx->ev1));
        float ms = 0.0f;
        BCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        if (bvh.n_tris != n) {
            rc = lt_fail(ctx, LT_ERR_HIP, "BVH collapse lost triangles");
            goto cleanup;
        }
        // traversal stack: one group entry per ancestor level; entries pack
        // child_base << 8, so node indices must stay below 2^24
        if (depth > (uint32_t)ltb::kStackTotal) {
            rc = lt_fail(ctx, LT_ERR_UNSUPPORTED,
                         "BVH8 depth " + std::to_string(depth) + " exceeds the traversal stack bound");
            goto cleanup;
        }
        if (bvh.n_nodes >= (1u << 24)) {
            rc = lt_fail(ctx, LT_ERR_UNSUPPORTED, "BVH8 has more than 2^24 nodes");
            goto cleanup;
        }
        {
            // structural check before any traversal runs: every internal child
            // index is larger than its parent's (so traversal terminates) and in
            // range, and every leaf range lies inside the triangle array
            std::vector<Node8> h(bvh.n_nodes);
            BCHK(hipMemcpy(h.data(), bvh.nodes, sizeof(Node8) * h.size(), hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < bvh.n_nodes; i++) {
                const Node8& nd = h[i];
                for (int c = 0; c < 8; c++) {
                    const uint32_t meta = nd.meta[c];
                    if ((nd.imask >> c) & 1u) {
                        const uint64_t ch = (uint64_t)nd.child_base + meta;
                        if (ch <= i || ch >= bvh.n_nodes) {
                            rc = lt_fail(ctx, LT_ERR_HIP, "BVH8 node " + std::to_string(i) + " has a bad child");
                            goto cleanup;
                        }
                    } else if (meta & 7u) {
                        const uint64_t end = (uint64_t)nd.tri_base + (meta >> 3) + (meta & 7u);
                        if (end > bvh.n_tris) {
                            rc = lt_fail(ctx, LT_ERR_HIP, "BVH8 node " + std::to_string(i) + " has a bad leaf");
                            goto cleanup;
                        }
                    }
                }
            }
        }
        bvh.built = true;
        if (stats) {
            stats->num_nodes = bvh.n_nodes;
            stats->num_tris = bvh.n_tris;
            stats->max_depth = depth;
            stats->binary_depth = 0;
            stats->build_ms = ms;
        }
    }
cleanup:
#undef BCHK
    void* tmp[] = {tris_unsorted, cb, keys, vals, keys2, vals2, left, right, pint, pleaf, rlo, rhi, flags, ibox,
                   lbox, q0, q1, ctr, sort_tmp, sah_R, sah_dec, cl_node0, cl_node1, cl_box0, cl_box1, nn, keep,
                   kpos, cnt, leaf_pos, vals3, pleaf3, mctr, scan_tmp};
    for (void* p : tmp)
        if (p) (void)hipFree(p);
    if (rc != LT_OK) {
        if (bvh.nodes) (void)hipFree(bvh.nodes);
        if (bvh.tris) (void)hipFree(bvh.tris);
        bvh = DeviceBVH{};
    }
    return rc;
}
