// bvh_layout.h — HBM layout of the BVH8 and of the triangle records
// (shared by the builder and the traversal kernels).
//
// BVH8 node: exactly 128 bytes (one gfx950 L2 cache line), 128-B aligned.
// Child boxes are stored per plane as fp16 values q in [0, 65504] against a
// per-node origin p and power-of-two scales 2^e (lower planes rounded down,
// upper planes rounded up), so decoding p + q * 2^e is exact in the product
// and monotone in the sum: a decoded box always contains the child.  fp16
// rather than 16-bit integers lets the slab test convert and multiply-add a
// plane in one v_fma_mix_f32.  Empty slots: lower 65504, upper 0.
//
//   bytes   0..11  p.x p.y p.z         origin (min corner of the node box)
//          12..14  e.x e.y e.z         int8 exponents
//          15      imask               bit i: child i is an internal node; the
//                                      internal children fill the low slots, so
//                                      imask = 2^n_internal - 1
//          16..19  child_base          node index of the first internal child
//          20..23  tri_base            index of the first triangle record
//          24..31  meta[8]             internal: its slot (node child_base + slot)
//                                      leaf: (tri offset << 3) | count, 0 = empty
//          32..79  qlo[3][8]           fp16 lower planes per axis
//          80..127 qhi[3][8]           fp16 upper planes per axis
//
// Triangle record: 48 bytes = v0, v1, v2 (world space, 36 B) + global
// triangle index + mesh-local primitive index + mesh (instance) index.
#pragma once
#include <stdint.h>

namespace ltb {

struct alignas(128) Node8 {
    float p[3];
    int8_t e[3];
    uint8_t imask;
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t meta[8];
    uint16_t qlo[3][8];
    uint16_t qhi[3][8];
};
static_assert(sizeof(Node8) == 128, "BVH8 node must be one 128-byte line");

struct alignas(16) TriRec {
    float v0[3];
    float v1[3];
    float v2[3];
    uint32_t gid;   // global triangle index (tie-break key)
    uint32_t prim;  // mesh-local primitive index (gl_PrimitiveID)
    uint32_t inst;  // mesh index (gl_InstanceCustomIndexEXT)
};
static_assert(sizeof(TriRec) == 48, "triangle record is 48 bytes");

// traversal stack entries per thread kept in LDS, and queued leaf entries per
// thread (LDS); build-time parameters of the autotuner (lumen_amd/autotune.py)
#ifndef LT_LDS_STACK
#define LT_LDS_STACK 16
#endif
#ifndef LT_LEAF_QUEUE
#define LT_LEAF_QUEUE 16
#endif
constexpr int kLdsStack = LT_LDS_STACK;
constexpr int kLeafQueue = LT_LEAF_QUEUE;
static_assert(kLdsStack >= 4 && kLdsStack <= 32 && kLeafQueue >= 8 && kLeafQueue <= 32, "LDS stack / queue sizes");
constexpr int kLdsWords = kLdsStack + kLeafQueue;  // LDS words per thread of a traversal
constexpr int kStackTotal = 192;  // LDS + global spill entries per thread
constexpr int kMaxLeafTris = 7;   // count field is 3 bits

}  // namespace ltb
