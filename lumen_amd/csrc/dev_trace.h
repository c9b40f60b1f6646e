// dev_trace.h — BVH8 traversal for gfx950: closest hit and any hit.
//
// Semantics of the reference's queries (traceRayEXT with gl_RayFlagsOpaqueEXT,
// and TerminateOnFirstHit|SkipClosestHit for shadows;
// src/shaders/integrators/path/path.rgen:48, pt_commons.glsl:20-23):
//   closest = the triangle with the smallest t in [tmin, tmax], ties to the
//             smaller global triangle index;
//   any     = some triangle with t in [tmin, tmax].
// Triangle test: Woop, Benthin & Wald, JCGT 2013 (watertight, with the
// double-precision edge fallback), evaluated without fused multiply-adds so
// that t, u, v are bitwise those of the host restatement.
//
// Stack: a per-thread short stack in LDS (kLdsStack entries, stride = block
// size so a wave's accesses are bank-conflict free), continued in a global
// spill area (kStackTotal - kLdsStack entries per resident thread).
#pragma once
#include "bvh_layout.h"
#include "context.h"
#include "dev_math.h"

namespace ltd {

struct RayPre {
    int kx, ky, kz;
    float Sx, Sy, Sz;
    f3 inv;
};

LT_DEV float sel(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

LT_DEV RayPre prep_ray(f3 d) {
#pragma clang fp contract(off)
    RayPre r;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    float dz = sel(d, kz);
    if (dz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.Sx = sel(d, kx) / dz;
    r.Sy = sel(d, ky) / dz;
    r.Sz = 1.0f / dz;
    // box-test reciprocal; exact zeros get a tiny same-signed stand-in
    const float tiny = 1e-30f;
    r.inv = {1.0f / (fabsf(d.x) < tiny ? copysignf(tiny, d.x) : d.x),
             1.0f / (fabsf(d.y) < tiny ? copysignf(tiny, d.y) : d.y),
             1.0f / (fabsf(d.z) < tiny ? copysignf(tiny, d.z) : d.z)};
    return r;
}

// Woop et al. 2013, listing 1 (no back-face culling)
LT_DEV bool isect_tri(f3 A0, f3 B0, f3 C0, f3 o, const RayPre& rp, float tmin, float tmax, float& t_out,
                      float& u_out, float& v_out) {
#pragma clang fp contract(off)
    const f3 A = A0 - o, B = B0 - o, C = C0 - o;
    const float Akz = sel(A, rp.kz), Bkz = sel(B, rp.kz), Ckz = sel(C, rp.kz);
    const float Ax = sel(A, rp.kx) - rp.Sx * Akz, Ay = sel(A, rp.ky) - rp.Sy * Akz;
    const float Bx = sel(B, rp.kx) - rp.Sx * Bkz, By = sel(B, rp.ky) - rp.Sy * Bkz;
    const float Cx = sel(C, rp.kx) - rp.Sx * Ckz, Cy = sel(C, rp.ky) - rp.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    // the double-precision edge test is rare (an edge function exactly 0):
    // a wave-uniform branch keeps it off the common path
    const bool edge0 = U == 0.0f || V == 0.0f || W == 0.0f;
    if (__ballot(edge0) != 0ull && edge0) {
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
    const float Az = rp.Sz * Akz, Bz = rp.Sz * Bkz, Cz = rp.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    const float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    u_out = V * rcp;
    v_out = W * rcp;
    return true;
}

struct HitRec {
    float t, u, v;
    uint32_t gid, prim, inst;  // gid = 0xFFFFFFFF on a miss
};

// Relative bound (2^-21) on the rounding of a child slab distance q * a + b:
// a = 2^e * inv and b = (p - o) * inv each carry one rounding, inv one more
// and the fused multiply-add one, i.e. |error| < 4 eps (65535 |a| + |b|)
// (planes are at most 65504).
constexpr float kSlabErr = 4.76837158e-7f;

// LDS-qualified so that stack accesses compile to ds_read / ds_write rather
// than flat accesses through the generic address space
typedef __attribute__((address_space(3))) uint32_t lds_u32;

struct TraceStack {
    lds_u32* lds;     // this thread's column: entry i at lds[i * kTraceBlock]
    lds_u32* q;       // its leaf queue, after the kLdsStack stack entries
    uint32_t* spill;  // this thread's spill area
    LT_DEV TraceStack(uint32_t* lds_col, uint32_t* spill_area)
        : lds((lds_u32*)lds_col), q((lds_u32*)lds_col + ltb::kLdsStack * kTraceBlock), spill(spill_area) {}
};

LT_DEV void stk_push(const TraceStack& s, uint32_t& sp, uint32_t v) {
    if (sp < (uint32_t)ltb::kLdsStack)
        s.lds[sp * kTraceBlock] = v;
    else
        s.spill[sp - ltb::kLdsStack] = v;
    sp++;
}
LT_DEV uint32_t stk_pop(const TraceStack& s, uint32_t& sp) {
    sp--;
    return sp < (uint32_t)ltb::kLdsStack ? s.lds[sp * kTraceBlock] : s.spill[sp - ltb::kLdsStack];
}
LT_DEV uint32_t stk_top(const TraceStack& s, uint32_t sp) {
    return sp - 1 < (uint32_t)ltb::kLdsStack ? s.lds[(sp - 1) * kTraceBlock] : s.spill[sp - 1 - ltb::kLdsStack];
}
LT_DEV void stk_set_top(const TraceStack& s, uint32_t sp, uint32_t v) {
    if (sp - 1 < (uint32_t)ltb::kLdsStack)
        s.lds[(sp - 1) * kTraceBlock] = v;
    else
        s.spill[sp - 1 - ltb::kLdsStack] = v;
}

// a lane takes its next node while at most this many leaf entries are queued
#ifndef LT_NODE_PEND_MAX
#define LT_NODE_PEND_MAX (ltb::kLeafQueue - 8)
#endif
// a node step queues up to 8 more entries: the queue must hold them
static_assert(LT_NODE_PEND_MAX >= 0 && LT_NODE_PEND_MAX + 8 <= (int)ltb::kLeafQueue,
              "LT_NODE_PEND_MAX + 8 leaf entries exceed the per-lane LDS leaf queue");

// A stack entry is a group of sibling internal nodes still to visit:
// (child_base << 8) | mask over their slots (the internal children fill the
// low slots, child slot c is node child_base + c).  One entry per tree level
// instead of one per child; the group's nodes are visited from the highest
// slot down.
constexpr uint32_t kNoNode = 0xFFFFFFFFu;

// The traversal of one query as a state machine, so that a lane can take a
// new query as soon as its current one finishes (trace_step returns true).
struct TraceState {
    f3 o, d;
    RayPre rp;
    float tmin, tmax, tcur;
    HitRec best;
    uint32_t sp;
    uint32_t next;  // node to visit next (the nearest hit child), or kNoNode
    uint32_t pend;  // leaf entries queued, (first triangle << 3) | count each
};

LT_DEV void trace_init(TraceState& s, const TraceStack& stk, f3 o, f3 d, float tmin, float tmax, bool has_nodes) {
    s.o = o;
    s.d = d;
    s.tmin = tmin;
    s.tmax = tmax;
    s.tcur = tmax;
    s.best.gid = 0xFFFFFFFFu;
    s.best.t = tmax;
    s.best.prim = s.best.inst = 0xFFFFFFFFu;
    s.best.u = s.best.v = 0.0f;
    s.rp = prep_ray(d);
    s.sp = 0;
    s.pend = 0;
    s.next = has_nodes ? 0u : kNoNode;
    (void)stk;
}

// Processes one BVH8 node (its children's boxes, then its leaf triangles,
// then pushes the hit internal children, nearest last).  any: finishes at the
// first hit in [tmin, tmax] (a per-lane flag; constant where trace_step's
// template argument gives it).  Counters are only updated when COUNT.
template <bool COUNT>
LT_DEV bool trace_step_dyn(const SceneView& sv, TraceState& s, const TraceStack& stk, bool any, uint32_t& n_nodes,
                           uint32_t& n_tris) {
    const bool has_node = s.next != kNoNode || s.sp != 0;
    if (!has_node && s.pend == 0) return true;
    const f3 o = s.o;
    const RayPre& rp = s.rp;
    const float tmin = s.tmin, tmax = s.tmax;
    float& tcur = s.tcur;
    HitRec& best = s.best;
    uint32_t& sp = s.sp;
    if (has_node && s.pend <= (uint32_t)LT_NODE_PEND_MAX) {
        uint32_t ni = s.next;
        if (ni == kNoNode) {
            const uint32_t e = stk_top(stk, sp);
            const uint32_t ranks = e & 0xFFu;
            const uint32_t r = 31u - (uint32_t)__builtin_clz(ranks);
            const uint32_t left = ranks & ~(1u << r);
            if (left)
                stk_set_top(stk, sp, (e & ~0xFFu) | left);
            else
                sp--;
            ni = (e >> 8) + r;
        }
        s.next = kNoNode;
        if (COUNT) n_nodes++;
        // 32-bit byte offsets from the node array (scalar base + vector offset
        // addressing; the builder caps a BVH at 2^24 nodes)
        const char* nbase = (const char*)sv.nodes;
        const uint32_t noff = ni * 128u;
        const float4 w0 = *(const float4*)(nbase + noff);
        const uint4 w1 = *(const uint4*)(nbase + (noff + 16u));
        // near / far planes by the ray octant: the plane vectors sit at
        // 16-byte slots 2..4 (qlo x, y, z) and 5..7 (qhi), so the octant picks
        // the load address instead of selecting between loaded words
        const bool negx = rp.inv.x < 0.0f, negy = rp.inv.y < 0.0f, negz = rp.inv.z < 0.0f;
        const uint4 qnx = *(const uint4*)(nbase + (noff + (negx ? 80u : 32u)));
        const uint4 qfx = *(const uint4*)(nbase + (noff + (negx ? 32u : 80u)));
        const uint4 qny = *(const uint4*)(nbase + (noff + (negy ? 96u : 48u)));
        const uint4 qfy = *(const uint4*)(nbase + (noff + (negy ? 48u : 96u)));
        const uint4 qnz = *(const uint4*)(nbase + (noff + (negz ? 112u : 64u)));
        const uint4 qfz = *(const uint4*)(nbase + (noff + (negz ? 64u : 112u)));
        const uint32_t ew = __float_as_uint(w0.w);  // e.x, e.y, e.z, imask
        const int ex = (int)(int8_t)(ew & 0xFF), ey = (int)(int8_t)((ew >> 8) & 0xFF),
                  ez = (int)(int8_t)((ew >> 16) & 0xFF);
        const uint32_t imask = ew >> 24;
        const uint32_t child_base = w1.x, tri_base = w1.y;
        const uint32_t meta_lo = w1.z, meta_hi = w1.w;
        // Slab distances of the children: t = q * a + b per axis with
        // a = 2^e / d and b = (p - o) / d.  E bounds the rounding of that
        // evaluation against the exact (p + q 2^e - o) / d, so
        // [t_near - E, t_far + E] contains every child's true slab interval.
        // 2^e * inv, exact (a power-of-two scaling), as one ldexp
        const float ax = ldexpf(rp.inv.x, ex);
        const float ay = ldexpf(rp.inv.y, ey);
        const float az = ldexpf(rp.inv.z, ez);
        const float bx = (w0.x - o.x) * rp.inv.x, by = (w0.y - o.y) * rp.inv.y, bz = (w0.z - o.z) * rp.inv.z;
        const float Ex = kSlabErr * fmaf(65535.0f, fabsf(ax), fabsf(bx));
        const float Ey = kSlabErr * fmaf(65535.0f, fabsf(ay), fabsf(by));
        const float Ez = kSlabErr * fmaf(65535.0f, fabsf(az), fabsf(bz));
        const float bnx = bx - Ex, bfx = bx + Ex, bny = by - Ey, bfy = by + Ey, bnz = bz - Ez, bfz = bz + Ez;
        // The watertight t of a triangle carries an absolute error of a few ulp
        // of the largest sheared z of its vertices, i.e. of the farthest t of
        // the enclosing box along the dominant axis kz.  Children are kept if
        // they reach [tmin - margin, tcur + margin], so a triangle whose computed
        // t undercuts the current best by that error is still tested.
        const float akz = rp.kz == 0 ? ax : (rp.kz == 1 ? ay : az);
        const float bkz = rp.kz == 0 ? bx : (rp.kz == 1 ? by : bz);
        const float far_kz = fmaxf(fabsf(bkz), fabsf(fmaf(65535.0f, akz, bkz)));
        const float margin = far_kz * (1.0f / 65536.0f);
        const float lo_lim = tmin - margin;
        const float hi_lim = tcur + margin;
        uint32_t hm = 0;
        float tn_c[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int sh = (c & 1) * 16;
            const int wd = c >> 1;
            auto lane16 = [&](const uint4& q) {  // the fp16 plane; the fma below takes it as is
                const uint32_t w = wd == 0 ? q.x : (wd == 1 ? q.y : (wd == 2 ? q.z : q.w));
                return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> sh));
            };
            float t0 = fmaxf(fmaxf(fmaf(lane16(qnx), ax, bnx), fmaf(lane16(qny), ay, bny)), fmaf(lane16(qnz), az, bnz));
            float t1 = fminf(fminf(fmaf(lane16(qfx), ax, bfx), fmaf(lane16(qfy), ay, bfy)), fmaf(lane16(qfz), az, bfz));
            t0 = fmaxf(t0, lo_lim);
            t1 = fminf(t1, hi_lim);
            tn_c[c] = t0;
            hm |= (t0 <= t1 ? 1u : 0u) << c;
        }
        const uint32_t hit_int = hm & imask;
        // hit slots that are not internal: leaves, and (rarely, within the
        // rounding margin) empty slots, whose count of 0 tests nothing
        uint32_t hit_leaf = hm & ~imask;
        // the hit leaves join the lane's queue; their triangles are tested
        // below, a few lanes' worth at a time
        while (hit_leaf) {
            const int c = __builtin_ctz(hit_leaf);
            hit_leaf &= hit_leaf - 1;
            const uint32_t meta = ((c < 4 ? meta_lo : meta_hi) >> ((c & 3) * 8)) & 0xFF;
            if (meta & 7u) {
                stk.q[s.pend * kTraceBlock] = ((tri_base + (meta >> 3)) << 3) | (meta & 7u);
                s.pend++;
            }
        }
        if (hit_int) {
            // nearest internal child is pushed last (popped first); it starts
            // at the first hit child, so that it names a hit child even if no
            // entry distance compares below +inf (an overflowed slab)
            int nearest = __builtin_ctz(hit_int);
            // an occlusion query takes its children in slot order: any hit
            // ends it, and without the nearest-first choice it also visits
            // fewer nodes (a host replay of the BVH8s: 9.3 -> 9.0 nodes per
            // shadow ray on a terrain, 9.5 -> 8.4 on the caustics scene)
            if (!any) {
                float tbest = INFINITY;
#pragma unroll
                for (int c = 0; c < 8; c++)
                    if (((hit_int >> c) & 1u) && tn_c[c] < tbest) {
                        tbest = tn_c[c];
                        nearest = c;
                    }
            }
            // the nearest is visited next; the others wait as one group and
            // are visited from the highest slot down (internal children fill
            // the low slots: child c is node child_base + c)
            const uint32_t rest = hit_int & ~(1u << nearest);
            if (rest) stk_push(stk, sp, (child_base << 8) | rest);
            s.next = child_base + (uint32_t)nearest;
        }
    }
    // One queued triangle per lane and step: a node step and a triangle test
    // share the step instead of the wave looping over the most leaf hits of
    // any lane.  The other queued triangles wait for the next steps, which
    // only delays the shrinking of tcur (culling stays conservative); an
    // any-hit query ends at its first hit.
    if (s.pend != 0u) {
        const uint32_t e = stk.q[(s.pend - 1) * kTraceBlock];
        const uint32_t cnt = e & 7u;
        if (cnt == 1u)
            s.pend--;
        else
            stk.q[(s.pend - 1) * kTraceBlock] = e - 1u;
        const float4* tp = (const float4*)(sv.tris + (e >> 3) + cnt - 1u);
        const float4 a = tp[0], b = tp[1], cc = tp[2];
        if (COUNT) n_tris++;
        float t, u, v;
        if (isect_tri(mk3(a.x, a.y, a.z), mk3(a.w, b.x, b.y), mk3(b.z, b.w, cc.x), o, rp, tmin, tmax, t, u, v)) {
            const uint32_t gid = __float_as_uint(cc.y);
            if (any) {
                best.t = t;
                best.u = u;
                best.v = v;
                best.gid = gid;
                best.prim = __float_as_uint(cc.z);
                best.inst = __float_as_uint(cc.w);
                return true;
            }
            if (t < best.t || (t == best.t && gid < best.gid)) {
                best.t = t;
                best.u = u;
                best.v = v;
                best.gid = gid;
                best.prim = __float_as_uint(cc.z);
                best.inst = __float_as_uint(cc.w);
                tcur = t;
            }
        }
    }
    return s.next == kNoNode && sp == 0 && s.pend == 0u;
}

template <bool ANY, bool COUNT>
LT_DEV bool trace_step(const SceneView& sv, TraceState& s, const TraceStack& stk, uint32_t& n_nodes,
                       uint32_t& n_tris) {
    return trace_step_dyn<COUNT>(sv, s, stk, ANY, n_nodes, n_tris);
}

// Streams the queries [0, n_items) of a queue through one wave: every lane
// steps its own query one node at a time, and lanes whose query finished
// take new ones (one atomic per refill, refills once sv.refill_min lanes are
// idle), so the wave stays full however uneven the traversal lengths are.
// load(item, o, d, tmin, tmax) returns a tag handed to finish(tag, state).
template <bool ANY, bool COUNT, class Load, class Finish>
LT_DEV void trace_stream(const SceneView& sv, const TraceStack& stk, uint32_t n_items, uint32_t* work, Load load,
                         Finish finish, uint32_t& n_nodes, uint32_t& n_tris) {
    const uint32_t lane = threadIdx.x & 63u;
    TraceState s;
    uint32_t tag = 0;
    bool has = false;
    bool exhausted = false;  // wave-uniform
    for (;;) {
        if (!exhausted) {
            const unsigned long long idle = __ballot(!has);
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            if (n_idle >= sv.refill_min || n_idle == 64u) {
                const int leader = __ffsll((long long)idle) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(work, n_idle);
                base = __shfl(base, leader);
                if (base + n_idle >= n_items) exhausted = true;
                if (!has) {
                    const uint32_t mine = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                    if (mine < n_items) {
                        f3 o, d;
                        float tmin, tmax;
                        tag = load(mine, o, d, tmin, tmax);
                        trace_init(s, stk, o, d, tmin, tmax, sv.num_nodes != 0);
                        has = true;
                    }
                }
            }
        }
        if (__ballot(has) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (has && trace_step<ANY, COUNT>(sv, s, stk, n_nodes, n_tris)) {
            finish(tag, s);
            has = false;
        }
    }
}

// trace_stream over a queue whose queries are closest-hit or any-hit one by
// one: load(item, o, d, tmin, tmax, any) also says which.
template <bool COUNT, class Load, class Finish>
LT_DEV void trace_stream_mixed(const SceneView& sv, const TraceStack& stk, uint32_t n_items, uint32_t* work, Load load,
                               Finish finish, uint32_t& n_nodes, uint32_t& n_tris) {
    const uint32_t lane = threadIdx.x & 63u;
    TraceState s;
    uint32_t tag = 0;
    bool has = false, any = false;
    bool exhausted = false;  // wave-uniform
    for (;;) {
        if (!exhausted) {
            const unsigned long long idle = __ballot(!has);
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            if (n_idle >= sv.refill_min || n_idle == 64u) {
                const int leader = __ffsll((long long)idle) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(work, n_idle);
                base = __shfl(base, leader);
                if (base + n_idle >= n_items) exhausted = true;
                if (!has) {
                    const uint32_t mine = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                    if (mine < n_items) {
                        f3 o, d;
                        float tmin, tmax;
                        tag = load(mine, o, d, tmin, tmax, any);
                        trace_init(s, stk, o, d, tmin, tmax, sv.num_nodes != 0);
                        has = true;
                    }
                }
            }
        }
        if (__ballot(has) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (has && trace_step_dyn<COUNT>(sv, s, stk, any, n_nodes, n_tris)) {
            finish(tag, s);
            has = false;
        }
    }
}

// Whole query on one lane.
template <bool ANY, bool COUNT>
LT_DEV bool trace_bvh(const SceneView& sv, f3 o, f3 d, float tmin, float tmax, HitRec& best, const TraceStack& stk,
                      uint32_t& n_nodes, uint32_t& n_tris) {
    TraceState s;
    trace_init(s, stk, o, d, tmin, tmax, sv.num_nodes != 0);
    while (!trace_step<ANY, COUNT>(sv, s, stk, n_nodes, n_tris)) {
    }
    best = s.best;
    return s.best.gid != 0xFFFFFFFFu;
}

}  // namespace ltd
