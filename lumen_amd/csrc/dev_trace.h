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

// gamma(3) widening of PBRT's robust slab test
constexpr float kBoxWiden = 1.0f + 2.0f * (3.0f * 5.9604645e-8f) / (1.0f - 3.0f * 5.9604645e-8f);

struct TraceStack {
    uint32_t* lds;    // this thread's column: entry i at lds[i * kTraceBlock]
    uint32_t* spill;  // this thread's spill area
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

// Traverses the BVH8.  ANY: returns at the first hit in [tmin, tmax].
// Counters (nodes popped, triangles tested) are only updated when COUNT.
template <bool ANY, bool COUNT>
LT_DEV bool trace_bvh(const SceneView& sv, f3 o, f3 d, float tmin, float tmax, HitRec& best, const TraceStack& stk,
                      uint32_t& n_nodes, uint32_t& n_tris) {
    best.gid = 0xFFFFFFFFu;
    best.t = tmax;
    best.prim = best.inst = 0xFFFFFFFFu;
    best.u = best.v = 0.0f;
    if (sv.num_nodes == 0) return false;
    const RayPre rp = prep_ray(d);
    float tcur = tmax;
    uint32_t sp = 0;
    stk_push(stk, sp, 0u);
    while (sp > 0) {
        const uint32_t ni = stk_pop(stk, sp);
        if (COUNT) n_nodes++;
        const ltb::Node8* np = sv.nodes + ni;
        const float4 w0 = *((const float4*)np + 0);
        const uint4 w1 = *((const uint4*)np + 1);
        const uint4 q0 = *((const uint4*)np + 2);
        const uint4 q1 = *((const uint4*)np + 3);
        const uint4 q2 = *((const uint4*)np + 4);
        const uint4 q3 = *((const uint4*)np + 5);
        const uint4 q4 = *((const uint4*)np + 6);
        const uint4 q5 = *((const uint4*)np + 7);
        const float px = w0.x, py = w0.y, pz = w0.z;
        const uint32_t ew = __float_as_uint(w0.w);  // e.x, e.y, e.z, imask
        const int ex = (int)(int8_t)(ew & 0xFF), ey = (int)(int8_t)((ew >> 8) & 0xFF),
                  ez = (int)(int8_t)((ew >> 16) & 0xFF);
        const uint32_t imask = ew >> 24;
        const float sx = __int_as_float((ex + 127) << 23), sy = __int_as_float((ey + 127) << 23),
                    sz = __int_as_float((ez + 127) << 23);
        const uint32_t child_base = w1.x, tri_base = w1.y;
        const uint32_t meta_lo = w1.z, meta_hi = w1.w;
        const uint32_t qlx[4] = {q0.x, q0.y, q0.z, q0.w}, qly[4] = {q1.x, q1.y, q1.z, q1.w},
                       qlz[4] = {q2.x, q2.y, q2.z, q2.w};
        const uint32_t qhx[4] = {q3.x, q3.y, q3.z, q3.w}, qhy[4] = {q4.x, q4.y, q4.z, q4.w},
                       qhz[4] = {q5.x, q5.y, q5.z, q5.w};
        // The watertight t of a triangle carries an absolute error of a few ulp
        // of the largest sheared z of its vertices, i.e. of the farthest t of
        // the enclosing box along the dominant axis kz.  Children are kept if
        // they reach [tmin - margin, tcur + margin], so a triangle whose computed
        // t undercuts the current best by that error is still tested.
        const float pkz = rp.kz == 0 ? px : (rp.kz == 1 ? py : pz);
        const float skz = rp.kz == 0 ? sx : (rp.kz == 1 ? sy : sz);
        const float okz = sel(o, rp.kz), ikz = sel(rp.inv, rp.kz);
        const float far_kz = fmaxf(fabsf((pkz - okz) * ikz), fabsf((pkz + 65535.0f * skz - okz) * ikz));
        const float margin = far_kz * (1.0f / 65536.0f);
        const float lo_lim = tmin - margin;
        const float hi_lim = tcur + margin;
        uint32_t hit_int = 0, hit_leaf = 0;
        float tn_c[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int sh = (c & 1) * 16;
            const float lx = px + (float)((qlx[c >> 1] >> sh) & 0xFFFF) * sx;
            const float ly = py + (float)((qly[c >> 1] >> sh) & 0xFFFF) * sy;
            const float lz = pz + (float)((qlz[c >> 1] >> sh) & 0xFFFF) * sz;
            const float hx = px + (float)((qhx[c >> 1] >> sh) & 0xFFFF) * sx;
            const float hy = py + (float)((qhy[c >> 1] >> sh) & 0xFFFF) * sy;
            const float hz = pz + (float)((qhz[c >> 1] >> sh) & 0xFFFF) * sz;
            float ta = (lx - o.x) * rp.inv.x, tb = (hx - o.x) * rp.inv.x;
            float t0 = fminf(ta, tb), t1 = fmaxf(ta, tb);
            ta = (ly - o.y) * rp.inv.y;
            tb = (hy - o.y) * rp.inv.y;
            t0 = fmaxf(t0, fminf(ta, tb));
            t1 = fminf(t1, fmaxf(ta, tb));
            ta = (lz - o.z) * rp.inv.z;
            tb = (hz - o.z) * rp.inv.z;
            t0 = fmaxf(t0, fminf(ta, tb));
            t1 = fminf(t1, fmaxf(ta, tb));
            t0 = t0 > 0.0f ? t0 * (1.0f / kBoxWiden) : t0 * kBoxWiden;
            t1 = t1 > 0.0f ? t1 * kBoxWiden : t1 * (1.0f / kBoxWiden);
            t0 = fmaxf(t0, lo_lim);
            t1 = fminf(t1, hi_lim);
            const bool h = (t0 <= t1) && (lx <= hx);
            tn_c[c] = t0;
            const uint32_t meta = ((c < 4 ? meta_lo : meta_hi) >> ((c & 3) * 8)) & 0xFF;
            const bool internal = (imask >> c) & 1u;
            if (h && internal) hit_int |= 1u << c;
            if (h && !internal && (meta & 7u) != 0) hit_leaf |= 1u << c;
        }
        // leaves first: shrink tcur before ordering the internal children
        while (hit_leaf) {
            const int c = __builtin_ctz(hit_leaf);
            hit_leaf &= hit_leaf - 1;
            const uint32_t meta = ((c < 4 ? meta_lo : meta_hi) >> ((c & 3) * 8)) & 0xFF;
            const uint32_t first = tri_base + (meta >> 3), cnt = meta & 7u;
            for (uint32_t k = 0; k < cnt; k++) {
                const float4* tp = (const float4*)(sv.tris + first + k);
                const float4 a = tp[0], b = tp[1], cc = tp[2];
                if (COUNT) n_tris++;
                float t, u, v;
                if (isect_tri(mk3(a.x, a.y, a.z), mk3(a.w, b.x, b.y), mk3(b.z, b.w, cc.x), o, rp, tmin, tmax, t, u,
                              v)) {
                    const uint32_t gid = __float_as_uint(cc.y);
                    if (ANY) {
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
        }
        if (hit_int) {
            // nearest internal child is pushed last (popped first)
            int nearest = -1;
            float tbest = INFINITY;
#pragma unroll
            for (int c = 0; c < 8; c++)
                if (((hit_int >> c) & 1u) && tn_c[c] < tbest) {
                    tbest = tn_c[c];
                    nearest = c;
                }
            uint32_t rest = hit_int & ~(1u << nearest);
            while (rest) {
                const int c = __builtin_ctz(rest);
                rest &= rest - 1;
                const uint32_t meta = ((c < 4 ? meta_lo : meta_hi) >> ((c & 3) * 8)) & 0xFF;
                stk_push(stk, sp, child_base + meta);
            }
            const uint32_t meta = ((nearest < 4 ? meta_lo : meta_hi) >> ((nearest & 3) * 8)) & 0xFF;
            stk_push(stk, sp, child_base + meta);
        }
    }
    return best.gid != 0xFFFFFFFFu;
}

}  // namespace ltd
