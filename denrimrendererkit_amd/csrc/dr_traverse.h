// dr_traverse.h -- software BVH traversal and triangle test for gfx950.
//
// CDNA4 has no ray-tracing units, so traversal is plain VALU work:
//   * one node fetch (4 x dwordx4, 64 B) tests both child boxes with a
//     fused-multiply-add slab test against precomputed 1/d and o/d;
//   * leaf children are intersected as soon as they are hit, nearer inner
//     child first, the farther one pushed;
//   * the traversal stack lives in LDS (DR_LDS_STACK entries per lane,
//     lane-interleaved so a wave's pushes hit 64 distinct banks) and spills
//     to a per-lane global area only past that depth.
// The triangle test restates PathTrace.metal:1801-1852 operation for
// operation (Moeller-Trumbore, |det| < 1e-7 reject, t > 0.0005) with no
// contraction, and the closest-hit update takes the lower triangle index on
// equal t, so the winner is the brute-force winner whatever the tree.
#pragma once

#include "dr_bvh.h"
#include "dr_math.h"
#include "dr_scene.h"

#ifndef DR_LDS_STACK
#define DR_LDS_STACK 24
#endif

struct DrStack {
    uint32_t *lds;   // &lds[threadIdx.x], stride blockDim.x
    uint32_t ldsStride;
    uint32_t *glob;  // &overflow[globalLane], stride gStride
    uint32_t gStride;

    __device__ __forceinline__ void put(uint32_t sp, uint32_t v) {
        if (sp < DR_LDS_STACK) lds[sp * ldsStride] = v;
        else glob[(sp - DR_LDS_STACK) * gStride] = v;
    }
    __device__ __forceinline__ uint32_t get(uint32_t sp) const {
        return sp < DR_LDS_STACK ? lds[sp * ldsStride] : glob[(sp - DR_LDS_STACK) * gStride];
    }
};

struct DrRayPre {
    vec3 o, d;
    vec3 inv;   // 1/d with zero components replaced by +-1e-20
    vec3 oinv;  // o * inv
};

__device__ __forceinline__ float safeInv(float d) {
    float a = fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d);
    return 1.0f / a;
}

__device__ __forceinline__ DrRayPre dr_prep(vec3 o, vec3 d) {
    DrRayPre r;
    r.o = o;
    r.d = d;
    r.inv = vec3{safeInv(d.x), safeInv(d.y), safeInv(d.z)};
    r.oinv = vec3{o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z};
    return r;
}

// Slab test of one box; returns entry distance or +inf on miss.
__device__ __forceinline__ float dr_box(const DrRayPre &r, float x0, float y0, float z0, float x1, float y1, float z1,
                                        float tmax) {
    float tx0 = __builtin_fmaf(x0, r.inv.x, -r.oinv.x);
    float tx1 = __builtin_fmaf(x1, r.inv.x, -r.oinv.x);
    float ty0 = __builtin_fmaf(y0, r.inv.y, -r.oinv.y);
    float ty1 = __builtin_fmaf(y1, r.inv.y, -r.oinv.y);
    float tz0 = __builtin_fmaf(z0, r.inv.z, -r.oinv.z);
    float tz1 = __builtin_fmaf(z1, r.inv.z, -r.oinv.z);
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    return tn <= tf ? tn : __builtin_inff();
}

// Moeller-Trumbore with precomputed edges (e1 = v1 - v0, e2 = v2 - v0 are
// computed on the host with the same float subtraction the reference does).
// The IEEE reciprocal is the expensive step, so the sign tests run first on
// the numerators: u = inv * un is negative exactly when un and det differ in
// sign and the product does not underflow to -0; only rejections that are
// certain (|un| > 2^-100 |det|) are taken early, everything else falls
// through to the reference arithmetic, so the accepted set and the values of
// u, v, t are unchanged.
__device__ __forceinline__ bool dr_tri(const DrRayPre &r, float4 A, float4 B, float4 C, float &t, float &u, float &v) {
    vec3 v0 = v3(A.x, A.y, A.z), e1 = v3(B.x, B.y, B.z), e2 = v3(C.x, C.y, C.z);
    vec3 p = cross(r.d, e2);
    float det = dot(e1, p);
    if (fabsf(det) < 1e-7f) return false;
    vec3 s = r.o - v0;
    float un = dot(s, p);
    const float tiny = 7.8886091e-31f * fabsf(det);  // 2^-100 |det|
    if ((un < 0.0f) != (det < 0.0f) && fabsf(un) > tiny) return false;
    vec3 q = cross(s, e1);
    float vn = dot(r.d, q);
    if ((vn < 0.0f) != (det < 0.0f) && fabsf(vn) > tiny) return false;
    float inv = 1.0f / det;
    float uu = inv * un;
    if (uu < 0.0f || uu > 1.0f) return false;
    float vv = inv * vn;
    if (vv < 0.0f || uu + vv > 1.0f) return false;
    float tt = inv * dot(e2, q);
    if (tt <= 0.0005f) return false;
    t = tt;
    u = uu;
    v = vv;
    return true;
}

struct DrHitRec {
    float t, u, v;
    uint32_t id;  // original triangle index, 0xffffffff = miss
};

// Closest hit (ANY = false) or any hit below tmax (ANY = true).
// On entry best.t is the search limit; best.id must be 0xffffffff for a
// closest-hit query and the result is the lowest-index triangle among those
// with minimal t < limit.
// Node fetches and triangle tests are added to nodeVisits / triTests.
template <bool ANY>
__device__ __forceinline__ bool dr_traverse(const DrSceneDev &S, const DrRayPre &r, DrHitRec &best, DrStack &stk,
                                            uint32_t &nodeVisits, uint32_t &triTests) {
    if (S.triangleCount == 0) return false;
    uint32_t sp = 0;
    uint32_t cur = 0;
    uint32_t nv = 0, nt = 0;
    bool found = false;
    while (true) {
        const DrNode n = S.nodes[cur];
        ++nv;
        float t0 = dr_box(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, best.t);
        float t1 = (n.d.y == DR_EMPTY_CHILD) ? __builtin_inff()
                                             : dr_box(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, best.t);
        bool h0 = t0 != __builtin_inff();
        bool h1 = t1 != __builtin_inff();
        // Leaf children: intersect now.
        for (int c = 0; c < 2; ++c) {
            bool hit = c == 0 ? h0 : h1;
            uint32_t cnt = c == 0 ? n.d.z : n.d.w;
            if (!hit || cnt == 0) continue;
            uint32_t first = c == 0 ? n.d.x : n.d.y;
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4 *tp = S.triPos + size_t(first + k) * 3;
                float4 A = tp[0], B = tp[1], C = tp[2];
                ++nt;
                float t, u, v;
                if (dr_tri(r, A, B, C, t, u, v)) {
                    uint32_t id = __float_as_uint(A.w);
                    if (t < best.t || (t == best.t && id < best.id)) {
                        best.t = t;
                        best.u = u;
                        best.v = v;
                        best.id = id;
                        found = true;
                        if (ANY) {
                            nodeVisits += nv;
                            triTests += nt;
                            return true;
                        }
                    }
                }
            }
        }
        bool g0 = h0 && n.d.z == 0;
        bool g1 = h1 && n.d.w == 0;
        if (g0 && g1) {
            uint32_t nearC = t0 <= t1 ? n.d.x : n.d.y;
            uint32_t farC = t0 <= t1 ? n.d.y : n.d.x;
            stk.put(sp++, farC);
            cur = nearC;
        } else if (g0) {
            cur = n.d.x;
        } else if (g1) {
            cur = n.d.y;
        } else {
            if (sp == 0) break;
            cur = stk.get(--sp);
        }
    }
    nodeVisits += nv;
    triTests += nt;
    return found;
}
