// dr_traverse.h -- software BVH traversal and triangle test for gfx950.
//
// CDNA4 has no ray-tracing units, so traversal is plain VALU work:
//   * 4-wide BVH: one node fetch (8 x dwordx4, 128 B) tests four child
//     boxes with a fused-multiply-add slab test against precomputed 1/d and
//     o/d, so a ray needs about half the dependent fetches of a binary tree;
//   * the nearest hit inner child is visited next, the others are pushed far
//     to near; hit leaves become pending (one per lane) or stack entries;
//   * two loop shapes (dr_trace_queue): while-while (inner nodes until every
//     lane of the wave holds a leaf, then the leaves converged) and stepwise
//     (one node or triangle per lane and iteration);
//   * the traversal stack lives in LDS (DR_LDS_STACK entries per lane,
//     lane-interleaved) and spills to a per-lane global area only past that
//     depth.
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

// The LDS part is addressed through an LDS (address space 3) pointer: with a
// generic pointer the compiler folds the LDS and the spill access into one
// flat instruction, which is slower than ds_read/ds_write and is counted
// with the vector-memory loads the node fetches wait on.
typedef __attribute__((address_space(3))) uint32_t dr_lds_u32;

__device__ __forceinline__ dr_lds_u32 *dr_lds_ptr(uint32_t *p) { return (dr_lds_u32 *)p; }

struct DrStack {
    dr_lds_u32 *lds;  // &lds[threadIdx.x], stride blockDim.x
    uint32_t ldsStride;
    uint32_t *glob;   // &overflow[globalLane], stride gStride
    uint32_t gStride;

    __device__ __forceinline__ void put(uint32_t sp, uint32_t v) {
        if (sp < DR_LDS_STACK) lds[sp * ldsStride] = v;
        else glob[(sp - DR_LDS_STACK) * gStride] = v;
    }
    __device__ __forceinline__ uint32_t get(uint32_t sp) const {
        // The LDS read is unconditional (clamped slot); only a spilled entry
        // takes the global load.
        uint32_t v = lds[min(sp, uint32_t(DR_LDS_STACK - 1)) * ldsStride];
        if (sp >= DR_LDS_STACK) v = glob[(sp - DR_LDS_STACK) * gStride];
        return v;
    }
};

struct DrRayPre {
    vec3 o, d;
    vec3 inv;  // 1/d with zero components replaced by +-1e-20
    // Slab offsets: a plane at x is crossed at fma(x, inv, c). cNear serves
    // the plane of each axis the ray enters by (the low plane for inv > 0,
    // the high one for inv < 0), cFar the one it leaves by: -o * inv moved by
    // e = 2^-21 |o * inv| toward the near side and the far side. The rounding
    // of o * inv and the part of the fma's and 1/d's relative error that
    // scales with |o| stay below e, so a ray far from the scene's origin
    // never culls a box the exact test keeps; the part that scales with the
    // box's own coordinates is covered by the builder's outward widening of
    // every child box.
    vec3 cNear, cFar;
    // Byte offset (0 or 16) of the near plane's float4 within each axis's
    // {lo, hi} pair of a DrNode4; the far plane's is the other one.
    uint32_t nearX, nearY, nearZ;
    // Box tests run against max(best t) * (1 + 2^-16) + slack, slack =
    // 2^-14 max|o| / max|d|: a triangle's computed hit distance carries the
    // rounding of o - v0 (about 2^-24 |o|, amplified by the triangle's shape
    // and the ray's incidence), so a box must not be culled by a competitor
    // whose computed t lies below the box's exact entry by that much. Keeping
    // a few more boxes near the current best changes no result.
    float slack;
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
    const vec3 oinv = vec3{o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z};
    const float k = 4.76837158e-7f;  // 2^-21
    // (|o * inv| is capped so that an infinite offset never meets its
    // opposite and makes a NaN.)
    const vec3 e = vec3{k * fminf(fabsf(oinv.x), 1e37f), k * fminf(fabsf(oinv.y), 1e37f),
                        k * fminf(fabsf(oinv.z), 1e37f)};
    r.cNear = vec3{-oinv.x - e.x, -oinv.y - e.y, -oinv.z - e.z};
    r.cFar = vec3{-oinv.x + e.x, -oinv.y + e.y, -oinv.z + e.z};
    r.nearX = (__float_as_uint(r.inv.x) >> 27) & 16u;  // sign of inv (never 0)
    r.nearY = (__float_as_uint(r.inv.y) >> 27) & 16u;
    r.nearZ = (__float_as_uint(r.inv.z) >> 27) & 16u;
    const float om = fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z));
    const float dm = fmaxf(fmaxf(fmaxf(fabsf(d.x), fabsf(d.y)), fabsf(d.z)), 1e-30f);
    r.slack = 6.10351562e-5f * om / dm;  // 2^-14
    return r;
}

static_assert(offsetof(DrNode4, lox) == 0 && offsetof(DrNode4, hix) == 16 && offsetof(DrNode4, loy) == 32 &&
                  offsetof(DrNode4, hiy) == 48 && offsetof(DrNode4, loz) == 64 && offsetof(DrNode4, hiz) == 80 &&
                  offsetof(DrNode4, child) == 96 && offsetof(DrNode4, meta) == 112,
              "node planes as the traversal fetches them");

// Slab test of one box given its planes in the order the ray crosses them
// per axis: true on a hit, with the entry distance in tn. For a valid box
// (lo <= hi) fma(lo, inv, c) and fma(hi, inv, c) are ordered by the sign of
// inv, so these are the min and max of each axis's pair: the same tn and
// exit distance as a test that orders each pair itself.
__device__ __forceinline__ bool dr_box(const DrRayPre &r, float nx, float ny, float nz, float fx, float fy, float fz,
                                       float tmax, float &tn) {
    const float txn = __builtin_fmaf(nx, r.inv.x, r.cNear.x);
    const float txf = __builtin_fmaf(fx, r.inv.x, r.cFar.x);
    const float tyn = __builtin_fmaf(ny, r.inv.y, r.cNear.y);
    const float tyf = __builtin_fmaf(fy, r.inv.y, r.cFar.y);
    const float tzn = __builtin_fmaf(nz, r.inv.z, r.cNear.z);
    const float tzf = __builtin_fmaf(fz, r.inv.z, r.cFar.z);
    tn = fmaxf(fmaxf(txn, tyn), fmaxf(tzn, 0.0f));
    const float tf = fminf(fminf(txf, tyf), fminf(tzf, tmax));
    return tn <= tf;
}

// Moeller-Trumbore from the exact vertices (e1 = v1 - v0, e2 = v2 - v0 with
// the same float subtractions as the reference). The IEEE reciprocal is the
// expensive step, so the sign tests run first on
// the numerators: u = inv * un is negative exactly when un and det differ in
// sign and the product does not underflow to -0; only rejections that are
// certain (|un| > 2^-100 |det|) are taken early, everything else falls
// through to the reference arithmetic, so the accepted set and the values of
// u, v, t are unchanged.
__device__ __forceinline__ bool dr_tri(const DrRayPre &r, float4 A, float4 B, float4 C, float &t, float &u, float &v) {
    vec3 v0 = v3(A.x, A.y, A.z), e1 = v3(B.x - A.x, B.y - A.y, B.z - A.z), e2 = v3(C.x - A.x, C.y - A.y, C.z - A.z);
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

__device__ __forceinline__ float dr_comp(vec3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// Watertight ray/triangle test (Woop, Benthin and Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013) with backface culling off:
// shear and scale so the ray is the +z axis (kz = the dominant direction
// axis, kx/ky swapped for a negative one to keep the winding), evaluate the
// three 2-D edge functions, re-evaluate them in double when one is exactly
// zero, and accept when they share a sign. No ray through a shared edge or
// vertex slips between the triangles. The reference's t > 0.0005 acceptance
// and u, v = the weights of v1, v2 are kept so hits shade as with the
// Moller-Trumbore test.
__device__ __forceinline__ bool dr_tri_wt(const DrRayPre &r, float4 A4, float4 B4, float4 C4, float &t, float &u,
                                          float &v) {
    const float ax = fabsf(r.d.x), ay = fabsf(r.d.y), az = fabsf(r.d.z);
    const int kz = ay > ax ? (az > ay ? 2 : 1) : (az > ax ? 2 : 0);
    int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    const float dz = dr_comp(r.d, kz);
    if (dz < 0.0f) {
        const int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    const float Sx = dr_comp(r.d, kx) / dz, Sy = dr_comp(r.d, ky) / dz, Sz = 1.0f / dz;
    const vec3 A = v3(A4.x - r.o.x, A4.y - r.o.y, A4.z - r.o.z);
    const vec3 B = v3(B4.x - r.o.x, B4.y - r.o.y, B4.z - r.o.z);
    const vec3 C = v3(C4.x - r.o.x, C4.y - r.o.y, C4.z - r.o.z);
    const float Az = dr_comp(A, kz), Bz = dr_comp(B, kz), Cz = dr_comp(C, kz);
    const float Ax = dr_comp(A, kx) - Sx * Az, Ay = dr_comp(A, ky) - Sy * Az;
    const float Bx = dr_comp(B, kx) - Sx * Bz, By = dr_comp(B, ky) - Sy * Bz;
    const float Cx = dr_comp(C, kx) - Sx * Cz, Cy = dr_comp(C, ky) - Sy * Cz;
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (Sz * Az) + V * (Sz * Bz) + W * (Sz * Cz);
    const float rcp = 1.0f / det;
    const float tt = T * rcp;
    if (!(tt > 0.0005f)) return false;
    t = tt;
    u = V * rcp;
    v = W * rcp;
    return true;
}

struct DrHitRec {
    float t, u, v;
    uint32_t id;  // original triangle index, 0xffffffff = miss
};

// Traversal state of one ray in the while-while form: inner nodes are
// visited until the lane holds a pending leaf (or finishes); leaves are
// tested when every lane of the wave holds one, so triangle tests run
// converged instead of serializing the wave once per lane.
// Stack entries: an inner node index, or DR_LEAF_REF | count << 27 | first
// for a hit leaf child still to be tested (count <= DR_LEAF_MAX, first <
// DR_MAX_TRIANGLES), so a popped leaf needs no memory access. A node visit
// pushes at most three inner and four leaf children, so a lane needs at
// most 7 entries per tree level plus two (the scene's stackMax).
#define DR_POP 0xffffffffu
#define DR_LEAF_REF 0x80000000u

struct DrTrav {
    uint32_t cur;        // inner node to visit next, or DR_POP
    uint32_t sp;
    uint32_t leafFirst;  // pending leaf: next triangle (leaf order)
    uint32_t leafCount;  // triangles left in the pending leaf
    bool done;
};

__device__ __forceinline__ void dr_trav_init(DrTrav &s) {
    s.cur = 0;
    s.sp = 0;
    s.leafFirst = 0;
    s.leafCount = 0;
    s.done = false;
}

__device__ __forceinline__ uint32_t dr_sel4(uint4 v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w)); }

// Orders (ka, pa), (kb, pb) by key; equal keys keep their order.
__device__ __forceinline__ void dr_cswap(float &ka, uint32_t &pa, float &kb, uint32_t &pb) {
    const bool sw = kb < ka;
    const float k = sw ? kb : ka;
    kb = sw ? ka : kb;
    ka = k;
    const uint32_t q = sw ? pb : pa;
    pb = sw ? pa : pb;
    pa = q;
}

// One inner-node step; requires !s.done. With a leaf pending the step is
// speculative: it visits the known next node, or pops an inner node from
// the top of the stack, and the hit leaves it finds go to the stack.
__device__ __forceinline__ void dr_node_step(const DrSceneDev &S, const DrRayPre &r, float tmax, DrTrav &s,
                                             DrStack &stk, uint32_t &nv) {
    if (s.cur == DR_POP) {
        // With a leaf pending (speculative step) only an inner node on top
        // of the stack is taken; a leaf entry waits.
        if (s.sp == 0u) {
            if (s.leafCount == 0u) s.done = true;
            return;
        }
        const uint32_t e = stk.get(s.sp - 1u);
        const bool leaf = (e & DR_LEAF_REF) != 0u;
        if (leaf && s.leafCount != 0u) return;
        --s.sp;
        if (leaf) {
            s.leafFirst = e & (DR_MAX_TRIANGLES - 1u);
            s.leafCount = (e >> 27) & 15u;
            return;
        }
        s.cur = e;
    }
    const uint32_t ni = s.cur;
    // The node's planes are fetched by a 32-bit byte offset from the node
    // array (ni < 2^25): each axis's near plane at its lo or hi slot by the
    // ray's sign, the far plane at the other (offset ^ 16).
    const char *nb = reinterpret_cast<const char *>(S.nodes);
    const uint32_t base = ni << 7;
    const uint32_t ox = base | r.nearX, oy = base | r.nearY, oz = base | r.nearZ;
    const float4 nX = *reinterpret_cast<const float4 *>(nb + ox);
    const float4 fX = *reinterpret_cast<const float4 *>(nb + (ox ^ 16u));
    const float4 nY = *reinterpret_cast<const float4 *>(nb + 32 + oy);
    const float4 fY = *reinterpret_cast<const float4 *>(nb + 32 + (oy ^ 16u));
    const float4 nZ = *reinterpret_cast<const float4 *>(nb + 64 + oz);
    const float4 fZ = *reinterpret_cast<const float4 *>(nb + 64 + (oz ^ 16u));
    struct {
        uint4 child;
        uint2 meta;  // leaf triangle counts, child masks (the rest is unused)
    } n;
    n.child = *reinterpret_cast<const uint4 *>(nb + 96 + base);
    n.meta = *reinterpret_cast<const uint2 *>(nb + 112 + base);
    ++nv;
    const float inf = __builtin_inff();
    tmax = __builtin_fmaf(tmax, 1.0000153f, r.slack);  // 1 + 2^-16
    float t0, t1, t2, t3;
    const bool h0 = dr_box(r, nX.x, nY.x, nZ.x, fX.x, fY.x, fZ.x, tmax, t0);
    const bool h1 = dr_box(r, nX.y, nY.y, nZ.y, fX.y, fY.y, fZ.y, tmax, t1);
    const bool h2 = dr_box(r, nX.z, nY.z, nZ.z, fX.z, fY.z, fZ.z, tmax, t2);
    const bool h3 = dr_box(r, nX.w, nY.w, nZ.w, fX.w, fY.w, fZ.w, tmax, t3);
    const uint32_t hit = ((h0 ? 1u : 0u) | (h1 ? 2u : 0u) | (h2 ? 4u : 0u) | (h3 ? 8u : 0u)) & n.meta.y;
    const uint32_t leafBits = (n.meta.y >> 4) & 15u;
    uint32_t leafMask = hit & leafBits;
    const uint32_t innerMask = hit & ~leafBits;
    if (leafMask) {
        const uint32_t cnt = n.meta.x;
        if (s.leafCount == 0u) {
            // The nearest hit leaf (ties to the lower child) becomes pending.
            float lk = inf;
            uint32_t lc = 0u;
            if ((leafMask & 1u) && t0 < lk) lk = t0, lc = 0u;
            if ((leafMask & 2u) && t1 < lk) lk = t1, lc = 1u;
            if ((leafMask & 4u) && t2 < lk) lk = t2, lc = 2u;
            if ((leafMask & 8u) && t3 < lk) lk = t3, lc = 3u;
            s.leafFirst = dr_sel4(n.child, int(lc));
            s.leafCount = (cnt >> (8u * lc)) & 0xffu;
            leafMask &= ~(1u << lc);
        }
        // The other hit leaves are pushed in child order.
        const uint32_t l0 = DR_LEAF_REF | (cnt & 0xffu) << 27 | n.child.x;
        const uint32_t l1 = DR_LEAF_REF | (cnt >> 8 & 0xffu) << 27 | n.child.y;
        const uint32_t l2 = DR_LEAF_REF | (cnt >> 16 & 0xffu) << 27 | n.child.z;
        const uint32_t l3 = DR_LEAF_REF | (cnt >> 24) << 27 | n.child.w;
        if (s.sp + 4u <= DR_LDS_STACK) {
            // Compaction by unconditional writes in child order: entry c
            // goes to sp + (hit leaves before c); a child that is not pushed
            // writes the slot the next pushed one then overwrites, or a
            // slot past the new top.
            const uint32_t b1 = leafMask & 1u, b2 = b1 + ((leafMask >> 1) & 1u), b3 = b2 + ((leafMask >> 2) & 1u);
            stk.lds[s.sp * stk.ldsStride] = l0;
            stk.lds[(s.sp + b1) * stk.ldsStride] = l1;
            stk.lds[(s.sp + b2) * stk.ldsStride] = l2;
            stk.lds[(s.sp + b3) * stk.ldsStride] = l3;
            s.sp += __builtin_popcount(leafMask);
        } else {
            if (leafMask & 1u) stk.put(s.sp++, l0);
            if (leafMask & 2u) stk.put(s.sp++, l1);
            if (leafMask & 4u) stk.put(s.sp++, l2);
            if (leafMask & 8u) stk.put(s.sp++, l3);
        }
    }
    // Inner children in order of entry distance (a 5-comparator network over
    // (distance, reference), misses keyed +inf): the nearest is next, the
    // others are pushed far to near so that nearer ones are popped first.
    float k0 = (innerMask & 1u) ? t0 : inf, k1 = (innerMask & 2u) ? t1 : inf;
    float k2 = (innerMask & 4u) ? t2 : inf, k3 = (innerMask & 8u) ? t3 : inf;
    uint32_t p0 = n.child.x, p1 = n.child.y, p2 = n.child.z, p3 = n.child.w;
    dr_cswap(k0, p0, k1, p1);
    dr_cswap(k2, p2, k3, p3);
    dr_cswap(k0, p0, k2, p2);
    dr_cswap(k1, p1, k3, p3);
    dr_cswap(k1, p1, k2, p2);
    const uint32_t nIn = __builtin_popcount(innerMask);
    s.cur = nIn ? p0 : DR_POP;
    if (nIn > 1u) {
        const uint32_t e0 = nIn == 4u ? p3 : (nIn == 3u ? p2 : p1);
        const uint32_t e1 = nIn == 4u ? p2 : p1;
        if (s.sp + 3u <= DR_LDS_STACK) {
            // All three slots in LDS: write them unconditionally (slots
            // past the new top are scratch) and advance by nIn - 1.
            stk.lds[s.sp * stk.ldsStride] = e0;
            stk.lds[(s.sp + 1u) * stk.ldsStride] = e1;
            stk.lds[(s.sp + 2u) * stk.ldsStride] = p1;
            s.sp += nIn - 1u;
        } else {
            stk.put(s.sp++, e0);
            if (nIn >= 3u) stk.put(s.sp++, e1);
            if (nIn == 4u) stk.put(s.sp++, p1);
        }
    }
}

// One triangle of the pending leaf; requires s.leafCount > 0. Returns true
// when the test hits and improves the record.
__device__ __forceinline__ bool dr_leaf_step(const DrSceneDev &S, const DrRayPre &r, DrTrav &s, DrHitRec &best,
                                             uint32_t &nt) {
    const float4 *tp = S.triPos + size_t(s.leafFirst) * 3;
    float4 A = tp[0], B = tp[1], C = tp[2];
    ++nt;
    ++s.leafFirst;
    --s.leafCount;
    float t, u, v;
    if (S.watertight ? dr_tri_wt(r, A, B, C, t, u, v) : dr_tri(r, A, B, C, t, u, v)) {
        uint32_t id = __float_as_uint(A.w);
        if (t < best.t || (t == best.t && id < best.id)) {
            best.t = t;
            best.u = u;
            best.v = v;
            best.id = id;
            return true;
        }
    }
    return false;
}

// Closest hit (ANY = false) or any hit below tmax (ANY = true) of one ray
// per lane. On entry best.t is the search limit; best.id must be 0xffffffff
// for a closest-hit query and the result is the lowest-index triangle among
// those with minimal t < limit (boxes are tested with t_near <= limit, so
// equal-t candidates stay reachable in any visiting order).
// Node fetches and triangle tests are added to nodeVisits / triTests.
// Every active lane of the wave must call it (the phases use ballots).
template <bool ANY>
__device__ __forceinline__ bool dr_traverse(const DrSceneDev &S, const DrRayPre &r, DrHitRec &best, DrStack &stk,
                                            uint32_t &nodeVisits, uint32_t &triTests) {
    DrTrav s;
    dr_trav_init(s);
    if (S.triangleCount == 0) s.done = true;
    uint32_t nv = 0, nt = 0;
    bool found = false;
    while (true) {
        // Node phase: until every lane holds a leaf; lanes that already hold
        // one keep descending (speculatively) while they know a next node or
        // an inner node is on top of their stack.
        while (__any(!s.done && s.leafCount == 0)) {
            if (!s.done && (s.leafCount == 0 || s.cur != DR_POP || s.sp != 0u))
                dr_node_step(S, r, best.t, s, stk, nv);
        }
        if (!__any(s.leafCount != 0u)) break;
        while (__any(s.leafCount != 0u)) {
            if (s.leafCount != 0u && dr_leaf_step(S, r, s, best, nt)) {
                found = true;
                if (ANY) {
                    s.done = true;
                    s.leafCount = 0;
                }
            }
        }
    }
    nodeVisits += nv;
    triTests += nt;
    return found;
}

// A wave's view of a queue's striped claim cursors (DR_CLAIM_STRIPES
// cursors DR_CLAIM_STRIDE words apart, zero at the start; cursor j deals
// the queue's chunks j, j + stripes, ... of DR_CLAIM_CHUNK entries). The
// wave starts on cursor (wave index mod stripes), moves to the next one when
// a claim reaches a cursor's end, and is done once it has seen every
// cursor's end; each entry is handed out exactly once.
struct DrClaim {
    uint32_t *cursors;
    uint32_t fullRounds, rest, stripe, end, left;
    __device__ __forceinline__ uint32_t stripeCount(uint32_t j) const {
        const uint32_t lo = j * DR_CLAIM_CHUNK;
        return fullRounds * DR_CLAIM_CHUNK + (rest > lo ? min(rest - lo, DR_CLAIM_CHUNK) : 0u);
    }
    __device__ __forceinline__ DrClaim(uint32_t *c, uint32_t count) : cursors(c) {
        const uint32_t round = DR_CLAIM_STRIPES * DR_CLAIM_CHUNK;
        fullRounds = count / round;
        rest = count - fullRounds * round;
        stripe = (blockIdx.x * ((blockDim.x + 63u) >> 6) + (threadIdx.x >> 6)) % DR_CLAIM_STRIPES;
        end = stripeCount(stripe);
        left = DR_CLAIM_STRIPES;
    }
    __device__ __forceinline__ bool more() const { return left != 0u; }
    // Claims take entries of the current cursor (one atomic, by lane 0; all
    // lanes of the wave call it): the wave's entries are cursor positions
    // base + [0, take) below end on cursor j.
    __device__ __forceinline__ void claim(uint32_t take, uint32_t &base, uint32_t &j, uint32_t &e) {
        uint32_t b = 0u;
        if (__lane_id() == 0u) b = atomicAdd(cursors + stripe * DR_CLAIM_STRIDE, take);
        base = __shfl(b, 0);
        j = stripe;
        e = end;
        if (base + take >= end) {
            --left;
            stripe = stripe + 1u == DR_CLAIM_STRIPES ? 0u : stripe + 1u;
            end = stripeCount(stripe);
        }
    }
    // Queue entry of cursor j's position c.
    __device__ static __forceinline__ uint32_t entry(uint32_t c, uint32_t j) {
        return ((c / DR_CLAIM_CHUNK) * DR_CLAIM_STRIPES + j) * DR_CLAIM_CHUNK + c % DR_CLAIM_CHUNK;
    }
};

// Persistent tracing of a queue with per-lane refill: a lane whose work
// item is finished takes the next queue entry as soon as at least
// DR_REFILL_MIN lanes of the wave are idle (one atomic per refill), so a
// wave is not held by its longest item. claim points to the queue's striped
// claim cursors (DrClaim). fetch(qi, r, best) sets up the
// first query of queue entry qi; retire(res, found, r, best) consumes a
// query's result and either sets up the item's next query in r / best and
// returns true (a shadow chain step, a subsurface-walk segment), or
// returns false when the item is done. All 64 lanes of the wave must enter;
// the loop exits when the queue is drained and every lane is idle.
// laneAny, when given, points to a per-lane flag that fetch / retire set:
// the lane's current query is then an any-hit query (ends at its first hit)
// even where ANY is false, so closest-hit and any-hit items can share a
// queue.
// Refill thresholds: idle lanes a wave waits for before it claims, for the
// stepwise loop (chained queries) and the while-while loop (extension rays).
#ifndef DR_REFILL_MIN
#define DR_REFILL_MIN 12
#endif
#ifndef DR_REFILL_MIN_WW
#define DR_REFILL_MIN_WW 32
#endif
#ifndef DR_RETIRE_MIN
#define DR_RETIRE_MIN 24
#endif
// STEPWISE selects the loop shape: false = while-while (node phase until
// every lane holds a leaf, then the leaf phase; best for the extension
// queue, whose items are single queries), true = one node or triangle step
// per lane and iteration with batched retires (best for queues of chained
// queries -- shadow chains, subsurface walks -- whose lanes desynchronise).
template <bool ANY, bool STEPWISE = false, class Fetch, class Retire>
__device__ __forceinline__ void dr_trace_queue(const DrSceneDev &S, uint32_t count, uint32_t *claim, DrStack &stk,
                                               uint32_t &nv, uint32_t &nt, Fetch fetch, Retire retire,
                                               const bool *laneAny = nullptr) {
    const uint32_t lane = __lane_id();
    // A wave claims at most its fair share of a short queue, so that short
    // queues (subsurface walks, late bounces) spread over every CU instead
    // of filling a few waves whose chains of dependent fetches then set the
    // kernel's duration.
    const uint32_t waves = gridDim.x * ((blockDim.x + 63u) >> 6);
    const uint32_t share = max(4u, (count + waves - 1u) / waves);
    DrClaim cl(claim, count);
    bool active = false, found = false, more = true;
    DrRayPre r;
    DrHitRec best;
    DrTrav s;
    dr_trav_init(s);
    s.done = true;
    while (true) {
        if (more) {
            uint64_t idle = __ballot(!active);
            uint32_t nIdle = __popcll(idle);
            if (nIdle >= min(STEPWISE ? uint32_t(DR_REFILL_MIN) : uint32_t(DR_REFILL_MIN_WW), share)) {
                const uint32_t take = min(nIdle, share);
                uint32_t base, j, end;
                cl.claim(take, base, j, end);
                more = cl.more();
                if (!active) {
                    const uint32_t rank = __popcll(idle & ((1ull << lane) - 1ull));
                    const uint32_t c = base + rank;
                    const uint32_t qi = DrClaim::entry(c, j);
                    if (rank < take && c < end) {
                        fetch(qi, r, best);
                        dr_trav_init(s);
                        if (S.triangleCount == 0) s.done = true;
                        found = false;
                        active = true;
                    }
                }
            }
        }
        if (!__any(active)) {
            if (!more) break;
            continue;
        }
        if constexpr (STEPWISE) {
            // One step per lane and iteration: a triangle of the pending leaf,
            // else an inner node; finished lanes retire together once
            // DR_RETIRE_MIN of them are waiting (or nothing else is tracing).
            const bool tracing = active && !(s.done && s.leafCount == 0u);
            if (tracing) {
                if (s.leafCount != 0u) {
                    if (dr_leaf_step(S, r, s, best, nt)) {
                        found = true;
                        if (ANY || (laneAny && *laneAny)) {
                            s.done = true;
                            s.leafCount = 0;
                        }
                    }
                } else {
                    dr_node_step(S, r, best.t, s, stk, nv);
                }
            }
            const bool waiting = active && s.done && s.leafCount == 0u;
            const uint64_t wmask = __ballot(waiting);
            if (!wmask || (__popcll(wmask) < DR_RETIRE_MIN && __any(active && !(s.done && s.leafCount == 0u)))) continue;
        } else {
            while (__any(active && !s.done && s.leafCount == 0u)) {
                if (active && !s.done && (s.leafCount == 0u || s.cur != DR_POP || s.sp != 0u))
                    dr_node_step(S, r, best.t, s, stk, nv);
            }
            while (__any(active && s.leafCount != 0u)) {
                if (active && s.leafCount != 0u && dr_leaf_step(S, r, s, best, nt)) {
                    found = true;
                    if (ANY || (laneAny && *laneAny)) {
                        s.done = true;
                        s.leafCount = 0;
                    }
                }
            }
        }
        if (active && s.done) {
            const DrHitRec res = best;
            if (retire(res, found, r, best)) {
                dr_trav_init(s);
                if (S.triangleCount == 0) s.done = true;
                found = false;
            } else {
                active = false;
            }
        }
    }
}
