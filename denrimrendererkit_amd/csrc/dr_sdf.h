// dr_sdf.h -- sparse signed-distance brick tracing on gfx950.
//
// Follows the reference's sparse-volume path step for step:
//   closestSparseVolumeBrickDDAHit  PathTrace.metal:2881-3088 (3-D DDA over
//                                   the brick grid with 4^3 macro-cell skips)
//   intersectVolumeBrickPrepared    :2608-2807 (back-track, sphere trace <= 192
//                                   steps, 8 bisections, cubic polish)
//   trilinear / Catmull-Rom gradients and volumeBrickNormal :2096-2469
// Storage is this backend's: distances as FP16 when every input distance is
// exactly representable in FP16 (widened to FP32 on load, so tracing is FP32),
// else as FP32; materials as one uint2 per sample (A | B << 16, bits of the
// blend weight), read once per hit; a 64-B brick header. The values seen here
// equal the reference-layout FP32 values, and the arithmetic (only + - * /
// floor round sqrt) is IEEE, so hits are bit-exact with an FP32 evaluation of
// the reference over the same field.
#pragma once

#include "dr_math.h"
#include "dr_scene.h"

// Per-lane sparse-volume counters, added to DR_CNT_SDF_* at the end of a kernel.
struct DrSdfCounters {
    uint32_t denseTests, denseSteps, gridCells, macroSkips, brickTests, invalid, rangeCulls, marches, marchSteps,
        hits;
};

struct DrVolHit {
    float t;
    vec3 normal;
    uint2 material;        // sparse hits: A | B << 16, bits(blend)
    uint32_t denseSample;  // dense hits: index of the nearest sample (0xffffffff: none)
    uint32_t volume;
    bool frontFacing;
    bool dense;
};

__device__ __forceinline__ float dr_half(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

__device__ __forceinline__ vec3 mulCols(const float4 *c, vec3 v, float w) {
    return vec3{c[0].x * v.x + c[1].x * v.y + c[2].x * v.z + c[3].x * w,
                c[0].y * v.x + c[1].y * v.y + c[2].y * v.z + c[3].y * w,
                c[0].z * v.x + c[1].z * v.y + c[2].z * v.z + c[3].z * w};
}

__device__ __forceinline__ bool sdfAABB(vec3 o, vec3 d, vec3 bmin, vec3 bmax, float closestT, float &tNear,
                                        float &tFar) {  // intersectAABB :1865-1881
    vec3 inv = 1.0f / d;
    vec3 t0 = (bmin - o) * inv;
    vec3 t1 = (bmax - o) * inv;
    vec3 tn = vmin(t0, t1), tf = vmax(t0, t1);
    tNear = fmaxf(fmaxf(tn.x, tn.y), fmaxf(tn.z, 0.0005f));
    tFar = fminf(fminf(tf.x, tf.y), fminf(tf.z, closestT));
    return tNear <= tFar;
}

struct DrBrickCtx {
    uint32_t dim[3];
    uint32_t offset, strideY, strideZ;
    vec3 smin, smax, extent;
    bool valid;
};

__device__ __forceinline__ DrBrickCtx brickCtx(const DrSdfDev &V, const DrBrick &b) {  // :2074-2094
    DrBrickCtx c;
    uint32_t packed = __float_as_uint(b.cmin.w);
    c.dim[0] = packed & 0xffu;
    c.dim[1] = (packed >> 8) & 0xffu;
    c.dim[2] = (packed >> 16) & 0xffu;
    c.offset = __float_as_uint(b.cmax.w);
    c.strideY = c.dim[0];
    c.strideZ = c.dim[0] * c.dim[1];
    uint32_t n = c.dim[0] * c.dim[1] * c.dim[2];
    c.smin = v3(b.smin.x, b.smin.y, b.smin.z);
    c.smax = v3(b.smax.x, b.smax.y, b.smax.z);
    c.extent = vmax(c.smax - c.smin, v3s(1e-6f));
    bool range = c.offset < V.sampleCount && n <= V.sampleCount - c.offset;
    c.valid = c.dim[0] >= 2u && c.dim[1] >= 2u && c.dim[2] >= 2u && range;
    return c;
}

template <bool HALF>
__device__ __forceinline__ float sdAt(const DrSdfDev &V, const DrBrickCtx &c, uint32_t x, uint32_t y, uint32_t z) {
    uint32_t i = c.offset + x + y * c.strideY + z * c.strideZ;
    return HALF ? dr_half(V.dist16[i]) : V.dist32[i];
}

__device__ __forceinline__ void brickGrid(const DrBrickCtx &c, vec3 p, uint32_t b[3], uint32_t n[3], vec3 &f) {
    vec3 uvw = vclamp((p - c.smin) / c.extent, v3s(0.0f), v3s(1.0f));
    float g[3] = {uvw.x * (float)(c.dim[0] - 1u), uvw.y * (float)(c.dim[1] - 1u), uvw.z * (float)(c.dim[2] - 1u)};
    float fr[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b[a] = (uint32_t)floorf(g[a]);
        n[a] = min(b[a] + 1u, c.dim[a] - 1u);
        fr[a] = g[a] - (float)b[a];
    }
    f = v3(fr[0], fr[1], fr[2]);
}

// The 8 FP16 distances of the lookup cell b (n = b + 1 clamped) in one
// 16-B load: d[i] = (x: i & 1 ? n : b, y: i & 2 ? n : b, z: i & 4 ? n : b).
__device__ __forceinline__ void cornerLoad(const DrSdfDev &V, const DrBrickCtx &c, const uint32_t b[3], float d[8]) {
    const uint4 q = V.corner16[c.offset + b[0] + b[1] * c.strideY + b[2] * c.strideZ];
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        d[2 * k] = dr_half((uint16_t)(w[k] & 0xffffu));
        d[2 * k + 1] = dr_half((uint16_t)(w[k] >> 16));
    }
}

template <bool HALF>
__device__ __forceinline__ float brickDistance(const DrSdfDev &V, const DrBrickCtx &c, vec3 p) {  // :2096-2127
    if (!c.valid) return __builtin_inff();
    uint32_t b[3], n[3];
    vec3 f;
    brickGrid(c, p, b, n, f);
    if constexpr (HALF) {
        float d[8];
        cornerLoad(V, c, b, d);
        float c00 = mixf(d[0], d[1], f.x);
        float c10 = mixf(d[2], d[3], f.x);
        float c01 = mixf(d[4], d[5], f.x);
        float c11 = mixf(d[6], d[7], f.x);
        float c0 = mixf(c00, c10, f.y);
        float c1 = mixf(c01, c11, f.y);
        return mixf(c0, c1, f.z);
    }
    float c00 = mixf(sdAt<HALF>(V, c, b[0], b[1], b[2]), sdAt<HALF>(V, c, n[0], b[1], b[2]), f.x);
    float c10 = mixf(sdAt<HALF>(V, c, b[0], n[1], b[2]), sdAt<HALF>(V, c, n[0], n[1], b[2]), f.x);
    float c01 = mixf(sdAt<HALF>(V, c, b[0], b[1], n[2]), sdAt<HALF>(V, c, n[0], b[1], n[2]), f.x);
    float c11 = mixf(sdAt<HALF>(V, c, b[0], n[1], n[2]), sdAt<HALF>(V, c, n[0], n[1], n[2]), f.x);
    float c0 = mixf(c00, c10, f.y);
    float c1 = mixf(c01, c11, f.y);
    return mixf(c0, c1, f.z);
}

// The last corner record a lane loaded, keyed by its index (a bisection's
// midpoints close in on one cell, so most of them reuse it).
struct DrCornerCache {
    uint32_t idx;
    uint4 rec;
};

// brickDistance<true> through a one-record cache (FP16 corner records).
__device__ __forceinline__ float brickDistanceCached(const DrSdfDev &V, const DrBrickCtx &c, vec3 p,
                                                     DrCornerCache &cache) {
    if (!c.valid) return __builtin_inff();
    uint32_t b[3], n[3];
    vec3 f;
    brickGrid(c, p, b, n, f);
    const uint32_t idx = c.offset + b[0] + b[1] * c.strideY + b[2] * c.strideZ;
    if (idx != cache.idx) {
        cache.rec = V.corner16[idx];
        cache.idx = idx;
    }
    const uint32_t w[4] = {cache.rec.x, cache.rec.y, cache.rec.z, cache.rec.w};
    float d[8];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        d[2 * k] = dr_half((uint16_t)(w[k] & 0xffffu));
        d[2 * k + 1] = dr_half((uint16_t)(w[k] >> 16));
    }
    float c00 = mixf(d[0], d[1], f.x);
    float c10 = mixf(d[2], d[3], f.x);
    float c01 = mixf(d[4], d[5], f.x);
    float c11 = mixf(d[6], d[7], f.x);
    float c0 = mixf(c00, c10, f.y);
    float c1 = mixf(c01, c11, f.y);
    return mixf(c0, c1, f.z);
}

struct DrDG {
    float distance;
    vec3 gradient;
};

template <bool HALF>
__device__ __forceinline__ DrDG brickDistanceGradient(const DrSdfDev &V, const DrBrickCtx &c, vec3 p) {  // :2161-2219
    DrDG r{__builtin_inff(), v3s(0.0f)};
    if (!c.valid) return r;
    uint32_t b[3], n[3];
    vec3 f;
    brickGrid(c, p, b, n, f);
    float d000, d100, d010, d110, d001, d101, d011, d111;
    if constexpr (HALF) {
        float d[8];
        cornerLoad(V, c, b, d);
        d000 = d[0], d100 = d[1], d010 = d[2], d110 = d[3], d001 = d[4], d101 = d[5], d011 = d[6], d111 = d[7];
    } else {
        d000 = sdAt<HALF>(V, c, b[0], b[1], b[2]), d100 = sdAt<HALF>(V, c, n[0], b[1], b[2]);
        d010 = sdAt<HALF>(V, c, b[0], n[1], b[2]), d110 = sdAt<HALF>(V, c, n[0], n[1], b[2]);
        d001 = sdAt<HALF>(V, c, b[0], b[1], n[2]), d101 = sdAt<HALF>(V, c, n[0], b[1], n[2]);
        d011 = sdAt<HALF>(V, c, b[0], n[1], n[2]), d111 = sdAt<HALF>(V, c, n[0], n[1], n[2]);
    }
    float c00 = mixf(d000, d100, f.x), c10 = mixf(d010, d110, f.x);
    float c01 = mixf(d001, d101, f.x), c11 = mixf(d011, d111, f.x);
    float c0 = mixf(c00, c10, f.y), c1 = mixf(c01, c11, f.y);
    r.distance = mixf(c0, c1, f.z);
    float dx = mixf(mixf(d100 - d000, d110 - d010, f.y), mixf(d101 - d001, d111 - d011, f.y), f.z);
    float dy = mixf(mixf(d010 - d000, d110 - d100, f.x), mixf(d011 - d001, d111 - d101, f.x), f.z);
    float dz = c1 - c0;
    vec3 g2l = v3((float)(c.dim[0] - 1u), (float)(c.dim[1] - 1u), (float)(c.dim[2] - 1u)) / c.extent;
    r.gradient = v3(dx, dy, dz) * g2l;
    return r;
}

__device__ __forceinline__ float catmullRom(float p0, float p1, float p2, float p3, float t) {  // :2141-2150
    float t2 = t * t;
    float t3 = t2 * t;
    return 0.5f * (2.0f * p1 + (-p0 + p2) * t + (2.0f * p0 - 5.0f * p1 + 4.0f * p2 - p3) * t2 +
                   (-p0 + 3.0f * p1 - 3.0f * p2 + p3) * t3);
}
__device__ __forceinline__ float catmullRomD(float p0, float p1, float p2, float p3, float t) {  // :2152-2159
    float t2 = t * t;
    return 0.5f * ((-p0 + p2) + 2.0f * (2.0f * p0 - 5.0f * p1 + 4.0f * p2 - p3) * t +
                   3.0f * (-p0 + 3.0f * p1 - 3.0f * p2 + p3) * t2);
}

template <bool HALF>
__device__ __forceinline__ DrDG brickDistanceGradientCubic(const DrSdfDev &V, const DrBrickCtx &c, vec3 p) {  // :2257-2314
    DrDG inv{__builtin_inff(), v3s(0.0f)};
    if (!c.valid) return inv;
    vec3 uvw = vclamp((p - c.smin) / c.extent, v3s(0.0f), v3s(1.0f));
    float g[3] = {uvw.x * (float)(c.dim[0] - 1u), uvw.y * (float)(c.dim[1] - 1u), uvw.z * (float)(c.dim[2] - 1u)};
    uint32_t b[3] = {(uint32_t)floorf(g[0]), (uint32_t)floorf(g[1]), (uint32_t)floorf(g[2])};
    if (b[0] < 1u || b[1] < 1u || b[2] < 1u || b[0] + 2u >= c.dim[0] || b[1] + 2u >= c.dim[1] || b[2] + 2u >= c.dim[2])
        return brickDistanceGradient<HALF>(V, c, p);
    vec3 f = v3(g[0] - (float)b[0], g[1] - (float)b[1], g[2] - (float)b[2]);
    float yr[4], dyr[4], dxyr[4];
    // FP16: the 4^3 neighbourhood is the 2^3 corner records of cells
    // b - 1 + {0, 2} per axis (no clamping inside this range), 8 loads at once.
    uint32_t q[8][4];
    if constexpr (HALF) {
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) {
            const uint4 r = V.corner16[c.offset + (b[0] - 1u + 2u * (k & 1u)) + (b[1] - 1u + (k & 2u)) * c.strideY +
                                       (b[2] - 1u + ((k >> 1) & 2u)) * c.strideZ];
            q[k][0] = r.x, q[k][1] = r.y, q[k][2] = r.z, q[k][3] = r.w;
        }
    }
    // Sample (x, y, z) of the neighbourhood, x, y, z in 0..3.
    auto nb = [&](uint32_t x, uint32_t y, uint32_t z) -> float {
        if constexpr (HALF) {
            const uint32_t k = (x >> 1) | (y & 2u) | ((z & 2u) << 1);
            const uint32_t w = q[k][((y & 1u) | ((z & 1u) << 1))];
            return dr_half((uint16_t)((x & 1u) ? (w >> 16) : (w & 0xffffu)));
        } else {
            return sdAt<HALF>(V, c, b[0] - 1u + x, b[1] - 1u + y, b[2] - 1u + z);
        }
    };
#pragma unroll
    for (uint32_t z = 0; z < 4; ++z) {
        float xr[4], dxr[4];
#pragma unroll
        for (uint32_t y = 0; y < 4; ++y) {
            float p0 = nb(0u, y, z), p1 = nb(1u, y, z);
            float p2 = nb(2u, y, z), p3 = nb(3u, y, z);
            xr[y] = catmullRom(p0, p1, p2, p3, f.x);
            dxr[y] = catmullRomD(p0, p1, p2, p3, f.x);
        }
        yr[z] = catmullRom(xr[0], xr[1], xr[2], xr[3], f.y);
        dyr[z] = catmullRomD(xr[0], xr[1], xr[2], xr[3], f.y);
        dxyr[z] = catmullRom(dxr[0], dxr[1], dxr[2], dxr[3], f.y);
    }
    DrDG r;
    r.distance = catmullRom(yr[0], yr[1], yr[2], yr[3], f.z);
    float dx = catmullRom(dxyr[0], dxyr[1], dxyr[2], dxyr[3], f.z);
    float dy = catmullRom(dyr[0], dyr[1], dyr[2], dyr[3], f.z);
    float dz = catmullRomD(yr[0], yr[1], yr[2], yr[3], f.z);
    vec3 g2l = v3((float)(c.dim[0] - 1u), (float)(c.dim[1] - 1u), (float)(c.dim[2] - 1u)) / c.extent;
    r.gradient = v3(dx, dy, dz) * g2l;
    return r;
}

template <bool HALF>
__device__ __forceinline__ vec3 brickNormal(const DrSdfDev &V, const DrVolume &vol, const DrBrickCtx &c, vec3 p,
                                         bool cubic, bool coarse) {  // :2420-2469
    if (!c.valid) return v3(0, 1, 0);
    vec3 ln = cubic ? brickDistanceGradientCubic<HALF>(V, c, p).gradient : brickDistanceGradient<HALF>(V, c, p).gradient;
    if (coarse) {
        vec3 ext = vmax(v3f(vol.localBoundsMax) - v3f(vol.localBoundsMin), v3s(1e-6f));
        vec3 dims = v3((float)vol.dimensions.x, (float)vol.dimensions.y, (float)vol.dimensions.z);
        vec3 sp = (ext / vmax(dims - v3s(1.0f), v3s(1.0f))) * 1.5f;
        vec3 mn = c.smin, mx = c.smax;
        vec3 x0 = p, x1 = p;
        x0.x = fmaxf(p.x - sp.x, mn.x);
        x1.x = fminf(p.x + sp.x, mx.x);
        float dx = (brickDistance<HALF>(V, c, x1) - brickDistance<HALF>(V, c, x0)) / fmaxf(x1.x - x0.x, 1e-6f);
        vec3 y0 = p, y1 = p;
        y0.y = fmaxf(p.y - sp.y, mn.y);
        y1.y = fminf(p.y + sp.y, mx.y);
        float dy = (brickDistance<HALF>(V, c, y1) - brickDistance<HALF>(V, c, y0)) / fmaxf(y1.y - y0.y, 1e-6f);
        vec3 z0 = p, z1 = p;
        z0.z = fmaxf(p.z - sp.z, mn.z);
        z1.z = fminf(p.z + sp.z, mx.z);
        float dz = (brickDistance<HALF>(V, c, z1) - brickDistance<HALF>(V, c, z0)) / fmaxf(z1.z - z0.z, 1e-6f);
        vec3 cn = v3(dx, dy, dz);
        if (dot(cn, cn) > 1e-10f) ln = normalize(mixv(normalize(ln), normalize(cn), 0.35f));
    }
    if (dot(ln, ln) <= 1e-10f) ln = v3(0, 1, 0);
    return normalize(mulCols(vol.nrm, normalize(ln), 0.0f));
}

// A sign change found by the march: the bracket [lowT, highT] that the
// bisection starts from (lowD: the distance at lowT).
struct DrCrossing {
    float lowT, highT, lowD;
};

// intersectVolumeBrickPrepared (:2608-2807) up to the first sign change of
// the sphere-traced march: AABB clip, validity and range culls, the
// back-tracking entry and the march. The bisection, polish, normal and
// material of a crossing are brickFinish, which the caller runs once the
// whole wave has left its grid walk (same operations per lane, so the same
// bits; the lanes that hit then refine together instead of stalling the
// others' walks once per hit).
template <bool HALF>
__device__ __forceinline__ bool brickMarch(const DrSdfDev &V, uint32_t bi, float closestT, vec3 lo, vec3 ld,
                                           float lupw, float cellSize, DrSdfCounters &sc, DrCrossing &cr) {
    // The whole header in one round of loads: every field feeds the one
    // branch below, so no load waits behind another.
    const float4 *hp = reinterpret_cast<const float4 *>(V.bricks + bi);
    DrBrick brick;
    brick.smin = hp[0];
    brick.smax = hp[1];
    brick.cmin = hp[2];
    brick.cmax = hp[3];
    float tNear, tFar;
    const bool inBox = sdfAABB(lo, ld, v3(brick.cmin.x, brick.cmin.y, brick.cmin.z),
                               v3(brick.cmax.x, brick.cmax.y, brick.cmax.z), closestT, tNear, tFar);
    DrBrickCtx ctx = brickCtx(V, brick);
    const bool cull = brick.smin.w > 0.0f || brick.smax.w < 0.0f;
    sc.invalid += (inBox && !ctx.valid) ? 1u : 0u;
    sc.rangeCulls += (inBox && ctx.valid && cull) ? 1u : 0u;
    if (!inBox || !ctx.valid || cull) return false;
    sc.marches++;
    float minimumStep = fmaxf(cellSize * 0.10f / lupw, 0.0005f);
    // Successive samples often fall in one cell: FP16 lookups keep the last
    // corner record.
    DrCornerCache cache{0xffffffffu, make_uint4(0u, 0u, 0u, 0u)};
    auto dist = [&](float at) {
        return HALF ? brickDistanceCached(V, ctx, lo + ld * at, cache) : brickDistance<HALF>(V, ctx, lo + ld * at);
    };
    float prevT = tNear;
    float prevD = dist(prevT);
    if (!isfinite(prevD)) return false;
    float originalEntryT = prevT;
    if (prevD <= 0.0f) {
        float backStep = fmaxf(cellSize / lupw * 0.5f, minimumStep);
        float searchT = prevT;
        #pragma unroll 1
        for (uint32_t bt = 0; bt < 12u; ++bt) {
            searchT = fmaxf(searchT - backStep, 0.0005f);
            float sd = dist(searchT);
            if (!isfinite(sd)) break;
            if (sd > 0.0f) {
                prevT = searchT;
                prevD = sd;
                break;
            }
            if (searchT <= 0.0005f) break;
        }
    }
    float t = fmaxf(originalEntryT, prevT + fmaxf(fabsf(prevD) / lupw * 0.9f, minimumStep));
    #pragma unroll 1
    for (uint32_t step = 0; step < 192u && t <= tFar && t < closestT; ++step) {
        sc.marchSteps++;
        float d = dist(t);
        if (!isfinite(d)) return false;
        bool crossed = (prevD > 0.0f && d <= 0.0f) || (prevD < 0.0f && d >= 0.0f);
        if (crossed) {
            cr = DrCrossing{prevT, t, prevD};
            return true;
        }
        prevT = t;
        prevD = d;
        t += fmaxf(fabsf(d) / lupw * 0.8f, minimumStep);
    }
    return false;
}

// The rest of intersectVolumeBrickPrepared for a crossing (:2700-2807):
// 8 bisections, the cubic Newton polish (fine grids) or the trilinear
// normal, front-facing flip and the nearest sample's material.
template <bool HALF>
__device__ __forceinline__ void brickFinish(const DrSdfDev &V, const DrVolume &vol, uint32_t bi, vec3 rd, vec3 lo,
                                            vec3 ld, float cellSize, DrCrossing cr, DrSdfCounters &sc, DrVolHit &hit) {
    const DrBrickCtx ctx = brickCtx(V, V.bricks[bi]);
    float lowT = cr.lowT, highT = cr.highT, lowD = cr.lowD;
    float refinedT = cr.highT;
    DrCornerCache cache{0xffffffffu, make_uint4(0u, 0u, 0u, 0u)};
    #pragma unroll 1
    for (uint32_t k = 0; k < 8u; ++k) {
        float midT = 0.5f * (lowT + highT);
        float md = HALF ? brickDistanceCached(V, ctx, lo + ld * midT, cache) : brickDistance<HALF>(V, ctx, lo + ld * midT);
        if ((lowD > 0.0f && md > 0.0f) || (lowD < 0.0f && md < 0.0f)) {
            lowT = midT;
            lowD = md;
        } else {
            highT = midT;
        }
        refinedT = 0.5f * (lowT + highT);
    }
    bool cubicPolish = cellSize <= 0.03f;
    bool coarse = V.renderQuality > 0u && cellSize > 0.03f;
    vec3 rp = lo + ld * refinedT;
    vec3 n;
    if (cubicPolish) {
        // Newton polish steps and the cubic normal share one evaluation
        // site: the last pass (or the pass that stops the polish) evaluates
        // the gradient at the final point, which is what volumeBrickNormal
        // computes there.
        uint32_t polishCount = V.renderQuality > 1u ? 2u : 1u;
        vec3 ln;
        #pragma unroll 1
        for (uint32_t k = 0;; ++k) {
            DrDG pg = brickDistanceGradientCubic<HALF>(V, ctx, rp);
            ln = pg.gradient;
            if (k == polishCount) break;
            float slope = dot(pg.gradient, ld);
            if (!isfinite(pg.distance) || fabsf(slope) <= 1e-6f) break;
            float projected = refinedT - pg.distance / slope;
            if (projected <= lowT || projected >= highT) break;
            refinedT = projected;
            rp = lo + ld * refinedT;
        }
        if (dot(ln, ln) <= 1e-10f) ln = v3(0, 1, 0);
        n = normalize(mulCols(vol.nrm, normalize(ln), 0.0f));
    } else {
        n = brickNormal<HALF>(V, vol, ctx, rp, false, coarse);
    }
    bool ff = dot(n, rd) <= 0.0f;
    if (!ff) n = -n;
    // Nearest sample's material (sampleVolumeBrickMaterial, :2350-2356).
    vec3 uvw = vclamp((rp - ctx.smin) / ctx.extent, v3s(0.0f), v3s(1.0f));
    uint32_t nx = min((uint32_t)roundf(uvw.x * (float)(ctx.dim[0] - 1u)), ctx.dim[0] - 1u);
    uint32_t ny = min((uint32_t)roundf(uvw.y * (float)(ctx.dim[1] - 1u)), ctx.dim[1] - 1u);
    uint32_t nz = min((uint32_t)roundf(uvw.z * (float)(ctx.dim[2] - 1u)), ctx.dim[2] - 1u);
    hit.t = refinedT;
    hit.normal = n;
    hit.material = V.material[ctx.offset + nx + ny * ctx.strideY + nz * ctx.strideZ];
    hit.frontFacing = ff;
    hit.dense = false;
    sc.hits++;
}

// ------------------------------------------------------- dense volumes --
// sampleVolumeDistance (PathTrace.metal:1915-1955).
__device__ __forceinline__ float denseDistance(const DrSdfDev &V, const DrVolume &vol, vec3 lp) {
    uint32_t dim[3] = {vol.dimensions.x, vol.dimensions.y, vol.dimensions.z};
    uint32_t off = vol.metadata.x;
    uint32_t n = dim[0] * dim[1] * dim[2];
    if (off >= V.denseCount || off + n > V.denseCount) return __builtin_inff();
    vec3 lmin = v3f(vol.localBoundsMin);
    vec3 ext = vmax(v3f(vol.localBoundsMax) - lmin, v3s(1e-6f));
    vec3 uvw = vclamp((lp - lmin) / ext, v3s(0.0f), v3s(1.0f));
    float g[3] = {uvw.x * (float)(dim[0] - 1u), uvw.y * (float)(dim[1] - 1u), uvw.z * (float)(dim[2] - 1u)};
    uint32_t b[3], nx[3];
    float f[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b[a] = (uint32_t)floorf(g[a]);
        nx[a] = min(b[a] + 1u, dim[a] - 1u);
        f[a] = g[a] - (float)b[a];
    }
    uint32_t sy = dim[0], sz = dim[0] * dim[1];
    const DrVolumeSample *s = V.dense + off;
    float c00 = mixf(s[b[0] + b[1] * sy + b[2] * sz].distance, s[nx[0] + b[1] * sy + b[2] * sz].distance, f[0]);
    float c10 = mixf(s[b[0] + nx[1] * sy + b[2] * sz].distance, s[nx[0] + nx[1] * sy + b[2] * sz].distance, f[0]);
    float c01 = mixf(s[b[0] + b[1] * sy + nx[2] * sz].distance, s[nx[0] + b[1] * sy + nx[2] * sz].distance, f[0]);
    float c11 = mixf(s[b[0] + nx[1] * sy + nx[2] * sz].distance, s[nx[0] + nx[1] * sy + nx[2] * sz].distance, f[0]);
    float c0 = mixf(c00, c10, f[1]);
    float c1 = mixf(c01, c11, f[1]);
    return mixf(c0, c1, f[2]);
}

// volumeNormal (:2029-2055): central differences one sample apart.
__device__ __forceinline__ vec3 denseNormal(const DrSdfDev &V, const DrVolume &vol, vec3 lp) {
    vec3 ext = vmax(v3f(vol.localBoundsMax) - v3f(vol.localBoundsMin), v3s(1e-6f));
    vec3 sp = ext / vmax(v3((float)vol.dimensions.x, (float)vol.dimensions.y, (float)vol.dimensions.z) - v3s(1.0f),
                         v3s(1.0f));
    float dx = denseDistance(V, vol, lp + v3(sp.x, 0, 0)) - denseDistance(V, vol, lp - v3(sp.x, 0, 0));
    float dy = denseDistance(V, vol, lp + v3(0, sp.y, 0)) - denseDistance(V, vol, lp - v3(0, sp.y, 0));
    float dz = denseDistance(V, vol, lp + v3(0, 0, sp.z)) - denseDistance(V, vol, lp - v3(0, 0, sp.z));
    vec3 ln = v3(dx, dy, dz);
    if (dot(ln, ln) <= 1e-10f) ln = v3(0, 1, 0);
    return normalize(mulCols(vol.nrm, normalize(ln), 0.0f));
}

// intersectVolume (:2471-2606): world-t march, <= 512 steps, 10 bisections.
__device__ __forceinline__ bool intersectDense(const DrSdfDev &V, const DrVolume &vol, vec3 ro, vec3 rd,
                                               float closestT, DrSdfCounters &sc, DrVolHit &hit) {
    float tNear, tFar;
    if (!sdfAABB(ro, rd, v3f(vol.worldBoundsMin), v3f(vol.worldBoundsMax), closestT, tNear, tFar)) return false;
    vec3 localStep = mulCols(vol.w2l, rd, 0.0f);
    float lupw = fmaxf(length(localStep), 1e-6f);
    vec3 ext = vmax(v3f(vol.localBoundsMax) - v3f(vol.localBoundsMin), v3s(1e-6f));
    float cellSize = fminf(ext.x / fmaxf((float)vol.dimensions.x - 1.0f, 1.0f),
                           fminf(ext.y / fmaxf((float)vol.dimensions.y - 1.0f, 1.0f),
                                 ext.z / fmaxf((float)vol.dimensions.z - 1.0f, 1.0f)));
    float minimumStep = fmaxf(cellSize * 0.10f / lupw, 0.0005f);
    float prevT = tNear;
    float prevD = denseDistance(V, vol, mulCols(vol.w2l, ro + rd * prevT, 1.0f));
    if (!isfinite(prevD)) return false;
    float t = prevT + fmaxf(fabsf(prevD) / lupw * 0.9f, minimumStep);
    #pragma unroll 1
    for (uint32_t step = 0; step < 512u && t <= tFar && t < closestT; ++step) {
        sc.denseSteps++;
        float d = denseDistance(V, vol, mulCols(vol.w2l, ro + rd * t, 1.0f));
        if (!isfinite(d)) return false;
        bool crossed = (prevD > 0.0f && d <= 0.0f) || (prevD < 0.0f && d >= 0.0f);
        if (crossed) {
            float lowT = prevT, highT = t, lowD = prevD, refinedT = t;
            #pragma unroll 1
            for (uint32_t k = 0; k < 10u; ++k) {
                float midT = 0.5f * (lowT + highT);
                float md = denseDistance(V, vol, mulCols(vol.w2l, ro + rd * midT, 1.0f));
                if ((lowD > 0.0f && md > 0.0f) || (lowD < 0.0f && md < 0.0f)) {
                    lowT = midT;
                    lowD = md;
                } else {
                    highT = midT;
                }
                refinedT = 0.5f * (lowT + highT);
            }
            vec3 rl = mulCols(vol.w2l, ro + rd * refinedT, 1.0f);
            vec3 n = denseNormal(V, vol, rl);
            bool ff = dot(n, rd) <= 0.0f;
            if (!ff) n = -n;
            // sampleVolumeMaterial (:1957-1990): the nearest sample.
            uint32_t dim[3] = {vol.dimensions.x, vol.dimensions.y, vol.dimensions.z};
            vec3 uvw = vclamp((rl - v3f(vol.localBoundsMin)) / ext, v3s(0.0f), v3s(1.0f));
            uint32_t qx = min((uint32_t)roundf(uvw.x * (float)(dim[0] - 1u)), dim[0] - 1u);
            uint32_t qy = min((uint32_t)roundf(uvw.y * (float)(dim[1] - 1u)), dim[1] - 1u);
            uint32_t qz = min((uint32_t)roundf(uvw.z * (float)(dim[2] - 1u)), dim[2] - 1u);
            hit.t = refinedT;
            hit.normal = n;
            hit.denseSample = vol.metadata.x + qx + qy * dim[0] + qz * dim[0] * dim[1];
            hit.frontFacing = ff;
            hit.dense = true;
            return true;
        }
        prevT = t;
        prevD = d;
        t += fmaxf(fabsf(d) / lupw * 0.8f, minimumStep);
    }
    return false;
}

// closestVolumeHit (PathTrace.metal:3090-3163): dense volumes, then the
// brick-grid branch for sparse volumes, below closestT.
// occludeBelow / certain: an occlusion test. When a sparse brick's march
// brackets a crossing entirely below occludeBelow, the walk stops there with
// *certain = true and out.t = the bracket's upper end (the refined hit, not
// computed, lies at or below it); otherwise the result is the closest hit
// as usual.
template <bool HALF>
__device__ __forceinline__ bool dr_trace_volumes(const DrSdfDev &V, vec3 ro, vec3 rd, float closestT, DrSdfCounters &sc,
                                              DrVolHit &out, float occludeBelow = -__builtin_inff(),
                                              bool *certain = nullptr) {
    bool any = false;
    float best = closestT;
    for (uint32_t vi = 0; vi < V.volumeCount; ++vi) {
        const DrVolume &vol = V.volumes[vi];
        if (vol.metadata.w != 0u) continue;
        sc.denseTests++;
        DrVolHit cand;
        if (intersectDense(V, vol, ro, rd, best, sc, cand) && cand.t < best) {
            out = cand;
            out.volume = vi;
            best = cand.t;
            any = true;
        }
    }
    if (V.gridCount == 0u || V.gridIndexCount == 0u) return any;
    for (uint32_t vi = 0; vi < V.volumeCount; ++vi) {
        const DrVolume &vol = V.volumes[vi];
        if (vol.metadata.w != 1u) continue;
        const DrBrickGrid *grid = nullptr;
        for (uint32_t g = 0; g < V.gridCount; ++g)
            if (V.grids[g].brickSizeAndVolume.w == vi) {
                grid = &V.grids[g];
                break;
            }
        if (!grid) continue;
        uint32_t gd[3] = {grid->dimensionsAndIndexOffset.x, grid->dimensionsAndIndexOffset.y,
                          grid->dimensionsAndIndexOffset.z};
        uint32_t gio = grid->dimensionsAndIndexOffset.w;
        uint32_t bs[3] = {max(grid->brickSizeAndVolume.x, 1u), max(grid->brickSizeAndVolume.y, 1u),
                          max(grid->brickSizeAndVolume.z, 1u)};
        if (gd[0] == 0u || gd[1] == 0u || gd[2] == 0u) continue;
        float tNear, tFar;
        if (!sdfAABB(ro, rd, v3f(vol.worldBoundsMin), v3f(vol.worldBoundsMax), best, tNear, tFar)) continue;
        vec3 lo = mulCols(vol.w2l, ro, 1.0f);
        vec3 ld = mulCols(vol.w2l, rd, 0.0f);
        vec3 lbmin = v3f(vol.localBoundsMin);
        vec3 extent = vmax(v3f(vol.localBoundsMax) - lbmin, v3s(1e-6f));
        vec3 den = vmax(v3((float)vol.dimensions.x, (float)vol.dimensions.y, (float)vol.dimensions.z) - v3s(1.0f),
                        v3s(1.0f));
        vec3 bsf = vmax(v3((float)bs[0], (float)bs[1], (float)bs[2]), v3s(1.0f));
        vec3 spacing = extent / den;
        float cellSize = fminf(spacing.x, fminf(spacing.y, spacing.z));
        float lupw = fmaxf(length(ld), 1e-6f);
        float t = tNear;
        vec3 lp = lo + ld * t;
        vec3 samplePos = ((lp - lbmin) / extent) * den;
        vec3 brickPos = (samplePos + v3s(0.5f)) / bsf;
        vec3 gdm1 = v3((float)gd[0] - 1.0f, (float)gd[1] - 1.0f, (float)gd[2] - 1.0f);
        vec3 fc = vclamp(v3(floorf(brickPos.x), floorf(brickPos.y), floorf(brickPos.z)), v3s(0.0f), gdm1);
        int32_t cell[3] = {(int32_t)fc.x, (int32_t)fc.y, (int32_t)fc.z};
        vec3 vel = (ld / extent) * den / bsf;
        float velv[3] = {vel.x, vel.y, vel.z};
        float bpv[3] = {brickPos.x, brickPos.y, brickPos.z};
        int32_t stepDir[3];
        float tMax[3], tDelta[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            stepDir[a] = velv[a] > 0.0f ? 1 : (velv[a] < 0.0f ? -1 : 0);
            float nb = stepDir[a] > 0 ? (float)cell[a] + 1.0f : (float)cell[a];
            tMax[a] = stepDir[a] == 0 ? __builtin_inff() : t + (nb - bpv[a]) / velv[a];
            tDelta[a] = stepDir[a] == 0 ? __builtin_inff() : fabsf(1.0f / velv[a]);
        }
        uint32_t md[3] = {grid->macroDimensionsAndIndexOffset.x, grid->macroDimensionsAndIndexOffset.y,
                          grid->macroDimensionsAndIndexOffset.z};
        uint32_t mo = grid->macroDimensionsAndIndexOffset.w;
        uint32_t ms[3] = {max(grid->macroSizeAndReserved.x, 1u), max(grid->macroSizeAndReserved.y, 1u),
                          max(grid->macroSizeAndReserved.z, 1u)};
        bool hasMacro = md[0] > 0u && md[1] > 0u && md[2] > 0u && mo < V.gridIndexCount;
        // Macro cells of a power-of-two size (4 in the builder) divide by shifting.
        const bool msPow2 = (ms[0] & (ms[0] - 1u)) == 0u && (ms[1] & (ms[1] - 1u)) == 0u && (ms[2] & (ms[2] - 1u)) == 0u;
        const uint32_t msh[3] = {(uint32_t)__builtin_ctz(ms[0]), (uint32_t)__builtin_ctz(ms[1]),
                                 (uint32_t)__builtin_ctz(ms[2])};
        uint32_t maxSteps = min(gd[0] + gd[1] + gd[2] + 8u, 1024u);
        bool volHit = false;
        DrVolHit closest;
        closest.t = best;
        // The end of a grid step (:3064-3086): false when the walk stops.
        auto advance = [&]() -> bool {
            float nextT = fminf(tMax[0], fminf(tMax[1], tMax[2]));
            if (volHit && closest.t <= nextT + 0.0005f) return false;
            if (nextT > tFar || nextT >= closest.t) return false;
            if (tMax[0] <= tMax[1] && tMax[0] <= tMax[2]) {
                cell[0] += stepDir[0];
                t = tMax[0];
                tMax[0] += tDelta[0];
            } else if (tMax[1] <= tMax[2]) {
                cell[1] += stepDir[1];
                t = tMax[1];
                tMax[1] += tDelta[1];
            } else {
                cell[2] += stepDir[2];
                t = tMax[2];
                tMax[2] += tDelta[2];
            }
            return true;
        };
        // The grid walk leaves its loop at a brick whose march crossed the
        // surface; that hit is finished outside the loop (so the wave's
        // hitting lanes refine together) and the walk resumes where it
        // stopped, exactly as the reference's single loop continues.
        uint32_t step = 0;
        #pragma unroll 1
        while (true) {
            bool found = false;
            uint32_t foundBrick = 0u;
            DrCrossing cr;
            #pragma unroll 1
            for (; step < maxSteps && t <= tFar && t < closest.t; ++step) {
                sc.gridCells++;
                if (cell[0] < 0 || cell[1] < 0 || cell[2] < 0 || cell[0] >= (int32_t)gd[0] ||
                    cell[1] >= (int32_t)gd[1] || cell[2] >= (int32_t)gd[2])
                    break;
                // The cell's brick index is fetched beside its macro word
                // (one round trip; it is unused when the macro cell is empty).
                const uint32_t slot =
                    gio + (uint32_t)cell[0] + (uint32_t)cell[1] * gd[0] + (uint32_t)cell[2] * gd[0] * gd[1];
                const uint32_t bi = slot < V.gridIndexCount ? V.gridIndices[slot] : 0xffffffffu;
                if (hasMacro) {
                    uint32_t mc[3];
                    if (msPow2) {
#pragma unroll
                        for (int a = 0; a < 3; ++a) mc[a] = min((uint32_t)cell[a] >> msh[a], md[a] - 1u);
                    } else {
#pragma unroll
                        for (int a = 0; a < 3; ++a) mc[a] = min((uint32_t)cell[a] / ms[a], md[a] - 1u);
                    }
                    uint32_t mslot = mo + mc[0] + mc[1] * md[0] + mc[2] * md[0] * md[1];
                    if (mslot < V.gridIndexCount && V.gridIndices[mslot] == 0u) {
                        vec3 cur = brickPos + vel * (t - tNear);
                        float cv[3] = {cur.x, cur.y, cur.z};
                        float mt[3];
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            float boundary =
                                stepDir[a] > 0 ? (float)min((mc[a] + 1u) * ms[a], gd[a]) : (float)(mc[a] * ms[a]);
                            mt[a] = stepDir[a] == 0 ? __builtin_inff() : t + (boundary - cv[a]) / velv[a];
                        }
                        float macroNextT = fminf(mt[0], fminf(mt[1], mt[2]));
                        if (macroNextT > t + 1e-6f) {
                            sc.macroSkips++;
                            if (macroNextT > tFar || macroNextT >= closest.t) break;
                            t = macroNextT + 1e-5f;
                            cur = brickPos + vel * (t - tNear);
                            vec3 fc2 = vclamp(v3(floorf(cur.x), floorf(cur.y), floorf(cur.z)), v3s(0.0f), gdm1);
                            cell[0] = (int32_t)fc2.x;
                            cell[1] = (int32_t)fc2.y;
                            cell[2] = (int32_t)fc2.z;
                            float cv2[3] = {cur.x, cur.y, cur.z};
#pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                float nb = stepDir[a] > 0 ? (float)cell[a] + 1.0f : (float)cell[a];
                                tMax[a] = stepDir[a] == 0 ? __builtin_inff() : t + (nb - cv2[a]) / velv[a];
                            }
                            continue;
                        }
                    }
                }
                {
                    if (bi != 0xffffffffu && bi < V.brickCount) {
                        sc.brickTests++;
                        if (brickMarch<HALF>(V, bi, closest.t, lo, ld, lupw, cellSize, sc, cr)) {
                            found = true;
                            foundBrick = bi;
                            break;
                        }
                    }
                }
                if (!advance()) break;
            }
            if (!found) break;
            if (cr.highT < occludeBelow) {
                // The caller only asks whether some hit lies below
                // occludeBelow: this crossing's final t is at most highT.
                out.t = cr.highT;
                out.volume = vi;
                out.dense = false;
                *certain = true;
                return true;
            }
            // A crossing's final t lies in its bracket, below closest.t, so
            // it always replaces the current closest hit.
            DrVolHit cand;
            brickFinish<HALF>(V, vol, foundBrick, rd, lo, ld, cellSize, cr, sc, cand);
            if (cand.t < closest.t) {
                closest = cand;
                closest.volume = vi;
                volHit = true;
            }
            if (!advance()) break;
            ++step;
        }
        if (volHit && closest.t < best) {
            out = closest;
            best = closest.t;
            any = true;
        }
    }
    return any;
}
