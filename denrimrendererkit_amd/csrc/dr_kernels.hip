// dr_kernels.hip -- wavefront progressive path tracer for gfx950.
//
// One progressive sample of pathTraceKernel (PathTrace.metal:5167-5586) is
// split into stages that run over compacted queues of path slots:
//
//   raygen   -- slot -> (pixel, sample); RNG seed and camera ray (:5212-5225)
//   extend   -- closest-hit traversal for every live path (persistent,
//               wave-granular work stealing)
//   shade    -- one bounce of the reference loop body (:5279-5517): miss and
//               emission terms, pass-through, volume / dielectric
//               continuation, NEE sample generation, lobe sampling, RR. Live
//               paths are appended to the next queue with a wave ballot and
//               one atomic per wave; NEE shadow rays go to a shadow queue.
//               Paths that pass the subsurface gate go to a walk queue.
//   sss      -- the subsurface random walks of the bounce (:3525-3621), all
//               lanes of a wave walking, so their nested traversals run
//               converged; a failed walk sends the path to
//   resume   -- shade again from the dielectric step on (:5365-5517).
//   shadow   -- occlusion / transmittance of each NEE ray (any-hit when no
//               material can pass light, else the reference's 8-step chain)
//   resolve  -- sums a path's NEE terms in the reference order and adds the
//               clamped total (:5394-5422)
//   finalize -- running-mean accumulation in sample order (:5520-5524) and
//               miss AOVs (:5578-5585)
//
// Path state is structure-of-arrays of float4/uint4 (one dwordx4 per lane,
// coalesced). The per-path RNG stream is keyed by (frameSeed, pixel,
// sampleIndex) only, so a path's result does not depend on which slot,
// queue position or batch it ran in.
#include <algorithm>

#include "dr_kernels.h"
#include "dr_shade.h"
#include "dr_traverse.h"
#include "dr_sdf.h"

#define DR_FLAG_PRIMARY 1u
#define DR_FLAG_NEE 2u

// Minimum waves per SIMD the register allocator is held to for the
// triangle-only kernels (1 = no constraint); the volume variants of extend
// and shadow are held to DR_VOL_WAVES, the other volume variants are free.
// With 32 samples in flight the triangle tracers and k_shade gain from
// waves over registers: extension and walk/shadow tracing at 5 waves
// (27 and 16 VGPRs spilled) and k_shade at 4 (10 spilled) took the 1080p
// terrain from 1724 to 1855 Mrays/s; 6 waves on the walk/shadow tracer
// (more spills) lost 20%, and the volume tracers stay at 3.
#ifndef DR_EXTEND_WAVES
#define DR_EXTEND_WAVES 5
#endif
#ifndef DR_SHADOW_WAVES
#define DR_SHADOW_WAVES 1
#endif
#ifndef DR_SSS_WAVES
#define DR_SSS_WAVES 5
#endif
#ifndef DR_VOL_WAVES
#define DR_VOL_WAVES 3
#endif
#define DR_EXTEND_OCC __attribute__((amdgpu_waves_per_eu(VOL ? DR_VOL_WAVES : DR_EXTEND_WAVES)))
#define DR_SHADOW_OCC __attribute__((amdgpu_waves_per_eu(VOL ? DR_VOL_WAVES : DR_SHADOW_WAVES)))
#define DR_SSS_OCC __attribute__((amdgpu_waves_per_eu(VOL ? 1 : DR_SSS_WAVES)))
// Loop shape of k_sss's persistent tracer (walk segments and shadow-chain
// steps): stepwise.
#ifndef DR_SSS_STEPWISE
#define DR_SSS_STEPWISE true
#endif
// Sort keys of k_shade's block-local ordering by what a path hit.
#ifndef DR_SHADE_KEYS
#define DR_SHADE_KEYS 16u
#endif
#ifndef DR_SHADE_WAVES
#define DR_SHADE_WAVES 4
#endif
#ifndef DR_VOL_SHADE_WAVES
#define DR_VOL_SHADE_WAVES 1
#endif
#define DR_SHADE_OCC __attribute__((amdgpu_waves_per_eu(VOL ? DR_VOL_SHADE_WAVES : DR_SHADE_WAVES)))
__device__ __forceinline__ uint32_t laneId() { return __lane_id(); }

// Sample index of the batch's first sample: a launch argument, or a device
// word advanced by the batch graph itself when batches are replayed.
__device__ __forceinline__ uint32_t batchBegin(const DrLaunch &L) { return L.sampleBase ? *L.sampleBase : L.sampleBegin; }

__global__ void k_set_base(uint32_t *base, uint32_t value) { *base = value; }
__global__ void k_advance_base(uint32_t *base, uint32_t samples) { *base += samples; }

// Wave-aggregated append: returns the position for this lane (valid only if
// keep). All lanes of the wave that are still active must call it.
__device__ __forceinline__ uint32_t waveAppend(bool keep, uint32_t *counter) {
    uint64_t mask = __ballot(keep);
    uint32_t total = __popcll(mask);
    uint32_t leader = __ffsll((unsigned long long)__ballot(1)) - 1;
    uint32_t base = 0;
    if (total && laneId() == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader);
    uint32_t below = __popcll(mask & ((1ull << laneId()) - 1ull));
    return base + below;
}

// Adds the wave's total of v to a 64-bit counter with one atomic; every lane
// of the wave must be active (the tracers' exit, where all 64 lanes meet).
__device__ __forceinline__ void waveAddTotal(unsigned long long *stats, int slot, uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (laneId() == 0 && v) atomicAdd(&stats[slot], (unsigned long long)v);
}

__device__ __forceinline__ void waveAddStat(unsigned long long *stats, int slot, uint32_t v) {
    // Counting runs only (statsEnabled); one relaxed 64-bit atomic per lane.
    if (v) atomicAdd(&stats[slot], (unsigned long long)v);
}
__device__ __forceinline__ void addSdfStats(unsigned long long *stats, const DrSdfCounters &c,
                                            int base = DR_CNT_SDF_BASE) {
    waveAddStat(stats, base + 0, c.denseTests);
    waveAddStat(stats, base + 1, c.denseSteps);
    waveAddStat(stats, base + 2, c.gridCells);
    waveAddStat(stats, base + 3, c.macroSkips);
    waveAddStat(stats, base + 4, c.brickTests);
    waveAddStat(stats, base + 5, c.invalid);
    waveAddStat(stats, base + 6, c.rangeCulls);
    waveAddStat(stats, base + 7, c.marches);
    waveAddStat(stats, base + 8, c.marchSteps);
    waveAddStat(stats, base + 9, c.hits);
}

// A closest hit as passed between stages. Triangles: (t, u, v, bits(id)).
// Sparse volumes: (t, bits(blend), bits(A | B << 16),
// bits(DR_VOL_HIT | frontFacing << 30 | volume)); dense volumes:
// (t, bits(nearest sample), 0, bits(DR_VOL_HIT | DR_VOL_DENSE |
// frontFacing << 30 | volume)); volume normals in hitN. Miss: id 0xffffffff.
#define DR_VOL_HIT 0x80000000u
#define DR_VOL_FRONT 0x40000000u
#define DR_VOL_DENSE 0x20000000u
#define DR_VOL_INDEX 0x1fffffffu
#define DR_MISS 0xffffffffu

// intersectScene (PathTrace.metal:3294-3421): closest triangle (lowest index
// on equal t), then the closest sparse-volume hit in front of it (:3397-3418).
// VOL = false compiles the triangle-only form (scenes without volumes).
// occludeBelow / certain (volumes): an occlusion test; *certain is set
// when some hit is known to lie below occludeBelow (a triangle, or a sparse
// crossing whose bracket does), and h.x is then an upper bound of the
// closest hit's t rather than the hit.
template <bool VOL>
__device__ __forceinline__ bool sceneClosest(const DrSceneDev &S, vec3 o, vec3 d, DrStack &stk, uint32_t &nv,
                                             uint32_t &nt, DrSdfCounters &sc, float4 &h, float4 &hn,
                                             float occludeBelow = -__builtin_inff(), bool *certain = nullptr) {
    DrHitRec hr{__builtin_inff(), 0.0f, 0.0f, DR_MISS};
    bool hit = false;
    if (S.triangleCount) {
        DrRayPre r = dr_prep(o, d);
        hit = dr_traverse<false>(S, r, hr, stk, nv, nt);
    }
    h = make_float4(hr.t, hr.u, hr.v, __uint_as_float(hr.id));
    if constexpr (VOL) {
        if (hit && hr.t < occludeBelow) {
            *certain = true;
            return true;
        }
        DrVolHit vh;
        if (S.sdf.distHalf ? dr_trace_volumes<true>(S.sdf, o, d, hr.t, sc, vh, occludeBelow, certain)
                          : dr_trace_volumes<false>(S.sdf, o, d, hr.t, sc, vh, occludeBelow, certain)) {
            if (certain && *certain) {
                h.x = vh.t;
                return true;
            }
            uint32_t id = DR_VOL_HIT | (vh.frontFacing ? DR_VOL_FRONT : 0u) | vh.volume;
            if (vh.dense)
                h = make_float4(vh.t, __uint_as_float(vh.denseSample), 0.0f, __uint_as_float(id | DR_VOL_DENSE));
            else
                h = make_float4(vh.t, __uint_as_float(vh.material.y), __uint_as_float(vh.material.x),
                                __uint_as_float(id));
            hn = make_float4(vh.normal.x, vh.normal.y, vh.normal.z, 0.0f);
            hit = true;
        }
    }
    return hit;
}

// ---------------------------------------------------------------- raygen --
// Path slot i of the batch: pixel, sample index, RNG seed and the camera
// ray (pathTraceKernel's set-up, PathTrace.metal:5204-5230), written to the
// path state for the stages that follow. The extension kernel of bounce 0
// runs this for each slot it claims, so rays are generated by the
// persistent tracing lanes that trace them.
// Slots of a batch: the tile's pixels in groups of DR_PIXEL_GROUP
// consecutive pixels (the last group may be shorter); a group's slots hold
// its pixels' samples, sample-major within the group. With k = 16 samples a
// wavefront of primary rays is 4 pixels x 16 samples, whose rays take
// nearly the same path through the BVH and the bricks, and the finalize
// loop's lanes still read 64-byte runs. Measured against sample-major slots
// (each sample a full-frame wavefront sweep): sparse SDF 1024^2 950 -> 1068
// Mrays/s, terrain 1219 -> 1239, room 1660 -> 1729; groups of 1 pixel
// trace alike but finalize reads 256-byte strides (5x slower), groups of 8
// or 16 trace 1-3% slower.
#ifndef DR_PIXEL_GROUP
#define DR_PIXEL_GROUP 4u
#endif
__device__ __forceinline__ void slotPixelSample(const DrLaunch &L, uint32_t i, uint32_t &p, uint32_t &s) {
    const uint32_t n = L.tileWidth * L.tileHeight, k = L.samplesInBatch;
    const uint32_t g = i / (DR_PIXEL_GROUP * k);
    const uint32_t first = g * DR_PIXEL_GROUP;
    const uint32_t width = min(DR_PIXEL_GROUP, n - first);
    const uint32_t r = i - first * k;
    s = r / width;
    p = first + (r - s * width);
}
__device__ __forceinline__ uint32_t pixelSampleSlot(const DrLaunch &L, uint32_t p, uint32_t s) {
    const uint32_t n = L.tileWidth * L.tileHeight, k = L.samplesInBatch;
    const uint32_t first = p / DR_PIXEL_GROUP * DR_PIXEL_GROUP;
    const uint32_t width = min(DR_PIXEL_GROUP, n - first);
    return first * k + s * width + (p - first);
}

__device__ __forceinline__ void generatePath(const DrLaunch &L, const DrCamera &cam, const DrPaths &P, uint32_t i,
                                             vec3 &o, vec3 &d) {
    uint32_t p, sLocal;
    slotPixelSample(L, i, p, sLocal);
    uint32_t px = L.tileX + p % L.tileWidth;
    uint32_t py = L.tileY + p / L.tileWidth;
    uint32_t sampleIndex = batchBegin(L) + sLocal;
    uint32_t state = L.frameSeed ^ dr_hash(px + py * L.width) ^ dr_hash(sampleIndex + 1u);
    float jx = randomFloat(state);
    float jy = randomFloat(state);
    float lx = randomFloat(state);
    float ly = randomFloat(state);
    makeCameraRay(cam, L.width, L.height, px, py, jx, jy, lx, ly, o, d);
    P.rayO[i] = make_float4(o.x, o.y, o.z, 0.0f);  // w: previousBSDFPDF
    P.rayD[i] = make_float4(d.x, d.y, d.z, 0.0f);
    P.thr[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);  // w: outputAlpha
    P.rad[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.meta[i] = make_uint4(px + py * L.width, sampleIndex, state, 0u);
    P.queue[0][i] = i;
}

// Start of a sample batch: every counter word and the claim cursors of the
// batch's bounces zeroed, bounce 0's queue holds all nSlots paths (generated
// by k_extend(0)).
__global__ void k_batch_begin(DrPaths P, uint32_t nSlots, uint32_t claimWords) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, n = gridDim.x * blockDim.x;
    if (blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < DR_CTR_WORDS; k += blockDim.x)
            P.counters[k] = k == DR_CTR_Q(0) ? nSlots : 0u;
    for (uint32_t k = i; k < claimWords; k += n) P.claims[k] = 0u;
}

// Fills the surface record of a hit (attribute half of intersectTriangle,
// PathTrace.metal:1841-1850, and the Hit fields set at :3372-3384).
struct Surf {
    vec3 position, normal, tangent, bitangent;
    vec2 uv;
    float t;
    uint32_t materialID, objectID, primitiveID;
    uint32_t materialID2;
    float materialBlend;
    uint32_t fieldSample;  // dense hits: sample whose material fields apply (~0: none)
    bool frontFacing;
};
// Volume hit record (closestVolumeHit's Hit fields, :3397-3418, and
// intersectVolume's, :2560-2596): the tangent frame is built from the
// normal with the helper axis rule.
__device__ __forceinline__ Surf volumeSurfaceAt(const DrSceneDev &S, vec3 o, vec3 d, float4 h, float4 hn) {
    Surf s;
    uint32_t id = __float_as_uint(h.w);
    const DrVolume &vol = S.sdf.volumes[id & DR_VOL_INDEX];
    s.t = h.x;
    s.position = o + d * h.x;
    s.normal = v3(hn.x, hn.y, hn.z);
    s.uv = vec2{0.0f, 0.0f};
    vec3 helper = fabsf(s.normal.y) < 0.999f ? v3(0, 1, 0) : v3(1, 0, 0);
    s.tangent = normalize(cross(helper, s.normal));
    s.bitangent = normalize(cross(s.normal, s.tangent));
    if (id & DR_VOL_DENSE) {
        uint32_t k = __float_as_uint(h.y);
        const DrVolumeSample &ds = S.sdf.dense[k];
        s.materialID = ds.materialA;
        s.materialID2 = ds.materialB;
        s.materialBlend = clampf(ds.materialBlend, 0.0f, 1.0f);
        s.fieldSample = k;
    } else {
        uint32_t ab = __float_as_uint(h.z);
        s.materialID = ab & 0xffffu;
        s.materialID2 = ab >> 16;
        s.materialBlend = clampf(h.y, 0.0f, 1.0f);
        s.fieldSample = DR_MISS;
    }
    s.objectID = vol.metadata.y;
    s.primitiveID = vol.metadata.z;
    s.frontFacing = (id & DR_VOL_FRONT) != 0u;
    return s;
}
template <bool VOL>
__device__ __forceinline__ Surf surfaceAt(const DrSceneDev &S, vec3 o, vec3 d, float4 h, float4 hn) {
    Surf s;
    uint32_t id = __float_as_uint(h.w);
    if constexpr (VOL) {
        if (id & DR_VOL_HIT) return volumeSurfaceAt(S, o, d, h, hn);
    }
    const DrTriangle &tri = S.triangles[id];
    float t = h.x, u = h.y, v = h.z;
    float w = 1.0f - u - v;
    vec3 n = normalize(w * xyz(tri.n0) + u * xyz(tri.n1) + v * xyz(tri.n2));
    s.frontFacing = dot(n, d) <= 0.0f;
    if (dot(n, d) > 0.0f) n = -n;
    s.t = t;
    s.position = o + t * d;
    s.normal = n;
    s.uv = vec2{w * tri.uv0.x + u * tri.uv1.x + v * tri.uv2.x, w * tri.uv0.y + u * tri.uv1.y + v * tri.uv2.y};
    s.tangent = normalize(xyz(tri.tangent));
    s.bitangent = normalize(xyz(tri.bitangent));
    s.materialID = tri.materialID;
    s.materialID2 = tri.materialID;
    s.materialBlend = 0.0f;
    s.fieldSample = DR_MISS;
    s.objectID = tri.objectID;
    s.primitiveID = id;
    return s;
}

__device__ __forceinline__ Mat materialFor(const DrSceneDev &S, const Surf &h) {  // materialForHit :1096-1137
    const DrMaterial *p = &S.materials[min(h.materialID, S.materialCount - 1u)];
    Mat m{p, p, 0.0f, p->s[M_BASE], 0u, {}, {}, {}};
    if (h.materialBlend > 0.0f) {
        m.q = &S.materials[min(h.materialID2, S.materialCount - 1u)];
        m.t = h.materialBlend;
        m.base = mix4(m.base, m.q->s[M_BASE], m.t);
    }
    if (h.fieldSample != DR_MISS) applyVolumeMaterialFields(m, S.sdf.dense[h.fieldSample]);
    return m;
}

__device__ __forceinline__ void applyNormalMap(const DrSceneDev &S, Surf &h, const Mat &m, vec3 rayDir) {  // :4021-4054
    uint32_t ti = (uint32_t)ms<M_PARAMS>(m).w;
    if (ti == 0u) return;
    vec3 mapped = xyz(sampleTexture(S, ti, h.uv, make_float4(0.5f, 0.5f, 1.0f, 1.0f)));
    mapped = normalize(mapped * 2.0f - v3s(1.0f));
    vec3 t = h.tangent - h.normal * dot(h.tangent, h.normal);
    if (length(t) <= 1e-5f) {
        vec3 helper = fabsf(h.normal.y) < 0.999f ? v3(0, 1, 0) : v3(1, 0, 0);
        t = cross(helper, h.normal);
    }
    t = normalize(t);
    vec3 b = h.bitangent - h.normal * dot(h.bitangent, h.normal);
    if (length(b) <= 1e-5f) b = cross(h.normal, t);
    b = normalize(b);
    h.normal = normalize(mapped.x * t + mapped.y * b + mapped.z * h.normal);
    if (dot(h.normal, rayDir) > 0.0f) h.normal = -h.normal;
}
__device__ __forceinline__ void applyBaseColorTexture(const DrSceneDev &S, Mat &m, const Surf &h) {  // :4005-4019
    uint32_t ti = (uint32_t)ms<M_PARAMS>(m).z;
    if (ti == 0u) return;
    float4 t = sampleTexture(S, ti, h.uv, make_float4(1.0f, 1.0f, 1.0f, 1.0f));
    m.base = make_float4(m.base.x * t.x, m.base.y * t.y, m.base.z * t.z, m.base.w * t.w);
}

__device__ __forceinline__ float lightPDFForHit(const DrSceneDev &S, const Surf &h, vec3 incoming) {  // :1580-1613
    if (h.primitiveID >= S.triangleCount) return 0.0f;
    uint32_t lp1 = S.triangles[h.primitiveID].padding2;
    if (lp1 == 0u) return 0.0f;
    uint32_t index = lp1 - 1u;
    if (index >= S.lightCount) return 0.0f;
    const DrLight &l = S.lights[index];
    if (l.triangleIndex != h.primitiveID || l.area <= 0.0f) return 0.0f;
    float cosLight = fmaxf(0.0f, dot(xyz(l.normal), -incoming));
    if (cosLight <= 0.0f) return 0.0f;
    return lightSelectionPDF(S.lights, index) * lightSolidAnglePDF(h.t * h.t, cosLight, l.area);
}

// Subsurface helpers (PathTrace.metal:3490-3523).
__device__ __forceinline__ float sampleFreeFlight(vec3 sigmaT, float ux, float uy) {
    float sigma = sigmaT.z;
    if (ux < 0.33333333333f) sigma = sigmaT.x;
    else if (ux < 0.66666666667f) sigma = sigmaT.y;
    return -logf(fmaxf(1.0f - uy, 1e-6f)) / fmaxf(sigma, 1e-4f);
}
__device__ __forceinline__ float collisionPDF(vec3 sigmaT, float d) {
    vec3 tr = vexp((-sigmaT) * d);
    return fmaxf((sigmaT.x * tr.x + sigmaT.y * tr.y + sigmaT.z * tr.z) * 0.33333333333f, 1e-6f);
}
__device__ __forceinline__ float boundaryPDF(vec3 sigmaT, float d) {
    vec3 tr = vexp((-sigmaT) * d);
    return fmaxf((tr.x + tr.y + tr.z) * 0.33333333333f, 1e-6f);
}


// The walk returns false before drawing anything unless the surface is
// entered from the front (:3534-3537); only then is it handed to k_sss.
__device__ __forceinline__ bool subsurfaceWalkStarts(const Mat &m, const Surf &entry) {
    return materialSubsurface(m) > 0.0f && entry.frontFacing;
}

template <bool VOL>
__device__ __forceinline__ bool continueSubsurfaceRandomWalk(const DrSceneDev &S, const Surf &entry, const Mat &m, vec3 &ro,
                                                          vec3 &rd, vec3 &T, float &prevPdf, uint32_t &state,
                                                          DrStack &stk, uint32_t &cBounce, uint32_t &cNodes,
                                                          uint32_t &cTris, uint32_t &cSss,
                                                          DrSdfCounters &sc) {  // :3525-3621
    float subsurface = materialSubsurface(m);
    if (subsurface <= 0.0f || !entry.frontFacing) return false;
    float e0 = randomFloat(state);
    float e1 = randomFloat(state);
    vec3 entryDir = orientHemisphere(cosineHemisphere(e0, e1), -entry.normal);
    vec3 mo = entry.position + entryDir * 0.001f;
    vec3 md = entryDir;
    vec3 sigmaT = v3s(1.0f) / materialSubsurfaceRadius(m);
    vec3 albedo = materialSubsurfaceColor(m);
    float g = materialSubsurfaceAnisotropy(m);
    for (uint32_t step = 0u; step < 8u; ++step) {
        cBounce++;
        cSss++;
        float4 hr, hn;
        if (!sceneClosest<VOL>(S, mo, md, stk, cNodes, cTris, sc, hr, hn)) return false;
        Surf b = surfaceAt<VOL>(S, mo, md, hr, hn);
        float f0 = randomFloat(state);
        float f1 = randomFloat(state);
        float ff = sampleFreeFlight(sigmaT, f0, f1);
        if (ff < b.t) {
            mo += md * ff;
            float h0 = randomFloat(state);
            float h1 = randomFloat(state);
            md = sampleHenyeyGreenstein(h0, h1, md, g);
            vec3 tr = vexp((-sigmaT) * ff);
            T *= albedo * sigmaT * tr / collisionPDF(sigmaT, ff);
            if (maxComponent(T) <= 1e-4f) return false;
            continue;
        }
        vec3 exitN = b.frontFacing ? b.normal : -b.normal;
        exitN = normalize(exitN);
        float x0 = randomFloat(state);
        float x1 = randomFloat(state);
        vec3 local = cosineHemisphere(x0, x1);
        rd = orientHemisphere(local, exitN);
        ro = b.position + exitN * 0.001f;
        vec3 tr = vexp((-sigmaT) * b.t);
        T *= tr / boundaryPDF(sigmaT, b.t);
        prevPdf = fmaxf(dot(exitN, rd), 0.0f) * 0.31830988618f;
        return true;
    }
    return false;
}

__device__ __forceinline__ bool continueTransmissionVolumeRandomWalk(const Surf &b, const Mat &m, vec3 &ro, vec3 &rd,
                                                                     vec3 &T, float &prevPdf,
                                                                     uint32_t &state) {  // :3685-3721
    if (b.frontFacing || !materialHasVolumeScattering(m)) return false;
    vec3 sigmaT = materialVolumeSigmaT(m);
    vec3 sigmaS = materialVolumeSigmaS(m);
    if (maxComponent(sigmaT) <= 1e-5f || maxComponent(sigmaS) <= 1e-5f) return false;
    float f0 = randomFloat(state);
    float f1 = randomFloat(state);
    float ff = sampleFreeFlight(sigmaT, f0, f1);
    if (ff >= b.t) {
        T *= materialVolumeTransmittance(m, b.t) / boundaryPDF(sigmaT, b.t);
        return false;
    }
    vec3 tr = vexp((-sigmaT) * ff);
    T *= sigmaS * tr / collisionPDF(sigmaT, ff);
    ro += rd * ff;
    float h0 = randomFloat(state);
    float h1 = randomFloat(state);
    rd = sampleHenyeyGreenstein(h0, h1, rd, clampf(ms<M_VOL_PARAMS>(m).y, -0.95f, 0.95f));
    ro += rd * 0.001f;
    prevPdf = 0.0f;
    return maxComponent(T) > 1e-4f;
}

__device__ __forceinline__ bool continueTransmissiveSurface(const Surf &h, const Mat &m, vec3 &ro, vec3 &rd, vec3 &T,
                                                         float &prevPdf, uint32_t &state) {  // :1723-1799
    float transmission = materialTransmission(m);
    if (transmission <= 0.0f) return false;
    float gate = randomFloat(state);
    if (gate > transmission) return false;
    vec3 incoming = rd;
    vec3 view = -incoming;
    float roughness = materialTransmissionRoughness(m);
    float ior = materialTransmissionIOR(m);
    vec3 micro = h.normal;
    if (roughness > 0.01f) {
        float a = randomFloat(state);
        float b = randomFloat(state);
        micro = sampleGGXHalfVector(a, b, h.normal, roughness);
        if (dot(micro, incoming) > 0.0f) micro = -micro;
    }
    float etaI = h.frontFacing ? 1.0f : ior;
    float etaT = h.frontFacing ? ior : 1.0f;
    float eta = etaI / etaT;
    float cosThetaI = fmaxf(dot(view, micro), 0.0f);
    float fresnel = dielectricFresnel(cosThetaI, etaI, etaT);
    if (materialThinWalled(m)) {
        fresnel = dielectricFresnel(cosThetaI, 1.0f, ior);
        float rp = clampf(fresnel, 0.02f, 0.98f);
        float choice = randomFloat(state);
        if (choice < rp) {
            rd = normalize(reflect(incoming, micro));
            ro = h.position + micro * 0.001f;
            T *= fresnel / fmaxf(rp, 1e-5f);
        } else {
            rd = incoming;
            ro = h.position + rd * 0.001f;
            float grazing = 1.0f - fabsf(dot(view, h.normal));
            float tint = mixf(0.35f, 0.95f, grazing * grazing);
            float w = (1.0f - fresnel) / fmaxf(1.0f - rp, 1e-5f);
            T *= materialTransmissionTint(m, tint) * w;
        }
        prevPdf = 0.0f;
        return true;
    }
    vec3 refracted = refract(incoming, micro, eta);
    bool tir = dot(refracted, refracted) <= 1e-7f;
    float rp = tir ? 1.0f : clampf(fresnel, 0.02f, 0.98f);
    float choice = randomFloat(state);
    if (choice < rp) {
        rd = normalize(reflect(incoming, micro));
        ro = h.position + micro * 0.001f;
        T *= tir ? 1.0f : fresnel / fmaxf(rp, 1e-5f);
    } else {
        rd = normalize(refracted);
        ro = h.position + rd * 0.001f;
        float grazing = 1.0f - fabsf(dot(view, h.normal));
        float tint = mixf(0.35f, 0.95f, grazing * grazing);
        float w = (1.0f - fresnel) / fmaxf(1.0f - rp, 1e-5f);
        vec3 absorption = materialTransmissionAbsorption(m, h.t, !h.frontFacing && !materialHasVolumeScattering(m));
        T *= materialTransmissionTint(m, tint) * absorption * w;
    }
    prevPdf = 0.0f;
    return true;
}

__device__ __forceinline__ void writeAOV(float4 *plane, uint32_t pix, float4 v) { plane[pix] = v; }

// ----------------------------------------------------------------- shade --
// The ray, throughput and prevPdf a path carries into its subsurface walk
// are stashed in two of its shadow records (free until NEE of the bounce),
// since the volume step before the gate may have moved them while the hit
// surface is rebuilt from the ray in rayO / rayD.
__device__ __forceinline__ void stashWalkRay(const DrPaths &P, uint32_t slot, vec3 ro, vec3 rd, float prevPdf) {
    P.shR[slot * 3u] = make_float4(ro.x, ro.y, ro.z, prevPdf);
    P.shR[slot * 3u + 1u] = make_float4(rd.x, rd.y, rd.z, 0.0f);
}

// RESUME = false: a bounce from the hit in P.hit. RESUME = true: paths from
// resumeQueue continue at continueTransmissiveSurface with the throughput
// and RNG state their failed walk left.
template <bool VOL, bool RESUME>
__global__ void __launch_bounds__(DR_SHADE_BLOCK) DR_SHADE_OCC k_shade(DrSceneDev S, DrLaunch L, DrCamera cam, DrCamera prevCam,
                                                         DrPaths P, DrOutputs O, uint32_t bounce) {
    const uint32_t *queue = RESUME ? P.resumeQueue : P.queue[bounce & 1u];
    uint32_t *queueOut = P.queue[(bounce + 1u) & 1u];
    const uint32_t count = P.counters[RESUME ? DR_CTR_RESUMEQ + bounce : DR_CTR_Q(bounce)];
    const uint32_t first = blockIdx.x * blockDim.x;
    if (first >= count) return;  // whole blocks past the end leave together
    uint32_t cPrimary = 0u, cBounce = 0u;
    // Each queued path was one extension query (k_extend): the launch adds
    // the queue's length once (one atomic per wave on a single word would
    // serialize a few hundred thousand waves on it).
    if (!RESUME && first + threadIdx.x == 0u) atomicAdd(&P.stats[DR_CNT_RAY_QUERIES], (unsigned long long)count);

    // The block's paths are shaded in the order of a key of what they hit
    // (miss, volume, or the shading class of the triangle's material),
    // a counting sort in LDS: after the first bounce the paths of a block
    // hit many materials, and a wave whose lanes share one takes its
    // material's branches once instead of once per kind. Each path's
    // computation is unchanged; only which lane does it.
    uint32_t slot;
    if constexpr (RESUME) {
        if (first + threadIdx.x >= count) return;
        slot = queue[first + threadIdx.x];
    } else {
        __shared__ uint32_t sBins[DR_SHADE_KEYS];
        __shared__ uint32_t sOrder[DR_SHADE_BLOCK];
        const uint32_t t = threadIdx.x;
        const uint32_t n = min(blockDim.x, count - first);
        if (t < DR_SHADE_KEYS) sBins[t] = 0u;
        __syncthreads();
        uint32_t mine = 0u, key = 0u, rank = 0u;
        if (t < n) {
            mine = queue[first + t];
            const uint32_t id = __float_as_uint(P.hit[mine].w);
            key = id == 0xffffffffu ? 0u
                  : (VOL && (id & DR_VOL_HIT)) ? DR_SHADE_KEYS - 1u
                  : 1u + S.matKey[min(S.triangles[id].materialID, S.materialCount - 1u)];
            rank = atomicAdd(&sBins[key], 1u);
        }
        __syncthreads();
        if (t == 0u) {
            uint32_t run = 0u;
            for (uint32_t k = 0; k < DR_SHADE_KEYS; ++k) {
                const uint32_t c = sBins[k];
                sBins[k] = run;
                run += c;
            }
        }
        __syncthreads();
        if (t < n) sOrder[sBins[key] + rank] = mine;
        __syncthreads();
        if (t >= n) return;
        slot = sOrder[t];
    }
    float4 o4 = P.rayO[slot], d4 = P.rayD[slot], t4 = P.thr[slot], l4 = P.rad[slot];
    uint4 meta = P.meta[slot];
    float4 h4 = P.hit[slot];
    vec3 ro = v3(o4.x, o4.y, o4.z), rd = v3(d4.x, d4.y, d4.z);
    const vec3 hitO = ro, hitD = rd;  // the ray of the hit in P.hit
    vec3 T = v3(t4.x, t4.y, t4.z), radiance = v3(l4.x, l4.y, l4.z);
    float outputAlpha = t4.w;
    float prevPdf = o4.w;
    if (RESUME) {
        float4 a = P.shR[slot * 3u], b = P.shR[slot * 3u + 1u];
        ro = v3(a.x, a.y, a.z);
        prevPdf = a.w;
        rd = v3(b.x, b.y, b.z);
    }
    uint32_t state = meta.z;
    uint32_t flags = meta.w & DR_FLAG_PRIMARY;
    const bool hasPrimary = (flags & DR_FLAG_PRIMARY) != 0u;
    const bool lastSample = meta.y == batchBegin(L) + L.samplesInBatch - 1u;
    const uint32_t pix = meta.x;
    const float clampValue = L.sampleRadianceClamp;
    bool alive = false;
    if (!RESUME) {
        if (bounce == 0u) cPrimary++;
        else cBounce++;
    }

    // NEE shadow records are written straight to the slot's three records;
    // neeMask holds which of them are valid.
    uint32_t neeMask = 0u;

    const uint32_t id = __float_as_uint(h4.w);
    if (id == 0xffffffffu) {
        // Miss (:5279-5305).
        if (L.transparentBackground != 0u && !hasPrimary) {
            outputAlpha = 0.0f;
        } else if (L.showsEnvironmentBackground == 0u && !hasPrimary) {
            outputAlpha = 1.0f;
            radiance += vmax(v3(L.backgroundColor[0], L.backgroundColor[1], L.backgroundColor[2]), v3s(0.0f));
        } else {
            float w = 1.0f;
            if (bounce > 0u && prevPdf > 0.0f) w = powerHeuristic(prevPdf, environmentPDF(S, rd));
            vec3 c = T * sampleEnvironment(S, rd) * w;
            radiance += clampSampleContribution(c, clampValue);
        }
    } else {
        float4 hn4 = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
        if (VOL && (id & DR_VOL_HIT)) hn4 = P.hitN[slot];
        Surf h = surfaceAt<VOL>(S, hitO, hitD, h4, hn4);
        Mat m = materialFor(S, h);
        applyNormalMap(S, h, m, hitD);
        applyBaseColorTexture(S, m, h);
        if (!RESUME && m.base.w <= 0.001f) {
            // Pass-through: consumes this bounce (:5318-5321).
            ro = h.position + rd * 0.001f;
            alive = true;
        } else {
            if (!hasPrimary) {
                flags |= DR_FLAG_PRIMARY;
                if (lastSample && L.denoiserEnabled == 0u) {
                    vec3 en = h.normal * 0.5f + v3s(0.5f);
                    float mid = (float)(h.materialID + 1u);
                    float oid = (float)(h.objectID + 1u);
                    writeAOV(O.depth, pix, make_float4(h.t, h.t, h.t, 1.0f));
                    writeAOV(O.normal, pix, make_float4(en.x, en.y, en.z, 1.0f));
                    writeAOV(O.albedo, pix, m.base);
                    writeAOV(O.materialID, pix, make_float4(mid, mid, mid, 1.0f));
                    writeAOV(O.objectID, pix, make_float4(oid, oid, oid, 1.0f));
                    vec2 cur, prv;
                    if (projectToScreen(cam, h.position, L.width, L.height, cur) &&
                        projectToScreen(prevCam, h.position, L.width, L.height, prv)) {
                        writeAOV(O.motion, pix, make_float4(prv.x - cur.x, prv.y - cur.y, 0.0f, 1.0f));
                    } else {
                        writeAOV(O.motion, pix, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
                    }
                    float tid = __uint_as_float(h.primitiveID + 1u);
                    writeAOV(O.triangleID, pix, make_float4(tid, tid, tid, tid));
                }
            }
            bool handled = false;
            if (!RESUME) {
                if (continueTransmissionVolumeRandomWalk(h, m, ro, rd, T, prevPdf, state)) {
                    handled = true;
                    alive = !terminatePathWithRussianRoulette(bounce, T, state);
                }
                if (!handled) {
                    float subsurface = materialSubsurface(m);
                    if (subsurface > 0.0f) {
                        float gate = randomFloat(state);
                        if (gate < subsurface && subsurfaceWalkStarts(m, h)) {
                            // The walk runs in k_sss; hand the path over.
                            stashWalkRay(P, slot, ro, rd, prevPdf);
                            P.thr[slot] = make_float4(T.x, T.y, T.z, outputAlpha);
                            P.meta[slot] = make_uint4(meta.x, meta.y, state, flags);
                            uint32_t pos = waveAppend(true, &P.counters[DR_CTR_SSSQ + bounce]);
                            P.sssQueue[pos] = slot;
                            if (L.statsEnabled) {
                                waveAddStat(P.stats, DR_CNT_PRIMARY, cPrimary);
                                waveAddStat(P.stats, DR_CNT_BOUNCE, cBounce);
                            }
                            return;
                        }
                    }
                }
            }
            if (!handled && continueTransmissiveSurface(h, m, ro, rd, T, prevPdf, state)) {
                handled = true;
                alive = !terminatePathWithRussianRoulette(bounce, T, state);
            }
            if (!handled && maxComponent(xyz(ms<M_EMISSION>(m))) > 0.0f) {
                // Emitter (:5376-5392): add and stop.
                float w = 1.0f;
                if (bounce > 0u && prevPdf > 0.0f) w = powerHeuristic(prevPdf, lightPDFForHit(S, h, rd));
                vec3 c = T * xyz(ms<M_EMISSION>(m)) * w;
                radiance += clampSampleContribution(c, clampValue);
                handled = true;
            }
            if (!handled) {
                vec3 view = -rd;
                // NEE sample generation (sampleDirectLighting, :4407-4585); the
                // visibility term is applied later by k_shadow/k_resolve.
                // The three samples (emissive triangle, two environment
                // directions) are drawn first, in the reference's RNG order,
                // then the bounce direction; the material's pdf and BRDF are
                // evaluated afterwards in one loop over the samples and the
                // bounce direction, so the kernel holds one copy of that code
                // and a wave runs it once per sample for all its lanes.
                vec3 nl[3];         // sample directions
                float npdf[3];      // light / environment pdf
                vec3 nrad[3];       // emission (light); the environment's is looked up in the loop
                float ng[3];        // G (light) or cos (environment)
                float now[3], ndw[3];  // shadow record .w of origin and direction
                bool nvalid[3] = {false, false, false};
                if (S.lightCount > 0u) {
                    float r0 = randomFloat(state);
                    uint32_t index = selectLightIndex(S.lights, S.lightCount, r0);
                    const DrLight &light = S.lights[index];
                    float selectionPDF = lightSelectionPDF(S.lights, index);
                    uint32_t lti = light.triangleIndex;
                    if (lti < S.triangleCount && selectionPDF > 0.0f) {
                        const DrTriangle &lt = S.triangles[lti];
                        const DrMaterial &lm = S.materials[min(light.materialIndex, S.materialCount - 1u)];
                        float area = light.area;
                        if (area > 0.0f) {
                            float ux = randomFloat(state);
                            float uy = randomFloat(state);
                            float su0 = sqrtf(ux);
                            float b0 = 1.0f - su0, b1 = su0 * (1.0f - uy), b2 = su0 * uy;
                            vec3 lp = b0 * xyz(lt.v0) + b1 * xyz(lt.v1) + b2 * xyz(lt.v2);
                            vec3 toLight = lp - h.position;
                            float d2 = dot(toLight, toLight);
                            float dist = sqrtf(d2);
                            vec3 ldir = toLight / dist;
                            float cosS = fmaxf(0.0f, dot(h.normal, ldir));
                            float cosL = fmaxf(0.0f, dot(xyz(light.normal), -ldir));
                            if (cosS > 0.0f && cosL > 0.0f) {
                                nvalid[0] = true;
                                nl[0] = ldir;
                                npdf[0] = selectionPDF * lightSolidAnglePDF(d2, cosL, area);
                                nrad[0] = xyz(lm.s[M_EMISSION]);
                                ng[0] = cosS * cosL * area / fmaxf(d2, 1e-6f);
                                now[0] = dist - 0.002f;
                                ndw[0] = selectionPDF;
                            }
                        }
                    }
                }
#pragma unroll
                for (uint32_t e = 0; e < 2u; ++e) {
                    float px = randomFloat(state);
                    float py = randomFloat(state);
                    float cdf = randomFloat(state);
                    const float coin = S.envAliasSampling ? randomFloat(state) : 0.0f;  // alias mode draws one more
                    float envPDF;
                    vec3 dir = sampleEnvironmentDirection(S, px, py, cdf, coin, envPDF);
                    float cosS = fmaxf(0.0f, dot(h.normal, dir));
                    nvalid[1u + e] = envPDF > 0.0f && cosS > 0.0f;
                    nl[1u + e] = dir;
                    npdf[1u + e] = envPDF;
                    ng[1u + e] = cosS;
                    now[1u + e] = 1.0e20f;
                    ndw[1u + e] = envPDF * 2.0f;
                }
                const vec3 Tnee = T;
                // Lobe sampling (:5423-5513). The clearcoat and specular lobes
                // sample the visible normals of one GGX distribution each;
                // one call serves both with the lobe's roughness and
                // anisotropy.
                float roughness = clampf(ms<M_PARAMS>(m).x, 0.02f, 1.0f);
                float anisotropy = materialSpecularAnisotropy(m);
                vec3 f0 = materialF0(m);
                float clearcoat = materialClearcoat(m);
                float ccr = materialClearcoatRoughness(m);
                float pd, ps, pc;
                materialLobeProbabilities(m, pd, ps, pc);
                float d0 = randomFloat(state);
                float d1 = randomFloat(state);
                vec3 diffuseDir = orientHemisphere(cosineHemisphere(d0, d1), h.normal);
                vec3 next = diffuseDir;
                float lobe = randomFloat(state);
                const bool coat = lobe < pc;
                bool diffuse = !(coat || lobe < pc + ps);
                if (!diffuse) {
                    float a = randomFloat(state);
                    float b = randomFloat(state);
                    vec3 hv = sampleGGXVisibleNormal(a, b, h.normal, h.tangent, h.bitangent, view, coat ? ccr : roughness,
                                                     coat ? 0.0f : anisotropy);
                    next = normalize(reflect(rd, hv));
                    if (dot(next, h.normal) <= 0.0f) {
                        next = diffuseDir;
                        diffuse = true;
                    } else {
                        float vDotH = fmaxf(dot(view, hv), 1e-5f);
                        vec3 fresnel = fresnelSchlickThinFilm(vDotH, coat ? materialClearcoatF0Color(m) : f0, m);
                        vec3 w;
                        float G, p;
                        if (coat) {
                            w = v3s(clearcoat);
                            G = geometrySmithG1GGX(fmaxf(dot(h.normal, next), 1e-5f), ccr);
                            p = pc;
                        } else {
                            w = materialClearcoatAttenuationForDirection(m, h.normal, view, next);
                            G = geometrySmithG1GGXAnisotropic(next, h.normal, h.tangent, h.bitangent, roughness, anisotropy);
                            p = ps;
                        }
                        T *= w * fresnel * G / fmaxf(p, 1e-5f);
                    }
                }
                if (diffuse) T *= materialDiffuseBounceWeight(m, h.normal, view, next) / fmaxf(pd, 1e-5f);
                const vec3 so = h.position + h.normal * 0.001f;
#pragma unroll 1
                for (uint32_t i = 0; i < 4u; ++i) {
                    const vec3 l = i == 0u ? nl[0] : (i == 1u ? nl[1] : (i == 2u ? nl[2] : next));
                    const bool v = i == 0u ? nvalid[0] : (i == 1u ? nvalid[1] : (i == 2u ? nvalid[2] : true));
                    if (!v) continue;
                    const float bsdfPDF = materialBSDFPDF(m, h.normal, h.tangent, h.bitangent, view, l);
                    if (i == 3u) {
                        prevPdf = bsdfPDF;
                        break;
                    }
                    const float lp = i == 0u ? npdf[0] : (i == 1u ? npdf[1] : npdf[2]);
                    const float g = i == 0u ? ng[0] : (i == 1u ? ng[1] : ng[2]);
                    float mis = powerHeuristic(lp, bsdfPDF);
                    vec3 brdf = evaluateMaterialBRDF(m, h.normal, h.tangent, h.bitangent, view, l);
                    vec3 rad = i == 0u ? nrad[0] : sampleEnvironment(S, l);
                    vec3 pre = brdf * rad * g * mis;
                    const uint32_t k = slot * 3u + i;
                    P.shO[k] = make_float4(so.x, so.y, so.z, i == 0u ? now[0] : now[1]);
                    P.shD[k] = make_float4(l.x, l.y, l.z, i == 0u ? ndw[0] : (i == 1u ? ndw[1] : ndw[2]));
                    // .w: the record's pdf denominator (> 0), so k_resolve
                    // reads one float4 per record for both.
                    P.shC[k] = make_float4(pre.x, pre.y, pre.z, i == 0u ? ndw[0] : (i == 1u ? ndw[1] : ndw[2]));
                    neeMask |= 1u << i;
                }
                if (neeMask != 0u) {
                    flags |= DR_FLAG_NEE;
                    P.neeT[slot] = make_float4(Tnee.x, Tnee.y, Tnee.z, 0.0f);
                }
                ro = so;
                rd = next;
                alive = !terminatePathWithRussianRoulette(bounce, T, state);
            }
        }
    }
    alive = alive && (bounce + 1u < L.maxBounces);

    P.rayO[slot] = make_float4(ro.x, ro.y, ro.z, prevPdf);
    P.rayD[slot] = make_float4(rd.x, rd.y, rd.z, 0.0f);
    P.thr[slot] = make_float4(T.x, T.y, T.z, outputAlpha);
    P.rad[slot] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
    P.meta[slot] = make_uint4(meta.x, meta.y, state, flags);

    uint32_t pos = waveAppend(alive, &P.counters[DR_CTR_Q(bounce + 1u)]);
    if (alive) queueOut[pos] = slot;
    // The wave's shadow records take one contiguous run of the shadow queue
    // (one atomic for all three kinds; an atomic per kind and wave on the one
    // counter word was most of this kernel's time): the light samples, then
    // each environment sample's.
    {
        const uint64_t m0 = __ballot(neeMask & 1u), m1 = __ballot(neeMask & 2u), m2 = __ballot(neeMask & 4u);
        const uint32_t c0 = __popcll(m0), c01 = c0 + __popcll(m1), total = c01 + __popcll(m2);
        const uint32_t leader = __ffsll((unsigned long long)__ballot(1)) - 1;
        uint32_t base = 0u;
        if (total && laneId() == leader) base = atomicAdd(&P.counters[DR_CTR_SQ(bounce)], total);
        base = __shfl(base, leader);
        const uint64_t below = (1ull << laneId()) - 1ull;
        const uint32_t at[3] = {base + uint32_t(__popcll(m0 & below)), base + c0 + uint32_t(__popcll(m1 & below)),
                                base + c01 + uint32_t(__popcll(m2 & below))};
#pragma unroll
        for (uint32_t j = 0; j < 3u; ++j) {
            if ((neeMask >> j) & 1u) P.shadowQueue[at[j]] = slot * 3u + j;
            else P.shC[slot * 3u + j].w = 0.0f;
        }
    }
    if (L.statsEnabled) {
        waveAddStat(P.stats, DR_CNT_PRIMARY, cPrimary);
        waveAddStat(P.stats, DR_CNT_BOUNCE, cBounce);
    }
}

// ------------------------------------------------------------------- sss --
// continueSubsurfaceRandomWalk for the paths k_shade queued this bounce.
// Persistent waves take 64 paths at a time; every lane of a wave walks, so
// the nested closest-hit queries run converged. A successful walk exits the
// surface (then RR, :5359-5363); a failed one goes to k_shade<RESUME> with
// its throughput and RNG state.
// ------------------------------------------------ refill-tracer items --
// A work item of the refill tracer: its current ray and the state it carries
// between queries. Shadow rays and subsurface walks share the fields so that
// one kernel can trace both without holding two sets of registers.
struct DrItem {
    uint32_t id;    // shadow record index, or the path slot of a walk
    uint32_t step;  // queries done
    vec3 w;         // shadow: transmittance; walk: path throughput
    float a, b;     // shadow: maxDistance, traveled; walk: anisotropy g
};
// The item's current ray is the tracer's DrRayPre (r.o, r.d).

// shadowOccluded (PathTrace.metal:4196-4274) for one queued NEE ray as an
// item: a chain of closest hits, one query per step (anyHit = false), or one
// any-hit query below maxDistance when no material can let light through
// (anyHit = true; the limit is exclusive: id 0 never wins a tie). The
// transmittance (0 if occluded) goes to shR[k].
// Hits at or beyond maxDistance - traveled end the chain, so each query is
// limited there: exactly (exclusive limit) on the first step, with a
// relative margin afterwards where the reference test
// traveled + t >= maxDistance decides.
__device__ __forceinline__ DrHitRec shadowLimit(const DrItem &it) {
    return it.step == 0u ? DrHitRec{it.a, 0.0f, 0.0f, 0u}
                         : DrHitRec{(it.a - it.b) * 1.000001f + 1e-30f, 0.0f, 0.0f, DR_MISS};
}
__device__ __forceinline__ void shadowBegin(const DrPaths &P, uint32_t k, DrItem &it, DrRayPre &r, DrHitRec &best) {
    float4 o4 = P.shO[k], d4 = P.shD[k];
    it.id = k;
    it.step = 0u;
    it.w = v3s(1.0f);
    it.a = o4.w;
    it.b = 0.0f;
    r = dr_prep(v3(o4.x, o4.y, o4.z), v3(d4.x, d4.y, d4.z));
    best = shadowLimit(it);
}
// Consumes a query's result; true when the chain continues with the next
// query (set up in r / best).
__device__ __forceinline__ bool shadowStep(const DrSceneDev &S, const DrPaths &P, bool anyHit, const DrHitRec &h,
                                           bool hit, DrItem &it, DrRayPre &r, DrHitRec &best, uint32_t &nq) {
    if (anyHit) {
        P.shR[it.id] = hit ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        return false;
    }
    bool occluded = true, next = false;
    vec3 o = r.o;
    const vec3 d = r.d;
    if (!hit || it.b + h.t >= it.a) {
        occluded = false;
    } else {
        // materialForHit of a triangle hit is its material record; the
        // interpolated surface is only needed when light can pass.
        const DrMaterial *mp = &S.materials[min(S.triangles[h.id].materialID, S.materialCount - 1u)];
        const Mat m{mp, mp, 0.0f, mp->s[M_BASE], 0u, {}, {}, {}};
        if (m.base.w <= 0.001f) {
            it.b += h.t;
            o = (o + h.t * d) + d * 0.001f;
            next = true;
        } else {
            float transmission = materialTransmission(m);
            if (transmission > 0.001f) {
                float4 h4 = make_float4(h.t, h.u, h.v, __uint_as_float(h.id));
                Surf sh = surfaceAt<false>(S, o, d, h4, h4);
                bool inside = !sh.frontFacing && !materialThinWalled(m);
                vec3 medium = materialHasVolumeScattering(m) && inside ? materialVolumeTransmittance(m, sh.t)
                                                                       : materialTransmissionAbsorption(m, sh.t, inside);
                it.w = it.w * materialTransmissionTint(m, 0.35f) * medium * transmission;
                if (maxComponent(it.w) >= 0.02f) {
                    it.b += sh.t;
                    o = sh.position + d * 0.001f;
                    next = true;
                }
            }
        }
    }
    if (next && ++it.step < 8u) {
        r = dr_prep(o, d);
        best = shadowLimit(it);
        ++nq;
        return true;
    }
    P.shR[it.id] = occluded ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : make_float4(it.w.x, it.w.y, it.w.z, 1.0f);
    return false;
}

// ---------------------------------------------------------------- extend --
// STATS: the counting variant (node, triangle and query counters); the
// other one keeps no counters at all.
//
// On triangles the launch also traces the previous bounce's remaining shadow
// rays (those k_shade<RESUME> queued after k_sss), claimed after the
// extension rays: they are short chains whose latency would otherwise set
// the length of a launch of their own. Their k_resolve runs after this
// kernel and before the next k_shade, so radiance is summed in the same
// order as before.
template <bool VOL, bool STATS>
__global__ void __launch_bounds__(DR_TRACE_BLOCK) DR_EXTEND_OCC k_extend(DrSceneDev S, DrLaunch L, DrCamera cam, DrPaths P,
                                                                         uint32_t bounce) {
    __shared__ uint32_t ldsStack[DR_LDS_STACK * DR_TRACE_BLOCK];
    const uint32_t *queue = P.queue[bounce & 1u];
    const uint32_t count = P.counters[DR_CTR_Q(bounce)];
    uint32_t *steal = P.claims + DR_CLAIM_AT(DR_CLAIM_EXTEND, bounce);
    uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    DrStack stk{dr_lds_ptr(&ldsStack[threadIdx.x]), DR_TRACE_BLOCK, P.overflow + gl, P.overflowStride};
    uint32_t nv = 0, nt = 0;
    uint32_t shQ = 0u, shNodes = 0u, shTris = 0u;
    DrSdfCounters sc{};
    if constexpr (!VOL) {
        const uint32_t shBegin = bounce ? P.counters[DR_CTR_SHBEGIN + bounce - 1u] : 0u;
        const uint32_t nShadow = bounce ? P.counters[DR_CTR_SQ(bounce - 1u)] - shBegin : 0u;
        const uint32_t *shQueue = P.shadowQueue + shBegin;
        const bool anyShadow = S.opaqueOnly != 0u;
        uint32_t slot = 0, markV = 0u, markT = 0u;
        bool shadow = false, laneAny = false;
        DrItem it{};
        dr_trace_queue<false>(
            S, count + nShadow, steal, stk, nv, nt,
            [&](uint32_t qi, DrRayPre &r, DrHitRec &best) {
                shadow = qi >= count;
                if (!shadow) {
                    if (bounce == 0u) {
                        slot = qi;
                        vec3 o, d;
                        generatePath(L, cam, P, slot, o, d);
                        r = dr_prep(o, d);
                    } else {
                        slot = queue[qi];
                        float4 o4 = P.rayO[slot], d4 = P.rayD[slot];
                        r = dr_prep(v3(o4.x, o4.y, o4.z), v3(d4.x, d4.y, d4.z));
                    }
                    best = DrHitRec{__builtin_inff(), 0.0f, 0.0f, DR_MISS};
                    laneAny = false;
                } else {
                    shadowBegin(P, shQueue[qi - count], it, r, best);
                    ++shQ;
                    laneAny = anyShadow;
                }
                markV = nv;
                markT = nt;
            },
            [&](const DrHitRec &h, bool hit, DrRayPre &r, DrHitRec &best) {
                if (!shadow) {
                    P.hit[slot] = make_float4(h.t, h.u, h.v, __uint_as_float(h.id));
                    return false;
                }
                shNodes += nv - markV;
                shTris += nt - markT;
                markV = nv;
                markT = nt;
                return shadowStep(S, P, anyShadow, h, hit, it, r, best, shQ);
            },
            &laneAny);
        nv -= shNodes;
        nt -= shTris;
    } else {
        DrClaim cl(steal, count);
        while (cl.more()) {
            uint32_t base, j, end;
            cl.claim(64u, base, j, end);
            const uint32_t c = base + laneId();
            if (c < end) {
                const uint32_t qi = DrClaim::entry(c, j);
                uint32_t slot;
                vec3 o, d;
                if (bounce == 0u) {
                    slot = qi;
                    generatePath(L, cam, P, slot, o, d);
                } else {
                    slot = queue[qi];
                    float4 o4 = P.rayO[slot], d4 = P.rayD[slot];
                    o = v3(o4.x, o4.y, o4.z);
                    d = v3(d4.x, d4.y, d4.z);
                }
                float4 h, hn;
                sceneClosest<true>(S, o, d, stk, nv, nt, sc, h, hn);
                P.hit[slot] = h;
                if (__float_as_uint(h.w) != DR_MISS && (__float_as_uint(h.w) & DR_VOL_HIT)) P.hitN[slot] = hn;
            }
        }
    }
    waveAddTotal(P.stats, DR_CNT_RAY_QUERIES, shQ);  // the extension queries are counted by k_shade
    if constexpr (STATS) {
        waveAddStat(P.stats, DR_CNT_NODES, nv);
        waveAddStat(P.stats, DR_CNT_TRIS, nt);
        waveAddStat(P.stats, DR_CNT_SHADOW, shQ);
        waveAddStat(P.stats, DR_CNT_SHADOW_NODES, shNodes);
        waveAddStat(P.stats, DR_CNT_SHADOW_TRIS, shTris);
        waveAddStat(P.stats, DR_CNT_EXT_SHADOW, shQ);
        waveAddStat(P.stats, DR_CNT_EXT_SHADOW_NODES, shNodes);
        waveAddStat(P.stats, DR_CNT_EXT_SHADOW_TRIS, shTris);
        if constexpr (VOL) addSdfStats(P.stats, sc);
    }
}

// continueSubsurfaceRandomWalk (PathTrace.metal:3525-3621) on triangles as
// an item: one closest-hit query per walk segment. The path's other state
// (output alpha, meta words) stays in memory until the walk ends.
struct DrWalk {
    vec3 sigmaT, albedo;
    uint32_t state;
};
__device__ __forceinline__ void walkBegin(const DrSceneDev &S, const DrPaths &P, uint32_t slot, DrItem &it, DrWalk &wk,
                                          DrRayPre &r, DrHitRec &best) {
    float4 o4 = P.rayO[slot], d4 = P.rayD[slot], t4 = P.thr[slot];
    float4 h4 = P.hit[slot];
    vec3 hitO = v3(o4.x, o4.y, o4.z), hitD = v3(d4.x, d4.y, d4.z);
    Surf h = surfaceAt<false>(S, hitO, hitD, h4, h4);
    Mat m = materialFor(S, h);
    applyNormalMap(S, h, m, hitD);
    applyBaseColorTexture(S, m, h);
    it.id = slot;
    it.w = v3(t4.x, t4.y, t4.z);
    wk.state = P.meta[slot].z;
    // Entry (:3534-3548): cosine-weighted into the surface.
    float e0 = randomFloat(wk.state);
    float e1 = randomFloat(wk.state);
    vec3 entryDir = orientHemisphere(cosineHemisphere(e0, e1), -h.normal);
    wk.sigmaT = v3s(1.0f) / materialSubsurfaceRadius(m);
    wk.albedo = materialSubsurfaceColor(m);
    it.a = materialSubsurfaceAnisotropy(m);
    it.step = 0u;
    r = dr_prep(h.position + entryDir * 0.001f, entryDir);
    best = DrHitRec{__builtin_inff(), 0.0f, 0.0f, DR_MISS};
}
// Consumes a segment's query; true when the walk continues (next segment
// set up in r / best). A finished walk writes the path back and queues it:
// for the next bounce when it exits alive, for k_shade<RESUME> when it fails.
__device__ __forceinline__ bool walkStep(const DrSceneDev &S, const DrLaunch &L, const DrPaths &P, uint32_t bounce,
                                         const DrHitRec &hr, bool hit, DrItem &it, DrWalk &wk, DrRayPre &r,
                                         DrHitRec &best, uint32_t &cBounce, uint32_t &cSss) {
    bool exited = false, more = false;
    vec3 ro = r.o, rd = r.d;
    float prevPdf = 0.0f;
    if (hit) {
        float f0 = randomFloat(wk.state);
        float f1 = randomFloat(wk.state);
        float ff = sampleFreeFlight(wk.sigmaT, f0, f1);
        if (ff < hr.t) {  // a collision inside: the boundary surface is not needed
            ro += rd * ff;
            float h0 = randomFloat(wk.state);
            float h1 = randomFloat(wk.state);
            rd = sampleHenyeyGreenstein(h0, h1, rd, it.a);
            vec3 tr = vexp((-wk.sigmaT) * ff);
            it.w *= wk.albedo * wk.sigmaT * tr / collisionPDF(wk.sigmaT, ff);
            more = !(maxComponent(it.w) <= 1e-4f || ++it.step >= 8u);
        } else {
            float4 h4 = make_float4(hr.t, hr.u, hr.v, __uint_as_float(hr.id));
            Surf b = surfaceAt<false>(S, ro, rd, h4, h4);
            vec3 exitN = normalize(b.frontFacing ? b.normal : -b.normal);
            float x0 = randomFloat(wk.state);
            float x1 = randomFloat(wk.state);
            rd = orientHemisphere(cosineHemisphere(x0, x1), exitN);
            ro = b.position + exitN * 0.001f;
            vec3 tr = vexp((-wk.sigmaT) * b.t);
            it.w *= tr / boundaryPDF(wk.sigmaT, b.t);
            prevPdf = fmaxf(dot(exitN, rd), 0.0f) * 0.31830988618f;
            exited = true;
        }
    }
    if (more) {
        cBounce++;
        cSss++;
        r = dr_prep(ro, rd);
        best = DrHitRec{__builtin_inff(), 0.0f, 0.0f, DR_MISS};
        return true;
    }
    const uint32_t slot = it.id;
    bool alive = false;
    if (exited) {
        alive = !terminatePathWithRussianRoulette(bounce, it.w, wk.state) && (bounce + 1u < L.maxBounces);
        P.rayO[slot] = make_float4(ro.x, ro.y, ro.z, prevPdf);
        P.rayD[slot] = make_float4(rd.x, rd.y, rd.z, 0.0f);
    }
    const float alpha = P.thr[slot].w;
    const uint4 meta = P.meta[slot];
    P.thr[slot] = make_float4(it.w.x, it.w.y, it.w.z, alpha);
    P.meta[slot] = make_uint4(meta.x, meta.y, wk.state, meta.w);
    uint32_t pos = waveAppend(alive, &P.counters[DR_CTR_Q(bounce + 1u)]);
    if (alive) P.queue[(bounce + 1u) & 1u][pos] = slot;
    const bool resume = !exited;
    uint32_t rpos = waveAppend(resume, &P.counters[DR_CTR_RESUMEQ + bounce]);
    if (resume) P.resumeQueue[rpos] = slot;
    return false;
}

// ------------------------------------------------------------ subsurface --
// The bounce's subsurface walks. On triangles the kernel also traces the
// bounce's shadow rays queued so far (by k_shade): walks are claimed first
// and are chains of up to 8 dependent queries, so the shadow rays fill the
// lanes around them instead of waiting for the longest walk; k_shadow then
// traces only the shadow rays k_shade<RESUME> adds after the failed walks.
// Results are the same as tracing them in separate kernels: each item
// writes only its own records.
template <bool VOL, bool STATS>
__global__ void __launch_bounds__(DR_TRACE_BLOCK) DR_SSS_OCC k_sss(DrSceneDev S, DrLaunch L, DrPaths P, uint32_t bounce) {
    __shared__ uint32_t ldsStack[DR_LDS_STACK * DR_TRACE_BLOCK];
    const uint32_t count = P.counters[DR_CTR_SSSQ + bounce];
    uint32_t *claim = P.claims + DR_CLAIM_AT(DR_CLAIM_SSS, bounce);
    uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    DrStack stk{dr_lds_ptr(&ldsStack[threadIdx.x]), DR_TRACE_BLOCK, P.overflow + gl, P.overflowStride};
    uint32_t cBounce = 0u, cNodes = 0u, cTris = 0u, cSss = 0u;
    uint32_t shQ = 0u, shNodes = 0u, shTris = 0u;
    DrSdfCounters sc{};
    if constexpr (!VOL) {
        const uint32_t nShadow = P.counters[DR_CTR_SQ(bounce)];
        if (blockIdx.x == 0 && threadIdx.x == 0) P.counters[DR_CTR_SHBEGIN + bounce] = nShadow;
        const bool anyShadow = S.opaqueOnly != 0u;
        DrItem it{};
        DrWalk wk{};
        bool walk = false, laneAny = false;
        uint32_t markV = 0u, markT = 0u;
        dr_trace_queue<false, DR_SSS_STEPWISE>(
            S, count + nShadow, claim, stk, cNodes, cTris,
            [&](uint32_t qi, DrRayPre &r, DrHitRec &best) {
                walk = qi < count;
                if (walk) {
                    walkBegin(S, P, P.sssQueue[qi], it, wk, r, best);
                    cBounce++;
                    cSss++;
                    laneAny = false;
                } else {
                    shadowBegin(P, P.shadowQueue[qi - count], it, r, best);
                    ++shQ;
                    laneAny = anyShadow;
                }
                markV = cNodes;
                markT = cTris;
            },
            [&](const DrHitRec &h, bool hit, DrRayPre &r, DrHitRec &best) {
                if (!walk) {
                    shNodes += cNodes - markV;
                    shTris += cTris - markT;
                }
                markV = cNodes;
                markT = cTris;
                return walk ? walkStep(S, L, P, bounce, h, hit, it, wk, r, best, cBounce, cSss)
                            : shadowStep(S, P, anyShadow, h, hit, it, r, best, shQ);
            },
            &laneAny);
        cNodes -= shNodes;
        cTris -= shTris;
    }
    for (DrClaim cl(claim, count); VOL && cl.more();) {
        uint32_t base, j, end;
        cl.claim(64u, base, j, end);
        const uint32_t c = base + laneId();
        const bool have = c < end;
        uint32_t slot = have ? P.sssQueue[DrClaim::entry(c, j)] : 0u;
        bool alive = false, resume = false;
        if (have) {
            float4 o4 = P.rayO[slot], d4 = P.rayD[slot], t4 = P.thr[slot];
            uint4 meta = P.meta[slot];
            float4 h4 = P.hit[slot];
            vec3 hitO = v3(o4.x, o4.y, o4.z), hitD = v3(d4.x, d4.y, d4.z);
            float4 a = P.shR[slot * 3u], b = P.shR[slot * 3u + 1u];
            vec3 ro = v3(a.x, a.y, a.z), rd = v3(b.x, b.y, b.z);
            float prevPdf = a.w;
            vec3 T = v3(t4.x, t4.y, t4.z);
            uint32_t state = meta.z;
            float4 hn4 = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
            if (VOL && (__float_as_uint(h4.w) & DR_VOL_HIT)) hn4 = P.hitN[slot];
            Surf h = surfaceAt<VOL>(S, hitO, hitD, h4, hn4);
            Mat m = materialFor(S, h);
            applyNormalMap(S, h, m, hitD);
            applyBaseColorTexture(S, m, h);
            if (continueSubsurfaceRandomWalk<VOL>(S, h, m, ro, rd, T, prevPdf, state, stk, cBounce, cNodes, cTris,
                                                  cSss, sc)) {
                alive = !terminatePathWithRussianRoulette(bounce, T, state) && (bounce + 1u < L.maxBounces);
                P.rayO[slot] = make_float4(ro.x, ro.y, ro.z, prevPdf);
                P.rayD[slot] = make_float4(rd.x, rd.y, rd.z, 0.0f);
            } else {
                resume = true;
            }
            P.thr[slot] = make_float4(T.x, T.y, T.z, t4.w);
            P.meta[slot] = make_uint4(meta.x, meta.y, state, meta.w);
        }
        uint32_t pos = waveAppend(alive, &P.counters[DR_CTR_Q(bounce + 1u)]);
        if (alive) P.queue[(bounce + 1u) & 1u][pos] = slot;
        uint32_t rpos = waveAppend(resume, &P.counters[DR_CTR_RESUMEQ + bounce]);
        if (resume) P.resumeQueue[rpos] = slot;
    }
    waveAddTotal(P.stats, DR_CNT_RAY_QUERIES, cSss + shQ);
    if constexpr (STATS) {
        waveAddStat(P.stats, DR_CNT_BOUNCE, cBounce);
        waveAddStat(P.stats, DR_CNT_SSS_NODES, cNodes);
        waveAddStat(P.stats, DR_CNT_SSS_TRIS, cTris);
        waveAddStat(P.stats, DR_CNT_SSS, cSss);
        waveAddStat(P.stats, DR_CNT_SHADOW, shQ);
        waveAddStat(P.stats, DR_CNT_SHADOW_NODES, shNodes);
        waveAddStat(P.stats, DR_CNT_SHADOW_TRIS, shTris);
        waveAddStat(P.stats, DR_CNT_SSS_SHADOW, shQ);
        waveAddStat(P.stats, DR_CNT_SSS_SHADOW_NODES, shNodes);
        waveAddStat(P.stats, DR_CNT_SSS_SHADOW_TRIS, shTris);
        if constexpr (VOL) addSdfStats(P.stats, sc);
    }
}

// shadowOccluded (PathTrace.metal:4196-4274) for one ray: up to 8 closest
// hits, stepping through nearly transparent surfaces and attenuating
// through transmissive ones; tr returns the transmittance. The triangle-only
// form limits each query to the remaining distance.
// OCCLUDE (volume scenes whose every hit blocks light, S.sdf.opaqueShadows):
// the first query only has to decide whether some hit lies closer than
// maxDistance, which a triangle below it or a sparse crossing bracketed
// below it settles without refining the hit; the counting kernels keep the
// full queries so their counters stay the reference's.
template <bool VOL, bool OCCLUDE = false>
__device__ __forceinline__ bool shadowChain(const DrSceneDev &S, vec3 o, vec3 d, float maxDistance, vec3 &tr,
                                            DrStack &stk, uint32_t &nq, uint32_t &nv, uint32_t &nt,
                                            DrSdfCounters &sc) {
    tr = v3s(1.0f);
    bool occluded;
    occluded = true;
    float traveled = 0.0f;
    if constexpr (VOL && OCCLUDE) {
        if (S.sdf.opaqueShadows) {
            ++nq;
            float4 h4, hn4;
            bool certain = false;
            const bool hit = sceneClosest<true>(S, o, d, stk, nv, nt, sc, h4, hn4, maxDistance, &certain);
            return certain || (hit && h4.x < maxDistance);
        }
    }
    for (uint32_t step = 0; step < 8u; ++step) {
        ++nq;
        float4 h4, hn4;
        bool hit;
        if constexpr (!VOL) {
            DrRayPre r = dr_prep(o, d);
            // Hits at or beyond maxDistance - traveled end the chain,
            // so the search is limited there: exactly (exclusive
            // limit) on the first step, with a relative margin
            // afterwards where the reference test
            // traveled + t >= maxDistance decides.
            float limit = step == 0u ? maxDistance : (maxDistance - traveled) * 1.000001f + 1e-30f;
            DrHitRec h{limit, 0.0f, 0.0f, step == 0u ? 0u : DR_MISS};
            hit = dr_traverse<false>(S, r, h, stk, nv, nt);
            h4 = make_float4(h.t, h.u, h.v, __uint_as_float(h.id));
        } else {
            // The volume search is bounded by the closest triangle,
            // wherever it lies, so the triangle query is unlimited.
            hit = sceneClosest<true>(S, o, d, stk, nv, nt, sc, h4, hn4);
        }
        if (!hit || traveled + h4.x >= maxDistance) {
            occluded = false;
            break;
        }
        Surf sh = surfaceAt<VOL>(S, o, d, h4, hn4);
        Mat m = materialFor(S, sh);
        if (m.base.w <= 0.001f) {
            traveled += sh.t;
            o = sh.position + d * 0.001f;
            continue;
        }
        float transmission = materialTransmission(m);
        if (transmission <= 0.001f) break;
        bool inside = !sh.frontFacing && !materialThinWalled(m);
        vec3 medium = materialHasVolumeScattering(m) && inside ? materialVolumeTransmittance(m, sh.t)
                                                               : materialTransmissionAbsorption(m, sh.t, inside);
        tr = tr * materialTransmissionTint(m, 0.35f) * medium * transmission;
        if (maxComponent(tr) < 0.02f) break;
        traveled += sh.t;
        o = sh.position + d * 0.001f;
    }
    return occluded;
}

// ---------------------------------------------------------------- shadow --
// shadowOccluded (PathTrace.metal:4196-4274) for the bounce's NEE rays not
// traced by k_sss (on triangles: those k_shade<RESUME> queued; with volumes:
// all of them); writes each ray's transmittance (0 if occluded) to shR[k].
template <bool VOL, bool STATS>
__global__ void __launch_bounds__(DR_TRACE_BLOCK) DR_SHADOW_OCC k_shadow(DrSceneDev S, DrPaths P, uint32_t bounce) {
    __shared__ uint32_t ldsStack[DR_LDS_STACK * DR_TRACE_BLOCK];
    const uint32_t begin = P.counters[DR_CTR_SHBEGIN + bounce];
    const uint32_t count = P.counters[DR_CTR_SQ(bounce)] - begin;
    const uint32_t *queue = P.shadowQueue + begin;
    uint32_t *steal = P.claims + DR_CLAIM_AT(DR_CLAIM_SHADOW, bounce);
    uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    DrStack stk{dr_lds_ptr(&ldsStack[threadIdx.x]), DR_TRACE_BLOCK, P.overflow + gl, P.overflowStride};
    uint32_t nv = 0, nt = 0, nq = 0;
    DrSdfCounters sc{};
    if constexpr (!VOL) {
        DrItem it{};
        auto fetch = [&](uint32_t qi, DrRayPre &r, DrHitRec &best) {
            shadowBegin(P, queue[qi], it, r, best);
            ++nq;
        };
        if (S.opaqueOnly) {
            dr_trace_queue<true, true>(S, count, steal, stk, nv, nt, fetch,
                                 [&](const DrHitRec &h, bool hit, DrRayPre &r, DrHitRec &best) {
                                     return shadowStep(S, P, true, h, hit, it, r, best, nq);
                                 });
        } else {
            // A lane continues its chain without waiting for the wave.
            dr_trace_queue<false, true>(S, count, steal, stk, nv, nt, fetch,
                                  [&](const DrHitRec &h, bool hit, DrRayPre &r, DrHitRec &best) {
                                      return shadowStep(S, P, false, h, hit, it, r, best, nq);
                                  });
        }
    }
    for (DrClaim cl(steal, count); VOL && cl.more();) {
        uint32_t base, j, end;
        cl.claim(64u, base, j, end);
        const uint32_t c = base + laneId();
        if (c >= end) continue;
        uint32_t k = queue[DrClaim::entry(c, j)];
        float4 o4 = P.shO[k], d4 = P.shD[k];
        vec3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
        float maxDistance = o4.w;
        vec3 tr;
        const bool occluded = shadowChain<VOL, !STATS>(S, o, d, maxDistance, tr, stk, nq, nv, nt, sc);
        P.shR[k] = occluded ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : make_float4(tr.x, tr.y, tr.z, 1.0f);
    }
    waveAddTotal(P.stats, DR_CNT_RAY_QUERIES, nq);
    if constexpr (STATS) {
        waveAddStat(P.stats, DR_CNT_SHADOW, nq);
        waveAddStat(P.stats, DR_CNT_SHADOW_NODES, nv);
        waveAddStat(P.stats, DR_CNT_SHADOW_TRIS, nt);
        if constexpr (VOL) addSdfStats(P.stats, sc, DR_CNT_SDF_SHADOW_BASE);
    }
}

// ------------------------------------------------------------- preview --
// The kernels of the preview and interactive qualities
// (DenrimRenderer.primaryPipeline, DenrimRenderer.swift:197-215): one pixel
// per thread, the pixel-centre ray, the accumulation and the AOVs written
// (not averaged) each sample. flatPreviewKernel (PathTrace.metal:4728-4868):
// base colour under a fixed key light plus emission; interactiveMaterialKernel
// (:4870-5165): direct light from up to 12 emitters (centroid, shadow
// chain), a six-lookup diffuse and five-lookup specular environment,
// contact occlusion from four short closest-hit probes, and the material
// terms of the reference.
__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

__device__ __forceinline__ vec3 interactiveDiffuseEnvironment(const DrSceneDev &S, vec3 n) {  // :3805-3823
    const vec3 up = sampleEnvironment(S, v3(0.0f, 1.0f, 0.0f));
    const vec3 down = sampleEnvironment(S, v3(0.0f, -1.0f, 0.0f));
    const vec3 side = (sampleEnvironment(S, v3(1.0f, 0.0f, 0.0f)) + sampleEnvironment(S, v3(-1.0f, 0.0f, 0.0f)) +
                       sampleEnvironment(S, v3(0.0f, 0.0f, 1.0f)) + sampleEnvironment(S, v3(0.0f, 0.0f, -1.0f))) *
                      0.25f;
    const float upW = smoothstepf(-0.15f, 0.85f, n.y);
    const float downW = smoothstepf(0.15f, -0.85f, n.y);
    const float sideW = fmaxf(0.0f, 1.0f - upW - downW);
    return up * upW + down * downW + side * sideW;
}

__device__ __forceinline__ vec3 interactiveHemisphereDirection(vec3 d, vec3 n) {  // :3825-3831
    const float nd = dot(n, d);
    if (nd < 0.0f) d = d - n * (2.0f * nd);
    return normalize(d);
}

__device__ __forceinline__ vec3 interactiveSpecularEnvironment(const DrSceneDev &S, vec3 r, vec3 n,
                                                               float roughness) {  // :3833-3880
    r = interactiveHemisphereDirection(normalize(r), n);
    roughness = clampf(roughness, 0.0f, 1.0f);
    const float spread = roughness * roughness;
    const vec3 helper = fabsf(r.y) < 0.999f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    const vec3 t = normalize(cross(helper, r));
    const vec3 b = cross(r, t);
    vec3 cone = sampleEnvironment(S, r) * 0.42f;
    cone += sampleEnvironment(S, interactiveHemisphereDirection(r + t * spread * 0.95f + n * spread * 0.18f, n)) * 0.145f;
    cone += sampleEnvironment(S, interactiveHemisphereDirection(r - t * spread * 0.95f + n * spread * 0.18f, n)) * 0.145f;
    cone += sampleEnvironment(S, interactiveHemisphereDirection(r + b * spread * 0.95f + n * spread * 0.18f, n)) * 0.145f;
    cone += sampleEnvironment(S, interactiveHemisphereDirection(r - b * spread * 0.95f + n * spread * 0.18f, n)) * 0.145f;
    return mixv(cone, interactiveDiffuseEnvironment(S, n), smoothstepf(0.62f, 0.95f, roughness));
}

template <bool VOL>
__device__ __forceinline__ float interactiveContactOcclusion(const DrSceneDev &S, const Surf &h, const Mat &m, vec3 n,
                                                             DrStack &stk, uint32_t &nq, uint32_t &nv, uint32_t &nt,
                                                             DrSdfCounters &sc) {  // :3882-3944
    const Frame f = makeTangentFrame(n, h.tangent, h.bitangent);
    const vec3 local[4] = {v3(0.000f, 0.000f, 1.000f), v3(0.780f, 0.180f, 0.600f), v3(-0.420f, 0.720f, 0.552f),
                           v3(-0.620f, -0.520f, 0.587f)};
    const float radius = 0.34f;
    float occlusion = 0.0f;
    for (int k = 0; k < 4; ++k) {
        const vec3 l = normalize(local[k]);
        const vec3 dir = normalize(l.x * f.t + l.y * f.b + l.z * n);
        const vec3 o = h.position + n * 0.003f;
        float4 hr, hn;
        ++nq;
        if (!sceneClosest<VOL>(S, o, dir, stk, nv, nt, sc, hr, hn) || hr.x >= radius) continue;
        const Surf oh = surfaceAt<VOL>(S, o, dir, hr, hn);
        const Mat om = materialFor(S, oh);
        const float blocker = clampf(om.base.w, 0.0f, 1.0f) * (1.0f - materialTransmission(om) * 0.72f);
        occlusion += blocker * (1.0f - smoothstepf(0.0f, radius, oh.t));
    }
    const float strength = 0.72f + (0.36f - 0.72f) * materialTransmission(m);
    return 1.0f - clampf((occlusion / 4.0f) * strength, 0.0f, 0.72f);
}

template <bool VOL, bool INTERACTIVE>
__global__ void __launch_bounds__(DR_SHADE_BLOCK) k_preview(DrSceneDev S, DrLaunch L, DrCamera cam, DrCamera prevCam,
                                                            DrPaths P, DrOutputs O) {
    __shared__ uint32_t ldsStack[DR_LDS_STACK * DR_SHADE_BLOCK];
    const uint32_t i = blockIdx.x * DR_SHADE_BLOCK + threadIdx.x;
    DrStack stk{dr_lds_ptr(&ldsStack[threadIdx.x]), DR_SHADE_BLOCK, P.overflow + i, P.overflowStride};
    if (i >= L.tileWidth * L.tileHeight) return;
    const uint32_t gx = L.tileX + i % L.tileWidth, gy = L.tileY + i / L.tileWidth;
    if (gx >= L.width || gy >= L.height) return;
    vec3 ro, rd;
    makeCameraRay(cam, L.width, L.height, gx, gy, 0.5f, 0.5f, 0.0f, 0.0f, ro, rd);
    uint32_t nq = 0u, nv = 0u, nt = 0u;
    DrSdfCounters sc{};
    float4 hr, hn;
    const bool hit = sceneClosest<VOL>(S, ro, rd, stk, nv, nt, sc, hr, hn);
    const size_t pix = size_t(gy) * L.width + gx;
    if (L.statsEnabled) {
        atomicAdd(&P.stats[DR_CNT_PRIMARY], 1ull);
        atomicAdd(&P.stats[DR_CNT_NODES], (unsigned long long)nv);
        atomicAdd(&P.stats[DR_CNT_TRIS], (unsigned long long)nt);
        if constexpr (VOL) addSdfStats(P.stats, sc);  // misses count their volume traversal too
    }
    nv = nt = 0u;
    sc = DrSdfCounters{};
    if (!hit) {
        vec3 color = vmax(v3(L.backgroundColor[0], L.backgroundColor[1], L.backgroundColor[2]), v3s(0.0f));
        if (!L.transparentBackground && L.showsEnvironmentBackground) color = sampleEnvironment(S, rd);
        O.accumulation[pix] = make_float4(color.x, color.y, color.z, L.transparentBackground ? 0.0f : 1.0f);
        O.depth[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        O.normal[pix] = make_float4(0.5f, 0.5f, 0.5f, 0.0f);
        O.albedo[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        O.materialID[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        O.objectID[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        O.motion[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        O.triangleID[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    Surf h = surfaceAt<VOL>(S, ro, rd, hr, hn);
    Mat m = materialFor(S, h);
    applyNormalMap(S, h, m, rd);
    applyBaseColorTexture(S, m, h);
    const vec3 n = normalize(h.normal);
    const vec3 view = normalize(-rd);
    const vec3 baseColor = vmax(xyz(m.base), v3s(0.0f));
    vec3 color;
    float alpha;
    float4 albedo;
    if constexpr (!INTERACTIVE) {  // :4838-4845
        const vec3 key = normalize(v3(0.35f, 0.72f, 0.58f));
        const float lambert = fmaxf(dot(n, key), 0.0f);
        const float facing = fmaxf(dot(n, view), 0.0f);
        const float rim = powf(1.0f - facing, 2.0f) * 0.08f;
        const float shade = 0.32f + 0.68f * lambert + rim;
        color = baseColor * shade + vmax(xyz(ms<M_EMISSION>(m)), v3s(0.0f));
        alpha = L.transparentBackground ? clampf(m.base.w, 0.0f, 1.0f) : 1.0f;
        albedo = m.base;
    } else {  // :4967-5140
        const float roughness = clampf(ms<M_PARAMS>(m).x, 0.02f, 1.0f);
        const float metallic = clampf(ms<M_PARAMS>(m).y, 0.0f, 1.0f);
        const float opacity = clampf(m.base.w, 0.0f, 1.0f);
        const float specular = clampf(ms<M_PARAMS2>(m).x, 0.0f, 1.0f);
        color = vmax(xyz(ms<M_EMISSION>(m)), v3s(0.0f));
        const uint32_t nl = min(S.lightCount, 12u);
        for (uint32_t li = 0; li < nl; ++li) {
            const DrLight light = S.lights[li];
            if (light.triangleIndex >= S.triangleCount || light.materialIndex >= S.materialCount) continue;
            const DrTriangle &lt = S.triangles[light.triangleIndex];
            const vec3 emission = vmax(xyz(S.materials[light.materialIndex].s[M_EMISSION]), v3s(0.0f));
            if (maxComponent(emission) <= 0.0f || light.area <= 0.0f) continue;
            const vec3 lp = (xyz(lt.v0) + xyz(lt.v1) + xyz(lt.v2)) / 3.0f;
            const vec3 toLight = lp - h.position;
            const float dist2 = fmaxf(dot(toLight, toLight), 1e-5f);
            const float dist = sqrtf(dist2);
            const vec3 ldir = toLight / dist;
            const float cosS = fmaxf(dot(n, ldir), 0.0f);
            const float cosL = fmaxf(dot(xyz(light.normal), -ldir), 0.0f);
            if (cosS <= 0.0f || cosL <= 0.0f) continue;
            vec3 tr;
            if (shadowChain<VOL>(S, h.position + n * 0.001f, ldir, dist - 0.002f, tr, stk, nq, nv, nt, sc)) continue;
            vec3 brdf = evaluateMaterialBRDF(m, n, h.tangent, h.bitangent, view, ldir);
            const float ccGloss = materialClearcoat(m) * powf(1.0f - materialClearcoatRoughness(m), 2.0f);
            const float gloss =
                clampf(metallic + ccGloss + (1.0f - metallic) * powf(1.0f - roughness, 4.0f) * specular, 0.0f, 1.0f);
            const vec3 diffuseOnly =
                (1.0f - metallic) * baseColor * vmax(v3s(0.0f), v3s(1.0f) - materialF0(m)) * 0.31830988618f;
            brdf = mixv(diffuseOnly, brdf, gloss);
            color += brdf * emission * (cosS * cosL * light.area / dist2) * tr;
        }
        const vec3 envDiffuse = interactiveDiffuseEnvironment(S, n);
        const vec3 refl = normalize(reflect(-view, n));
        const vec3 baseRefl = interactiveSpecularEnvironment(S, refl, n, roughness);
        const float ccRough = materialClearcoatRoughness(m);
        const vec3 ccRefl = interactiveSpecularEnvironment(S, refl, n, ccRough);
        const float nDotV = fmaxf(dot(n, view), 0.0f);
        const vec3 fresnel = fresnelSchlickThinFilm(nDotV, materialF0(m), m);
        const float glossyW = powf(1.0f - roughness, 4.0f);
        const float metalW = metallic * powf(1.0f - roughness, 2.0f);
        const float ccW = materialClearcoat(m) * powf(1.0f - ccRough, 3.0f);
        const float dielW = (1.0f - metallic) * glossyW * specular;
        const vec3 roughFresnel = mixv(fresnel, materialF0(m), smoothstepf(0.45f, 0.95f, roughness));
        const vec3 ccFresnel = fresnelSchlickThinFilm(nDotV, materialClearcoatF0Color(m), m);
        const float horizon = 0.58f + 0.42f * clampf(n.y * 0.5f + 0.5f, 0.0f, 1.0f);
        const float contact = interactiveContactOcclusion<VOL>(S, h, m, n, stk, nq, nv, nt, sc);
        const float ao = horizon * contact;
        color += baseColor * envDiffuse * (0.58f * (1.0f - metallic) * ao);
        color += baseRefl * roughFresnel * (metalW * 1.05f + dielW * 0.32f);
        color += ccRefl * ccFresnel * ccW * 0.42f;
        const vec3 key = normalize(v3(-0.32f, 0.78f, 0.52f));
        color += baseColor * ((1.0f - metallic) * fmaxf(dot(n, key), 0.0f) * 0.16f * contact);
        const float sss = materialSubsurface(m);
        if (sss > 0.0f) {
            const float wrap = powf(1.0f - nDotV, 2.0f);
            color += materialSubsurfaceColor(m) * baseColor * envDiffuse * (sss * (0.12f + 0.22f * wrap) * contact);
        }
        const float transmission = materialTransmission(m);
        if (transmission > 0.0f) {
            const vec3 transmitted = sampleEnvironment(S, rd);
            color = mixv(color, transmitted * materialTransmissionTint(m, 0.65f), transmission * 0.55f);
            color += baseRefl * fresnel * (transmission * 0.35f);
        }
        color = vclamp(color, v3s(0.0f), v3s(fmaxf(L.sampleRadianceClamp, 1.0f)));
        alpha = L.transparentBackground ? opacity : 1.0f;
        albedo = make_float4(baseColor.x, baseColor.y, baseColor.z, opacity);
        if (L.statsEnabled) {
            atomicAdd(&P.stats[DR_CNT_SHADOW], (unsigned long long)nq);
            atomicAdd(&P.stats[DR_CNT_SHADOW_NODES], (unsigned long long)nv);
            atomicAdd(&P.stats[DR_CNT_SHADOW_TRIS], (unsigned long long)nt);
        }
    }
    if constexpr (VOL) {
        if (L.statsEnabled) addSdfStats(P.stats, sc);
    }
    O.accumulation[pix] = make_float4(color.x, color.y, color.z, alpha);
    const vec3 en = n * 0.5f + v3s(0.5f);
    O.depth[pix] = make_float4(h.t, h.t, h.t, 1.0f);
    O.normal[pix] = make_float4(en.x, en.y, en.z, 1.0f);
    O.albedo[pix] = albedo;
    const float mid = (float)(h.materialID + 1u), oid = (float)(h.objectID + 1u);
    O.materialID[pix] = make_float4(mid, mid, mid, 1.0f);
    O.objectID[pix] = make_float4(oid, oid, oid, 1.0f);
    vec2 cur, prv;
    if (projectToScreen(cam, h.position, L.width, L.height, cur) && projectToScreen(prevCam, h.position, L.width, L.height, prv))
        O.motion[pix] = make_float4(prv.x - cur.x, prv.y - cur.y, 0.0f, 1.0f);
    else
        O.motion[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float tid = __uint_as_float(h.primitiveID + 1u);
    O.triangleID[pix] = make_float4(tid, tid, tid, tid);
}

// --------------------------------------------------------------- resolve --
__global__ void k_resolve(DrLaunch L, DrPaths P, uint32_t bounce) {
    const uint32_t count = P.counters[DR_CTR_Q(bounce)];
    uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= count) return;
    uint32_t slot = P.queue[bounce & 1u][qi];
    uint4 meta = P.meta[slot];
    if ((meta.w & DR_FLAG_NEE) == 0u) return;
    // direct += pre * T / denom, light sample then the two HDRI samples
    // (:4508, :4578-4579), then radiance += clamp(throughput * direct).
    vec3 direct = v3s(0.0f);
    for (uint32_t j = 0; j < 3u; ++j) {
        uint32_t k = slot * 3u + j;
        float4 c = P.shC[k];  // .w: pdf denominator, 0 for no record
        if (c.w == 0.0f) continue;
        float4 r = P.shR[k];
        if (r.w == 0.0f) continue;
        direct += v3(c.x, c.y, c.z) * v3(r.x, r.y, r.z) / c.w;
    }
    float4 t = P.neeT[slot];
    float4 l = P.rad[slot];
    vec3 add = clampSampleContribution(v3(t.x, t.y, t.z) * direct, L.sampleRadianceClamp);
    P.rad[slot] = make_float4(l.x + add.x, l.y + add.y, l.z + add.z, l.w);
    P.meta[slot] = make_uint4(meta.x, meta.y, meta.z, meta.w & ~DR_FLAG_NEE);
}

// ---------------------------------------------------------------- finish --
// Primary AOV ray used when a denoiser is enabled (tracePrimaryAOV,
// PathTrace.metal:4093-4160).
template <bool VOL>
__device__ __forceinline__ bool tracePrimaryAOV(const DrSceneDev &S, vec3 o, vec3 d, DrStack &stk, Surf &out, Mat &outM,
                                DrSdfCounters &sc) {
    for (uint32_t pass = 0; pass < 32u; ++pass) {
        float4 hr, hn;
        uint32_t nv = 0u, nt = 0u;
        if (!sceneClosest<VOL>(S, o, d, stk, nv, nt, sc, hr, hn)) return false;
        Surf h = surfaceAt<VOL>(S, o, d, hr, hn);
        Mat m = materialFor(S, h);
        applyNormalMap(S, h, m, d);
        applyBaseColorTexture(S, m, h);
        if (m.base.w <= 0.001f) {
            o = h.position + d * 0.001f;
            continue;
        }
        out = h;
        outM = m;
        return true;
    }
    return false;
}

template <bool VOL>
__global__ void __launch_bounds__(DR_SHADE_BLOCK) k_finalize(DrSceneDev S, DrLaunch L, DrCamera cam, DrCamera prevCam,
                                                            DrPaths P, DrOutputs O) {
    __shared__ uint32_t ldsStack[DR_LDS_STACK * DR_SHADE_BLOCK];
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t tilePixels = L.tileWidth * L.tileHeight;
    if (p >= tilePixels) return;
    uint32_t px = L.tileX + p % L.tileWidth;
    uint32_t py = L.tileY + p / L.tileWidth;
    uint32_t pix = px + py * L.width;
    float4 acc = O.accumulation[pix];
    uint32_t lastFlags = 0;
    const uint32_t sampleBegin = batchBegin(L);
    for (uint32_t s = 0; s < L.samplesInBatch; ++s) {
        uint32_t slot = pixelSampleSlot(L, p, s);
        float4 l = P.rad[slot];
        float alpha = P.thr[slot].w;
        float w = 1.0f / ((float)(sampleBegin + s) + 1.0f);
        acc.x = mixf(acc.x, l.x, w);
        acc.y = mixf(acc.y, l.y, w);
        acc.z = mixf(acc.z, l.z, w);
        acc.w = mixf(acc.w, alpha, w);
        lastFlags = P.meta[slot].w;
    }
    O.accumulation[pix] = acc;
    bool has = (lastFlags & DR_FLAG_PRIMARY) != 0u;
    if (L.denoiserEnabled != 0u) {
        DrStack stk{dr_lds_ptr(&ldsStack[threadIdx.x]), DR_SHADE_BLOCK, P.overflow + p, P.overflowStride};
        vec3 o, d;
        makeCameraRay(cam, L.width, L.height, px, py, 0.5f, 0.5f, 0.0f, 0.0f, o, d);
        Surf h;
        Mat m;
        DrSdfCounters sc{};
        has = tracePrimaryAOV<VOL>(S, o, d, stk, h, m, sc);
        if (VOL && L.statsEnabled) addSdfStats(P.stats, sc);
        if (has) {
            vec3 en = h.normal * 0.5f + v3s(0.5f);
            float mid = (float)(h.materialID + 1u), oid = (float)(h.objectID + 1u);
            O.depth[pix] = make_float4(h.t, h.t, h.t, 1.0f);
            O.normal[pix] = make_float4(en.x, en.y, en.z, 1.0f);
            O.albedo[pix] = m.base;
            O.materialID[pix] = make_float4(mid, mid, mid, 1.0f);
            O.objectID[pix] = make_float4(oid, oid, oid, 1.0f);
            vec2 cur, prv;
            if (projectToScreen(cam, h.position, L.width, L.height, cur) &&
                projectToScreen(prevCam, h.position, L.width, L.height, prv))
                O.motion[pix] = make_float4(prv.x - cur.x, prv.y - cur.y, 0.0f, 1.0f);
            else
                O.motion[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float tid = __uint_as_float(h.primitiveID + 1u);
            O.triangleID[pix] = make_float4(tid, tid, tid, tid);
        }
    }
    if (!has) {
        O.depth[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        O.normal[pix] = make_float4(0.5f, 0.5f, 0.5f, 0.0f);
        O.albedo[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        O.materialID[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        O.objectID[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        O.motion[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        O.triangleID[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// --------------------------------------------------------------- launches --
namespace dr {

// One ray per lane through the same while-while traversal and triangle
// tests as the path tracer's extension queries (triangle geometry only).
// The launch covers rays [first, first + lanes); lane k's stack spill is
// overflow[k + j * lanes].
__global__ __launch_bounds__(DR_TRACE_BLOCK) void k_intersect(DrSceneDev S, const float4 *__restrict__ rays,
                                                              uint32_t first, uint32_t count,
                                                              float *__restrict__ tOut,
                                                              uint32_t *__restrict__ idOut, uint32_t *overflow) {
    __shared__ uint32_t ldsStack[DR_LDS_STACK * DR_TRACE_BLOCK];
    const uint32_t lane = blockIdx.x * DR_TRACE_BLOCK + threadIdx.x;
    const uint32_t i = first + lane;
    DrStack stk{dr_lds_ptr(&ldsStack[threadIdx.x]), DR_TRACE_BLOCK, overflow + lane, gridDim.x * DR_TRACE_BLOCK};
    const bool live = i < count;
    // Lanes past the end trace a ray that misses every box.
    const float4 o = live ? rays[2 * i] : make_float4(3e38f, 3e38f, 3e38f, 0.f);
    const float4 d = live ? rays[2 * i + 1] : make_float4(0.f, 0.f, 1.f, 0.f);
    DrRayPre r = dr_prep(v3(o.x, o.y, o.z), v3(d.x, d.y, d.z));
    DrHitRec best{live ? __builtin_inff() : 0.f, 0.f, 0.f, 0xffffffffu};
    uint32_t nv = 0, nt = 0;
    dr_traverse<false>(S, r, best, stk, nv, nt);
    if (live) {
        tOut[i] = best.id != 0xffffffffu ? best.t : __builtin_inff();
        idOut[i] = best.id;
    }
}

hipError_t launchIntersect(const DrSceneDev &S, const float4 *rays, uint32_t count, float *t, uint32_t *ids,
                           uint32_t *overflow, hipStream_t st, uint32_t maxLanes) {
    const uint32_t chunk = maxLanes ? maxLanes : (count + DR_TRACE_BLOCK - 1) / DR_TRACE_BLOCK * DR_TRACE_BLOCK;
    for (uint32_t first = 0; first < count; first += chunk) {
        const uint32_t n = std::min(chunk, count - first);
        const uint32_t blocks = (n + DR_TRACE_BLOCK - 1) / DR_TRACE_BLOCK;
        hipLaunchKernelGGL(k_intersect, dim3(blocks), dim3(DR_TRACE_BLOCK), 0, st, S, rays, first, count, t, ids,
                           overflow);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Traversal-stack spill of the allocation-free ray queries
// (dr_hip_scene_intersect_device): a fixed array each device holds until the
// process ends, addressed through hipGetSymbolAddress once per device (the
// array's name in host code is not its device address).
static __device__ __align__(256) uint32_t scratch[DR_INTERSECT_SCRATCH_WORDS];

hipError_t intersectScratchAddress(uint32_t **ptr) {
    return hipGetSymbolAddress(reinterpret_cast<void **>(ptr), HIP_SYMBOL(scratch));
}

// Corner records of the FP16 brick distances (DrSdfDev::corner16), one
// workgroup per brick, written at the brick's own sample indices; bricks
// that brickCtx rejects get none.
__global__ __launch_bounds__(256) void k_brick_corners(const DrBrick *bricks, uint32_t brickCount,
                                                       const uint16_t *dist, uint32_t sampleCount, uint4 *corner) {
    const uint32_t bi = blockIdx.x;
    if (bi >= brickCount) return;
    const DrBrick b = bricks[bi];
    const uint32_t packed = __float_as_uint(b.cmin.w), offset = __float_as_uint(b.cmax.w);
    const uint32_t dx = packed & 0xffu, dy = (packed >> 8) & 0xffu, dz = (packed >> 16) & 0xffu;
    const uint32_t n = dx * dy * dz;
    if (dx < 2u || dy < 2u || dz < 2u || offset >= sampleCount || n > sampleCount - offset) return;
    const uint16_t *src = dist + offset;
    uint4 *dst = corner + offset;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t x = i % dx, y = (i / dx) % dy, z = i / (dx * dy);
        const uint32_t x1 = min(x + 1u, dx - 1u), y1 = min(y + 1u, dy - 1u), z1 = min(z + 1u, dz - 1u);
        auto at = [&](uint32_t xx, uint32_t yy, uint32_t zz) -> uint32_t { return src[xx + yy * dx + zz * dx * dy]; };
        dst[i] = make_uint4(at(x, y, z) | at(x1, y, z) << 16, at(x, y1, z) | at(x1, y1, z) << 16,
                            at(x, y, z1) | at(x1, y, z1) << 16, at(x, y1, z1) | at(x1, y1, z1) << 16);
    }
}

hipError_t launchBrickCorners(const DrBrick *bricks, uint32_t brickCount, const uint16_t *dist, uint32_t sampleCount,
                              uint4 *corner, hipStream_t st) {
    if (brickCount == 0u) return hipSuccess;
    hipLaunchKernelGGL(k_brick_corners, dim3(brickCount), dim3(256), 0, st, bricks, brickCount, dist, sampleCount,
                       corner);
    return hipGetLastError();
}

hipError_t launchSetSampleBase(uint32_t *base, uint32_t value, hipStream_t st) {
    hipLaunchKernelGGL(k_set_base, dim3(1), dim3(1), 0, st, base, value);
    return hipGetLastError();
}
hipError_t launchAdvanceSampleBase(uint32_t *base, uint32_t samples, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_base, dim3(1), dim3(1), 0, st, base, samples);
    return hipGetLastError();
}

hipError_t launchStage(DrStage stage, uint32_t bounce, const DrSceneDev &S, const DrLaunch &L, const DrCamera &cam,
                       const DrCamera &prevCam, const DrPaths &P, const DrOutputs &O, uint32_t nSlots,
                       uint32_t traceGrid, hipStream_t st) {
    const uint32_t bThreads = 256;
    const bool vol = S.sdf.volumeCount != 0u;
    switch (stage) {
    case DR_STAGE_RAYGEN:
    {
        const uint32_t words = std::max(1u, L.maxBounces) * DR_CLAIM_PER_BOUNCE;
        hipLaunchKernelGGL(k_batch_begin, dim3((words + bThreads - 1) / bThreads), dim3(bThreads), 0, st, P, nSlots,
                           words);
    }
        break;
    case DR_STAGE_EXTEND:
        if (vol && L.statsEnabled)
            hipLaunchKernelGGL((k_extend<true, true>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, L, cam, P, bounce);
        else if (vol)
            hipLaunchKernelGGL((k_extend<true, false>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, L, cam, P, bounce);
        else if (L.statsEnabled)
            hipLaunchKernelGGL((k_extend<false, true>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, L, cam, P, bounce);
        else
            hipLaunchKernelGGL((k_extend<false, false>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, L, cam, P, bounce);
        break;
    case DR_STAGE_SHADE:
    case DR_STAGE_RESUME: {
        const dim3 grid((nSlots + DR_SHADE_BLOCK - 1) / DR_SHADE_BLOCK), block(DR_SHADE_BLOCK);
        const bool resume = stage == DR_STAGE_RESUME;
        if (vol && resume)
            hipLaunchKernelGGL((k_shade<true, true>), grid, block, 0, st, S, L, cam, prevCam, P, O, bounce);
        else if (vol)
            hipLaunchKernelGGL((k_shade<true, false>), grid, block, 0, st, S, L, cam, prevCam, P, O, bounce);
        else if (resume)
            hipLaunchKernelGGL((k_shade<false, true>), grid, block, 0, st, S, L, cam, prevCam, P, O, bounce);
        else
            hipLaunchKernelGGL((k_shade<false, false>), grid, block, 0, st, S, L, cam, prevCam, P, O, bounce);
        break;
    }
    case DR_STAGE_SSS:
        if (vol && L.statsEnabled)
            hipLaunchKernelGGL((k_sss<true, true>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, L, P, bounce);
        else if (vol)
            hipLaunchKernelGGL((k_sss<true, false>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, L, P, bounce);
        else if (L.statsEnabled)
            hipLaunchKernelGGL((k_sss<false, true>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, L, P, bounce);
        else
            hipLaunchKernelGGL((k_sss<false, false>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, L, P, bounce);
        break;
    case DR_STAGE_SHADOW:
        if (vol && L.statsEnabled)
            hipLaunchKernelGGL((k_shadow<true, true>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, P, bounce);
        else if (vol)
            hipLaunchKernelGGL((k_shadow<true, false>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, P, bounce);
        else if (L.statsEnabled)
            hipLaunchKernelGGL((k_shadow<false, true>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, P, bounce);
        else
            hipLaunchKernelGGL((k_shadow<false, false>), dim3(traceGrid), dim3(DR_TRACE_BLOCK), 0, st, S, P, bounce);
        break;
    case DR_STAGE_PREVIEW: {
        const uint32_t px = L.tileWidth * L.tileHeight;
        const dim3 grid((px + DR_SHADE_BLOCK - 1) / DR_SHADE_BLOCK), block(DR_SHADE_BLOCK);
        const bool interactive = S.sdf.renderQuality == 1u;
        if (vol && interactive)
            hipLaunchKernelGGL((k_preview<true, true>), grid, block, 0, st, S, L, cam, prevCam, P, O);
        else if (vol)
            hipLaunchKernelGGL((k_preview<true, false>), grid, block, 0, st, S, L, cam, prevCam, P, O);
        else if (interactive)
            hipLaunchKernelGGL((k_preview<false, true>), grid, block, 0, st, S, L, cam, prevCam, P, O);
        else
            hipLaunchKernelGGL((k_preview<false, false>), grid, block, 0, st, S, L, cam, prevCam, P, O);
        break;
    }
    case DR_STAGE_RESOLVE:
        hipLaunchKernelGGL(k_resolve, dim3((nSlots + bThreads - 1) / bThreads), dim3(bThreads), 0, st, L, P, bounce);
        break;
    case DR_STAGE_FINALIZE: {
        uint32_t tilePixels = L.tileWidth * L.tileHeight;
        if (vol)
            hipLaunchKernelGGL(k_finalize<true>, dim3((tilePixels + DR_SHADE_BLOCK - 1) / DR_SHADE_BLOCK),
                               dim3(DR_SHADE_BLOCK), 0, st, S, L, cam, prevCam, P, O);
        else
            hipLaunchKernelGGL(k_finalize<false>, dim3((tilePixels + DR_SHADE_BLOCK - 1) / DR_SHADE_BLOCK),
                               dim3(DR_SHADE_BLOCK), 0, st, S, L, cam, prevCam, P, O);
        break;
    }
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dr
