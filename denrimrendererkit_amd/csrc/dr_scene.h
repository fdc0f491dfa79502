// dr_scene.h -- device-resident scene and per-launch settings.
//
// The scene is held in HBM in two forms:
//   * the caller's flat records in the reference layouts (GPUTriangle 192 B,
//     GPUMaterial 272 B, GPULightRecord 32 B, GPUEnvironmentSample 8 B,
//     texture descriptors and RGBA32F texels; Metal/GPUTypes.swift), read
//     only at shading time;
//   * a traversal form built by this library: a four-wide BVH whose nodes
//     hold their four child boxes (DrNode4, 128 B, one 8 x dwordx4 fetch
//     tests four boxes) and a position-only triangle stream in leaf order,
//     48 B per triangle {v0 | original index, v1, v2}, so a leaf's triangles
//     are contiguous and the traversal never touches the 192-B records.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// GPUTriangle (PathTrace.metal:14-30).
struct DrTriangle {
    float4 v0, v1, v2, n0, n1, n2, uv0, uv1, uv2, tangent, bitangent;
    uint32_t materialID, objectID, primitiveID, padding2;
};
static_assert(sizeof(DrTriangle) == 192, "GPUTriangle layout");

// GPUMaterial (PathTrace.metal:32-50): 17 float4 slots.
enum DrMatSlot {
    M_BASE = 0, M_EMISSION, M_PARAMS, M_PARAMS2, M_SPECULAR, M_SHEEN, M_TRANS, M_PARAMS3,
    M_CLEARCOAT, M_CC_ATTEN, M_TRANS_ABS, M_THINFILM, M_SSS_COLOR, M_SSS_RADIUS, M_SSS_PARAMS,
    M_VOL_SCATTER, M_VOL_PARAMS, M_SLOTS
};
struct DrMaterial {
    float4 s[M_SLOTS];
};
static_assert(sizeof(DrMaterial) == 272, "GPUMaterial layout");

// GPULightRecord (PathTrace.metal:62-68).
struct DrLight {
    uint32_t triangleIndex, materialIndex;
    float area, selectionCDF;
    float4 normal;
};
static_assert(sizeof(DrLight) == 32, "GPULightRecord layout");

// Traversal node of the 4-wide BVH, one 128-B L2 line: the four child
// boxes as planes of x/y/z bounds (component c = child c), the child
// references, then meta.x = the four leaf triangle counts (8 bits each,
// child c at bits 8c; 0 = inner node) and meta.y = the mask of present
// children (bits 0-3) and of leaf children (bits 4-7). A leaf child's
// reference is its first triangle in leaf order, an inner child's its node
// index.
struct DrNode4 {
    float4 lox, hix, loy, hiy, loz, hiz;
    uint4 child;
    uint4 meta;
};
static_assert(sizeof(DrNode4) == 128, "node layout");

// Compact brick header (64 B).
struct DrBrick {
    float4 smin;  // sampleBoundsMin.xyz, min distance
    float4 smax;  // sampleBoundsMax.xyz, max distance
    float4 cmin;  // core localBoundsMin.xyz, bits(dims.x | dims.y << 8 | dims.z << 16)
    float4 cmax;  // core localBoundsMax.xyz, bits(sample offset)
};
static_assert(sizeof(DrBrick) == 64, "brick header");

// GPUVolumeDescriptor (PathTrace.metal:74-90).
struct DrVolume {
    float4 worldBoundsMin, worldBoundsMax, localBoundsMin, localBoundsMax;
    uint4 dimensions, metadata, materialProgram;
    float4 w2l[4];
    float4 nrm[4];
};
static_assert(sizeof(DrVolume) == 240, "GPUVolumeDescriptor layout");

// GPUVolumeBrickGrid (PathTrace.metal:143-148).
struct DrBrickGrid {
    uint4 dimensionsAndIndexOffset, brickSizeAndVolume, macroDimensionsAndIndexOffset, macroSizeAndReserved;
};

// GPUVolumeSample (PathTrace.metal:92-101): one dense-volume sample.
struct DrVolumeSample {
    float distance;
    uint32_t materialA, materialB;
    float materialBlend;
    float4 baseColorOpacity, emissionTransmission, surface;
    uint4 materialFieldFlags;
};
static_assert(sizeof(DrVolumeSample) == 80, "GPUVolumeSample layout");

struct DrSdfDev {
    const DrVolume *volumes;
    const DrVolumeSample *dense;  // dense-volume samples
    uint32_t denseCount;
    const DrBrick *bricks;
    const uint16_t *dist16;     // FP16 bits (distHalf != 0)
    const float *dist32;        // FP32 (distHalf == 0)
    // FP16 only: for sample i = (x, y, z) of a valid brick (valid bricks'
    // sample ranges are disjoint), the 8 distances a trilinear lookup of the
    // cell at i reads, (x|x+1, y, z), (x|x+1, y+1, z), (x|x+1, y, z+1),
    // (x|x+1, y+1, z+1) with the upper index clamped to the brick: one 16-B
    // record per sample.
    const uint4 *corner16;
    const uint2 *material;      // x: A | B << 16, y: bits(blend)
    const DrBrickGrid *grids;
    const uint32_t *gridIndices;
    uint32_t volumeCount, brickCount, sampleCount, gridCount, gridIndexCount;
    uint32_t renderQuality;     // 2 = final
    uint32_t distHalf;
    // 1: every material is opaque and lets no light through (as
    // DrSceneDev::opaqueOnly, with a margin that blends of two materials
    // keep) and no dense-volume sample overrides material fields, so any hit
    // in front of a light blocks it.
    uint32_t opaqueShadows;
};

struct DrSceneDev {
    const DrNode4 *nodes;
    const float4 *triPos;        // 3 float4 per triangle in leaf order
    const DrTriangle *triangles; // caller's records, original order
    const DrMaterial *materials;
    const uint32_t *matKey;      // per material: shading class (which branches its paths take), 0..13
    const DrLight *lights;
    const float2 *env;           // (cdf, pdf) per environment texel
    const uint32_t *envGuide;    // envGuideSize + 1 entries: first texel with cdf >= k / envGuideSize
    const uint2 *envAlias;       // alias table per texel: (bits(probability), alias texel)
    const uint4 *texDesc;        // (offset, width, height, mode)
    const float4 *texels;
    uint32_t triangleCount, materialCount, lightCount, envCount, texCount, envGuideSize;
    uint32_t envTexIndexPlusOne;
    float envIntensity, envRotationY, envMaxRadiance;
    uint32_t opaqueOnly;         // 1: no material can let a shadow ray through
    uint32_t watertight;         // 1: watertight triangle test, 0: Moller-Trumbore (the reference's)
    uint32_t envAliasSampling;   // 1: HDRI texels drawn from the alias table, 0: CDF search (the reference's)
    DrSdfDev sdf;                // sparse distance volumes (volumeCount 0: none)
};

// Per-launch settings: the fields of GPURenderConstants
// (PathTrace.metal:150-189) that the triangle path reads.
struct DrLaunch {
    uint32_t width, height;
    uint32_t maxBounces;
    uint32_t frameSeed;
    uint32_t transparentBackground, showsEnvironmentBackground, denoiserEnabled;
    float sampleRadianceClamp;
    float backgroundColor[4];
    uint32_t tileX, tileY, tileWidth, tileHeight;
    uint32_t sampleBegin;      // sample index of the first sample in this batch
    const uint32_t *sampleBase;  // if set, the device word holding sampleBegin (graph replays)
    uint32_t samplesInBatch;   // samples per pixel in flight together
    uint32_t statsEnabled;
};

struct DrCamera {
    float4 origin, lowerLeft, horizontal, vertical, lens;
};

// Traversal and query counters (the reference's SDF counter slots 10-12,
// PathTrace.metal:201-203, plus node fetches and triangle tests for the
// roofline: extension and subsurface queries in NODES/TRIS, NEE shadow
// queries in SHADOW_NODES/SHADOW_TRIS).
enum {
    DR_CNT_PRIMARY = 0,
    DR_CNT_BOUNCE = 1,
    DR_CNT_SHADOW = 2,
    DR_CNT_NODES = 3,
    DR_CNT_TRIS = 4,
    DR_CNT_SHADOW_NODES = 5,
    DR_CNT_SHADOW_TRIS = 6,
    DR_CNT_SSS = 7,  // subsurface-walk queries (also counted in BOUNCE)
    // SDF counters, the reference's slots 0-9 (PathTrace.metal:191-200) at
    // DR_CNT_SDF_BASE + slot; the same ten for NEE shadow queries (k_shadow)
    // at DR_CNT_SDF_SHADOW_BASE. The reference's totals are the sums.
    DR_CNT_SDF_BASE = 8,
    DR_CNT_SDF_SHADOW_BASE = 18,
    // Node fetches and triangle tests of subsurface-walk queries (k_sss);
    // NODES / TRIS count extension and AOV queries.
    DR_CNT_SSS_NODES = 28,
    DR_CNT_SSS_TRIS = 29,
    // The part of the SHADOW / SHADOW_NODES / SHADOW_TRIS counts traced
    // inside k_sss (the shadow rays queued by k_shade on triangle scenes).
    DR_CNT_SSS_SHADOW = 30,
    DR_CNT_SSS_SHADOW_NODES = 31,
    DR_CNT_SSS_SHADOW_TRIS = 32,
    // The part traced inside k_extend (the previous bounce's shadow rays
    // queued by k_shade<RESUME>, triangle scenes).
    DR_CNT_EXT_SHADOW = 33,
    DR_CNT_EXT_SHADOW_NODES = 34,
    DR_CNT_EXT_SHADOW_TRIS = 35,
    // Every scene query (extension, NEE shadow-chain step, subsurface-walk
    // segment), counted by every kernel variant, not only in counting runs:
    // one atomic per wave, so timed runs know their own query count.
    DR_CNT_RAY_QUERIES = 36,
    DR_CNT_SLOTS = 37
};
