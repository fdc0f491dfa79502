/*
 * oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is a plain C++ restatement of the
 * reference progressive path tracer (pathTraceKernel and its callees in
 * Sources/DenrimRendererKit/Resources/Shaders/PathTrace.metal, plus the
 * count-median BVH of Acceleration/BVHBuilder.swift). Only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load it, and only as the
 * checker; the product never links it.
 *
 * Struct layouts below restate Sources/DenrimRendererKit/Metal/GPUTypes.swift
 * (GPUTriangle :12, GPUMaterial :30, GPULightRecord, GPUEnvironmentSample,
 * GPUTextureDescriptor, GPUCamera :4) byte for byte, because the oracle
 * consumes the same flat arrays as the reference kernel.
 */
#ifndef DENRIM_ORACLE_H
#define DENRIM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z, w; } or_float4;
typedef struct { uint32_t x, y, z, w; } or_uint4;

typedef struct {
    or_float4 origin, lowerLeft, horizontal, vertical, lens;
} or_camera;                                  /* 80 B */

typedef struct {
    or_float4 v0, v1, v2, n0, n1, n2, uv0, uv1, uv2, tangent, bitangent;
    uint32_t materialID, objectID, primitiveID, padding2;
} or_triangle;                                /* 192 B */

typedef struct {
    or_float4 baseColor, emission, parameters, parameters2, specularColor,
        sheenColor, transmissionColor, parameters3, clearcoatColor,
        clearcoatAttenuation, transmissionAbsorption, thinFilm,
        subsurfaceColor, subsurfaceRadius, subsurfaceParameters,
        volumeScattering, volumeParameters;
} or_material;                                /* 272 B */

typedef struct {
    uint32_t triangleIndex, materialIndex;
    float area, selectionCDF;
    or_float4 normal;
} or_light;                                   /* 32 B */

typedef struct { float cdf, pdf; } or_env_sample;   /* 8 B */

/* GPUVolumeDescriptor (PathTrace.metal:74-90), 240 B. */
typedef struct {
    or_float4 worldBoundsMin, worldBoundsMax, localBoundsMin, localBoundsMax;
    or_uint4 dimensions, metadata, materialProgram;
    or_float4 worldToLocal0, worldToLocal1, worldToLocal2, worldToLocal3;
    or_float4 normalTransform0, normalTransform1, normalTransform2, normalTransform3;
} or_volume;

/* GPUVolumeBrickDescriptor (PathTrace.metal:132-141), 128 B. */
typedef struct {
    or_float4 worldBoundsMin, worldBoundsMax, localBoundsMin, localBoundsMax, sampleBoundsMin, sampleBoundsMax;
    or_uint4 gridOriginAndVolume, dimensionsAndSampleOffset;
} or_brick;

/* GPUVolumeBrickSample (PathTrace.metal:103-108), 16 B. */
typedef struct { float distance; uint32_t materialA, materialB; float materialBlend; } or_brick_sample;
/* GPUVolumeSample (PathTrace.metal:92-101), dense volumes: 80 B. */
typedef struct {
    float distance; uint32_t materialA, materialB; float materialBlend;
    or_float4 baseColorOpacity, emissionTransmission, surface;
    uint32_t materialFieldFlags[4];
} or_volume_sample;

/* GPUVolumeBrickGrid (PathTrace.metal:143-148), 64 B. */
typedef struct {
    or_uint4 dimensionsAndIndexOffset, brickSizeAndVolume, macroDimensionsAndIndexOffset, macroSizeAndReserved;
} or_brick_grid;
typedef struct { uint32_t offset, width, height, mode; } or_texture_desc; /* 16 B */

/* The reference's per-dispatch constants that the triangle path reads
 * (GPURenderConstants, PathTrace.metal:150-189). */
typedef struct {
    uint32_t width, height;
    uint32_t maxBounces;
    uint32_t frameSeed;
    uint32_t transparentBackground;
    uint32_t showsEnvironmentBackground;
    uint32_t environmentTextureIndexPlusOne;
    uint32_t denoiserEnabled;
    float environmentIntensity;
    float environmentRotationY;
    float environmentMaxRadiance;
    float sampleRadianceClamp;
    float backgroundColor[4];
    uint32_t tileX, tileY, tileWidth, tileHeight;
    uint32_t renderQuality;  // RenderSettings.shaderQualityLevel: 0 preview, 1 interactive, 2 final
    uint32_t intersection;   // 0 Moller-Trumbore (the reference's), 1 watertight (this build's option)
    uint32_t environmentSampling;  // 0 CDF search (the reference's), 1 alias table (this build's option)
} or_settings;

typedef struct {
    const or_triangle *triangles; uint32_t triangleCount;
    const or_material *materials; uint32_t materialCount;
    const or_light *lights; uint32_t lightCount;
    const or_env_sample *envSamples; uint32_t envSampleCount;
    const or_texture_desc *textureDescs; uint32_t textureCount;
    const or_float4 *texturePixels; uint64_t texturePixelCount;
    const or_volume *volumes; uint32_t volumeCount;
    const or_volume_sample *volumeSamples; uint32_t volumeSampleCount;
    const or_brick *bricks; uint32_t brickCount;
    const or_brick_sample *brickSamples; uint32_t brickSampleCount;
    const or_brick_grid *brickGrids; uint32_t brickGridCount;
    const uint32_t *brickGridIndices; uint32_t brickGridIndexCount;
} or_scene_desc;

typedef struct or_scene or_scene;

/* Counter slots written by or_render (traversal work, for the roofline). */
enum {
    OR_CNT_PRIMARY_QUERIES = 0,
    OR_CNT_BOUNCE_QUERIES = 1,
    OR_CNT_SHADOW_QUERIES = 2,
    OR_CNT_NODE_VISITS = 3,
    OR_CNT_TRIANGLE_TESTS = 4,
    /* sparse-SDF counters at 8 + the reference's slot (PathTrace.metal:191-200) */
    OR_CNT_SDF_BASE = 8,
    OR_CNT_COUNT = 24
};

/* Scene: copies the arrays and builds the count-median BVH. */
int or_scene_create(const or_scene_desc *desc, uint32_t maxLeaf, or_scene **out);
void or_scene_destroy(or_scene *scene);
uint32_t or_scene_node_count(const or_scene *scene);
/* Flat BVH export in the reference layout: nodes as 12 x u32 (48 B), indices u32. */
int or_scene_export_bvh(const or_scene *scene, void *nodes48, uint32_t *primitiveIndices);

/*
 * Render `sampleCount` progressive samples starting at `sampleBegin` into the
 * caller's buffers, following pathTraceKernel (PathTrace.metal:5167-5586).
 *   precision 32: accumulation/aovs are float[W*H*4]; 64: double[W*H*4].
 *   traversal 0: flat BVH; 1: brute force over all triangles.
 *   aovs: 7 planes of W*H*4 (depth, normal, albedo, materialID, objectID,
 *         motion, triangleID) in the caller's precision; may be NULL.
 *   counters: OR_CNT_COUNT u64 (may be NULL), accumulated.
 *   threads: OpenMP threads (0 = default).
 */
int or_render(const or_scene *scene, const or_settings *settings,
              const or_camera *camera, const or_camera *previousCamera,
              uint32_t sampleBegin, uint32_t sampleCount, int precision,
              int traversal, void *accumulation, void *aovs,
              uint64_t *counters, int threads);

/* Closest-hit queries for arbitrary rays (fp32 path): rays are 8 floats
 * (ox,oy,oz,_,dx,dy,dz,_); out: t (float), primitive index (u32, ~0 = miss),
 * and, if hitOut is not NULL, 8 floats per ray: shading normal xyz,
 * materialID, materialID2, materialBlend, objectID, frontFacing. */
int or_intersect(const or_scene *scene, const float *rays, uint32_t count,
                 int traversal, int intersection, float *tOut, uint32_t *idOut, float *hitOut);

/* The alias table built for the scene's environment (count = environment
 * sample count, > 1), for tests. */
int or_environment_alias(const or_scene *scene, float *prob, uint32_t *alias, uint32_t count);

/* RNG restatement (PathTrace.metal:264-276). */
uint32_t or_hash(uint32_t value);
float or_random_float(uint32_t *state);

#ifdef __cplusplus
}
#endif

#endif
