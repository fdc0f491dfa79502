// dr_kernels.h -- path-state buffers and the per-sample launch sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dr_scene.h"

#ifndef DR_TRACE_BLOCK
#define DR_TRACE_BLOCK 128
#endif
#ifndef DR_SHADE_BLOCK
#define DR_SHADE_BLOCK 256
#endif
#define DR_MAX_BOUNCES 64
// Size of that scratch: 16 MiB per device.
#define DR_INTERSECT_SCRATCH_WORDS (4u << 20)
// Traversal stack entries kept in LDS per lane; deeper entries spill to a
// per-lane global area of (stackMax - DR_LDS_STACK) words, stackMax being
// the scene's bound of 7 entries per BVH level plus two.
#ifndef DR_LDS_STACK
#define DR_LDS_STACK 24
#endif

// Device counter words: queue lengths per bounce (index b = queue consumed
// by bounce b), work-stealing cursors of extend and shadow per bounce, and
// shadow-queue lengths per bounce. Zeroed by k_batch_begin at the start of each
// sample batch, so no counter is reset while a kernel reads it.
#define DR_CTR_QUEUE 0
#define DR_CTR_SHADOWQ (DR_CTR_QUEUE + DR_MAX_BOUNCES + 1)
#define DR_CTR_Q(b) (DR_CTR_QUEUE + (b))
#define DR_CTR_SQ(b) (DR_CTR_SHADOWQ + (b))
// Subsurface-walk queue lengths per bounce, and the queue of paths whose
// walk failed and resume surface shading.
#define DR_CTR_SSSQ (DR_CTR_SHADOWQ + DR_MAX_BOUNCES)
#define DR_CTR_RESUMEQ (DR_CTR_SSSQ + DR_MAX_BOUNCES)
// First shadow-queue entry of the bounce left to k_shadow: the triangle-only
// k_sss traces the entries queued before it ran (those of k_shade) beside
// its walks and stores their count here; k_shadow takes the rest (those of
// k_shade<RESUME>). Zero when k_sss does not trace shadow rays.
#define DR_CTR_SHBEGIN (DR_CTR_RESUMEQ + DR_MAX_BOUNCES)
#define DR_CTR_WORDS (DR_CTR_SHBEGIN + DR_MAX_BOUNCES)

// Claim cursors of the persistent tracers, per bounce and queue kind: the
// kind's queue is dealt in chunks of DR_CLAIM_CHUNK items to
// DR_CLAIM_STRIPES cursors in turn (chunk c to cursor c mod stripes), the
// cursors DR_CLAIM_STRIDE words apart. A wave claims from its own cursor
// and moves on to the next when that one's chunks are gone, so the claims
// of a launch's few thousand waves spread over eight words instead of
// queueing on one. Zeroed by k_batch_begin for the batch's bounces.
#define DR_CLAIM_STRIPES 8u
#define DR_CLAIM_CHUNK 64u
#define DR_CLAIM_STRIDE 64u
#define DR_CLAIM_EXTEND 0u
#define DR_CLAIM_SHADOW 1u
#define DR_CLAIM_SSS 2u
#define DR_CLAIM_KINDS 3u
#define DR_CLAIM_PER_BOUNCE (DR_CLAIM_KINDS * DR_CLAIM_STRIPES * DR_CLAIM_STRIDE)
#define DR_CLAIM_WORDS (DR_MAX_BOUNCES * DR_CLAIM_PER_BOUNCE)
#define DR_CLAIM_AT(kind, b) ((b) * DR_CLAIM_PER_BOUNCE + (kind) * DR_CLAIM_STRIPES * DR_CLAIM_STRIDE)

struct DrPaths {
    float4 *rayO;   // origin.xyz, previousBSDFPDF
    float4 *rayD;   // direction.xyz, -
    float4 *thr;    // throughput.xyz, outputAlpha
    float4 *rad;    // radiance.xyz, -
    uint4 *meta;    // pixel, sampleIndex, rng state, flags
    float4 *hit;    // t, u, v, bits(triangle id); volume hits: see DR_VOL_HIT
    float4 *hitN;   // normal of a sparse-volume hit
    float4 *neeT;   // throughput at the NEE vertex
    float4 *shO;    // 3 per slot: shadow origin.xyz, max distance
    float4 *shD;    // 3 per slot: direction.xyz, pdf denominator
    float4 *shC;    // 3 per slot: unshadowed contribution.xyz, pdf denominator (0: no record)
    float4 *shR;    // 3 per slot: transmittance.xyz, visible
    uint32_t *queue[2];
    uint32_t *shadowQueue;
    uint32_t *sssQueue;     // paths entering a subsurface walk this bounce
    uint32_t *resumeQueue;  // paths whose walk failed: surface shading resumes
    uint32_t *counters;  // DR_CTR_WORDS
    uint32_t *claims;    // DR_CLAIM_WORDS
    unsigned long long *stats;  // DR_CNT_SLOTS, accumulated over launches
    uint32_t *overflow;  // traversal stack spill, overflowStride lanes
    uint32_t overflowStride;
};

struct DrOutputs {
    float4 *accumulation;
    float4 *depth, *normal, *albedo, *materialID, *objectID, *motion, *triangleID;
};

enum DrStage {
    DR_STAGE_RAYGEN = 0,  // batch start (counters); the rays come from EXTEND of bounce 0
    DR_STAGE_EXTEND = 1,
    DR_STAGE_SHADE = 2,
    DR_STAGE_SHADOW = 3,
    DR_STAGE_RESOLVE = 4,
    DR_STAGE_FINALIZE = 5,
    DR_STAGE_SSS = 6,     // subsurface random walks of the bounce
    DR_STAGE_RESUME = 7,  // surface shading after a failed walk
    DR_STAGE_PREVIEW = 8, // flat preview / interactive material kernels (quality 0 / 1)
    DR_STAGE_COUNT = 9
};

namespace dr {
// Enqueues one stage of a sample batch on `st`. One batch is RAYGEN, then
// for b in [0, maxBounces): EXTEND, SHADE, SSS, RESUME, SHADOW, RESOLVE;
// then FINALIZE.
hipError_t launchStage(DrStage stage, uint32_t bounce, const DrSceneDev &S, const DrLaunch &L, const DrCamera &cam,
                       const DrCamera &prevCam, const DrPaths &P, const DrOutputs &O, uint32_t nSlots,
                       uint32_t traceGrid, hipStream_t st);
// Device word of the first sample index of the next batch: set it, or add
// the batch's sample count after a batch (the last node of a batch graph).
hipError_t launchSetSampleBase(uint32_t *base, uint32_t value, hipStream_t st);
// Closest triangle hit of `count` rays (2 float4 each: origin, direction):
// t (+inf on a miss) and the original triangle index (~0 on a miss);
// overflow holds the traversal-stack spill of ceil(count / 128) * 128 lanes.
// maxLanes > 0 splits the rays into launches of at most maxLanes lanes (a
// multiple of the block size) that share one overflow area of maxLanes lanes.
hipError_t launchIntersect(const DrSceneDev &S, const float4 *rays, uint32_t count, float *t, uint32_t *ids,
                           uint32_t *overflow, hipStream_t st, uint32_t maxLanes = 0);
// Device address of the static traversal-stack scratch of the allocation-free
// ray queries (DR_INTERSECT_SCRATCH_WORDS words); call with the device current.
hipError_t intersectScratchAddress(uint32_t **ptr);
hipError_t launchAdvanceSampleBase(uint32_t *base, uint32_t samples, hipStream_t st);
// Fills the FP16 corner records (DrSdfDev::corner16) of every valid brick.
hipError_t launchBrickCorners(const DrBrick *bricks, uint32_t brickCount, const uint16_t *dist, uint32_t sampleCount,
                              uint4 *corner, hipStream_t st);
}
