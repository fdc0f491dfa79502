// dr_bvh.h -- host-side builder of the traversal BVH.
#pragma once

#include <stdint.h>
#include <vector>

#include "dr_scene.h"

namespace dr {

struct BvhBuildStats {
    uint32_t nodeCount = 0;
    uint32_t leafCount = 0;
    uint32_t maxDepth = 0;
    double sahCost = 0.0;
    double buildSeconds = 0.0;
};

// Builds a binned-SAH binary BVH over the triangles' world-space vertices
// (reference GPUTriangle records), collapses it to a 4-wide BVH (each wide
// node takes the cut of its binary subtree into at most four children of
// least SAH cost, a bottom-up dynamic program over (node, slots)) and
// writes the traversal layout:
//   nodes  -- DrNode4 per inner node, depth-first pre-order, root at 0;
//   triPos -- 3 float4 per triangle in leaf order: {v0, bits(orig index)},
//             {v1, 0}, {v2, 0} (exact vertices: the watertight test needs
//             them; the Moller-Trumbore test forms v1 - v0, v2 - v0 itself,
//             the same float subtractions as the reference).
// Child boxes are widened outward by 1e-6 of their coordinates' magnitude
// plus extent, which covers the fused-multiply-add slab test's error that
// scales with the box (the part that scales with the ray origin is covered
// per ray by the traversal's slab offsets), so no box that the exact test
// keeps is culled and the closest hit equals a brute-force search.
// Depth is capped at DR_MAX_DEPTH; stats->maxDepth is the 4-wide tree's
// depth, and a traversal lane needs at most 7 stack entries per level
// (three inner children and four leaf children) plus two. Leaves hold at
// most DR_LEAF_MAX triangles (larger ones become nodes of leaf chunks).
// A traversal stack entry of a leaf packs its count (4 bits) and first
// triangle (27 bits), so a scene holds fewer than DR_MAX_TRIANGLES.
void buildBvh(const DrTriangle *tris, uint32_t count, uint32_t maxLeaf, std::vector<DrNode4> &nodes,
              std::vector<float4> &triPos, BvhBuildStats *stats);

}  // namespace dr

#define DR_MAX_DEPTH 64
#define DR_LEAF_MAX 15u
#define DR_MAX_TRIANGLES (1u << 27)
