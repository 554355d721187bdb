// Host-built BVH with the surface-area heuristic (binned SAH).
//
// The reference indexes its scene with an octree (rt/raytrace.c:595-760,
// common/octree.h); here the index is a binary BVH whose internal node holds
// both children's FP32 boxes (BVHNode).  This builder produces the same node
// format as the device LBVH (rg_build.hip), with leaf refs holding ~primitive.
#pragma once
#include <stdint.h>

#include <vector>

#include "rg_geom.hpp"

namespace rg {

// lo/hi: 3 floats per box (already widened); ids: primitive index of each box.
// Fills `nodes` (n-1 internal nodes, root first, depth-first order) and
// returns the root ref: 0, or ~0 when n == 1 (the single primitive is
// ids[0]).  The tree depth is kept at most `max_depth` by falling back to
// median splits below that depth.
int32_t build_sah_bvh(const float* lo, const float* hi, const int32_t* ids, uint32_t n, std::vector<BVHNode>& nodes,
                      int max_depth);

// Four-wide form of a binary tree: one node per binary node a four-wide
// walk visits (the root, then the grandchildren of visited nodes), holding
// that node's grandchildren (a leaf child kept in place), numbered in
// depth-first preorder from 0 (the root).  Returns the largest number of
// four-wide nodes on a path from `root` to a leaf.
int collapse_bvh4(const std::vector<BVHNode>& nodes, int32_t root, std::vector<BVH4Node>& out);

}  // namespace rg
