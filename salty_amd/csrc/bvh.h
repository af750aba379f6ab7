// bvh.h -- host side of the acceleration structure: the reference binned-SAH
// tree (BVH8 with BVH4/Leaf fall-backs, or BVH4 / BVH2) built over triangle
// bounding boxes, whose depth-first order ranks tied hits; the tree the kernels
// traverse; and the re-layout of either into compressed 128-byte 8-wide nodes.
#pragma once
#include <stdint.h>
#include <vector>

#include "gpu_types.h"

namespace s3dg {

struct Aabb {
    float lo[3];
    float hi[3];
};

// Node of the reference-shaped tree.  kind: 8, 4, 2 or 0 (leaf).
struct RefNode {
    int32_t kind;
    int32_t first_child;   // index into RefTree::children
    int32_t child_count;
    int32_t begin;         // range of the depth-first prim order covered
    int32_t count;
    Aabb box;              // exact union of the prims' boxes
};

struct RefTree {
    std::vector<RefNode> nodes;
    std::vector<int32_t> children;
    std::vector<int32_t> order;   // depth-first position -> input prim index
    int32_t root = -1;
};

// Builds the tree the reference's BVH8::Create / BVH4::Create / BVH2::Create
// would build over prims with the given boxes (src/src/s3d_bvh8.cpp:254-479,
// src/src/s3d_bvh4.cpp:237-423, src/src/s3d_bvh2.cpp:174-274).  Prim centres
// are box centres, as Triangle::GetCenter returns (s3d_triangle.cpp:131-132).
void build_reference_tree(const std::vector<Aabb> &boxes, int kind, RefTree &out);

// Builds the tree the kernels traverse: binned SAH on all three axes (32
// bins), binary, leaves of at most max_leaf prims (more only where every centre
// coincides and the range cannot be split).  RefNode kinds are 2 and 0; order
// is the depth-first prim order of this tree.
// node_cost is the SAH cost of a binary node relative to one prim test.
void build_traversal_tree(const std::vector<Aabb> &boxes, int max_leaf, RefTree &out, float node_cost = 0.5f);

// Output of the re-layout.
struct Compressed {
    std::vector<GNode> nodes;
    int32_t max_depth = 0;
    uint32_t node_offset = 0;   // device index of nodes[0] (node indices written are offset by it)
};

// Appends the compressed form of `tree` (whose prims have the boxes
// prim_boxes, in input order) to out.nodes.  The prims are given a new order
// in which every node's leaves are contiguous: leaf_order[j] is the
// depth-first position of the prim stored at tri_offset + j (triangles, or
// shape records for top-level trees).  Returns the root node index (offset by
// out.node_offset, like every node index written).  `margin`
// is an absolute amount by which every child box is grown before
// quantisation (outward to the 8-bit grid); the scenes pass 0, since the
// kernels' slab tests widen their distances by a bound on their rounding.
uint32_t compress_tree(const RefTree &tree, const std::vector<Aabb> &prim_boxes, uint32_t tri_offset,
                       bool top_level, float margin, Compressed &out, std::vector<int32_t> &leaf_order);

}  // namespace s3dg
