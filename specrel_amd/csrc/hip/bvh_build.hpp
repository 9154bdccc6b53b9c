// Device build of the BVH (Bvh4D::build, bvh4d.rs:122-427) that emits the host builder's tree
// bit for bit (same nodes in the same preorder, same leaf order), for specrel_build_bvh.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "specrel_types.h"

namespace dev {

struct BvhBuildStats {
    float ms;                  // device time of the build (kernels and the few host decisions)
    uint32_t levels;           // breadth-first levels of large segments
    uint32_t large_segments;   // segments partitioned level-parallel
    uint32_t small_subtrees;   // subtrees built one per thread
    uint32_t fallback_segments;
    uint32_t jump_passes;      // pointer-jumping passes of the partitions
};

// Builds the tree over n records whose padded bounds are d_boxes (2 float4 per record: min, max,
// on the device). d_nodes (capacity 2n - 1 nodes) and d_order (n) receive the nodes in preorder
// and the leaf order; *n_nodes the node count. Returns an empty string or an error message.
std::string build_bvh(const float4 *d_boxes, uint32_t n, specrel_node *d_nodes, uint32_t *d_order,
                      uint32_t *n_nodes, hipStream_t s, BvhBuildStats *stats);

}  // namespace dev
