// Generic bottom-up refit for arbitrary topologies (reference kFixAABBTree,
// lib/builder/LBVH_CUDA.cu:341-374, and kFixTree1,
// lib/builder/Bittner13CUDA.cu:372-404).
#pragma once

#include "runtime.hpp"

namespace bvht {

// Every listed leaf (leafList, or the range [leafFirst, leafFirst + count))
// walks to the root; the second arrival at a node merges its children.
// counters must hold numCounters zeroes; they are zero again on return.
bool refitFromLeaves(bvht_BVHNode *nodes, uint32_t leafFirst, uint32_t count, const uint32_t *leafList,
                     uint32_t *counters, uint32_t numCounters);

}  // namespace bvht
