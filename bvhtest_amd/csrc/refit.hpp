// Generic bottom-up refit for arbitrary topologies (reference kFixAABBTree,
// lib/builder/LBVH_CUDA.cu:341-374, and kFixTree1,
// lib/builder/Bittner13CUDA.cu:372-404).
#pragma once

#include "runtime.hpp"
#include "wave.hpp"

namespace bvht {

// A node's box (and area) read and written with device-coherent accesses
// (relaxed agent-scope atomics per word: loads read past the XCD's L2,
// stores write through it).  The bottom-up refits pass boxes between
// threads on different XCDs this way: the writer waits for its stores
// (waitVectorMemory) before its arrival count, the reader loads after its
// count returned, and neither needs an L2-wide write-back or invalidation.
__device__ __forceinline__ bvht_AABB loadBoxAgent(const bvht_BVHNode *nd) {
  bvht_AABB b;
  b.minMax[0].x = loadRelaxedAgent(&nd->bbox.minMax[0].x);
  b.minMax[0].y = loadRelaxedAgent(&nd->bbox.minMax[0].y);
  b.minMax[0].z = loadRelaxedAgent(&nd->bbox.minMax[0].z);
  b.minMax[1].x = loadRelaxedAgent(&nd->bbox.minMax[1].x);
  b.minMax[1].y = loadRelaxedAgent(&nd->bbox.minMax[1].y);
  b.minMax[1].z = loadRelaxedAgent(&nd->bbox.minMax[1].z);
  return b;
}
__device__ __forceinline__ void storeBoxAgent(bvht_BVHNode *nd, const bvht_AABB &b, float sa) {
  storeRelaxedAgent(&nd->bbox.minMax[0].x, b.minMax[0].x);
  storeRelaxedAgent(&nd->bbox.minMax[0].y, b.minMax[0].y);
  storeRelaxedAgent(&nd->bbox.minMax[0].z, b.minMax[0].z);
  storeRelaxedAgent(&nd->bbox.minMax[1].x, b.minMax[1].x);
  storeRelaxedAgent(&nd->bbox.minMax[1].y, b.minMax[1].y);
  storeRelaxedAgent(&nd->bbox.minMax[1].z, b.minMax[1].z);
  storeRelaxedAgent(&nd->surfaceArea, sa);
}


// Every listed leaf (leafList, or the range [leafFirst, leafFirst + count))
// walks to the root; the second arrival at a node merges its children.
// counters must hold numCounters zeroes; they are zero again on return.
bool refitFromLeaves(bvht_BVHNode *nodes, uint32_t leafFirst, uint32_t count, const uint32_t *leafList,
                     uint32_t *counters, uint32_t numCounters);

}  // namespace bvht
