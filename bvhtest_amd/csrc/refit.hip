// Generic bottom-up refit (see refit.hpp).
//
// Arrival counting uses a relaxed agent-scope counter; the box a thread
// writes for a node must be visible to the sibling's thread, which may run
// on another XCD with its own L2, so boxes move through device-coherent
// loads and stores (loadBoxAgent / storeBoxAgent) and the stores are waited
// for before the next count.  Every walk is bounded by the node count, so a
// malformed tree cannot hang the kernel.
#include "refit.hpp"
#include "wave.hpp"

namespace bvht {

namespace {

__global__ __launch_bounds__(256) void k_refitUp(bvht_BVHNode *__restrict__ nodes, uint32_t leafFirst, uint32_t count,
                                                 const uint32_t *__restrict__ leafList, uint32_t *__restrict__ counters,
                                                 uint32_t numNodes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t leaf = leafList ? leafList[i] : leafFirst + i;
  uint32_t node = nodes[leaf].parent;
  if (node == leaf || node >= numNodes) return;
  for (uint32_t steps = 0; steps < numNodes; ++steps) {
    // the leaf boxes come from an earlier kernel; every box written below is
    // complete (waited for) before the next arrival count.  The topology is
    // not written here: its loads go out beside the count.
    const uint32_t l = loadRelaxedAgent(&nodes[node].left), r = loadRelaxedAgent(&nodes[node].right);
    const uint32_t parent = loadRelaxedAgent(&nodes[node].parent);
    const uint32_t seen = __hip_atomic_fetch_add(counters + node, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (seen == 0) return;
    bvht_AABB m = loadBoxAgent(nodes + l);
    mergeInto(m, loadBoxAgent(nodes + r));
    storeBoxAgent(nodes + node, m, surfaceArea(m));
    waitVectorMemory();
    if (parent == node || parent >= numNodes) return;
    node = parent;
  }
}

}  // namespace

bool refitFromLeaves(bvht_BVHNode *nodes, uint32_t leafFirst, uint32_t count, const uint32_t *leafList,
                     uint32_t *counters, uint32_t numCounters) {
  if (count == 0) return true;
  k_refitUp<<<divUp(count, 256), 256, 0, stream()>>>(nodes, leafFirst, count, leafList, counters, numCounters);
  if (!checkLaunch("k_refitUp")) return false;
  return fillU32(counters, 0u, numCounters);
}

}  // namespace bvht
