// Bittner et al. 2013 parallel SAH optimizer for gfx950
// (reference: lib/builder/Bittner13CUDA.cu, lib/base/BVHPatch.hpp).
//
// One step (bn13_doAlgorithmStep):
//   k_cost      removal priority of every inner node, SA^3*2/((L+R)*min(L,R))
//               (Bittner13CUDA.cu:480-500), as a descending sort key;
//   onesweep    stable sort of (priority, node id) -- the processing order;
//   per chunk of K nodes (bn13_rmAndReinsChunk):
//     k_search  one 64-lane wave per node: remove the node (patch), then two
//               reinsertions, each found by a wave-parallel exact branch and
//               bound over the patched view (the reference's findNode1 keeps
//               a 4096-entry heap per thread in local memory); the patch
//               reserves its nodes (round 1);
//     k_commitFirst / k_resolveTail  deterministic reservations: exactly the
//               patches a rank-ordered first-come locking applies (the
//               reference's kCheckConflicts, timing-dependent there);
//     k_contain a winner whose insertion point lies in a subtree another
//               winner moves is dropped (two such patches would form a cycle;
//               the reference has no guard);
//     k_apply   winners write their topology, all locks are released;
//     k_refit   two-phase targeted refit from the 3 changed nodes of every
//               winner (kFixTree3_1 / kFixTree3_2 protocol, agent-scope
//               acquire/release atomics).  Boxes are exact min/max, so this
//               equals the reference's full refit (kFixTree1).
// The root never moves (reinsertion at the root is refused, :310).
#include "mlp.hpp"
#include "radix_sort.hpp"
#include "refit.hpp"
#include "wave.hpp"

#include <cstring>
#include <new>
#include <vector>

namespace bvht {

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int kNNode = 10;  // BVHPatch.hpp:40
constexpr uint32_t P_GRAND_PARENT = 0, P_1ST_ROOT = 1, P_1ST_BEST = 2, P_NODE = 3, P_1ST_INSERT = 4, P_2ND_ROOT = 5,
                   P_2ND_BEST = 6, P_2ND_INSERT = 7, P_SIBLING = 8, P_PARENT = 9;  // BVHPatch.hpp:44-53
constexpr int kStackCap = 1024;  // per-search frontier (findNode1 heap: 4096)
constexpr uint32_t kAltSlots = 16;  // findNode2's array (CUDA_ALT_QUEUE_SIZE, Bittner13CUDA.hpp:28)

// Stats words (device)
enum { ST_CONFLICT = 0, ST_FAILED, ST_CONTAIN, ST_OVERFLOW, ST_APPLIED, ST_UNRESOLVED, ST_RUNAWAY, ST_SORTERR, ST_WORDS = 8 };

// Patch states and node locks.  A lock holds key(round, rank): the upper word
// is ~round, so later rounds and applied patches (round kRoundWin, upper word
// 0) compare smaller.
constexpr uint64_t kLockFree = ~0ull;
constexpr uint32_t kRoundWin = 0xFFFFFFFFu;
enum : uint32_t { REC_NONE = 0, REC_PENDING = 1, REC_LOST = 2, REC_CONTAINED = 3, REC_WON = 4 };

__device__ __forceinline__ uint64_t lockKey(uint32_t round, uint32_t k) {
  return ((uint64_t)(0xFFFFFFFFu - round) << 32) | k;
}

struct PatchRec {
  uint32_t node[kNNode];
  uint32_t parent[kNNode];
  uint32_t left[kNNode];
  uint32_t right[kNNode];
  uint32_t isLeft[kNNode];
  uint32_t moved[2];  // reinserted subtree roots
  uint32_t best[2];   // insertion points
  uint32_t fix[3];    // refit starts: grandparent, unused1, unused2
  uint32_t ok;
};

// BVHPatch (ENABLE_AABB_PATH == 0) in LDS, manipulated by lane 0.
struct Patch {
  uint32_t vPatch[kNNode];
  uint32_t parent[kNNode], left[kNNode], right[kNNode];
  uint16_t isLeaf[kNNode], isLeft[kNNode];
  uint32_t path[4];
  float box[4][6];
};

struct NodeView {
  uint32_t parent, left, right, isLeaf, isLeft;
};

__device__ __forceinline__ NodeView globalView(const bvht_BVHNode *nodes, uint32_t n) {
  const uint4 w = reinterpret_cast<const uint4 *>(nodes + n)[1];  // maxY maxZ parent left
  const uint4 z = reinterpret_cast<const uint4 *>(nodes + n)[2];  // right flags level sa
  return NodeView{w.z, w.w, z.x, z.y & 0xFFFFu, z.y >> 16};
}

__device__ __forceinline__ void globalBox(const bvht_BVHNode *nodes, uint32_t n, float b[6]) {
  const uint4 a = reinterpret_cast<const uint4 *>(nodes + n)[0];
  const uint4 w = reinterpret_cast<const uint4 *>(nodes + n)[1];
  b[0] = __uint_as_float(a.x);
  b[1] = __uint_as_float(a.y);
  b[2] = __uint_as_float(a.z);
  b[3] = __uint_as_float(a.w);
  b[4] = __uint_as_float(w.x);
  b[5] = __uint_as_float(w.y);
}

__device__ __forceinline__ float sa6(const float b[6]) { return surfaceArea(b[0], b[1], b[2], b[3], b[4], b[5]); }
__device__ __forceinline__ void merge6(float a[6], const float b[6]) {
  a[0] = mergeMin(a[0], b[0]);
  a[1] = mergeMin(a[1], b[1]);
  a[2] = mergeMin(a[2], b[2]);
  a[3] = mergeMax(a[3], b[3]);
  a[4] = mergeMax(a[4], b[4]);
  a[5] = mergeMax(a[5], b[5]);
}

// ---- patch operations (lane 0) -------------------------------------------

__device__ void patchClear(Patch &P) {
  for (int i = 0; i < kNNode; ++i) P.vPatch[i] = NONE;
  for (int i = 0; i < 4; ++i) P.path[i] = NONE;
}
__device__ uint32_t patchIndex(const Patch &P, uint32_t n) {
  for (int i = 0; i < kNNode; ++i)
    if (P.vPatch[i] == n) return i;
  return NONE;
}
// patchIndexSubset -- BVHPatch.hpp:134-142
__device__ __forceinline__ uint32_t patchIndexSubset(const Patch &P, uint32_t n) {
  if (P.vPatch[0] == n) return 0;
  if (P.vPatch[1] == NONE) return NONE;
  if (P.vPatch[1] == n) return 1;
  if (P.vPatch[2] == n) return 2;
  if (P.vPatch[3] == n) return 3;
  return NONE;
}
__device__ NodeView patchGet(const Patch &P, const bvht_BVHNode *nodes, uint32_t n) {
  uint32_t i = patchIndex(P, n);
  if (i == NONE) return globalView(nodes, n);
  return NodeView{P.parent[i], P.left[i], P.right[i], P.isLeaf[i], P.isLeft[i]};
}
__device__ __forceinline__ NodeView patchGetSubset(const Patch &P, const bvht_BVHNode *nodes, uint32_t n) {
  uint32_t i = patchIndexSubset(P, n);
  if (i == NONE) return globalView(nodes, n);
  return NodeView{P.parent[i], P.left[i], P.right[i], P.isLeaf[i], P.isLeft[i]};
}
__device__ void patchNode(Patch &P, const bvht_BVHNode *nodes, uint32_t n, uint32_t slot) {
  NodeView v = globalView(nodes, n);
  P.vPatch[slot] = n;
  P.parent[slot] = v.parent;
  P.left[slot] = v.left;
  P.right[slot] = v.right;
  P.isLeaf[slot] = (uint16_t)v.isLeaf;
  P.isLeft[slot] = (uint16_t)v.isLeft;
}
__device__ void movePatch(Patch &P, uint32_t from, uint32_t to) {
  P.vPatch[to] = P.vPatch[from];
  P.vPatch[from] = NONE;
  P.parent[to] = P.parent[from];
  P.left[to] = P.left[from];
  P.right[to] = P.right[from];
  P.isLeaf[to] = P.isLeaf[from];
  P.isLeft[to] = P.isLeft[from];
}
// getAABB -- BVHPatch.hpp:319-325 (first cached path box that matches)
__device__ __forceinline__ void patchBox(const Patch &P, const bvht_BVHNode *nodes, uint32_t n, float b[6]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (P.path[i] == n) {
#pragma unroll
      for (int k = 0; k < 6; ++k) b[k] = P.box[i][k];
      return;
    }
  globalBox(nodes, n, b);
}
// patchAABBFrom -- BVHPatch.hpp:293-316
__device__ void patchAABBFrom(Patch &P, const bvht_BVHNode *nodes, uint32_t node, uint32_t pIdx, uint32_t root) {
  uint32_t idx = node;
  NodeView cur = patchGetSubset(P, nodes, idx);
  float box[6], s[6];
  patchBox(P, nodes, cur.left, box);
  bool lastWasLeft = true;
  for (uint32_t i = 0; i < 2; ++i) {
    uint32_t sib = lastWasLeft ? cur.right : cur.left;
    patchBox(P, nodes, sib, s);
    merge6(box, s);
    for (int k = 0; k < 6; ++k) P.box[pIdx * 2 + i][k] = box[k];
    P.path[pIdx * 2 + i] = idx;
    lastWasLeft = cur.isLeft != 0;
    if (idx == root) break;
    idx = cur.parent;
    cur = patchGetSubset(P, nodes, idx);
  }
}

// removeNode -- Bittner13CUDA.cu:267-305 (a patch that cannot be built is
// dropped whole instead of being applied half done).
__device__ bool removeNode(Patch &P, const bvht_BVHNode *nodes, uint32_t node, uint32_t root, uint32_t ins[2],
                           uint32_t unused[2], uint32_t &gpOut) {
  NodeView nv = globalView(nodes, node);
  if (nv.isLeaf || node == root) return false;
  if (nv.parent == root) return false;
  patchNode(P, nodes, node, P_NODE);
  NodeView pn = patchGet(P, nodes, P.parent[P_NODE]);
  uint32_t sib = P.isLeft[P_NODE] == 0 ? pn.left : pn.right;
  uint32_t parentIdx = P.parent[P_NODE];
  patchNode(P, nodes, sib, P_SIBLING);
  patchNode(P, nodes, parentIdx, P_PARENT);
  uint32_t gp = P.parent[P_PARENT];
  patchNode(P, nodes, gp, P_GRAND_PARENT);
  if (P.isLeft[P_PARENT] != 0) {
    P.left[P_GRAND_PARENT] = sib;
    P.isLeft[P_SIBLING] = 1;
  } else {
    P.right[P_GRAND_PARENT] = sib;
    P.isLeft[P_SIBLING] = 0;
  }
  P.parent[P_SIBLING] = gp;
  patchAABBFrom(P, nodes, gp, 0, root);
  const float lsa = nodes[P.left[P_NODE]].surfaceArea, rsa = nodes[P.right[P_NODE]].surfaceArea;
  if (lsa > rsa) {
    ins[0] = P.left[P_NODE];
    ins[1] = P.right[P_NODE];
  } else {
    ins[0] = P.right[P_NODE];
    ins[1] = P.left[P_NODE];
  }
  unused[0] = node;
  unused[1] = parentIdx;
  gpOut = gp;
  return true;
}

// Splice of reinsert -- Bittner13CUDA.cu:312-355 (the search is done before).
__device__ void spliceAt(Patch &P, const bvht_BVHNode *nodes, uint32_t node, uint32_t unusedIdx, uint32_t bestIdx,
                         bool update, uint32_t root) {
  uint32_t bp = patchIndex(P, bestIdx);
  uint32_t bs;
  if (bp == NONE) {
    bs = update ? P_1ST_BEST : P_2ND_BEST;
    patchNode(P, nodes, bestIdx, bs);
  } else if (bp == P_GRAND_PARENT) {
    bs = P_GRAND_PARENT;
  } else {
    bs = update ? P_1ST_BEST : P_2ND_BEST;
    movePatch(P, bp, bs);
  }
  const uint32_t ns = update ? P_1ST_INSERT : P_2ND_INSERT;
  patchNode(P, nodes, node, ns);
  const uint32_t us = update ? P_NODE : P_PARENT;
  const uint32_t rootIdx = P.parent[bs];
  uint32_t rp = patchIndex(P, rootIdx);
  uint32_t rs = update ? P_1ST_ROOT : P_2ND_ROOT;
  if (rp == NONE)
    patchNode(P, nodes, rootIdx, rs);
  else
    movePatch(P, rp, rs);
  // a move may have relocated the best node's slot
  if (P.vPatch[bs] != bestIdx) bs = patchIndex(P, bestIdx);
  if (P.isLeft[bs] != 0) {
    P.left[rs] = unusedIdx;
    P.isLeft[us] = 1;
  } else {
    P.right[rs] = unusedIdx;
    P.isLeft[us] = 0;
  }
  P.parent[us] = rootIdx;
  P.left[us] = bestIdx;
  P.right[us] = node;
  P.parent[bs] = unusedIdx;
  P.isLeft[bs] = 1;
  P.parent[ns] = unusedIdx;
  P.isLeft[ns] = 0;
  if (update) patchAABBFrom(P, nodes, unusedIdx, 1, root);
}

// ---- wave-parallel branch and bound (findNode1, Bittner13CUDA.cu:141-191) --
//
// Exact: an entry is expanded only while induced + SA(x) < best, which is a
// lower bound of the cost of every node below it, so the returned node is a
// cost minimiser of the patched view (ties: the first step that reaches the
// minimum, smallest node id inside it).  Frontier: LDS stack, up to 64
// entries evaluated per step.  Every step pops at least one entry and a tree
// pushes each node at most once, so a search of a tree ends within `maxSteps`
// = node count steps; a search still running then (a patched view that is
// not a tree) returns NONE and its patch is dropped.
__device__ uint32_t waveSearch(const Patch &P, const bvht_BVHNode *__restrict__ nodes, uint32_t x, uint32_t root,
                               uint32_t *s_node, float *s_cost, uint32_t *s_top, int cap, uint32_t maxSteps,
                               uint32_t *overflow, float &bestCost) {
  const uint32_t lane = threadIdx.x;
  float xb[6];
  globalBox(nodes, x, xb);  // getOrig(_n)->bbox
  const float sa = nodes[x].surfaceArea;
  if (lane == 0) {
    s_node[0] = root;
    s_cost[0] = 0.0f;
    *s_top = 1;
  }
  __syncthreads();
  float best = HUGE_VALF;
  uint32_t bestNode = 0;
  const uint64_t lt = __lanemask_lt();
  uint32_t top = 1;
  for (uint32_t step = 0; top > 0; ++step) {
    if (step >= maxSteps) return NONE;  // uniform: step and top are wave-uniform
    const uint32_t m = top < 64 ? top : 64;
    const bool active = lane < m;
    uint32_t node = 0;
    float ind = 0.0f;
    if (active) {
      node = s_node[top - 1 - lane];
      ind = s_cost[top - 1 - lane];
    }
    __syncthreads();
    const uint32_t newTop = top - m;
    bool eval = active && (ind + sa) < best;
    float total = HUGE_VALF, newInd = 0.0f;
    NodeView nv = {0, 0, 0, 1, 0};
    if (eval) {
      nv = patchGetSubset(P, nodes, node);
      float b[6];
      patchBox(P, nodes, node, b);
      newInd = -1.0f * sa6(b);
      merge6(b, xb);
      total = ind + sa6(b);
      newInd += total;
    }
    const float mt = waveMin(total);
    if (mt < best) {
      uint64_t tied = __ballot(eval && total == mt);
      uint32_t cand = (eval && total == mt) ? node : NONE;
      // smallest node id among the tied lanes
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        uint32_t y = __shfl_xor(cand, o, 64);
        cand = y < cand ? y : cand;
      }
      (void)tied;
      best = mt;
      bestNode = cand;
    }
    const bool push = eval && nv.isLeaf == 0 && (newInd + sa) < best;
    const uint64_t pm = __ballot(push);
    const uint32_t pos = newTop + 2u * (uint32_t)__popcll(pm & lt);
    const uint32_t cnt = 2u * (uint32_t)__popcll(pm);
    if (push) {
      if (pos + 1 < (uint32_t)cap) {
        s_node[pos] = nv.left;
        s_cost[pos] = newInd;
        s_node[pos + 1] = nv.right;
        s_cost[pos + 1] = newInd;
      }
    }
    uint32_t nt = newTop + cnt;
    if (nt > (uint32_t)cap) {
      if (lane == 0) atomicAdd(overflow, 1u);
      nt = (uint32_t)cap & ~1u;
    }
    top = nt;
    __syncthreads();
  }
  bestCost = best;
  return bestNode;
}

// ---- findNode2 (altFindNode), Bittner13CUDA.cu:194-264 ---------------------
//
// The reference's approximate search, exactly: a best-first search over a
// fixed array of 16 (node, induced cost) slots, held here by lanes 0..15 of
// the wave (one slot per lane).  Each iteration pops the slot of minimum
// cost; an expanded node's left child takes that slot and its right child
// the slot of maximum cost found by the previous scan (an empty slot holds
// HUGE_VALF, so it is the maximum while one is free; once the array is full
// the costliest entry is overwritten).  Both scans keep the first slot of
// their extreme (the reference's strict < and > loops); the maximum scan
// starts from 0 and keeps the previous slot when no cost is above 0.  NaN
// costs never compare, so they are mapped out of both scans.  One entry is
// popped per iteration and a tree pushes each node at most once, so a
// search of a tree ends within `maxIters` = node count iterations; a longer
// one (a patched view that is not a tree) returns NONE.
__device__ uint32_t altSearch(const Patch &P, const bvht_BVHNode *__restrict__ nodes, uint32_t x, uint32_t root,
                              uint32_t maxIters, float &bestCost) {
  const uint32_t lane = threadIdx.x;
  const bool slot = lane < kAltSlots;
  float xb[6];
  globalBox(nodes, x, xb);  // getOrig(_n)->bbox
  const float sa = nodes[x].surfaceArea;
  uint32_t slotNode = root;
  float slotCost = lane == 0 ? 0.0f : HUGE_VALF;
  uint32_t minIdx = 0, maxIdx = 1;
  float best = HUGE_VALF, lMin = 0.0f;
  uint32_t bestNode = 0;
  for (uint32_t it = 0; lMin < HUGE_VALF; ++it) {  // lMin, minIdx, maxIdx are wave-uniform
    if (it >= maxIters) return NONE;
    const uint32_t cur = (uint32_t)__shfl((int)slotNode, (int)minIdx, 64);
    const float curCost = __shfl(slotCost, (int)minIdx, 64);
    if (lane == minIdx) slotCost = HUGE_VALF;
    if ((curCost + sa) >= best) break;
    const NodeView nv = patchGetSubset(P, nodes, cur);
    float b[6];
    patchBox(P, nodes, cur, b);
    float newInd = -1.0f * sa6(b);
    merge6(b, xb);
    const float total = curCost + sa6(b);
    if (total < best) {
      best = total;
      bestNode = cur;
    }
    newInd += total;
    if ((newInd + sa) < best && nv.isLeaf == 0) {
      if (lane == minIdx) {
        slotNode = nv.left;
        slotCost = newInd;
      }
      if (lane == maxIdx) {  // written second: takes the slot if both are one
        slotNode = nv.right;
        slotCost = newInd;
      }
    }
    const bool num = slot && !__builtin_isnan(slotCost);
    const float cMin = num ? slotCost : HUGE_VALF, cMax = num ? slotCost : 0.0f;
    lMin = waveMin(cMin);
    const float lMax = waveMax(cMax);
    if (lMin < HUGE_VALF) minIdx = (uint32_t)__builtin_ctzll(__ballot(num && cMin == lMin));
    if (lMax > 0.0f) maxIdx = (uint32_t)__builtin_ctzll(__ballot(num && cMax == lMax));
  }
  bestCost = best;
  return bestNode;
}

// Destination of SAH-gain records (feat == nullptr: not recording).
struct RecordSink {
  const float *feat;  // features by inner-list position
  float *recFeat;     // [recMax][12]
  float *recGain;     // [recMax]
  uint32_t *recCount;
  uint32_t recMax;
};

// Gain of a patch in unit-cost SAH relative to the root: the areas of the
// two removed nodes plus the shrink of the two path boxes above the removal
// point, minus the two insertion costs the searches found.
__device__ float patchGain(const Patch &P, const bvht_BVHNode *nodes, uint32_t node, uint32_t parent, uint32_t root,
                           float c1, float c2) {
  float shrink = 0.0f;
  for (int i = 0; i < 2; ++i)
    if (P.path[i] != NONE) shrink += nodes[P.path[i]].surfaceArea - sa6(P.box[i]);
  const float g = nodes[node].surfaceArea + nodes[parent].surfaceArea + shrink - c1 - c2;
  const float sr = nodes[root].surfaceArea;
  return sr > 0.0f ? g / sr : 0.0f;
}

// PROBE: evaluate the patches of the nodes at inner-list positions
// order[blockIdx.x] and record their gains; nothing is reserved or applied.
template <bool PROBE>
__global__ __launch_bounds__(64) void k_search(const bvht_BVHNode *__restrict__ nodes, uint32_t root, uint32_t N,
                                               const uint32_t *__restrict__ order, const uint32_t *__restrict__ inner,
                                               uint32_t numInner, uint32_t chunk, uint32_t numChunks, uint32_t K,
                                               uint32_t offsetAccess, uint32_t alt, PatchRec *__restrict__ recs,
                                               uint64_t *locks, uint32_t *__restrict__ stats, RecordSink sink,
                                               const uint32_t *__restrict__ sortErr) {
  __shared__ Patch P;
  __shared__ uint32_t s_node[kStackCap];
  __shared__ float s_cost[kStackCap];
  __shared__ uint32_t s_top;
  __shared__ uint32_t s_ok, s_ins[2], s_unused[2], s_gp, s_best[2];
  const uint32_t k = blockIdx.x, lane = threadIdx.x;
  if (sortErr && k == 0 && lane == 0 && *sortErr) atomicAdd(stats + ST_SORTERR, 1u);
  PatchRec *rec = recs + k;
  const uint32_t pos = PROBE ? order[k] : order[offsetAccess ? k * numChunks + chunk : chunk * K + k];
  if (pos >= numInner) {  // block-uniform; only a bad probe list gets here
    if (!PROBE && lane == 0) rec->ok = REC_NONE;
    return;
  }
  const uint32_t nodeIdx = inner[pos];
  float c1 = 0.0f, c2 = 0.0f;
  if (lane == 0) {
    patchClear(P);
    s_ok = removeNode(P, nodes, nodeIdx, root, s_ins, s_unused, s_gp) ? 1u : 0u;
  }
  __syncthreads();
  if (s_ok) {
    const uint32_t b1 = alt ? altSearch(P, nodes, s_ins[0], root, N, c1)
                            : waveSearch(P, nodes, s_ins[0], root, s_node, s_cost, &s_top, kStackCap, N,
                                         stats + ST_OVERFLOW, c1);
    if (lane == 0) {
      if (b1 == root || b1 == NONE) {
        if (b1 == NONE) atomicAdd(stats + ST_RUNAWAY, 1u);
        s_ok = 0;
      } else {
        s_best[0] = b1;
        spliceAt(P, nodes, s_ins[0], s_unused[0], b1, true, root);
      }
    }
    __syncthreads();
  }
  if (s_ok) {
    const uint32_t b2 = alt ? altSearch(P, nodes, s_ins[1], root, N, c2)
                            : waveSearch(P, nodes, s_ins[1], root, s_node, s_cost, &s_top, kStackCap, N,
                                         stats + ST_OVERFLOW, c2);
    if (lane == 0) {
      if (b2 == root || b2 == NONE) {
        if (b2 == NONE) atomicAdd(stats + ST_RUNAWAY, 1u);
        s_ok = 0;
      } else {
        s_best[1] = b2;
        spliceAt(P, nodes, s_ins[1], s_unused[1], b2, false, root);
      }
    }
    __syncthreads();
  }
  if (!s_ok) {
    if (!PROBE && lane == 0) {
      rec->ok = REC_NONE;
      atomicAdd(stats + ST_FAILED, 1u);
    }
    return;
  }
  if (sink.feat && lane == 0) {
    const uint32_t slot = atomicAdd(sink.recCount, 1u);
    if (slot < sink.recMax) {
      for (int f = 0; f < kFeat; ++f) sink.recFeat[(size_t)slot * kFeat + f] = sink.feat[(size_t)pos * kFeat + f];
      sink.recGain[slot] = patchGain(P, nodes, nodeIdx, s_unused[1], root, c1, c2);
    }
  }
  if (PROBE) return;
  if (lane < (uint32_t)kNNode) {
    const uint32_t n = P.vPatch[lane];
    rec->node[lane] = n;
    rec->parent[lane] = P.parent[lane];
    rec->left[lane] = P.left[lane];
    rec->right[lane] = P.right[lane];
    rec->isLeft[lane] = P.isLeft[lane];
    // round-1 reservation of the deterministic conflict resolution
    if (n != NONE) atomicMin(reinterpret_cast<unsigned long long *>(locks + n), (unsigned long long)lockKey(1, k));
  }
  if (lane == 0) {
    rec->moved[0] = s_ins[0];
    rec->moved[1] = s_ins[1];
    rec->best[0] = s_best[0];
    rec->best[1] = s_best[1];
    rec->fix[0] = s_gp;
    rec->fix[1] = s_unused[0];
    rec->fix[2] = s_unused[1];
    rec->ok = REC_PENDING;
  }
}

// ---- conflict resolution: deterministic reservations ---------------------
//
// kCheckConflicts (Bittner13CUDA.cu:567-609) lets patch k lock its nodes
// first come, first served; serialised in rank order this is the greedy rule
// "k is applied iff no node of k belongs to an applied patch of lower rank".
// That rule is computed here exactly and independently of timing: in round r
// every undecided patch reserves its nodes with atomicMin(key(r, k)); a patch
// that holds every reservation is applied (its nodes get key(WIN, k), smaller
// than any reservation), a patch that meets an applied patch's node loses,
// the rest retry in round r + 1.  The undecided patch of lowest rank always
// decides, so a round never leaves the set unchanged.
__device__ __forceinline__ void reserve(const PatchRec *r, uint32_t k, uint32_t round, uint64_t *locks) {
  const uint64_t key = lockKey(round, k);
  for (int i = 0; i < kNNode; ++i) {
    const uint32_t n = r->node[i];
    if (n != NONE) atomicMin(reinterpret_cast<unsigned long long *>(locks + n), (unsigned long long)key);
  }
}

// Returns REC_WON, REC_LOST or REC_PENDING and records a decision.
__device__ uint32_t decide(PatchRec *r, uint32_t k, uint32_t round, uint64_t *locks, uint32_t *marks,
                           uint32_t *skipped, uint32_t *stats) {
  const uint64_t mine = lockKey(round, k);
  bool holdAll = true, lost = false;
  for (int i = 0; i < kNNode; ++i) {
    const uint32_t n = r->node[i];
    if (n == NONE) continue;
    const uint64_t v = __hip_atomic_load(locks + n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((v >> 32) == 0) lost = true;  // node of an applied patch
    if (v != mine) holdAll = false;
  }
  if (lost) {
    r->ok = REC_LOST;
    skipped[k] += 1;
    atomicAdd(stats + ST_CONFLICT, 1u);
    return REC_LOST;
  }
  if (!holdAll) return REC_PENDING;
  const uint64_t win = lockKey(kRoundWin, k);
  for (int i = 0; i < kNNode; ++i) {
    const uint32_t n = r->node[i];
    if (n != NONE) atomicMin(reinterpret_cast<unsigned long long *>(locks + n), (unsigned long long)win);
  }
  r->ok = REC_WON;
  marks[r->moved[0]] = k;  // moved roots are patch nodes: one applied patch each
  marks[r->moved[1]] = k;
  return REC_WON;
}

// Round 1 over every patch (the reservations were made by k_search).
__global__ __launch_bounds__(256) void k_commitFirst(PatchRec *__restrict__ recs, uint32_t K, uint64_t *locks,
                                                     uint32_t *__restrict__ marks, uint32_t *__restrict__ skipped,
                                                     uint32_t *__restrict__ stats, uint32_t *__restrict__ pending,
                                                     uint32_t *__restrict__ numPending) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= K) return;
  PatchRec *r = recs + k;
  if (r->ok != REC_PENDING) return;
  if (decide(r, k, 1, locks, marks, skipped, stats) == REC_PENDING) pending[atomicAdd(numPending, 1u)] = k;
}

// Rounds 2.. over the (few) undecided patches, in one workgroup.  At most
// `num + 1` rounds are needed; patches still undecided after that bound
// (never, unless memory ordering is broken) are dropped and counted.
//
// Up to kTailBlock patches (one per thread), the rounds run in LDS and
// registers: the nodes of the undecided patches go into an LDS hash table
// holding, per node, whether a winner took it (round 1's from its global
// lock, read once) and the lowest rank reserving it this round, and each
// thread keeps its patch's table slots in registers.  A round is then LDS
// atomics and barriers instead of global atomics and loads.  The decisions
// are those of the global rounds: in round r a patch reserves every node no
// winner took, wins if it holds them all, loses if a winner took one, and
// waits otherwise.  Global locks keep their round-1 values (k_apply
// releases them all).
constexpr int kTailBlock = 1024;
constexpr uint32_t kTailSlotsLog = 14, kTailSlots = 1u << kTailSlotsLog;  // LDS hash slots (load <= 0.63)
constexpr uint32_t kTailEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t tailHash(uint32_t n) { return (n * 0x9E3779B1u) >> (32 - kTailSlotsLog); }

// Slot of node n, inserted if absent (the table never fills).
__device__ __forceinline__ uint32_t tailInsert(uint32_t *keys, uint32_t n) {
  uint32_t h = tailHash(n);
  for (uint32_t probe = 0; probe < kTailSlots; ++probe) {
    const uint32_t prev = atomicCAS(keys + h, kTailEmpty, n);
    if (prev == kTailEmpty || prev == n) return h;
    h = (h + 1) & (kTailSlots - 1);
  }
  return kTailEmpty;  // unreachable
}

__global__ __launch_bounds__(kTailBlock) void k_resolveTail(PatchRec *__restrict__ recs, uint64_t *locks,
                                                           uint32_t *__restrict__ marks,
                                                           uint32_t *__restrict__ skipped, uint32_t *__restrict__ stats,
                                                           uint32_t *listA, uint32_t *listB,
                                                           const uint32_t *__restrict__ numPending) {
  __shared__ uint32_t s_key[kTailSlots];   // node id, kTailEmpty
  __shared__ uint32_t s_min[kTailSlots];   // lowest reserving rank this round
  __shared__ uint8_t s_taken[kTailSlots];  // taken by a winner
  __shared__ uint32_t s_cnt;
  uint32_t num = *numPending;
  const uint32_t maxRound = num + 2;
  uint32_t round = 2;
  if (num <= (uint32_t)kTailBlock) {
    for (uint32_t i = threadIdx.x; i < kTailSlots; i += kTailBlock) s_key[i] = kTailEmpty;
    __syncthreads();
    const bool mine = threadIdx.x < num;
    const uint32_t k = mine ? listA[threadIdx.x] : 0u;
    uint32_t sl[kNNode];
    for (int j = 0; j < kNNode; ++j) {
      const uint32_t n = mine ? recs[k].node[j] : NONE;
      sl[j] = n == NONE ? kTailEmpty : tailInsert(s_key, n);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kTailSlots; i += kTailBlock) {
      s_min[i] = kTailEmpty;
      const uint32_t n = s_key[i];
      s_taken[i] = n != kTailEmpty &&
                   (__hip_atomic_load(locks + n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) == 0;
    }
    bool open = mine;  // undecided
    __syncthreads();
    for (; num > 0 && round <= maxRound; ++round) {
      if (open)  // reserve
        for (int j = 0; j < kNNode; ++j)
          if (sl[j] != kTailEmpty && !s_taken[sl[j]]) atomicMin(s_min + sl[j], k);
      if (threadIdx.x == 0) s_cnt = 0;
      __syncthreads();
      if (open) {  // decide
        bool holdAll = true, lost = false;
        for (int j = 0; j < kNNode; ++j) {
          if (sl[j] == kTailEmpty) continue;
          if (s_taken[sl[j]]) lost = true;
          if (s_min[sl[j]] != k) holdAll = false;
        }
        PatchRec *r = recs + k;
        if (lost) {
          r->ok = REC_LOST;
          skipped[k] += 1;
          atomicAdd(stats + ST_CONFLICT, 1u);
          open = false;
        } else if (holdAll) {
          for (int j = 0; j < kNNode; ++j)
            if (sl[j] != kTailEmpty) s_taken[sl[j]] = 1;
          r->ok = REC_WON;
          marks[r->moved[0]] = k;  // moved roots are patch nodes: one applied patch each
          marks[r->moved[1]] = k;
          open = false;
        } else {
          atomicAdd(&s_cnt, 1u);
        }
      }
      __syncthreads();
      num = s_cnt;
      for (uint32_t i = threadIdx.x; i < kTailSlots; i += kTailBlock) s_min[i] = kTailEmpty;
      __syncthreads();
    }
    if (open) {
      recs[k].ok = REC_LOST;
      skipped[k] += 1;
      atomicAdd(stats + ST_CONFLICT, 1u);
      atomicAdd(stats + ST_UNRESOLVED, 1u);
    }
    return;
  }
  // more undecided patches than the LDS table takes: the rounds on the
  // global locks
  uint32_t *cur = listA, *nxt = listB;
  for (; num > 0 && round <= maxRound; ++round) {
    for (uint32_t i = threadIdx.x; i < num; i += kTailBlock) reserve(recs + cur[i], cur[i], round, locks);
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < num; i += kTailBlock) {
      const uint32_t k = cur[i];
      if (decide(recs + k, k, round, locks, marks, skipped, stats) == REC_PENDING) nxt[atomicAdd(&s_cnt, 1u)] = k;
    }
    __syncthreads();
    num = s_cnt;
    uint32_t *t = cur;
    cur = nxt;
    nxt = t;
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < num; i += kTailBlock) {
    PatchRec *r = recs + cur[i];
    r->ok = REC_LOST;
    skipped[cur[i]] += 1;
    atomicAdd(stats + ST_CONFLICT, 1u);
    atomicAdd(stats + ST_UNRESOLVED, 1u);
  }
}

// Drops a winner whose insertion point lies below a node another winner moves.
__global__ __launch_bounds__(256) void k_contain(PatchRec *__restrict__ recs, uint32_t K,
                                                 const bvht_BVHNode *__restrict__ nodes, uint32_t root, uint32_t N,
                                                 const uint32_t *__restrict__ marks, uint32_t *__restrict__ skipped,
                                                 uint32_t *__restrict__ stats) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= K) return;
  PatchRec *r = recs + k;
  if (r->ok != REC_WON) return;
  bool bad = false;
  for (int w = 0; w < 2 && !bad; ++w) {
    uint32_t x = r->best[w];
    for (uint32_t steps = 0; steps < N; ++steps) {
      const uint32_t m = marks[x];
      if (m != NONE && m != k) {
        bad = true;
        break;
      }
      if (x == root) break;
      x = nodes[x].parent;
    }
  }
  if (bad) {
    r->ok = REC_CONTAINED;
    skipped[k] += 1;
    atomicAdd(stats + ST_CONTAIN, 1u);
  }
}

// Winners write their topology; every patch releases its locks and marks.
__global__ __launch_bounds__(256) void k_apply(PatchRec *__restrict__ recs, uint32_t K, bvht_BVHNode *__restrict__ nodes,
                                               uint64_t *__restrict__ locks, uint32_t *__restrict__ marks,
                                               uint32_t *__restrict__ fixStarts, uint32_t *__restrict__ numPending,
                                               uint32_t *__restrict__ stats) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= K) return;
  if (k == 0) *numPending = 0;
  PatchRec *r = recs + k;
  const uint32_t ok = r->ok;
  const bool win = ok == REC_WON;
  for (int i = 0; i < 3; ++i) fixStarts[3 * k + i] = win ? r->fix[i] : NONE;
  if (ok == REC_NONE) return;
  for (int i = 0; i < kNNode; ++i) {
    const uint32_t n = r->node[i];
    if (n == NONE) continue;
    locks[n] = kLockFree;
    if (win) {  // MiniPatch::applyOne -- BVHPatch.hpp:72-80
      nodes[n].parent = r->parent[i];
      nodes[n].left = r->left[i];
      nodes[n].right = r->right[i];
      nodes[n].isLeft = (uint16_t)r->isLeft[i];
    }
  }
  marks[r->moved[0]] = NONE;
  marks[r->moved[1]] = NONE;
  r->ok = REC_NONE;
  if (win) atomicAdd(stats + ST_APPLIED, 1u);
}

// kFixTree3_1 -- Bittner13CUDA.cu:407-423: count the arrivals per node.
__global__ __launch_bounds__(256) void k_refitCount(const uint32_t *__restrict__ starts, uint32_t num,
                                                    const bvht_BVHNode *__restrict__ nodes, uint32_t N,
                                                    uint32_t *__restrict__ cnt) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num) return;
  uint32_t node = starts[i];
  if (node >= N) return;
  for (uint32_t steps = 0; steps < N; ++steps) {
    const uint32_t p = loadRelaxedAgent(&nodes[node].parent);  // issued beside the count: one round trip per level
    if (atomicAdd(cnt + node, 1u) != 0) return;
    if (p == node) return;
    node = p;
  }
}

// kFixTree3_2 -- Bittner13CUDA.cu:425-467: the last arrival recomputes.
// With leafCnt/subSA (the learned priority's subtree features) the same walk
// keeps them exact: the nodes it recomputes are exactly those whose subtree
// changed.
__global__ __launch_bounds__(256) void k_refitApply(const uint32_t *__restrict__ starts, uint32_t num,
                                                    bvht_BVHNode *__restrict__ nodes, uint32_t N,
                                                    uint32_t *__restrict__ cnt, uint32_t *leafCnt, float *subSA) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num) return;
  uint32_t node = starts[i];
  if (node >= N) return;
  if (nodes[node].isLeafFlag != 0) {
    // a leaf start (dynamic refit) is not recomputed; k_refitCount counted
    // every occurrence of it and walked on once, so only the last
    // occurrence walks on here (a list may name a leaf more than once)
    if (__hip_atomic_fetch_sub(cnt + node, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 1) return;
    const uint32_t p = nodes[node].parent;
    if (p == node || p >= N) return;
    node = p;
  }
  for (uint32_t steps = 0; steps < N; ++steps) {
    // The topology does not change during the refit, so its load goes out
    // beside the count (the chain of last arrivals near the root is serial:
    // each level costs the count, the child boxes and the stores).  Relaxed
    // atomic loads keep the compiler from sinking them below the count.
    const uint32_t p = loadRelaxedAgent(&nodes[node].parent), l = loadRelaxedAgent(&nodes[node].left),
                   r = loadRelaxedAgent(&nodes[node].right);
    const uint32_t left = __hip_atomic_fetch_sub(cnt + node, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (left != 1) return;
    // Boxes and features written by this walk move between XCDs through
    // device-coherent accesses: the loads below read past the L2 (issued
    // after the count returned), the stores write through it and are waited
    // for before the next count.  This replaces an acquire fence (an L2
    // invalidation) and a release fence (an L2 write-back) per level.
    bvht_AABB b = loadBoxAgent(nodes + l);
    const bvht_AABB br = loadBoxAgent(nodes + r);
    // the children's subtree features load with their boxes, ahead of the
    // node's stores (which may alias them), not one round trip later
    uint32_t cl = 0, cr = 0;
    float sl = 0.0f, sr = 0.0f;
    if (leafCnt) {
      cl = loadRelaxedAgent(leafCnt + l);
      cr = loadRelaxedAgent(leafCnt + r);
      sl = loadRelaxedAgent(subSA + l);
      sr = loadRelaxedAgent(subSA + r);
    }
    mergeInto(b, br);
    const float sa = surfaceArea(b);
    storeBoxAgent(nodes + node, b, sa);
    if (leafCnt) {
      storeRelaxedAgent(leafCnt + node, cl + cr);
      storeRelaxedAgent(subSA + node, sl + sr + sa);
    }
    waitVectorMemory();
    if (p == node) return;
    node = p;
  }
}

// kCalcCost -- Bittner13CUDA.cu:480-500, written as an ascending sort key
// (~bits of the non-negative cost: descending cost, ties by input order).
// keys2 (optional) receives a second copy that the sort leaves untouched.
__global__ __launch_bounds__(256) void k_cost(const bvht_BVHNode *__restrict__ nodes, const uint32_t *__restrict__ ids,
                                              uint32_t num, uint32_t *__restrict__ keys, uint32_t *__restrict__ keys2,
                                              uint32_t *__restrict__ hist) {
  __shared__ uint32_t s_hist[4][kSortRadix];
  for (int p = 0; p < 4; ++p) s_hist[p][threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < num; i += gridDim.x * 256) {
    const bvht_BVHNode nd = nodes[ids[i]];
    float cost = 0.0f;
    if (nd.isLeafFlag == 0) {
      const float s = nd.surfaceArea, l = nodes[nd.left].surfaceArea, r = nodes[nd.right].surfaceArea;
      cost = (s * s * s * 2.0f) / ((l + r) * (l < r ? l : r));
    }
    const uint32_t key = ~__float_as_uint(cost);
    keys[i] = key;
    if (keys2) keys2[i] = key;
#pragma unroll
    for (int p = 0; p < 4; ++p) atomicAdd(&s_hist[p][(key >> (8 * p)) & 255u], 1u);
  }
  __syncthreads();
  for (int p = 0; p < 4; ++p) {
    const uint32_t c = s_hist[p][threadIdx.x];
    if (c) atomicAdd(hist + p * kSortRadix + threadIdx.x, c);
  }
}

// Ordered-selection tiles: 16 consecutive positions per thread.
constexpr uint32_t kSelBlock = 256, kSelItems = 16, kSelTile = kSelBlock * kSelItems;

// ---- altSort selection (bn13_selectNodes with altSort, :682-699) ----------
//
// The costs sorted descending give the threshold t = sorted[used] (used =
// numChunks * chunkSize); every inner node with cost >= t is flagged
// (kGenerateFlags, :93-97) and the flagged inner-list positions are
// compacted in order (DeviceSelect::Flagged).  When `used` reaches the list
// length every node is flagged.  Costs are compared as floats, decoded from
// the ~bits sort keys.  Three launches: per-tile counts, one workgroup scans
// the tile counts (and fills the entries a step would take beyond the
// selected count with NONE), per-tile ordered writes.
constexpr uint32_t kAltTile = kSelTile;  // 4096 positions per workgroup

__device__ __forceinline__ bool altFlag(uint32_t key, bool all, float thr) {
  return all || __uint_as_float(~key) >= thr;
}

__device__ __forceinline__ void altThreshold(const uint32_t *sortedKeys, uint32_t num, uint32_t used, bool &all,
                                             float &thr) {
  all = used >= num;
  thr = all ? 0.0f : __uint_as_float(~sortedKeys[used]);
}

__global__ __launch_bounds__(kSelBlock) void k_altCount(const uint32_t *__restrict__ keys,
                                                        const uint32_t *__restrict__ sortedKeys, uint32_t num,
                                                        uint32_t used, uint32_t *__restrict__ counts) {
  __shared__ uint32_t s_scan[kSelBlock / 64];
  bool all;
  float thr;
  altThreshold(sortedKeys, num, used, all, thr);
  uint32_t c = 0;
  const uint32_t base = blockIdx.x * kAltTile + threadIdx.x * kSelItems;
  for (uint32_t j = 0; j < kSelItems; ++j)
    if (base + j < num && altFlag(keys[base + j], all, thr)) c++;
  const uint32_t ex = blockExclusiveScan<kSelBlock>(c, s_scan);
  if (threadIdx.x == kSelBlock - 1) counts[blockIdx.x] = ex + c;
}

constexpr int kAltScanBlock = 1024;
__global__ __launch_bounds__(kAltScanBlock) void k_altScan(uint32_t *__restrict__ counts, uint32_t numTiles,
                                                           uint32_t *__restrict__ order, uint32_t used) {
  __shared__ uint32_t s_scan[kAltScanBlock / 64];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < numTiles; b += kAltScanBlock) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t v = i < numTiles ? counts[i] : 0u;
    const uint32_t ex = blockExclusiveScan<kAltScanBlock>(v, s_scan);
    const uint32_t carry = s_carry;
    if (i < numTiles) counts[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == kAltScanBlock - 1) s_carry = carry + ex + v;
    __syncthreads();
  }
  for (uint32_t i = s_carry + threadIdx.x; i < used; i += kAltScanBlock) order[i] = NONE;
}

__global__ __launch_bounds__(kSelBlock) void k_altWrite(const uint32_t *__restrict__ keys,
                                                        const uint32_t *__restrict__ sortedKeys, uint32_t num,
                                                        uint32_t used, const uint32_t *__restrict__ offsets,
                                                        uint32_t *__restrict__ order) {
  __shared__ uint32_t s_scan[kSelBlock / 64];
  bool all;
  float thr;
  altThreshold(sortedKeys, num, used, all, thr);
  uint32_t c = 0;
  const uint32_t base = blockIdx.x * kAltTile + threadIdx.x * kSelItems;
  for (uint32_t j = 0; j < kSelItems; ++j)
    if (base + j < num && altFlag(keys[base + j], all, thr)) c++;
  uint32_t pos = offsets[blockIdx.x] + blockExclusiveScan<kSelBlock>(c, s_scan);
  for (uint32_t j = 0; j < kSelItems; ++j)
    if (base + j < num && altFlag(keys[base + j], all, thr)) order[pos++] = base + j;
}

// Order-preserving compaction of node ids by leaf flag (initData's two
// DeviceSelect::If calls, Bittner13CUDA.cu:962-978): per-block counts, a
// host-side scan of the (few) block counts, then an ordered write.

__global__ __launch_bounds__(kSelBlock) void k_selCount(const bvht_BVHNode *__restrict__ nodes, uint32_t N,
                                                        uint32_t *__restrict__ counts) {
  __shared__ uint32_t s_scan[kSelBlock / 64];
  uint32_t c = 0;
  const uint32_t base = blockIdx.x * kSelTile + threadIdx.x * kSelItems;
  for (uint32_t j = 0; j < kSelItems; ++j)
    if (base + j < N && nodes[base + j].isLeafFlag == 0) c++;
  const uint32_t ex = blockExclusiveScan<kSelBlock>(c, s_scan);
  if (threadIdx.x == kSelBlock - 1) counts[blockIdx.x] = ex + c;
}

__global__ __launch_bounds__(kSelBlock) void k_selWrite(const bvht_BVHNode *__restrict__ nodes, uint32_t N,
                                                        const uint32_t *__restrict__ innerOffsets,
                                                        uint32_t *__restrict__ inner, uint32_t *__restrict__ leaves) {
  __shared__ uint32_t s_scan[kSelBlock / 64];
  uint32_t c = 0;
  const uint32_t base = blockIdx.x * kSelTile + threadIdx.x * kSelItems;
  for (uint32_t j = 0; j < kSelItems; ++j)
    if (base + j < N && nodes[base + j].isLeafFlag == 0) c++;
  uint32_t pi = innerOffsets[blockIdx.x] + blockExclusiveScan<kSelBlock>(c, s_scan);
  uint32_t pl = base - pi;  // leaves before this thread's first item
  for (uint32_t j = 0; j < kSelItems; ++j) {
    const uint32_t id = base + j;
    if (id >= N) break;
    if (nodes[id].isLeafFlag == 0)
      inner[pi++] = id;
    else
      leaves[pl++] = id;
  }
}

// Deterministic SA sums: one f64 partial per block, summed on the host.
__global__ __launch_bounds__(256) void k_sahPartials(const bvht_BVHNode *__restrict__ nodes, uint32_t N, uint32_t root,
                                                     float cInner, float cLeaf, int weighted,
                                                     double *__restrict__ partials) {
  __shared__ double s[256];
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
    const bvht_BVHNode &nd = nodes[i];
    if (weighted) {
      if (i == root) continue;
      float c = nd.surfaceArea;
      c *= nd.isLeafFlag != 0 ? cLeaf : cInner;
      acc += c;
    } else {
      acc += nd.surfaceArea;
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < (uint32_t)o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

// selectDynamicLeafNodes (CUDAKernels.cu:456-484): leaves whose first face
// is at or after `begin`, in leaf-list order (per-block counts, host scan,
// ordered write -- the order DeviceSelect::If keeps).
__global__ __launch_bounds__(kSelBlock) void k_dynCount(const bvht_BVHNode *__restrict__ nodes,
                                                        const uint32_t *__restrict__ leaves, uint32_t num,
                                                        uint32_t begin, uint32_t *__restrict__ counts) {
  __shared__ uint32_t s_scan[kSelBlock / 64];
  uint32_t c = 0;
  const uint32_t base = blockIdx.x * kSelTile + threadIdx.x * kSelItems;
  for (uint32_t j = 0; j < kSelItems; ++j)
    if (base + j < num && nodes[leaves[base + j]].left >= begin) c++;
  const uint32_t ex = blockExclusiveScan<kSelBlock>(c, s_scan);
  if (threadIdx.x == kSelBlock - 1) counts[blockIdx.x] = ex + c;
}

__global__ __launch_bounds__(kSelBlock) void k_dynWrite(const bvht_BVHNode *__restrict__ nodes,
                                                        const uint32_t *__restrict__ leaves, uint32_t num,
                                                        uint32_t begin, const uint32_t *__restrict__ offsets,
                                                        uint32_t *__restrict__ out) {
  __shared__ uint32_t s_scan[kSelBlock / 64];
  uint32_t c = 0;
  const uint32_t base = blockIdx.x * kSelTile + threadIdx.x * kSelItems;
  for (uint32_t j = 0; j < kSelItems; ++j)
    if (base + j < num && nodes[leaves[base + j]].left >= begin) c++;
  uint32_t pos = offsets[blockIdx.x] + blockExclusiveScan<kSelBlock>(c, s_scan);
  for (uint32_t j = 0; j < kSelItems; ++j)
    if (base + j < num && nodes[leaves[base + j]].left >= begin) out[pos++] = leaves[base + j];
}

struct Internal {
  uint32_t N = 0, numLeaves = 0, numInner = 0, K = 0;
  uint32_t root = 0;
  uint32_t *inner = nullptr, *leaves = nullptr;
  uint32_t *keys[2] = {nullptr, nullptr};
  uint32_t *vals[2] = {nullptr, nullptr};
  uint32_t *keysAlt = nullptr;    // altSort: the keys the sort consumes (keys[0] keeps the unsorted costs)
  uint32_t *altCounts = nullptr;  // altSort: per-tile counts / offsets
  uint32_t *sortScratch = nullptr;
  SortPlan plan;
  uint32_t *sorted = nullptr;  // -> vals[result]
  PatchRec *recs = nullptr;
  uint64_t *locks = nullptr;
  uint32_t *marks = nullptr, *cnt = nullptr, *refitCnt = nullptr;
  uint32_t *pending[2] = {nullptr, nullptr}, *numPending = nullptr;
  uint32_t *fixStarts = nullptr;
  uint32_t *skipped = nullptr;
  uint32_t *stats = nullptr;
  uint32_t *selCounts = nullptr;
  double *partials = nullptr;
  bool selected = false;
  // learned priority and SAH-gain records
  float *feat = nullptr;  // [numInner][12], inner-list order
  uint32_t *leafCnt = nullptr;
  float *subSA = nullptr;
  bool subValid = false;  // leafCnt/subSA match the tree (kept by the refit)
  uint32_t *anc[2] = {nullptr, nullptr}, *dep[2] = {nullptr, nullptr};
  PriorityModel *model = nullptr;
  float *recFeat = nullptr, *recGain = nullptr;
  uint32_t *recCount = nullptr;
  uint32_t recMax = 0;
};

bool features(Internal *in, const bvht_CUDAMemoryBVHPointer *bvh) {
  if (!in->feat) {
    if (!devAlloc(&in->feat, (size_t)in->numInner * kFeat, "bn13 features") ||
        !devAlloc(&in->leafCnt, in->N, "bn13 leaf counts") || !devAlloc(&in->subSA, in->N, "bn13 subtree SA") ||
        !devAlloc(&in->anc[0], in->N, "bn13 ancestors") || !devAlloc(&in->anc[1], in->N, "bn13 ancestors") ||
        !devAlloc(&in->dep[0], in->N, "bn13 depths") || !devAlloc(&in->dep[1], in->N, "bn13 depths"))
      return false;
  }
  if (!in->subValid) {
    if (!computeSubtree(bvh->nodes, in->N, in->leaves, in->numLeaves, in->leafCnt, in->subSA, in->refitCnt))
      return false;
    in->subValid = true;
  }
  const int d = computeDepth(bvh->nodes, in->N, in->root, in->anc, in->dep);
  if (d < 0) return false;
  return computeFeatures(bvh->nodes, in->N, in->root, in->inner, in->numInner, in->leafCnt, in->subSA, in->dep[d],
                         in->feat);
}

// Refit arguments that keep the subtree features (when they are in use).
uint32_t *refitLeafCnt(const Internal *in) { return in->subValid ? in->leafCnt : nullptr; }
float *refitSubSA(const Internal *in) { return in->subValid ? in->subSA : nullptr; }

RecordSink sinkOf(const Internal *in) {
  RecordSink r{nullptr, in->recFeat, in->recGain, in->recCount, in->recMax};
  if (in->recMax) r.feat = in->feat;
  return r;
}

constexpr uint32_t kSahBlocks = 512;

Internal *inOf(bvht_GPUWorkingMemory *d) { return d ? reinterpret_cast<Internal *>(d->internal) : nullptr; }

bool validBVH(const bvht_CUDAMemoryBVHPointer *b, const char *what) {
  if (!b || !b->nodes || b->numNodes == 0) {
    setError("%s: no device BVH", what);
    return false;
  }
  return true;
}

double sahSum(const bvht_CUDAMemoryBVHPointer *b, double *partials, uint32_t root, float ci, float cl, int weighted) {
  k_sahPartials<<<kSahBlocks, 256, 0, stream()>>>(b->nodes, b->numNodes, root, ci, cl, weighted, partials);
  if (!checkLaunch("k_sahPartials")) return -1.0;
  std::vector<double> h(kSahBlocks);
  if (!check(hipMemcpyAsync(h.data(), partials, sizeof(double) * kSahBlocks, hipMemcpyDeviceToHost, stream()),
             "SAH partials") ||
      !check(hipStreamSynchronize(stream()), "SAH sync"))
    return -1.0;
  double s = 0.0;
  for (double x : h) s += x;
  return s;
}

double *sahScratch() {
  static double *p = nullptr;
  if (!p) devMalloc(reinterpret_cast<void **>(&p), sizeof(double) * kSahBlocks, "SAH scratch");
  return p;
}

float rootSA(const bvht_CUDAMemoryBVHPointer *b, uint32_t root) {
  float sa = 0.0f;
  if (check(hipMemcpyAsync(&sa, &b->nodes[root].surfaceArea, sizeof(float), hipMemcpyDeviceToHost, stream()), "root SA"))
    check(hipStreamSynchronize(stream()), "root SA sync");
  return sa;
}

}  // namespace

}  // namespace bvht

using namespace bvht;

extern "C" {

bvht_GPUWorkingMemory allocateMemory(bvht_CUDAMemoryBVHPointer *bvh, uint32_t batchSize, uint32_t numFaces) {
  bvht_GPUWorkingMemory m;
  std::memset(&m, 0, sizeof(m));
  if (!validBVH(bvh, "allocateMemory")) return m;
  if (bvh->numNodes != 2 * numFaces - 1 || numFaces < 2) {
    setError("allocateMemory: %u nodes do not form a binary BVH over %u faces", bvh->numNodes, numFaces);
    return m;
  }
  if (batchSize == 0) {
    setError("allocateMemory: batch size 0");
    return m;
  }
  Internal *in = new (std::nothrow) Internal();
  if (!in) {
    setError("allocateMemory: host allocation failed");
    return m;
  }
  m.internal = in;
  in->N = bvh->numNodes;
  in->numLeaves = numFaces;
  in->numInner = bvh->numNodes - numFaces;
  in->K = batchSize;
  in->plan = makeSortPlan(in->numInner, 4, 32);
  const uint32_t selBlocks = divUp(in->N, kSelTile);
  m.numLeafNodes = in->numLeaves;
  m.numInnerNodes = in->numInner;
  m.numPatches = batchSize;
  m.numSkipped = batchSize;
  m.numNodesToFix = 3 * batchSize;
  m.numFlags = in->N;
  bool ok = devAlloc(&in->inner, in->numInner, "bn13 inner") && devAlloc(&in->leaves, in->numLeaves, "bn13 leaves") &&
            devAlloc(&in->keys[0], in->numInner, "bn13 keys") && devAlloc(&in->keys[1], in->numInner, "bn13 keys") &&
            devAlloc(&in->vals[0], in->numInner, "bn13 vals") && devAlloc(&in->vals[1], in->numInner, "bn13 vals") &&
            devMalloc(reinterpret_cast<void **>(&in->sortScratch), sortScratchBytes(in->plan), "bn13 sort") &&
            devAlloc(&in->recs, batchSize, "bn13 patches") && devAlloc(&in->locks, in->N, "bn13 locks") &&
            devAlloc(&in->pending[0], batchSize, "bn13 pending") && devAlloc(&in->pending[1], batchSize, "bn13 pending") &&
            devAlloc(&in->numPending, 1, "bn13 pending count") &&
            devAlloc(&in->marks, in->N, "bn13 marks") && devAlloc(&in->cnt, in->N, "bn13 refit counters") &&
            devAlloc(&in->refitCnt, in->N, "bn13 refit counters") &&
            devAlloc(&in->fixStarts, 3 * (size_t)batchSize, "bn13 fix starts") &&
            devAlloc(&in->skipped, batchSize, "bn13 skipped") && devAlloc(&in->stats, ST_WORDS, "bn13 stats") &&
            devAlloc(&in->selCounts, selBlocks, "bn13 select") && devAlloc(&in->partials, kSahBlocks, "bn13 SAH");
  if (!ok) {
    freeMemory(&m);
    return m;
  }
  m.result = true;
  return m;
}

void freeMemory(bvht_GPUWorkingMemory *data) {
  Internal *in = inOf(data);
  if (in) {
    devFree(in->inner);
    devFree(in->leaves);
    devFree(in->keys[0]);
    devFree(in->keys[1]);
    devFree(in->keysAlt);
    devFree(in->altCounts);
    devFree(in->vals[0]);
    devFree(in->vals[1]);
    devFree(in->sortScratch);
    devFree(in->recs);
    devFree(in->locks);
    devFree(in->pending[0]);
    devFree(in->pending[1]);
    devFree(in->numPending);
    devFree(in->marks);
    devFree(in->cnt);
    devFree(in->refitCnt);
    devFree(in->fixStarts);
    devFree(in->skipped);
    devFree(in->stats);
    devFree(in->selCounts);
    devFree(in->partials);
    devFree(in->feat);
    devFree(in->leafCnt);
    devFree(in->subSA);
    for (int b = 0; b < 2; ++b) {
      devFree(in->anc[b]);
      devFree(in->dep[b]);
    }
    devFree(in->recFeat);
    devFree(in->recGain);
    devFree(in->recCount);
    delete in;
  }
  if (data) {
    std::memset(data, 0, sizeof(*data));
  }
}

void initData(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, uint32_t blockSize) {
  (void)blockSize;
  Internal *in = inOf(data);
  if (!in || !validBVH(bvh, "initData")) {
    if (!in) setError("initData: working memory not allocated");
    return;
  }
  // root from the device header (the GPU optimizer keeps it fixed)
  bvht_BVH hdr;
  if (!check(hipMemcpyAsync(&hdr, bvh->bvh, sizeof(hdr), hipMemcpyDeviceToHost, stream()), "initData header") ||
      !check(hipStreamSynchronize(stream()), "initData sync"))
    return;
  in->root = hdr.vRootIndex;
  in->subValid = false;
  const uint32_t selBlocks = divUp(in->N, kSelTile);
  k_selCount<<<selBlocks, kSelBlock, 0, stream()>>>(bvh->nodes, in->N, in->selCounts);
  if (!checkLaunch("k_selCount")) return;
  std::vector<uint32_t> counts(selBlocks);
  if (!check(hipMemcpyAsync(counts.data(), in->selCounts, 4ull * selBlocks, hipMemcpyDeviceToHost, stream()),
             "initData counts") ||
      !check(hipStreamSynchronize(stream()), "initData sync"))
    return;
  uint32_t total = 0;
  for (auto &c : counts) {
    uint32_t t = c;
    c = total;
    total += t;
  }
  if (total != in->numInner) {
    setError("initData: %u inner nodes found, expected %u", total, in->numInner);
    return;
  }
  if (!check(hipMemcpyAsync(in->selCounts, counts.data(), 4ull * selBlocks, hipMemcpyHostToDevice, stream()),
             "initData offsets"))
    return;
  k_selWrite<<<selBlocks, kSelBlock, 0, stream()>>>(bvh->nodes, in->N, in->selCounts, in->inner, in->leaves);
  if (!checkLaunch("k_selWrite")) return;
  fillU32(reinterpret_cast<uint32_t *>(in->locks), NONE, 2ull * in->N);  // kLockFree
  fillU32(in->numPending, 0u, 1);
  fillU32(in->marks, NONE, in->N);
  fillU32(in->cnt, 0u, in->N);
  fillU32(in->refitCnt, 0u, in->N);
  fillU32(in->skipped, 0u, in->K);
  fillU32(in->stats, 0u, ST_WORDS);
  fillU32(in->fixStarts, NONE, 3ull * in->K);
  check(hipMemsetAsync(in->recs, 0, sizeof(PatchRec) * in->K, stream()), "initData patches");
  check(hipStreamSynchronize(stream()), "initData sync");
}

void fixTree1(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, uint32_t blockSize) {
  (void)blockSize;
  Internal *in = inOf(data);
  if (!in || !validBVH(bvh, "fixTree1")) return;
  in->subValid = false;  // boxes may change; the subtree features are recomputed when next needed
  refitFromLeaves(bvh->nodes, 0, in->numLeaves, in->leaves, in->refitCnt, in->N);
}

void fixTree3(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, uint32_t blockSize) {
  (void)blockSize;
  Internal *in = inOf(data);
  if (!in || !validBVH(bvh, "fixTree3")) return;
  const uint32_t num = 3 * in->K;
  k_refitCount<<<divUp(num, 256), 256, 0, stream()>>>(in->fixStarts, num, bvh->nodes, in->N, in->cnt);
  if (!checkLaunch("k_refitCount")) return;
  k_refitApply<<<divUp(num, 256), 256, 0, stream()>>>(in->fixStarts, num, bvh->nodes, in->N, in->cnt,
                                                    refitLeafCnt(in), refitSubSA(in));
  checkLaunch("k_refitApply");
}

void selectDynamicLeafNodes(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, uint32_t *out,
                            uint32_t *numOut, uint32_t beginDynamicTriangle) {
  Internal *in = inOf(data);
  if (numOut) *numOut = 0;
  if (!in || !validBVH(bvh, "selectDynamicLeafNodes") || !out || !numOut) {
    if (!in) setError("selectDynamicLeafNodes: working memory not allocated");
    else if (!out || !numOut) setError("selectDynamicLeafNodes: no output");
    return;
  }
  const uint32_t blocks = divUp(in->numLeaves, kSelTile);
  k_dynCount<<<blocks, kSelBlock, 0, stream()>>>(bvh->nodes, in->leaves, in->numLeaves, beginDynamicTriangle,
                                                 in->selCounts);
  if (!checkLaunch("k_dynCount")) return;
  std::vector<uint32_t> counts(blocks);
  if (!check(hipMemcpyAsync(counts.data(), in->selCounts, 4ull * blocks, hipMemcpyDeviceToHost, stream()),
             "selectDynamicLeafNodes counts") ||
      !check(hipStreamSynchronize(stream()), "selectDynamicLeafNodes sync"))
    return;
  uint32_t total = 0;
  for (auto &c : counts) {
    const uint32_t t = c;
    c = total;
    total += t;
  }
  if (!check(hipMemcpyAsync(in->selCounts, counts.data(), 4ull * blocks, hipMemcpyHostToDevice, stream()),
             "selectDynamicLeafNodes offsets"))
    return;
  k_dynWrite<<<blocks, kSelBlock, 0, stream()>>>(bvh->nodes, in->leaves, in->numLeaves, beginDynamicTriangle,
                                                 in->selCounts, out);
  if (!checkLaunch("k_dynWrite") || !check(hipStreamSynchronize(stream()), "selectDynamicLeafNodes sync")) return;
  *numOut = total;
}

bool bvht_fixTreeFrom(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, const uint32_t *devStarts,
                      uint32_t num) {
  Internal *in = inOf(data);
  if (!in || !validBVH(bvh, "bvht_fixTreeFrom")) {
    if (!in) setError("bvht_fixTreeFrom: working memory not allocated");
    return false;
  }
  if (num == 0) return true;
  if (!devStarts) {
    setError("bvht_fixTreeFrom: no start nodes");
    return false;
  }
  k_refitCount<<<divUp(num, 256), 256, 0, stream()>>>(devStarts, num, bvh->nodes, in->N, in->cnt);
  if (!checkLaunch("k_refitCount")) return false;
  in->subValid = false;  // leaf boxes may have changed (dynamic refit)
  k_refitApply<<<divUp(num, 256), 256, 0, stream()>>>(devStarts, num, bvh->nodes, in->N, in->cnt, nullptr, nullptr);
  return checkLaunch("k_refitApply");
}

void bn13_selectNodes(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, bvht_AlgoCFG cfg) {
  Internal *in = inOf(data);
  if (!in || !validBVH(bvh, "bn13_selectNodes")) return;
  if (!cfg.sort) {  // Bittner13CUDA.cu:719
    setError("bn13_selectNodes: topKThElement is not supported (set option sort)");
    return;
  }
  if (cfg.priority > 1 || (cfg.priority == 1 && !in->model)) {
    setError("bn13_selectNodes: priority %u needs %s", cfg.priority,
             cfg.priority == 1 ? "a model (bvht_bn13_setModel)" : "0 (heuristic) or 1 (learned)");
    return;
  }
  if ((cfg.priority == 1 || in->recMax) && !features(in, bvh)) return;
  const bool alt = cfg.altSort != 0;
  if (alt && cfg.priority == 1) {
    setError("bn13_selectNodes: altSort selects by the heuristic priority only (priority 0)");
    return;
  }
  const uint64_t used = (uint64_t)cfg.numChunks * cfg.chunkSize;
  if (alt && used > in->numInner) {
    setError("bn13_selectNodes: altSort: %u chunks x %u exceed %u inner nodes", cfg.numChunks, cfg.chunkSize,
             in->numInner);
    return;
  }
  const uint32_t altTiles = divUp(in->numInner, kAltTile);
  if (alt && !in->keysAlt &&
      (!devAlloc(&in->keysAlt, in->numInner, "bn13 altSort keys") ||
       !devAlloc(&in->altCounts, altTiles, "bn13 altSort counts")))
    return;
  uint32_t *scratch = in->sortScratch;
  if (!sortReset(in->plan, scratch)) return;
  if (cfg.priority == 1) {
    if (!mlpInfer(*in->model, in->feat, in->numInner, in->keys[0], sortHist(in->plan, scratch), nullptr)) return;
  } else {
    uint32_t blocks = divUp(in->numInner, 256 * 8);
    if (blocks > (uint32_t)numCUs() * 8) blocks = numCUs() * 8;
    k_cost<<<blocks, 256, 0, stream()>>>(bvh->nodes, in->inner, in->numInner, alt ? in->keysAlt : in->keys[0],
                                         alt ? in->keys[0] : nullptr, sortHist(in->plan, scratch));
    if (!checkLaunch("k_cost")) return;
  }
  if (alt) {
    // SortKeysDescending (:682-686): keysAlt and keys[1] ping-pong, keys[0]
    // keeps the unsorted costs for the flags
    void *keys[2] = {in->keysAlt, in->keys[1]};
    uint32_t *vals[2] = {in->vals[0], in->vals[1]};
    const int r = sortRun(in->plan, keys, vals, true, true, scratch);
    if (r < 0) return;
    const uint32_t *sortedKeys = reinterpret_cast<const uint32_t *>(keys[r]);
    uint32_t *order = in->vals[0];
    k_altCount<<<altTiles, kSelBlock, 0, stream()>>>(in->keys[0], sortedKeys, in->numInner, (uint32_t)used,
                                                     in->altCounts);
    if (!checkLaunch("k_altCount")) return;
    k_altScan<<<1, kAltScanBlock, 0, stream()>>>(in->altCounts, altTiles, order, (uint32_t)used);
    if (!checkLaunch("k_altScan")) return;
    k_altWrite<<<altTiles, kSelBlock, 0, stream()>>>(in->keys[0], sortedKeys, in->numInner, (uint32_t)used,
                                                     in->altCounts, order);
    if (!checkLaunch("k_altWrite")) return;
    in->sorted = order;
    in->selected = true;
    return;
  }
  void *keys[2] = {in->keys[0], in->keys[1]};
  uint32_t *vals[2] = {in->vals[0], in->vals[1]};
  // values: inner-list positions (identity); k_search maps them to node ids
  const int r = sortRun(in->plan, keys, vals, true, false, scratch);
  if (r < 0) return;
  in->sorted = in->vals[r];
  in->selected = true;
}

void bn13_rmAndReinsChunk(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, bvht_AlgoCFG cfg, uint32_t chunk) {
  Internal *in = inOf(data);
  if (!in || !validBVH(bvh, "bn13_rmAndReinsChunk")) return;
  if (!in->selected) {
    setError("bn13_rmAndReinsChunk: call bn13_selectNodes first");
    return;
  }
  const uint32_t K = cfg.chunkSize;
  if (K == 0 || K > in->K || (uint64_t)cfg.numChunks * K > in->numInner || chunk >= cfg.numChunks) {
    setError("bn13_rmAndReinsChunk: chunk %u of %u x %u does not fit %u patches / %u inner nodes", chunk, cfg.numChunks,
             K, in->K, in->numInner);
    return;
  }
  // the first chunk of a step also counts a failed priority sort (its error
  // word is cleared with the sort state at the next selection)
  k_search<false><<<K, 64, 0, stream()>>>(bvh->nodes, in->root, in->N, in->sorted, in->inner, in->numInner, chunk,
                                          cfg.numChunks, K, cfg.offsetAccess ? 1u : 0u, cfg.altFindNode ? 1u : 0u,
                                          in->recs, in->locks, in->stats, sinkOf(in),
                                          chunk == 0 ? in->sortScratch + in->plan.errOff : nullptr);
  if (!checkLaunch("k_search")) return;
  k_commitFirst<<<divUp(K, 256), 256, 0, stream()>>>(in->recs, K, in->locks, in->marks, in->skipped, in->stats,
                                                     in->pending[0], in->numPending);
  if (!checkLaunch("k_commitFirst")) return;
  k_resolveTail<<<1, kTailBlock, 0, stream()>>>(in->recs, in->locks, in->marks, in->skipped, in->stats, in->pending[0],
                                                in->pending[1], in->numPending);
  if (!checkLaunch("k_resolveTail")) return;
  k_contain<<<divUp(K, 256), 256, 0, stream()>>>(in->recs, K, bvh->nodes, in->root, in->N, in->marks, in->skipped,
                                                 in->stats);
  if (!checkLaunch("k_contain")) return;
  k_apply<<<divUp(K, 256), 256, 0, stream()>>>(in->recs, K, bvh->nodes, in->locks, in->marks, in->fixStarts,
                                               in->numPending, in->stats);
  if (!checkLaunch("k_apply")) return;
  const uint32_t num = 3 * K;
  k_refitCount<<<divUp(num, 256), 256, 0, stream()>>>(in->fixStarts, num, bvh->nodes, in->N, in->cnt);
  if (!checkLaunch("k_refitCount")) return;
  k_refitApply<<<divUp(num, 256), 256, 0, stream()>>>(in->fixStarts, num, bvh->nodes, in->N, in->cnt,
                                                    refitLeafCnt(in), refitSubSA(in));
  checkLaunch("k_refitApply");
}

void bn13_doAlgorithmStep(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, bvht_AlgoCFG cfg) {
  bn13_selectNodes(data, bvh, cfg);
  for (uint32_t i = 0; i < cfg.numChunks; ++i) bn13_rmAndReinsChunk(data, bvh, cfg, i);
}

float CUDAcalcSAH(bvht_CUDAMemoryBVHPointer *bvh) {
  if (!validBVH(bvh, "CUDAcalcSAH")) return 0.0f;
  double *p = sahScratch();
  if (!p) return 0.0f;
  const double s = sahSum(bvh, p, 0, 1.0f, 1.0f, 0);
  const float r = rootSA(bvh, 0);  // the reference reads node 0 (Bittner13CUDA.cu:794)
  return (float)((s - (double)r) / (double)r);
}

double bvht_calcSAHWeighted(bvht_CUDAMemoryBVHPointer *bvh, float cInner, float cLeaf) {
  if (!validBVH(bvh, "bvht_calcSAHWeighted")) return 0.0;
  bvht_BVH hdr;
  if (!check(hipMemcpyAsync(&hdr, bvh->bvh, sizeof(hdr), hipMemcpyDeviceToHost, stream()), "SAH header") ||
      !check(hipStreamSynchronize(stream()), "SAH sync"))
    return 0.0;
  double *p = sahScratch();
  if (!p) return 0.0;
  const double s = sahSum(bvh, p, hdr.vRootIndex, cInner, cLeaf, 1);
  return s / (double)rootSA(bvh, hdr.vRootIndex);
}

void doCudaDevSync(void) { check(hipStreamSynchronize(stream()), "doCudaDevSync"); }

uint32_t calcNumSkipped(bvht_GPUWorkingMemory *data) {
  Internal *in = inOf(data);
  if (!in) return 0;
  std::vector<uint32_t> h(in->K);
  if (!check(hipMemcpyAsync(h.data(), in->skipped, 4ull * in->K, hipMemcpyDeviceToHost, stream()), "calcNumSkipped") ||
      !check(hipStreamSynchronize(stream()), "calcNumSkipped sync"))
    return 0;
  uint32_t s = 0;
  for (uint32_t x : h) s += x;
  return s;
}

// Counters of the optimizer (extension): conflicts, failed patches,
// containment drops, frontier overflows, applied patches, patches left
// undecided by the conflict resolution (always 0), searches stopped at the
// step bound (always 0 for a tree).
int bvht_bn13_stats(bvht_GPUWorkingMemory *data, uint32_t *out, int maxWords) {
  Internal *in = inOf(data);
  if (!in || !out) return 0;
  uint32_t h[ST_WORDS];
  if (!check(hipMemcpyAsync(h, in->stats, sizeof(h), hipMemcpyDeviceToHost, stream()), "bn13 stats") ||
      !check(hipStreamSynchronize(stream()), "bn13 stats sync"))
    return 0;
  int n = maxWords < 8 ? maxWords : 8;
  for (int i = 0; i < n; ++i) out[i] = h[i];
  return n;
}

// ---- learned priority and SAH-gain records (extension) -------------------

bool bvht_bn13_setModel(bvht_GPUWorkingMemory *data, bvht_PriorityModel *model) {
  Internal *in = inOf(data);
  if (!in) {
    setError("bvht_bn13_setModel: working memory not allocated");
    return false;
  }
  in->model = reinterpret_cast<PriorityModel *>(model);
  return true;
}

bool bvht_bn13_features(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, float *devFeat) {
  Internal *in = inOf(data);
  if (!in || !validBVH(bvh, "bvht_bn13_features") || !devFeat) {
    if (!in) setError("bvht_bn13_features: working memory not allocated");
    if (!devFeat) setError("bvht_bn13_features: no output buffer");
    return false;
  }
  if (!features(in, bvh)) return false;
  return check(hipMemcpyAsync(devFeat, in->feat, sizeof(float) * kFeat * (size_t)in->numInner,
                              hipMemcpyDeviceToDevice, stream()),
               "bvht_bn13_features");
}

bool bvht_bn13_setRecording(bvht_GPUWorkingMemory *data, uint32_t maxRecords) {
  Internal *in = inOf(data);
  if (!in) {
    setError("bvht_bn13_setRecording: working memory not allocated");
    return false;
  }
  devFree(in->recFeat);
  devFree(in->recGain);
  devFree(in->recCount);
  in->recFeat = in->recGain = nullptr;
  in->recCount = nullptr;
  in->recMax = 0;
  if (maxRecords == 0) return true;
  if (!devAlloc(&in->recFeat, (size_t)maxRecords * kFeat, "bn13 record features") ||
      !devAlloc(&in->recGain, maxRecords, "bn13 record gains") || !devAlloc(&in->recCount, 1, "bn13 record count") ||
      !fillU32(in->recCount, 0u, 1))
    return false;
  in->recMax = maxRecords;
  return true;
}

uint32_t bvht_bn13_records(bvht_GPUWorkingMemory *data, float **devFeat, float **devGain) {
  Internal *in = inOf(data);
  if (!in || !in->recMax) return 0;
  uint32_t n = 0;
  if (!check(hipMemcpyAsync(&n, in->recCount, 4, hipMemcpyDeviceToHost, stream()), "bvht_bn13_records") ||
      !check(hipStreamSynchronize(stream()), "bvht_bn13_records sync"))
    return 0;
  if (devFeat) *devFeat = in->recFeat;
  if (devGain) *devGain = in->recGain;
  return n < in->recMax ? n : in->recMax;
}

bool bvht_bn13_probe(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, bvht_AlgoCFG cfg,
                     const uint32_t *devPositions, uint32_t num) {
  Internal *in = inOf(data);
  if (!in || !validBVH(bvh, "bvht_bn13_probe")) {
    if (!in) setError("bvht_bn13_probe: working memory not allocated");
    return false;
  }
  if (!in->recMax) {
    setError("bvht_bn13_probe: enable recording first (bvht_bn13_setRecording)");
    return false;
  }
  if (num == 0) return true;
  if (!devPositions) {
    setError("bvht_bn13_probe: no positions");
    return false;
  }
  if (!features(in, bvh)) return false;
  k_search<true><<<num, 64, 0, stream()>>>(bvh->nodes, in->root, in->N, devPositions, in->inner, in->numInner, 0, 1,
                                          num, 0u, cfg.altFindNode ? 1u : 0u, in->recs, in->locks, in->stats,
                                          sinkOf(in), nullptr);
  return checkLaunch("k_search (probe)");
}

}  // extern "C"
