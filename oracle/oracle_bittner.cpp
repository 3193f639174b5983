// ORACLE -- test infrastructure, not product code (see oracle.cpp).
//
// Bittner et al. 2013 BVH optimisation, restated twice:
//  * orc_bittner13_seq    -- the sequential CPU optimizer
//                            (lib/builder/Bittner13.cpp:71-312), the SAH oracle;
//  * orc_bittner13_gpuref -- the GPU algorithm of lib/builder/Bittner13CUDA.cu
//                            (patches, root rule, lock conflicts, per-chunk
//                            refit) run as one serial interleaving: the
//                            threads of a kernel execute in index order.
//
// Deviations from the reference (bugs of the reference that would corrupt
// the tree; documented so results are comparable):
//  (1) a patch whose removal or reinsertion fails is cleared instead of
//      being locked and applied half done (Bittner13CUDA.cu:520-525 leaves
//      the half-built patch in place; kApplyPatches then applies it, which
//      detaches the subtrees of the removed node);
//  (2) a patch whose insertion point lies inside a subtree that another
//      applied patch of the same chunk moves is skipped (two such patches can
//      otherwise form a cycle);
//  (3) the heap of findNode has capacity checks (the reference asserts).
#include "oracle.hpp"

#include <cstdio>
#include <cstdlib>
#include <tuple>
#include <utility>

namespace orc {
namespace {

inline bool isLeaf(const Node &n) { return n.isLeafFlag != 0; }

// Removal priority of an inner node, SA^3 * 2 / ((L + R) * min(L, R)), in the
// f32 operation order of kCalcCost (Bittner13CUDA.cu:480-500, min as
// `L < R ? L : R`) or of the sequential runMetric (Bittner13.cpp:266-271,
// std::min: the two differ only for a NaN area); leaves get 0
// (Bittner13.cpp:264, Bittner13CUDA.cu:495).
inline float priorityOf(const std::vector<Node> &b, uint32_t j, bool seqMin = false) {
  const Node &n = b[j];
  if (isLeaf(n)) return 0.0f;
  const float sa = n.surfaceArea, l = b[n.left].surfaceArea, r = b[n.right].surfaceArea;
  const float m = seqMin ? std::min(l, r) : (l < r ? l : r);
  return (sa * sa * sa * 2.0f) / ((l + r) * m);
}

// Batch size of both optimizers: (uint32_t)((batchPercent / 100.0f) * (float)N)
// with batchPercent clamped to [0.01, 75] (Bittner13.cpp:227,
// Bittner13GPU.cpp:59-60, :79).
inline uint32_t batchCount(float batchPercent, uint32_t N) {
  if (batchPercent <= 0.01f) batchPercent = 0.01f;
  if (batchPercent >= 75.0f) batchPercent = 75.0f;
  return static_cast<uint32_t>((batchPercent / 100.0f) * static_cast<float>(N));
}

// The GPU optimizer's processing order for one step (bn13_selectNodes,
// Bittner13CUDA.cu:659-726) over the inner-node list `inner` (node ids in
// ascending order, initData's DeviceSelect::If, :962-978).
//  * sort (default): SortPairsDescending of (cost, id) -- stable, so equal
//    costs keep inner-list order;
//  * altSort: SortKeysDescending of the costs, threshold = the sorted cost
//    at index `used` (= numChunks * chunkSize), flags cost >= threshold
//    (kGenerateFlags, :93-97), DeviceSelect::Flagged keeps inner-list order.
//    When `used` equals the list length the threshold index is past the end
//    of the reference's array; every node is flagged here.
// Entries past the selected count are NONE (no patch).
std::vector<uint32_t> gpuSelect(const std::vector<Node> &b, const std::vector<uint32_t> &inner, uint32_t used,
                                bool altSort) {
  const size_t n = inner.size();
  std::vector<float> cost(n);
  for (size_t i = 0; i < n; ++i) cost[i] = priorityOf(b, inner[i]);
  std::vector<uint32_t> out;
  if (!altSort) {
    std::vector<uint32_t> ord(n);
    for (size_t i = 0; i < n; ++i) ord[i] = (uint32_t)i;
    // CUB's descending float radix order: by the bit pattern of a
    // non-negative float (every priority is >= 0 or NaN)
    auto key = [&](uint32_t i) {
      uint32_t u;
      std::memcpy(&u, &cost[i], 4);
      return u;
    };
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t c) { return key(a) > key(c); });
    out.resize(n);
    for (size_t i = 0; i < n; ++i) out[i] = inner[ord[i]];
    return out;
  }
  std::vector<uint32_t> bits(n);
  for (size_t i = 0; i < n; ++i) std::memcpy(&bits[i], &cost[i], 4);
  std::vector<uint32_t> sorted = bits;
  std::sort(sorted.begin(), sorted.end(), [](uint32_t a, uint32_t c) { return a > c; });
  const bool all = used >= n;
  float thr = 0.0f;
  if (!all) std::memcpy(&thr, &sorted[used], 4);
  for (size_t i = 0; i < n; ++i)
    if (all || cost[i] >= thr) out.push_back(inner[i]);
  if (out.size() < used) out.resize(used, UINT32_MAX);
  return out;
}

// BVH::sibling -- lib/base/BVH.hpp:149-151
inline uint32_t sibling(const std::vector<Node> &b, uint32_t n) {
  return b[n].isLeft == 0 ? b[b[n].parent].left : b[b[n].parent].right;
}

// ---------------------------- sequential ------------------------------------

struct FindNodeStruct {  // Bittner13.cpp:64-69
  uint32_t node;
  float cost;
  bool operator<(const FindNodeStruct &b) const noexcept { return cost > b.cost; }
};

struct SeqOpt {
  std::vector<Node> &b;
  uint32_t root;
  uint64_t heapOverflows = 0;
  static constexpr size_t QUEUE_SIZE = 16384;  // Bittner13.hpp:28
  std::vector<FindNodeStruct> pq = std::vector<FindNodeStruct>(QUEUE_SIZE);

  // Bittner13::fixTree -- Bittner13.cpp:198-220
  void fixTree(uint32_t node) {
    bool lastWasLeft = true;
    Box box = b[b[node].left].bbox;
    while (true) {
      uint32_t s = lastWasLeft ? b[node].right : b[node].left;
      mergeWith(box, b[s].bbox);
      b[node].bbox = box;
      b[node].surfaceArea = boxSA(box);
      if (node == root) return;
      lastWasLeft = b[node].isLeft != 0;
      node = b[node].parent;
    }
  }

  // Bittner13::findNodeForReinsertion -- Bittner13.cpp:71-113 (heap order of
  // libstdc++ push_heap/pop_heap, which CUDAHeap.hpp:33-112 copies).
  uint32_t findNode(uint32_t n) { return findNodeBox(b[n].bbox, b[n].surfaceArea); }

  uint32_t findNodeBox(const Box xbox, float sa) {
    float best = std::numeric_limits<float>::infinity();
    uint32_t bestIdx = 0;
    size_t size = 1;
    pq[0] = {root, 0.0f};
    while (size > 0) {
      FindNodeStruct curr = pq[0];
      const Node &cn = b[curr.node];
      std::pop_heap(pq.begin(), pq.begin() + size);
      size--;
      if ((curr.cost + sa) >= best) break;
      Box m = xbox;  // directCost(lNode, lCurrNode), Bittner13.hpp:38-42
      mergeWith(m, cn.bbox);
      float direct = boxSA(m);
      float total = curr.cost + direct;
      if (total < best) {
        best = total;
        bestIdx = curr.node;
      }
      float newInduced = total - cn.surfaceArea;
      if ((newInduced + sa) < best && !isLeaf(cn)) {
        if (size + 2 > QUEUE_SIZE) {
          heapOverflows++;
          continue;
        }
        pq[size + 0] = {cn.left, newInduced};
        pq[size + 1] = {cn.right, newInduced};
        std::push_heap(pq.begin(), pq.begin() + size + 1);
        std::push_heap(pq.begin(), pq.begin() + size + 2);
        size += 2;
      }
    }
    return bestIdx;
  }

  // Bittner13::removeNode -- Bittner13.cpp:115-158
  bool removeNode(uint32_t node, uint32_t ins[2], uint32_t unused[2]) {
    if (isLeaf(b[node]) || node == root) return false;
    uint32_t sib = sibling(b, node);
    uint32_t parent = b[node].parent;
    uint32_t gp = b[parent].parent;
    float lSA = b[b[node].left].surfaceArea, rSA = b[b[node].right].surfaceArea;
    if (parent == root) return false;
    if (b[parent].isLeft != 0) {
      b[gp].left = sib;
      b[sib].isLeft = 1;
      b[sib].parent = gp;
    } else {
      b[gp].right = sib;
      b[sib].isLeft = 0;
      b[sib].parent = gp;
    }
    fixTree(gp);
    if (lSA > rSA) {
      ins[0] = b[node].left;
      ins[1] = b[node].right;
    } else {
      ins[0] = b[node].right;
      ins[1] = b[node].left;
    }
    unused[0] = node;
    unused[1] = parent;
    return true;
  }

  // Bittner13::reinsert -- Bittner13.cpp:161-196
  void reinsert(uint32_t node, uint32_t unused) {
    uint32_t best = findNode(node);
    uint32_t rootIdx = b[best].parent;
    if (best == root) {
      root = unused;
      rootIdx = unused;
    } else {
      if (b[best].isLeft != 0) {
        b[rootIdx].left = unused;
        b[unused].isLeft = 1;
      } else {
        b[rootIdx].right = unused;
        b[unused].isLeft = 0;
      }
    }
    b[unused].parent = rootIdx;
    b[unused].left = best;
    b[unused].right = node;
    b[best].parent = unused;
    b[best].isLeft = 1;
    b[node].parent = unused;
    b[node].isLeft = 0;
    fixTree(unused);
  }
};

double sahOf(const std::vector<Node> &b, uint32_t root, float cInner, float cLeaf) {
  double s = 0.0;
  for (uint32_t i = 0; i < b.size(); ++i) {
    if (i == root) continue;
    float c = b[i].surfaceArea;
    c *= isLeaf(b[i]) ? cLeaf : cInner;
    s += c;
  }
  return s / b[root].surfaceArea;
}

// ------------------------------ GPU reference --------------------------------

// BVHNodePatch -- lib/base/BVHPatch.hpp:26-38
struct PNode {
  uint32_t parent, left, right;
  uint16_t isLeafFlag, isLeft;
};

constexpr uint32_t NNode = 10, NPath = 2, NAABB = 2;  // BVHPatch.hpp:40-42
constexpr uint32_t P_GRAND_PARENT = 0, P_1ST_ROOT = 1, P_1ST_BEST = 2, P_NODE = 3, P_1ST_INSERT = 4,
                   P_2ND_ROOT = 5, P_2ND_BEST = 6, P_2ND_INSERT = 7, P_SIBLING = 8, P_PARENT = 9;  // :44-53
constexpr uint32_t NONE = UINT32_MAX;

// BVHPatch (ENABLE_AABB_PATH == 0) -- lib/base/BVHPatch.hpp:87-335
struct Patch {
  PNode vNodes[NNode];
  uint32_t vPatch[NNode];
  uint32_t vAABBPath[NPath * NAABB];
  Box vAABBs[NPath * NAABB];
  std::vector<Node> *bvh;
  uint32_t rootIdx;

  static PNode toP(const Node &n) { return {n.parent, n.left, n.right, n.isLeafFlag, n.isLeft}; }
  void clear() {
    for (uint32_t i = 0; i < NNode; ++i) vPatch[i] = NONE;
    for (uint32_t i = 0; i < NAABB * NPath; ++i) vAABBPath[i] = NONE;
  }
  uint32_t patchIndexSubset(uint32_t n) const {
    if (vPatch[0] == n) return 0;
    if (vPatch[1] == NONE) return NONE;
    if (vPatch[1] == n) return 1;
    if (vPatch[2] == n) return 2;
    if (vPatch[3] == n) return 3;
    return NONE;
  }
  uint32_t patchIndex(uint32_t n) const {
    for (uint32_t i = 0; i < NNode; ++i)
      if (vPatch[i] == n) return i;
    return NONE;
  }
  PNode get(uint32_t n) const {
    uint32_t i = patchIndex(n);
    return i == NONE ? toP((*bvh)[n]) : vNodes[i];
  }
  PNode getSubset(uint32_t n) const {
    uint32_t i = patchIndexSubset(n);
    return i == NONE ? toP((*bvh)[n]) : vNodes[i];
  }
  const Node &getOrig(uint32_t n) const { return (*bvh)[n]; }
  PNode *patchNode(uint32_t n, uint32_t slot) {
    vPatch[slot] = n;
    vNodes[slot] = toP((*bvh)[n]);
    return &vNodes[slot];
  }
  PNode *movePatch(uint32_t from, uint32_t to) {
    vPatch[to] = vPatch[from];
    vPatch[from] = NONE;
    vNodes[to] = vNodes[from];
    return &vNodes[to];
  }
  uint32_t sibling(const PNode &n) const { return n.isLeft == 0 ? get(n.parent).left : get(n.parent).right; }
  Box getAABB(uint32_t n) const {
    for (uint32_t i = 0; i < NPath * NAABB; ++i)
      if (vAABBPath[i] == n) return vAABBs[i];
    return (*bvh)[n].bbox;
  }
  // patchAABBFrom -- BVHPatch.hpp:293-316
  void patchAABBFrom(uint32_t node, uint32_t pIdx) {
    uint32_t idx = node;
    PNode cur = getSubset(idx);
    Box box = getAABB(cur.left);
    bool lastWasLeft = true;
    for (uint32_t i = 0; i < NAABB; ++i) {
      uint32_t s = lastWasLeft ? cur.right : cur.left;
      mergeWith(box, getAABB(s));
      vAABBs[pIdx * NPath + i] = box;
      vAABBPath[pIdx * NPath + i] = idx;
      lastWasLeft = cur.isLeft != 0;
      if (idx == rootIdx) break;
      idx = cur.parent;
      cur = getSubset(idx);
    }
  }
};

struct HelperStruct {  // CUDAHelperStruct -- Bittner13CUDA.cu:110-116
  uint32_t node;
  float cost;
  uint32_t level;
  bool operator<(const HelperStruct &b) const noexcept { return cost > b.cost; }
};

struct GpuRef {
  std::vector<Node> &b;
  uint32_t root = 0;  // the GPU optimizer never moves the root
  uint64_t heapOverflows = 0;
  static constexpr size_t CUDA_QUEUE_SIZE = 4096;  // Bittner13CUDA.hpp:27
  std::vector<HelperStruct> pq = std::vector<HelperStruct>(CUDA_QUEUE_SIZE);
  float lastBest = 0.0f;  // cost of the node the last search returned

  // findNode1 -- Bittner13CUDA.cu:141-191
  uint32_t findNode1(uint32_t n, const Patch &p) {
    float best = HUGE_VALF;
    uint32_t bestIdx = 0;
    const Node &x = p.getOrig(n);
    Box xbox = x.bbox;
    float sa = x.surfaceArea;
    size_t size = 1;
    pq[0] = {root, 0.0f, 0};
    while (size > 0) {
      HelperStruct curr = pq[0];
      if ((curr.cost + sa) >= best) break;
      std::pop_heap(pq.begin(), pq.begin() + size);
      size--;
      PNode cn = p.getSubset(curr.node);
      Box box = p.getAABB(curr.node);
      float newInduced = -1 * boxSA(box);
      mergeWith(box, xbox);
      float direct = boxSA(box);
      float total = curr.cost + direct;
      if (total < best) {
        best = total;
        bestIdx = curr.node;
      }
      newInduced += total;
      if ((newInduced + sa) < best) {
        if (cn.isLeafFlag == 0) {
          if (size + 2 >= CUDA_QUEUE_SIZE) {
            heapOverflows++;
            continue;
          }
          pq[size + 0] = {cn.left, newInduced, curr.level + 1};
          pq[size + 1] = {cn.right, newInduced, curr.level + 1};
          std::push_heap(pq.begin(), pq.begin() + size + 1);
          std::push_heap(pq.begin(), pq.begin() + size + 2);
          size += 2;
        }
      }
    }
    lastBest = best;
    return bestIdx;
  }

  // findNode2 -- Bittner13CUDA.cu:194-264: best-first search over a fixed
  // array of `altSlots` entries (CUDA_ALT_QUEUE_SIZE = 16,
  // Bittner13CUDA.hpp:28).  The popped entry's slot (the minimum) takes the
  // left child and the slot of the maximum cost of the previous scan takes
  // the right child; an empty slot holds HUGE_VALF, so it is the maximum
  // while one is free, and once all slots are full the costliest entry is
  // overwritten (the search becomes approximate).  Both scans keep the first
  // index of their extreme (strict < and >); the maximum scan starts from 0
  // and keeps the old index when no slot is above 0.  Exactly one entry is
  // popped per iteration and a tree pushes each node at most once, so a
  // search over a tree ends within node-count iterations; a longer one (a
  // patched view that is not a tree) is stopped and fails.
  uint32_t altSlots = 16;
  uint64_t altRunaway = 0;
  uint32_t findNode2(uint32_t n, const Patch &p) {
    float best = HUGE_VALF;
    uint32_t bestIdx = 0;
    const Node &x = p.getOrig(n);
    const Box xbox = x.bbox;
    const float sa = x.surfaceArea;
    float lMin = 0.0f, lMax = HUGE_VALF;
    uint32_t minIdx = 0, maxIdx = 1;
    std::vector<uint32_t> nl(altSlots, 0);
    std::vector<float> co(altSlots, HUGE_VALF);
    nl[0] = root;
    co[0] = 0.0f;
    (void)lMax;
    for (uint64_t it = 0; lMin < HUGE_VALF; ++it) {
      if (it >= b.size()) {
        altRunaway++;
        return NONE;
      }
      const uint32_t curN = nl[minIdx];
      const float curC = co[minIdx];
      co[minIdx] = HUGE_VALF;
      if ((curC + sa) >= best) break;
      PNode cn = p.getSubset(curN);
      Box box = p.getAABB(curN);
      float newInduced = -1 * boxSA(box);
      mergeWith(box, xbox);
      const float total = curC + boxSA(box);
      if (total < best) {
        best = total;
        bestIdx = curN;
      }
      newInduced += total;
      if ((newInduced + sa) < best && cn.isLeafFlag == 0) {
        nl[minIdx] = cn.left;
        nl[maxIdx] = cn.right;
        co[minIdx] = newInduced;
        co[maxIdx] = newInduced;
      }
      lMin = HUGE_VALF;
      lMax = 0.0f;
      for (uint32_t i = 0; i < altSlots; ++i) {
        if (co[i] < lMin) {
          lMin = co[i];
          minIdx = i;
        }
        if (co[i] > lMax) {
          lMax = co[i];
          maxIdx = i;
        }
      }
    }
    lastBest = best;
    return bestIdx;
  }

  // The device search of this project's Bittner13 kernel, restated step for
  // step (not a reference algorithm): an exact branch and bound over a LIFO
  // frontier, taking up to 64 entries per step (entry top-1-j goes to lane
  // j); the step minimum replaces the best cost when strictly smaller, ties
  // inside the step go to the smallest node id; children are pushed in lane
  // order, the frontier is truncated at `waveCap` entries.  Used to replay the
  // device trajectory on the CPU (searchMode 1 and 2).
  int searchMode = 0;
  uint32_t waveCap = 1024;
  uint64_t waveOverflows = 0, waveRunaway = 0;
  uint64_t waveIters = 0, waveMaxIters = 0, chunkMaxIters = 0, sumChunkMax = 0;
  std::vector<uint32_t> sNode = std::vector<uint32_t>(1024);
  std::vector<float> sCost = std::vector<float>(1024);
  uint32_t findNodeWave(uint32_t n, const Patch &p) {
    const Node &x = p.getOrig(n);
    const Box xbox = x.bbox;
    const float sa = x.surfaceArea;
    sNode[0] = root;
    sCost[0] = 0.0f;
    uint32_t top = 1;
    float best = HUGE_VALF;
    uint32_t bestNode = 0;
    uint32_t lnode[64], left[64], right[64];
    float total[64], newInd[64];
    bool eval[64], leaf[64];
    // the device bound: one step per node (each step pops one entry or more,
    // a tree pushes each node once); a view that is not a tree fails the search
    const uint64_t limit = b.size();
    for (uint64_t it = 0; top > 0; ++it) {
      if (it >= limit) {
        waveRunaway++;
        return NONE;
      }
      waveIters++;
      if (it + 1 > chunkMaxIters) chunkMaxIters = it + 1;
      if (it + 1 > waveMaxIters) waveMaxIters = it + 1;
      const uint32_t m = top < 64 ? top : 64;
      const uint32_t newTop = top - m;
      float mt = HUGE_VALF;
      for (uint32_t j = 0; j < m; ++j) {
        lnode[j] = sNode[top - 1 - j];
        const float ind = sCost[top - 1 - j];
        eval[j] = (ind + sa) < best;
        total[j] = HUGE_VALF;
        if (!eval[j]) continue;
        PNode cn = p.getSubset(lnode[j]);
        Box box = p.getAABB(lnode[j]);
        float ni = -1.0f * boxSA(box);
        mergeWith(box, xbox);
        total[j] = ind + boxSA(box);
        newInd[j] = ni + total[j];
        leaf[j] = cn.isLeafFlag != 0;
        left[j] = cn.left;
        right[j] = cn.right;
        if (total[j] < mt) mt = total[j];
      }
      if (mt < best) {
        uint32_t cand = NONE;
        for (uint32_t j = 0; j < m; ++j)
          if (eval[j] && total[j] == mt && lnode[j] < cand) cand = lnode[j];
        best = mt;
        bestNode = cand;
      }
      uint32_t pos = newTop;
      for (uint32_t j = 0; j < m; ++j) {
        if (!(eval[j] && !leaf[j] && (newInd[j] + sa) < best)) continue;
        if (pos + 1 < waveCap) {
          sNode[pos] = left[j];
          sCost[pos] = newInd[j];
          sNode[pos + 1] = right[j];
          sCost[pos + 1] = newInd[j];
        }
        pos += 2;
      }
      if (pos > waveCap) {
        waveOverflows++;
        pos = waveCap & ~1u;
      }
      top = pos;
    }
    lastBest = best;
    return bestNode;
  }

  // removeNode -- Bittner13CUDA.cu:267-305.  Returns false without patching
  // anything when the node cannot be removed (deviation (1)).
  bool removeNode(uint32_t node, Patch &p, uint32_t ins[2], uint32_t unused[2], uint32_t gpSib[2]) {
    if (isLeaf(p.getOrig(node)) || node == root) return false;
    if (p.getOrig(node).parent == root) return false;
    PNode *lNode = p.patchNode(node, P_NODE);
    uint32_t sibIdx = p.sibling(*lNode);
    uint32_t parentIdx = lNode->parent;
    PNode *lSib = p.patchNode(sibIdx, P_SIBLING);
    PNode *lPar = p.patchNode(parentIdx, P_PARENT);
    uint32_t gpIdx = lPar->parent;
    PNode *lGP = p.patchNode(gpIdx, P_GRAND_PARENT);
    if (lPar->isLeft != 0) {
      lGP->left = sibIdx;
      lSib->isLeft = 1;
      lSib->parent = gpIdx;
    } else {
      lGP->right = sibIdx;
      lSib->isLeft = 0;
      lSib->parent = gpIdx;
    }
    p.patchAABBFrom(gpIdx, 0);
    if (p.getOrig(lNode->left).surfaceArea > p.getOrig(lNode->right).surfaceArea) {
      ins[0] = lNode->left;
      ins[1] = lNode->right;
    } else {
      ins[0] = lNode->right;
      ins[1] = lNode->left;
    }
    unused[0] = node;
    unused[1] = parentIdx;
    gpSib[0] = gpIdx;
    gpSib[1] = sibIdx;
    return true;
  }

  // reinsert -- Bittner13CUDA.cu:308-358.  searchMode 0: findNode1, 1: the
  // device wave search, 2: findNode2 (altFindNode).
  bool reinsert(uint32_t node, uint32_t unusedIdx, Patch &p, bool update, uint32_t &bestOut) {
    uint32_t bestIdx = searchMode == 1 ? findNodeWave(node, p) : searchMode == 2 ? findNode2(node, p) : findNode1(node, p);
    if (bestIdx == root || bestIdx == NONE) return false;
    uint32_t bp = p.patchIndex(bestIdx);
    PNode *best;
    if (bp == NONE)
      best = p.patchNode(bestIdx, update ? P_1ST_BEST : P_2ND_BEST);
    else if (bp == P_GRAND_PARENT)
      best = &p.vNodes[P_GRAND_PARENT];
    else
      best = p.movePatch(bp, update ? P_1ST_BEST : P_2ND_BEST);
    PNode *lNode = p.patchNode(node, update ? P_1ST_INSERT : P_2ND_INSERT);
    PNode *lUnused = &p.vNodes[update ? P_NODE : P_PARENT];
    uint32_t rootIdx = best->parent;
    uint32_t rp = p.patchIndex(rootIdx);
    PNode *lRoot = rp == NONE ? p.patchNode(rootIdx, update ? P_1ST_ROOT : P_2ND_ROOT)
                              : p.movePatch(rp, update ? P_1ST_ROOT : P_2ND_ROOT);
    if (best->isLeft != 0) {
      lRoot->left = unusedIdx;
      lUnused->isLeft = 1;
    } else {
      lRoot->right = unusedIdx;
      lUnused->isLeft = 0;
    }
    lUnused->parent = rootIdx;
    lUnused->left = bestIdx;
    lUnused->right = node;
    best->parent = unusedIdx;
    best->isLeft = 1;
    lNode->parent = unusedIdx;
    lNode->isLeft = 0;
    if (update) p.patchAABBFrom(unusedIdx, 1);
    bestOut = bestIdx;
    return true;
  }

  // A failure model, not a reference algorithm: the touched-path refit
  // (kFixTree3, Bittner13CUDA.cu:407-467, from the grandparent and the two
  // freed nodes of every applied patch) in which each read of a child box
  // that this refit itself rewrites returns the box from before the refit
  // with probability staleProb (a store that had not landed when the last
  // arrival read it).  Used to show that the topology the optimizer builds,
  // and the length of its searches, do not depend on box values.
  double staleProb = 0.0;
  uint64_t staleState = 0x9E3779B97F4A7C15ull;
  uint64_t staleReads = 0;
  bool staleDraw() {
    staleState ^= staleState << 13;
    staleState ^= staleState >> 7;
    staleState ^= staleState << 17;
    return (double)(staleState >> 11) * (1.0 / 9007199254740992.0) < staleProb;
  }
  void fixTreeStale(const std::vector<uint32_t> &starts) {
    const uint32_t N = (uint32_t)b.size();
    std::vector<std::pair<uint32_t, uint32_t>> touched;  // (depth, node)
    std::vector<uint8_t> seen(N, 0);
    for (uint32_t st : starts) {
      uint32_t n = st;
      for (uint32_t guard = 0; guard < N && !seen[n]; ++guard) {
        seen[n] = 1;
        uint32_t d = 0;
        for (uint32_t a = n; a != root && d < N; a = b[a].parent) ++d;
        touched.push_back({d, n});
        if (n == root) break;
        n = b[n].parent;
      }
    }
    std::sort(touched.begin(), touched.end(), [](const std::pair<uint32_t, uint32_t> &x,
                                                 const std::pair<uint32_t, uint32_t> &y) { return x.first > y.first; });
    std::vector<Box> old(touched.size());
    std::vector<uint32_t> slot(N, UINT32_MAX);
    for (size_t i = 0; i < touched.size(); ++i) {
      old[i] = b[touched[i].second].bbox;
      slot[touched[i].second] = (uint32_t)i;
    }
    auto childBox = [&](uint32_t c) {
      if (slot[c] != UINT32_MAX && staleDraw()) {
        staleReads++;
        return old[slot[c]];
      }
      return b[c].bbox;
    };
    for (auto &t : touched) {
      Node &n = b[t.second];
      if (isLeaf(n)) continue;
      Box box = childBox(n.left);
      mergeWith(box, childBox(n.right));
      n.bbox = box;
      n.surfaceArea = boxSA(box);
    }
  }

  // kFixTree1 -- Bittner13CUDA.cu:372-404 (every leaf walks up; the second
  // arrival at a node merges its children).
  void fixTree1() {
    const uint32_t N = (uint32_t)b.size();
    std::vector<uint32_t> flags(N, 0);
    for (uint32_t i = 0; i < N; ++i) {
      if (!isLeaf(b[i])) continue;
      uint32_t node = b[i].parent;
      uint64_t guard = 0;
      while (true) {
        if (++guard > N) return;
        if (flags[node] == 0) {
          flags[node] = 1;
          break;
        }
        Box box = b[b[node].left].bbox;
        mergeWith(box, b[b[node].right].bbox);
        b[node].bbox = box;
        b[node].surfaceArea = boxSA(box);
        if (node == b[node].parent) break;
        node = b[node].parent;
      }
    }
  }
};

// Nodes reachable from the root and whether every node is reached exactly
// once through consistent parent links.
bool treeIsValid(const std::vector<Node> &b, uint32_t root) {
  const uint32_t N = (uint32_t)b.size();
  std::vector<uint8_t> seen(N, 0);
  std::vector<uint32_t> st = {root};
  uint32_t count = 0;
  while (!st.empty()) {
    uint32_t n = st.back();
    st.pop_back();
    if (n >= N || seen[n]) return false;
    seen[n] = 1;
    count++;
    if (!isLeaf(b[n])) {
      uint32_t l = b[n].left, r = b[n].right;
      if (l >= N || r >= N || b[l].parent != n || b[r].parent != n || b[l].isLeft == 0 || b[r].isLeft != 0) return false;
      st.push_back(l);
      st.push_back(r);
    }
  }
  return count == N;
}

// Bittner13::runMetric selection (Bittner13.cpp:224-283): every node is a
// candidate (leaves with cost 0), nth_element puts the numNodes largest
// first, sort orders them when the batch is sorted.
typedef std::tuple<uint32_t, float> TUP;
void seqSelect(const std::vector<Node> &b, uint32_t numNodes, bool sortBatch, std::vector<TUP> &todo) {
  const uint32_t N = (uint32_t)b.size();
  todo.resize(N);
  for (uint32_t j = 0; j < N; ++j) todo[j] = TUP{j, priorityOf(b, j, true)};
  auto comp = [](TUP const &l, TUP const &r) -> bool { return std::get<1>(l) > std::get<1>(r); };
  std::nth_element(todo.begin(), todo.begin() + numNodes, todo.end(), comp);
  if (sortBatch) std::sort(todo.begin(), todo.begin() + numNodes, comp);
}

}  // namespace
}  // namespace orc

using namespace orc;

extern "C" {

// Bittner13::runMetric -- lib/builder/Bittner13.cpp:222-312 (sort on, not
// strictly sequential), then fixLevels at teardown (:392).
//   sahOut[steps + 1]: SAH(1, 1) before the first step and after each step.
// Returns 0 on success; *root receives the (possibly moved) root.
int orc_bittner13_seq(Node *nodes, uint32_t N, uint32_t *root, float batchPercent, uint32_t steps, int sortBatch,
                      double *sahOut, uint64_t *heapOverflows) {
  std::vector<Node> b(nodes, nodes + N);
  SeqOpt opt{b, *root};
  const uint32_t numNodes = batchCount(batchPercent, N);
  std::vector<TUP> todo(N);
  if (sahOut) sahOut[0] = sahOf(b, opt.root, 1.0f, 1.0f);
  for (uint32_t s = 0; s < steps; ++s) {
    seqSelect(b, numNodes, sortBatch != 0, todo);
    for (uint32_t j = 0; j < numNodes; ++j) {
      uint32_t ins[2], unused[2];
      if (!opt.removeNode(std::get<0>(todo[j]), ins, unused)) continue;
      opt.reinsert(ins[0], unused[0]);
      opt.reinsert(ins[1], unused[1]);
    }
    if (sahOut) sahOut[s + 1] = sahOf(b, opt.root, 1.0f, 1.0f);
  }
  *root = opt.root;
  std::copy(b.begin(), b.end(), nodes);
  if (heapOverflows) *heapOverflows = opt.heapOverflows;
  return 0;
}

// The GPU optimizer of lib/builder/Bittner13CUDA.cu with the defaults of
// lib/builder/Bittner13GPU.hpp:28-37 (sort on, kFixTree1), serialised thread
// by thread; patches are kept local until they succeed (localPatchCPY,
// kRemoveAndReinsert2, :533-564, which is deviation (1)).
//   searchMode: 0 findNode1, 1 this library's wave search, 2 findNode2
//   (altFindNode); altSort: the threshold selection of bn13_selectNodes.
//   sahOut[steps + 1]: CUDAcalcSAH (Bittner13CUDA.cu:776-803) before the
//   first step and after each step (f64 sum).
//   stats[0] = skipped (lock conflicts), [1] = patches dropped by the
//   reinsert-at-root rule or a failed removal, [2] = patches dropped by
//   deviation (2), [3] = heap / frontier overflows, [4] = searches stopped
//   because the patched view was not a tree.
//   chunkTrees (optional, steps * numChunks * N nodes): the tree after every
//   chunk.
// Returns 0, or 2 if the tree became invalid.
// Failure model for orc_bittner13_gpuref (see GpuRef::fixTreeStale): when
// prob > 0 the per-chunk refit is the touched-path refit with stale child
// reads.  stats[4] then still counts runaway searches; *staleReads receives
// the number of stale reads of the last run.
static double gStaleProb = 0.0;
static uint64_t gStaleSeed = 1, gStaleReads = 0, gMaxSearchSteps = 0;
void orc_set_stale_refit(double prob, uint64_t seed) {
  gStaleProb = prob;
  gStaleSeed = seed ? seed : 1;
}
void orc_last_run_info(uint64_t *staleReads, uint64_t *maxSearchSteps) {
  *staleReads = gStaleReads;
  *maxSearchSteps = gMaxSearchSteps;
}

int orc_bittner13_gpuref(Node *nodes, uint32_t N, float batchPercent, uint32_t numChunks, uint32_t steps,
                         int offsetAccess, int searchMode, int altSort, Node *chunkTrees, double *sahOut,
                         uint64_t *stats) {
  std::vector<Node> b(nodes, nodes + N);
  GpuRef g{b};
  g.searchMode = searchMode;
  g.waveCap = 1024;
  // Bittner13GPU::setup/runImpl -- Bittner13GPU.cpp:79-103
  const uint32_t numNodes = batchCount(batchPercent, N);
  const uint32_t K = numNodes / numChunks;
  std::vector<uint32_t> inner;  // initData: DeviceSelect::If over iota (:962-978)
  for (uint32_t i = 0; i < N; ++i)
    if (!isLeaf(b[i])) inner.push_back(i);
  std::vector<uint32_t> sortedIds;
  std::vector<Patch> patches(K);
  std::vector<uint32_t> flags(N, 0);
  std::vector<int32_t> mark(N, -1);
  std::vector<uint32_t> bestOf(2 * (size_t)K);
  std::vector<uint32_t> movedOf(2 * (size_t)K);
  std::vector<uint8_t> ok(K);
  std::vector<uint32_t> fixOf(3 * (size_t)K);
  g.staleProb = gStaleProb;
  g.staleState = gStaleSeed * 0x9E3779B97F4A7C15ull + 1;
  uint64_t st[5] = {0, 0, 0, 0, 0};
  auto sahGpu = [&]() {
    double s = 0.0;
    for (uint32_t i = 0; i < N; ++i) s += b[i].surfaceArea;
    s -= b[0].surfaceArea;
    return s / b[0].surfaceArea;
  };
  if (sahOut) sahOut[0] = sahGpu();
  for (uint32_t s = 0; s < steps; ++s) {
    // kCalcCost (:480-500) + the selection of bn13_selectNodes (:659-726)
    sortedIds = gpuSelect(b, inner, numChunks * K, altSort != 0);
    for (uint32_t c = 0; c < numChunks; ++c) {
      // kRemoveAndReinsert2 (:533-564)
      for (uint32_t k = 0; k < K; ++k) {
        Patch &p = patches[k];
        p.bvh = &b;
        p.rootIdx = g.root;
        p.clear();
        ok[k] = 0;
        uint32_t nodeIdx = sortedIds[offsetAccess ? k * numChunks + c : c * K + k];
        if (nodeIdx == NONE) continue;  // fewer nodes selected than the chunks take
        uint32_t ins[2], unused[2], gpSib[2], b1 = 0, b2 = 0;
        if (!g.removeNode(nodeIdx, p, ins, unused, gpSib)) {
          st[1]++;
          continue;
        }
        if (!g.reinsert(ins[0], unused[0], p, true, b1) || !g.reinsert(ins[1], unused[1], p, false, b2)) {
          p.clear();
          st[1]++;
          continue;
        }
        ok[k] = 1;
        bestOf[2 * k] = b1;
        bestOf[2 * k + 1] = b2;
        movedOf[2 * k] = ins[0];
        movedOf[2 * k + 1] = ins[1];
        fixOf[3 * k] = gpSib[0];
        fixOf[3 * k + 1] = unused[0];
        fixOf[3 * k + 2] = unused[1];
      }
      // kCheckConflicts (:567-609): thread k locks with id k + 1; a thread
      // that fails releases its own locks, counts a skip, clears its patch.
      for (uint32_t k = 0; k < K; ++k) {
        if (!ok[k]) continue;
        Patch &p = patches[k];
        bool failed = false;
        for (uint32_t i = 0; i < NNode && !failed; ++i)
          if (p.vPatch[i] != NONE) {
            if (flags[p.vPatch[i]] != 0) failed = true;
            else flags[p.vPatch[i]] = k + 1;
          }
        if (failed) {
          for (uint32_t i = 0; i < NNode; ++i)
            if (p.vPatch[i] != NONE && flags[p.vPatch[i]] == k + 1) flags[p.vPatch[i]] = 0;
          st[0]++;
          p.clear();
          ok[k] = 0;
        }
      }
      // deviation (2): no insertion into a subtree moved by another patch
      for (uint32_t k = 0; k < K; ++k)
        if (ok[k]) {
          mark[movedOf[2 * k]] = (int32_t)k;
          mark[movedOf[2 * k + 1]] = (int32_t)k;
        }
      for (uint32_t k = 0; k < K; ++k) {
        if (!ok[k]) continue;
        bool bad = false;
        for (int w = 0; w < 2 && !bad; ++w) {
          uint32_t n = bestOf[2 * k + w];
          for (uint32_t guard = 0; guard < N; ++guard) {
            if (mark[n] >= 0 && mark[n] != (int32_t)k) {
              bad = true;
              break;
            }
            if (n == g.root) break;
            n = b[n].parent;
          }
        }
        if (bad) {
          Patch &p = patches[k];
          for (uint32_t i = 0; i < NNode; ++i)
            if (p.vPatch[i] != NONE) flags[p.vPatch[i]] = 0;
          p.clear();
          ok[k] = 0;
          st[2]++;
        }
      }
      for (uint32_t k = 0; k < K; ++k) {
        mark[movedOf[2 * k]] = -1;
        mark[movedOf[2 * k + 1]] = -1;
      }
      // kApplyPatches (:612-627)
      for (uint32_t k = 0; k < K; ++k) {
        Patch &p = patches[k];
        for (uint32_t l = 0; l < NNode; ++l) {
          if (p.vPatch[l] == NONE) continue;
          flags[p.vPatch[l]] = 0;
          Node &n = b[p.vPatch[l]];
          n.parent = p.vNodes[l].parent;
          n.left = p.vNodes[l].left;
          n.right = p.vNodes[l].right;
          n.isLeft = p.vNodes[l].isLeft;
        }
        p.clear();
      }
      if (!treeIsValid(b, g.root)) {
        std::copy(b.begin(), b.end(), nodes);
        return 2;
      }
      if (g.staleProb > 0.0) {
        std::vector<uint32_t> starts;
        for (uint32_t k = 0; k < K; ++k)
          if (ok[k])
            for (int j = 0; j < 3; ++j) starts.push_back(fixOf[3 * k + j]);
        g.fixTreeStale(starts);
      } else {
        g.fixTree1();
      }
      g.sumChunkMax += g.chunkMaxIters;
      g.chunkMaxIters = 0;
      if (chunkTrees) std::copy(b.begin(), b.end(), chunkTrees + ((size_t)s * numChunks + c) * N);
    }
    if (sahOut) sahOut[s + 1] = sahGpu();
  }
  st[3] = g.searchMode == 1 ? g.waveOverflows : g.searchMode == 0 ? g.heapOverflows : 0;
  st[4] = g.waveRunaway + g.altRunaway;
  gStaleReads = g.staleReads;
  gMaxSearchSteps = g.waveMaxIters;
  if (std::getenv("ORC_WAVE_TRACE"))
    std::fprintf(stderr, "wave iterations: total %llu, max per search %llu, sum of per-chunk maxima %llu\n",
                 (unsigned long long)g.waveIters, (unsigned long long)g.waveMaxIters,
                 (unsigned long long)g.sumChunkMax);
  if (stats) std::copy(st, st + 5, stats);
  std::copy(b.begin(), b.end(), nodes);
  return 0;
}

// ---- single operations, for the known-answer and brute-force tests -------

// Removal priority of every node: GPU form (seq == 0, kCalcCost) or the
// sequential form (seq != 0).
void orc_bn13_priorities(const Node *nodes, uint32_t N, int seq, float *out) {
  std::vector<Node> b(nodes, nodes + N);
  for (uint32_t j = 0; j < N; ++j) out[j] = priorityOf(b, j, seq != 0);
}

// The node order one step processes.  mode 0: GPU sort, 1: GPU altSort,
// 2: sequential (nth_element + sort), 3: sequential without the sort.
// out (capacity N) receives the order; *count its length; *used the number
// of entries a step takes (numChunks * chunkSize for the GPU optimizer, the
// batch size for the sequential one).
void orc_bn13_select(const Node *nodes, uint32_t N, float batchPercent, uint32_t numChunks, int mode, uint32_t *out,
                     uint32_t *count, uint32_t *used) {
  std::vector<Node> b(nodes, nodes + N);
  const uint32_t numNodes = batchCount(batchPercent, N);
  if (mode <= 1) {
    std::vector<uint32_t> inner;
    for (uint32_t i = 0; i < N; ++i)
      if (!isLeaf(b[i])) inner.push_back(i);
    const uint32_t u = (numNodes / numChunks) * numChunks;
    std::vector<uint32_t> o = gpuSelect(b, inner, u, mode == 1);
    std::copy(o.begin(), o.end(), out);
    *count = (uint32_t)o.size();
    *used = u;
    return;
  }
  std::vector<TUP> todo;
  seqSelect(b, numNodes, mode == 2, todo);
  for (uint32_t j = 0; j < N; ++j) out[j] = std::get<0>(todo[j]);
  *count = N;
  *used = numNodes;
}

// One patch of the GPU optimizer (removeNode + both reinsertions,
// Bittner13CUDA.cu:267-358) on its own, applied (kApplyPatches) and refit
// (kFixTree1).  info[0]: 0 applied, 1 removal refused (leaf, root or child
// of the root), 2 a reinsertion refused (best == root, :310) or failed;
// info[1..2] insertion points, [3..4] reinserted nodes, [5..6] the freed
// node and its parent, [7] the grandparent.  The tree is unchanged unless
// info[0] == 0.
void orc_gpuref_patch(Node *nodes, uint32_t N, uint32_t nodeIdx, int searchMode, uint32_t altSlots, uint32_t *info) {
  std::vector<Node> b(nodes, nodes + N);
  GpuRef g{b};
  g.searchMode = searchMode;
  g.altSlots = altSlots < 2 ? 2 : altSlots;
  Patch p;
  p.bvh = &b;
  p.rootIdx = g.root;
  p.clear();
  uint32_t ins[2] = {0, 0}, unused[2] = {0, 0}, gpSib[2] = {0, 0}, b1 = 0, b2 = 0;
  for (int i = 0; i < 8; ++i) info[i] = 0;
  if (!g.removeNode(nodeIdx, p, ins, unused, gpSib)) {
    info[0] = 1;
    return;
  }
  info[3] = ins[0];
  info[4] = ins[1];
  info[5] = unused[0];
  info[6] = unused[1];
  info[7] = gpSib[0];
  if (!g.reinsert(ins[0], unused[0], p, true, b1) || !g.reinsert(ins[1], unused[1], p, false, b2)) {
    info[0] = 2;
    return;
  }
  info[1] = b1;
  info[2] = b2;
  for (uint32_t l = 0; l < NNode; ++l) {
    if (p.vPatch[l] == NONE) continue;
    Node &n = b[p.vPatch[l]];
    n.parent = p.vNodes[l].parent;
    n.left = p.vNodes[l].left;
    n.right = p.vNodes[l].right;
    n.isLeft = p.vNodes[l].isLeft;
  }
  g.fixTree1();
  std::copy(b.begin(), b.end(), nodes);
}

// The search alone: the insertion point the GPU optimizer finds for the
// first (stage 0) or the second (stage 1) reinsertion after removing
// `nodeIdx`, on the patched view of that moment.  Returns the node (NONE
// when the removal is refused or, for stage 1, the first reinsertion is),
// *cost its induced cost.
uint32_t orc_gpuref_search(const Node *nodes, uint32_t N, uint32_t nodeIdx, int stage, int searchMode,
                           uint32_t altSlots, float *cost) {
  std::vector<Node> b(nodes, nodes + N);
  GpuRef g{b};
  g.searchMode = searchMode;
  g.altSlots = altSlots < 2 ? 2 : altSlots;
  Patch p;
  p.bvh = &b;
  p.rootIdx = g.root;
  p.clear();
  uint32_t ins[2], unused[2], gpSib[2], b1 = 0;
  if (!g.removeNode(nodeIdx, p, ins, unused, gpSib)) return NONE;
  if (stage == 1 && !g.reinsert(ins[0], unused[0], p, true, b1)) return NONE;
  const uint32_t x = ins[stage == 1 ? 1 : 0];
  uint32_t best = searchMode == 1 ? g.findNodeWave(x, p) : searchMode == 2 ? g.findNode2(x, p) : g.findNode1(x, p);
  if (cost) *cost = g.lastBest;
  return best;
}

// One removal and both reinsertions of the sequential optimizer
// (Bittner13.cpp:115-220).  info[0]: 0 done, 1 refused; info[1..2] the
// reinserted nodes, [3..4] the freed node and its parent.  *root follows a
// reinsertion at the root (:170-173).
void orc_seq_patch(Node *nodes, uint32_t N, uint32_t *root, uint32_t nodeIdx, uint32_t *info) {
  std::vector<Node> b(nodes, nodes + N);
  SeqOpt opt{b, *root};
  uint32_t ins[2], unused[2];
  for (int i = 0; i < 5; ++i) info[i] = 0;
  if (!opt.removeNode(nodeIdx, ins, unused)) {
    info[0] = 1;
    return;
  }
  opt.reinsert(ins[0], unused[0]);
  opt.reinsert(ins[1], unused[1]);
  info[1] = ins[0];
  info[2] = ins[1];
  info[3] = unused[0];
  info[4] = unused[1];
  *root = opt.root;
  std::copy(b.begin(), b.end(), nodes);
}

// kFixTree1 -- Bittner13CUDA.cu:372-404: every inner node reachable from a
// leaf is recomputed from its children (box union and surface area).
void orc_fix_tree1(Node *nodes, uint32_t N) {
  std::vector<Node> b(nodes, nodes + N);
  GpuRef g{b};
  g.fixTree1();
  std::copy(b.begin(), b.end(), nodes);
}

// findNodeForReinsertion for a box that is not in the tree (tests compare
// it with an exhaustive search over every insertion position).
uint32_t orc_seq_find_node_box(Node *nodes, uint32_t N, uint32_t root, const float *box6, float sa) {
  std::vector<Node> b(nodes, nodes + N);
  SeqOpt opt{b, root};
  Box x = {{box6[0], box6[1], box6[2]}, {box6[3], box6[4], box6[5]}};
  return opt.findNodeBox(x, sa);
}

// Tree-shape check used by tests: every node reached once from the root
// with consistent parent / isLeft links.
int orc_tree_is_valid(const Node *nodes, uint32_t N, uint32_t root) {
  std::vector<Node> b(nodes, nodes + N);
  return treeIsValid(b, root) ? 1 : 0;
}

}  // extern "C"
