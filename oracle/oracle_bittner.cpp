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

  // reinsert -- Bittner13CUDA.cu:308-358 (findNode1 path).
  bool reinsert(uint32_t node, uint32_t unusedIdx, Patch &p, bool update, uint32_t &bestOut) {
    uint32_t bestIdx = searchMode ? findNodeWave(node, p) : findNode1(node, p);
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
  if (batchPercent <= 0.01f) batchPercent = 0.01f;
  if (batchPercent >= 75.0f) batchPercent = 75.0f;
  std::vector<Node> b(nodes, nodes + N);
  SeqOpt opt{b, *root};
  typedef std::tuple<uint32_t, float> TUP;
  std::vector<TUP> todo(N);
  uint32_t numNodes = static_cast<uint32_t>((batchPercent / 100.0f) * static_cast<float>(N));
  auto comp = [](TUP const &l, TUP const &r) -> bool { return std::get<1>(l) > std::get<1>(r); };
  if (sahOut) sahOut[0] = sahOf(b, opt.root, 1.0f, 1.0f);
  for (uint32_t s = 0; s < steps; ++s) {
    for (uint32_t j = 0; j < N; ++j) {
      const Node &n = b[j];
      float sa = n.surfaceArea;
      if (isLeaf(n)) {
        todo[j] = TUP{j, 0};
        continue;
      }
      float l = b[n.left].surfaceArea, r = b[n.right].surfaceArea;
      float cost = (sa * sa * sa * 2.0f) / ((l + r) * std::min(l, r));
      todo[j] = TUP{j, cost};
    }
    std::nth_element(todo.begin(), todo.begin() + numNodes, todo.end(), comp);
    if (sortBatch) std::sort(todo.begin(), todo.begin() + numNodes, comp);
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
// lib/builder/Bittner13GPU.hpp:28-37 (sort on, findNode1, kFixTree1,
// patches in global memory), serialised thread by thread.
//   sahOut[steps + 1]: CUDAcalcSAH (Bittner13CUDA.cu:776-803) before the
//   first step and after each step (f64 sum).
//   stats[0] = skipped (lock conflicts), [1] = patches dropped by the
//   reinsert-at-root rule or a failed removal, [2] = patches dropped by
//   deviation (2), [3] = heap / frontier overflows, [4] = replayed device
//   searches stopped because the patched view was not a tree.
//   chunkTrees (optional, steps * numChunks * N nodes): the tree after every
//   chunk.
// Returns 0, or 2 if the tree became invalid.
int orc_bittner13_gpuref(Node *nodes, uint32_t N, float batchPercent, uint32_t numChunks, uint32_t steps,
                         int offsetAccess, int searchMode, Node *chunkTrees, double *sahOut, uint64_t *stats) {
  if (batchPercent <= 0.01f) batchPercent = 0.01f;
  if (batchPercent >= 75.0f) batchPercent = 75.0f;
  std::vector<Node> b(nodes, nodes + N);
  GpuRef g{b};
  g.searchMode = searchMode;
  g.waveCap = searchMode == 2 ? 16 : 1024;
  // Bittner13GPU::runImpl -- Bittner13GPU.cpp:682-708
  uint32_t numNodes = (uint32_t)((batchPercent / 100.0f) * (float)N);
  uint32_t K = numNodes / numChunks;
  std::vector<uint32_t> inner;  // initData: DeviceSelect::If over iota (:962-978)
  for (uint32_t i = 0; i < N; ++i)
    if (!isLeaf(b[i])) inner.push_back(i);
  std::vector<float> cost(inner.size());
  std::vector<uint32_t> sortedIds(inner.size());
  std::vector<Patch> patches(K);
  std::vector<uint32_t> flags(N, 0);
  std::vector<int32_t> mark(N, -1);
  std::vector<uint32_t> bestOf(2 * (size_t)K);
  std::vector<uint32_t> movedOf(2 * (size_t)K);
  std::vector<uint8_t> ok(K);
  uint64_t st[5] = {0, 0, 0, 0, 0};
  auto sahGpu = [&]() {
    double s = 0.0;
    for (uint32_t i = 0; i < N; ++i) s += b[i].surfaceArea;
    s -= b[0].surfaceArea;
    return s / b[0].surfaceArea;
  };
  if (sahOut) sahOut[0] = sahGpu();
  for (uint32_t s = 0; s < steps; ++s) {
    // kCalcCost (:480-500) + DeviceRadixSort::SortPairsDescending (stable)
    for (size_t i = 0; i < inner.size(); ++i) {
      const Node &n = b[inner[i]];
      float sa = n.surfaceArea, l = b[n.left].surfaceArea, r = b[n.right].surfaceArea;
      cost[i] = (sa * sa * sa * 2.0f) / ((l + r) * (l < r ? l : r));
    }
    std::vector<uint32_t> ord(inner.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (uint32_t)i;
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t c) { return cost[a] > cost[c]; });
    for (size_t i = 0; i < ord.size(); ++i) sortedIds[i] = inner[ord[i]];
    for (uint32_t c = 0; c < numChunks; ++c) {
      // kRemoveAndReinsert1 (:504-531)
      for (uint32_t k = 0; k < K; ++k) {
        Patch &p = patches[k];
        p.bvh = &b;
        p.rootIdx = g.root;
        p.clear();
        ok[k] = 0;
        uint32_t nodeIdx = sortedIds[offsetAccess ? k * numChunks + c : c * K + k];
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
      g.fixTree1();
      g.sumChunkMax += g.chunkMaxIters;
      g.chunkMaxIters = 0;
      if (chunkTrees) std::copy(b.begin(), b.end(), chunkTrees + ((size_t)s * numChunks + c) * N);
    }
    if (sahOut) sahOut[s + 1] = sahGpu();
  }
  st[3] = g.searchMode ? g.waveOverflows : g.heapOverflows;
  st[4] = g.waveRunaway;
  if (std::getenv("ORC_WAVE_TRACE"))
    std::fprintf(stderr, "wave iterations: total %llu, max per search %llu, sum of per-chunk maxima %llu\n",
                 (unsigned long long)g.waveIters, (unsigned long long)g.waveMaxIters,
                 (unsigned long long)g.sumChunkMax);
  if (stats) std::copy(st, st + 5, stats);
  std::copy(b.begin(), b.end(), nodes);
  return 0;
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
