// Device-side Validate (lib/misc/Validate.cpp:42-263) for trees that stay on
// the GPU.  Error classes (counts):
//   0 tree:  root not its own parent; a node whose parent does not point back
//            to it (isLeft picks the slot); an inner node with equal, self or
//            out-of-range children, or children whose parent / isLeft links
//            disagree; a node whose parent chain does not reach the root
//            (a cycle or a detached part; decided by pointer jumping, O(N
//            log N)).  Levels are not checked:
//            the device builders do not maintain them (the reference checks
//            them after the host-side fixLevels of copyToHost).
//   1 bbox:        inner box does not contain both children (checkBBoxes)
//   2 bbox strict: inner box is not exactly their union (checkBBoxesStrict)
//   3 SA:          |SA - SA(box)| >= 1e-6 (checkSurfaceArea)
//   4 triangles:   leaves with a face index past the mesh or a vertex outside
//                  the leaf box (checkTris)
//   5 orphans:     faces no leaf references
#include "runtime.hpp"
#include "types.hpp"

#include <cstring>

namespace bvht {

namespace {

enum { V_TREE = 0, V_BOX, V_STRICT, V_SA, V_TRI, V_ORPHAN, V_WORDS };

__device__ __forceinline__ bool contains(const bvht_AABB &b, const bvht_AABB &c) {
  return b.minMax[0].x <= c.minMax[0].x && b.minMax[0].y <= c.minMax[0].y && b.minMax[0].z <= c.minMax[0].z &&
         b.minMax[1].x >= c.minMax[1].x && b.minMax[1].y >= c.minMax[1].y && b.minMax[1].z >= c.minMax[1].z;
}
__device__ __forceinline__ bool inBox(const bvht_AABB &b, const bvht_vec3 &v) {
  return b.minMax[0].x <= v.x && b.minMax[0].y <= v.y && b.minMax[0].z <= v.z && b.minMax[1].x >= v.x &&
         b.minMax[1].y >= v.y && b.minMax[1].z >= v.z;
}

// anc[i] = parent of i (i itself when the link leaves the array).
__global__ __launch_bounds__(256) void k_ancInit(const bvht_BVHNode *__restrict__ nodes, uint32_t N,
                                                 uint32_t *__restrict__ anc) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const uint32_t p = nodes[i].parent;
  anc[i] = p < N ? p : i;
}

// anc[i] <- anc[anc[i]]: after k rounds anc[i] is an ancestor 2^k steps up
// (or the root).  Updating in place only jumps further.
__global__ __launch_bounds__(256) void k_ancJump(uint32_t N, uint32_t *anc) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  anc[i] = anc[anc[i]];
}

__global__ __launch_bounds__(256) void k_validateNodes(const bvht_BVHNode *__restrict__ nodes, uint32_t N,
                                                       const uint32_t *__restrict__ anc,
                                                       uint32_t root, const bvht_vec3 *__restrict__ vert,
                                                       const bvht_Triangle *__restrict__ faces, uint32_t nf,
                                                       uint32_t *__restrict__ covered, uint32_t *__restrict__ err) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const bvht_BVHNode n = nodes[i];
  bool treeBad = false;
  if (i == root) {
    treeBad |= n.parent != root;
  } else {
    const uint32_t p = n.parent;
    treeBad |= p >= N || nodes[p].isLeafFlag != 0 || (n.isLeft ? nodes[p].left : nodes[p].right) != i;
    treeBad |= anc[i] != root;
  }
  if (n.isLeafFlag == 0) {
    const uint32_t l = n.left, r = n.right;
    const bool ok = l < N && r < N && l != r && l != i && r != i;
    treeBad |= !ok;
    if (ok) {
      const bvht_BVHNode a = nodes[l], b = nodes[r];
      treeBad |= a.parent != i || b.parent != i || a.isLeft == 0 || b.isLeft != 0;
      if (!(contains(n.bbox, a.bbox) && contains(n.bbox, b.bbox))) atomicAdd(err + V_BOX, 1u);
      const bool strict = n.bbox.minMax[0].x == fminf(a.bbox.minMax[0].x, b.bbox.minMax[0].x) &&
                          n.bbox.minMax[0].y == fminf(a.bbox.minMax[0].y, b.bbox.minMax[0].y) &&
                          n.bbox.minMax[0].z == fminf(a.bbox.minMax[0].z, b.bbox.minMax[0].z) &&
                          n.bbox.minMax[1].x == fmaxf(a.bbox.minMax[1].x, b.bbox.minMax[1].x) &&
                          n.bbox.minMax[1].y == fmaxf(a.bbox.minMax[1].y, b.bbox.minMax[1].y) &&
                          n.bbox.minMax[1].z == fmaxf(a.bbox.minMax[1].z, b.bbox.minMax[1].z);
      if (!strict) atomicAdd(err + V_STRICT, 1u);
    }
  } else {
    const uint64_t end = (uint64_t)n.left + n.right;
    bool triBad = end > nf;  // faces past the mesh (never read)
    const uint64_t stop = end < nf ? end : nf;
    for (uint64_t j = n.left; j < stop; ++j) {
      covered[j] = 1;
      const bvht_Triangle t = faces[j];
      triBad |= !(inBox(n.bbox, vert[t.v1]) && inBox(n.bbox, vert[t.v2]) && inBox(n.bbox, vert[t.v3]));
    }
    if (triBad) atomicAdd(err + V_TRI, 1u);
  }
  if (fabsf(n.surfaceArea - surfaceArea(n.bbox)) >= 0.000001f) atomicAdd(err + V_SA, 1u);
  if (treeBad) atomicAdd(err + V_TREE, 1u);
}

__global__ __launch_bounds__(256) void k_validateOrphans(const uint32_t *__restrict__ covered, uint32_t nf,
                                                         uint32_t *__restrict__ err) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j < nf && covered[j] == 0) atomicAdd(err + V_ORPHAN, 1u);
}

}  // namespace

}  // namespace bvht

using namespace bvht;

extern "C" {

uint32_t bvht_validate(const bvht_CUDAMemoryBVHPointer *bvh, uint32_t root, bvht_MeshRaw mesh, uint32_t *errOut) {
  if (!bvh || !bvh->nodes || bvh->numNodes == 0 || root >= bvh->numNodes || !mesh.faces || !mesh.vert) {
    setError("bvht_validate: missing tree or mesh, or root out of range");
    return UINT32_MAX;
  }
  uint32_t *buf = nullptr;
  const uint32_t nf = mesh.numFaces, N = bvh->numNodes;
  if (!devAlloc(&buf, (size_t)nf + V_WORDS + N, "bvht_validate scratch")) return UINT32_MAX;
  uint32_t *err = buf + nf, *anc = err + V_WORDS, h[V_WORDS] = {};
  bool ok = fillU32(buf, 0u, (size_t)nf + V_WORDS);
  if (ok) {
    k_ancInit<<<divUp(N, 256), 256, 0, stream()>>>(bvh->nodes, N, anc);
    ok = checkLaunch("k_ancInit");
    for (uint32_t span = 1; ok && span < 2 * N; span *= 2) {  // ceil(log2 N) + 1 rounds
      k_ancJump<<<divUp(N, 256), 256, 0, stream()>>>(N, anc);
      ok = checkLaunch("k_ancJump");
      if (span >= (1u << 31)) break;
    }
  }
  if (ok) {
    k_validateNodes<<<divUp(N, 256), 256, 0, stream()>>>(bvh->nodes, N, anc, root, mesh.vert, mesh.faces, nf, buf,
                                                          err);
    ok = checkLaunch("k_validateNodes");
  }
  if (ok && nf) {
    k_validateOrphans<<<divUp(nf, 256), 256, 0, stream()>>>(buf, nf, err);
    ok = checkLaunch("k_validateOrphans");
  }
  ok = ok && check(hipMemcpyAsync(h, err, sizeof(h), hipMemcpyDeviceToHost, stream()), "bvht_validate") &&
       check(hipStreamSynchronize(stream()), "bvht_validate sync");
  devFree(buf);
  if (!ok) return UINT32_MAX;
  uint32_t total = 0;
  for (int k = 0; k < V_WORDS; ++k) {
    if (errOut) errOut[k] = h[k];
    total += h[k];
  }
  return total;
}

}  // extern "C"
