// Device kernels shared by the single-mesh and the batched LBVH builders.
#pragma once

#include "radix_sort.hpp"
#include "wave.hpp"

namespace bvht {

constexpr int kBuildBlock = 256;

// One mesh of a (possibly batched) build.  Sorted values hold global face
// indices; a segment's faces hold mesh-local vertex indices.
struct SegDesc {
  uint32_t n;         // triangles
  uint32_t faceBase;  // first global face / sorted position
  uint32_t nodeBase;  // first node of the segment's 2n-1 nodes
  uint32_t vertBase;  // first vertex of the mesh
};

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }

// Writes one full 48-byte node with three 16-byte stores.
__device__ __forceinline__ void storeNode(bvht_BVHNode *dst, const bvht_AABB &b, uint32_t parent, uint32_t left,
                                          uint32_t right, uint32_t isLeaf, uint32_t isLeft, uint32_t level, float sa) {
  uint4 *d = reinterpret_cast<uint4 *>(dst);
  d[0] = make_uint4(f2u(b.minMax[0].x), f2u(b.minMax[0].y), f2u(b.minMax[0].z), f2u(b.minMax[1].x));
  d[1] = make_uint4(f2u(b.minMax[1].y), f2u(b.minMax[1].z), parent, left);
  d[2] = make_uint4(right, isLeaf | (isLeft << 16), level, f2u(sa));
}

// Parent index and isLeft flag of the node covering [l, r] (closed form of
// the reference topology, types.hpp).
__device__ __forceinline__ void parentOf(uint32_t l, uint32_t r, uint32_t N1, uint32_t &parent, uint32_t &isLeft) {
  if (l == 0 && r == N1) {
    parent = 0;  // the root is its own parent
    isLeft = 0;
    return;
  }
  uint32_t pl, pr;
  rangeParent(l, r, N1, pl, pr);
  parent = rangeNodeIndex(pl, pr, N1);
  isLeft = rangeIsLeft(l, r, N1) ? 1u : 0u;
}

// Float <-> order-preserving u32 (for atomic min/max of boxes).
__device__ __forceinline__ uint32_t f2o(float f) {
  uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// One level of the tile hierarchy.  Units are aligned runs of 2^unitLog leaf
// positions of one segment; a workgroup owns 2^TLOG consecutive units and
// writes every node whose leaf range is a union of 2..2^TLOG of its units
// (level 0 also gathers each sorted triangle and writes its leaf).  Boxes
// are reduced level by level in LDS: no global atomics.
template <int TLOG, bool LEVEL0, bool BATCH>
__global__ __launch_bounds__(kBuildBlock) void k_buildLevel(bvht_BVHNode *__restrict__ nodesAll,
                                                            const uint32_t *__restrict__ sorted,
                                                            const bvht_Triangle *__restrict__ faces,
                                                            const bvht_vec3 *__restrict__ vertAll, SegDesc single,
                                                            const SegDesc *__restrict__ segs,
                                                            const uint2 *__restrict__ tiles, uint32_t unitLog) {
  constexpr uint32_t T = 1u << TLOG;
  __shared__ float s_box[6][T];
  SegDesc sd;
  uint32_t group;
  if (BATCH) {
    const uint2 t = tiles[blockIdx.x];
    sd = segs[t.x];
    group = t.y;
  } else {
    sd = single;
    group = blockIdx.x;
  }
  const uint32_t n = sd.n, N1 = n - 1;
  bvht_BVHNode *__restrict__ nodes = nodesAll + sd.nodeBase;
  const bvht_vec3 *__restrict__ vert = vertAll + sd.vertBase;
  const uint64_t U = (((uint64_t)n) + (1ull << unitLog) - 1) >> unitLog;
  const uint64_t u0 = (uint64_t)group << TLOG;
  if (u0 >= U) return;  // block-uniform: a group past the last unit owns nothing
  const uint32_t cnt = (uint32_t)min((uint64_t)T, U - u0);

  if (LEVEL0) {
    // Every sorted index, face and vertex gather of the thread's leaves is
    // issued before any is used: one chain of dependent loads per thread
    // instead of one per leaf.
    constexpr int R = (T + kBuildBlock - 1) / kBuildBlock;
    uint32_t f[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const uint32_t k = threadIdx.x + i * kBuildBlock;
      f[i] = k < cnt ? sorted[sd.faceBase + (uint32_t)(u0 + k)] - sd.faceBase : 0u;
      if (f[i] >= n) f[i] = 0;  // never read outside the mesh, even from a corrupt permutation
    }
    bvht_Triangle t[R];
#pragma unroll
    for (int i = 0; i < R; ++i) t[i] = faces[sd.faceBase + f[i]];
    bvht_vec3 va[R], vb[R], vc[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      va[i] = vert[t[i].v1];
      vb[i] = vert[t[i].v2];
      vc[i] = vert[t[i].v3];
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const uint32_t k = threadIdx.x + i * kBuildBlock;
      if (k >= cnt) continue;
      const uint32_t p = (uint32_t)(u0 + k);
      const bvht_AABB b = triangleBox(va[i], vb[i], vc[i]);
      uint32_t parent, isLeft;
      parentOf(p, p, N1, parent, isLeft);
      // kGenLeafNodes (LBVH_CUDA.cu:273-292): left = face, right = 1, level = UINT8_MAX
      storeNode(nodes + N1 + p, b, parent, f[i], 1u, 1u, isLeft, 255u, surfaceArea(b));
      s_box[0][k] = b.minMax[0].x;
      s_box[1][k] = b.minMax[0].y;
      s_box[2][k] = b.minMax[0].z;
      s_box[3][k] = b.minMax[1].x;
      s_box[4][k] = b.minMax[1].y;
      s_box[5][k] = b.minMax[1].z;
    }
  }
  for (uint32_t k = threadIdx.x; !LEVEL0 && k < cnt; k += kBuildBlock) {
    const uint64_t u = u0 + k;
    bvht_AABB b;
    {
      const uint32_t l = (uint32_t)(u << unitLog);
      const uint32_t r = (uint32_t)min((u + 1) << unitLog, (uint64_t)n) - 1;
      const float *src = reinterpret_cast<const float *>(nodes + rangeNodeIndex(l, r, N1));
      b.minMax[0] = {src[0], src[1], src[2]};
      b.minMax[1] = {src[3], src[4], src[5]};
    }
    s_box[0][k] = b.minMax[0].x;
    s_box[1][k] = b.minMax[0].y;
    s_box[2][k] = b.minMax[0].z;
    s_box[3][k] = b.minMax[1].x;
    s_box[4][k] = b.minMax[1].y;
    s_box[5][k] = b.minMax[1].z;
  }
  __syncthreads();

  // Level j merges entries 2a and 2a + 1 of level j - 1 into entry a (a
  // compact layout: lanes read with stride 2 and write contiguously, where
  // an in-place tree reduction would read with stride 2^j and serialise on
  // LDS banks).  Reads go to registers before a barrier, then the writes.
  constexpr int RH = (T / 2 + kBuildBlock - 1) / kBuildBlock;
  for (uint32_t j = 1; j <= (uint32_t)TLOG; ++j) {
    const uint32_t half = 1u << (j - 1);
    const uint32_t nb = (cnt + (1u << j) - 1) >> j;
    const uint32_t nPrev = (cnt + half - 1) >> (j - 1);
    float res[RH][6];
#pragma unroll
    for (int q = 0; q < RH; ++q) {
      const uint32_t a = threadIdx.x + q * kBuildBlock;
      if (a >= nb) continue;
      const uint32_t c0 = 2 * a, c1 = c0 + 1;
#pragma unroll
      for (int c = 0; c < 6; ++c) res[q][c] = s_box[c][c0];
      if (c1 >= nPrev) continue;  // right half empty: same range as the left half
      bvht_AABB b;
      b.minMax[0] = {mergeMin(res[q][0], s_box[0][c1]), mergeMin(res[q][1], s_box[1][c1]),
                     mergeMin(res[q][2], s_box[2][c1])};
      b.minMax[1] = {mergeMax(res[q][3], s_box[3][c1]), mergeMax(res[q][4], s_box[4][c1]),
                     mergeMax(res[q][5], s_box[5][c1])};
      res[q][0] = b.minMax[0].x;
      res[q][1] = b.minMax[0].y;
      res[q][2] = b.minMax[0].z;
      res[q][3] = b.minMax[1].x;
      res[q][4] = b.minMax[1].y;
      res[q][5] = b.minMax[1].z;
      const uint32_t k0 = a << j, km = k0 + half;
      const uint32_t l = (uint32_t)((u0 + k0) << unitLog);
      const uint32_t m = (uint32_t)((u0 + km) << unitLog);
      const uint32_t r = (uint32_t)min((u0 + k0 + (1u << j)) << unitLog, (uint64_t)n) - 1;
      uint32_t parent, isLeft;
      parentOf(l, r, N1, parent, isLeft);
      storeNode(nodes + rangeNodeIndex(l, r, N1), b, parent, rangeNodeIndex(l, m - 1, N1), rangeNodeIndex(m, r, N1), 0u,
                isLeft, 0u, surfaceArea(b));
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RH; ++q) {
      const uint32_t a = threadIdx.x + q * kBuildBlock;
      if (a >= nb) continue;
#pragma unroll
      for (int c = 0; c < 6; ++c) s_box[c][a] = res[q][c];
    }
    __syncthreads();
  }
}

// Launches the tile hierarchy of one segment (BATCH == false) with tile log2
// `tlog`; returns false on a launch error.
bool launchTreeLevels(int tlog, bvht_BVHNode *nodes, const uint32_t *sorted, const bvht_Triangle *faces,
                      const bvht_vec3 *vert, uint32_t n, void (*afterLevel)(bool level0));

// Batched variant: tilesPerLevel[L] lists (segment, group) pairs of level L.
bool launchTreeLevelsBatch(int tlog, bvht_BVHNode *nodes, const uint32_t *sorted, const bvht_Triangle *faces,
                           const bvht_vec3 *vert, const SegDesc *segs, const uint2 *const *tilesPerLevel,
                           const uint32_t *tilesCount, int levels);

}  // namespace bvht
