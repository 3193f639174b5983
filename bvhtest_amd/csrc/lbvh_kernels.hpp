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

// The three 16-byte words of a 48-byte node.
struct NodeWords {
  uint4 w[3];
};
__device__ __forceinline__ NodeWords packNode(const bvht_AABB &b, uint32_t parent, uint32_t left, uint32_t right,
                                              uint32_t isLeaf, uint32_t isLeft, uint32_t level, float sa) {
  NodeWords o;
  o.w[0] = make_uint4(__float_as_uint(b.minMax[0].x), __float_as_uint(b.minMax[0].y), __float_as_uint(b.minMax[0].z),
                      __float_as_uint(b.minMax[1].x));
  o.w[1] = make_uint4(__float_as_uint(b.minMax[1].y), __float_as_uint(b.minMax[1].z), parent, left);
  o.w[2] = make_uint4(right, isLeaf | (isLeft << 16), level, __float_as_uint(sa));
  return o;
}

// The 64 lanes of a wave write the nodes dst[0..63] (lane l's node, if
// `write`) through an LDS staging buffer of 64 nodes, so that every store
// instruction covers 1 KB of consecutive bytes instead of 16 bytes every 48
// (three instructions per node, each touching 24 lines in part).  All
// threads of the workgroup call it (barriers: lanes read other lanes'
// words, and the next call reuses the buffer).
__device__ __forceinline__ void storeNodesStaged(uint4 *stage, uint8_t *stageOk, bvht_BVHNode *dst, const NodeWords &nw,
                                                 bool write) {
  const uint32_t lane = threadIdx.x & 63;
  stage[3 * lane] = nw.w[0];
  stage[3 * lane + 1] = nw.w[1];
  stage[3 * lane + 2] = nw.w[2];
  stageOk[lane] = write ? 1 : 0;
  __syncthreads();
  uint4 *d = reinterpret_cast<uint4 *>(dst);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const uint32_t c = lane + 64 * q;
    if (stageOk[c / 3]) d[c] = stage[c];
  }
  __syncthreads();
}

// Writes one full 48-byte node with three 16-byte stores.
__device__ __forceinline__ void storeNode(bvht_BVHNode *dst, const bvht_AABB &b, uint32_t parent, uint32_t left,
                                          uint32_t right, uint32_t isLeaf, uint32_t isLeft, uint32_t level, float sa) {
  const NodeWords nw = packNode(b, parent, left, right, isLeaf, isLeft, level, sa);
  uint4 *d = reinterpret_cast<uint4 *>(dst);
  d[0] = nw.w[0];
  d[1] = nw.w[1];
  d[2] = nw.w[2];
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

// Inner node whose index is position p of a segment with N1 + 1 leaves
// (p < N1), in the closed form of the reference topology: an odd p is the
// last position of a left child (a full aligned block), an even p > 0 the
// first position of a right child (an aligned block cut at N1), p = 0 the
// root.  Returns the node's range [l, r].
__device__ __forceinline__ void nodeAtIndex(uint32_t p, uint32_t N1, uint32_t &l, uint32_t &r) {
  if (p == 0) {
    l = 0;
    r = N1;
  } else if (p & 1u) {
    l = p + 1 - (1u << __builtin_ctz(p + 1));
    r = p;
  } else {
    l = p;
    r = (uint32_t)min((uint64_t)p + (1ull << __builtin_ctz(p)) - 1, (uint64_t)N1);
  }
}

// One butterfly step of the in-wave box reduction: b holds the box of the
// aligned block of `half` leaves around the lane's leaf k and becomes that
// of the block of 2 * half.  The lower half always comes first in the
// merge, as a node merges its left child with its right; an upper half
// starting at or past cnt (no leaves) leaves the box unchanged.
__device__ __forceinline__ void boxStep(float b[6], uint32_t k, uint32_t half, uint32_t cnt) {
  float o[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) o[c] = __shfl_xor(b[c], (int)half);
  if (((k & ~(2 * half - 1)) | half) >= cnt) return;
  const bool upper = (k & half) != 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float lo = upper ? o[c] : b[c], hi = upper ? b[c] : o[c];
    b[c] = mergeMin(lo, hi);
  }
#pragma unroll
  for (int c = 3; c < 6; ++c) {
    const float lo = upper ? o[c] : b[c], hi = upper ? b[c] : o[c];
    b[c] = mergeMax(lo, hi);
  }
}

// Node words and index of the inner node covering [l, r].
__device__ __forceinline__ NodeWords innerWords(uint32_t l, uint32_t r, uint32_t N1, const float b[6]) {
  bvht_AABB box;
  box.minMax[0] = {b[0], b[1], b[2]};
  box.minMax[1] = {b[3], b[4], b[5]};
  const uint32_t m = rangeSplit(l, r);
  uint32_t parent, isLeft;
  parentOf(l, r, N1, parent, isLeft);
  return packNode(box, parent, rangeNodeIndex(l, m - 1, N1), rangeNodeIndex(m, r, N1), 0u, isLeft, 0u, surfaceArea(box));
}

__device__ __forceinline__ void storeInner(bvht_BVHNode *nodes, uint32_t l, uint32_t r, uint32_t N1, const float b[6]) {
  bvht_AABB box;
  box.minMax[0] = {b[0], b[1], b[2]};
  box.minMax[1] = {b[3], b[4], b[5]};
  const uint32_t m = rangeSplit(l, r);
  uint32_t parent, isLeft;
  parentOf(l, r, N1, parent, isLeft);
  storeNode(nodes + rangeNodeIndex(l, r, N1), box, parent, rangeNodeIndex(l, m - 1, N1), rangeNodeIndex(m, r, N1), 0u,
            isLeft, 0u, surfaceArea(box));
}

// Level 0 of the tile hierarchy: a workgroup owns 2^TLOG consecutive sorted
// positions of one segment (a tile), gathers each sorted triangle, writes
// its leaf, and writes every inner node whose range lies in the tile.  The
// inner node written by the lane of position p is the one whose index is p,
// so a wave's inner nodes are as contiguous as its leaves (an index of a
// node in the tile lies in the tile).  Boxes of aligned blocks of up to 64
// positions come from butterfly shuffles inside the wave; the few larger
// blocks of the tile are reduced in LDS from the 64-position blocks.
// Six workgroups per CU (at most 80 VGPRs) measure faster than four (the
// compiler's 105 VGPRs without the bound) and than eight (64 VGPRs, with
// spills).
// REC: the leaf boxes are read from `rec` (three float2 per triangle, by
// global face index, written by the key pass) instead of being made from a
// face and three vertices: one gather per leaf instead of four.
template <int TLOG, bool BATCH, bool REC>
__global__ __launch_bounds__(kBuildBlock, 6) void k_buildLevel0(bvht_BVHNode *__restrict__ nodesAll,
                                                             const uint32_t *__restrict__ sorted,
                                                             const bvht_Triangle *__restrict__ faces,
                                                             const bvht_vec3 *__restrict__ vertAll,
                                                             const float2 *__restrict__ rec, SegDesc single,
                                                             const SegDesc *__restrict__ segs,
                                                             const uint2 *__restrict__ tiles) {
  constexpr uint32_t T = 1u << TLOG;
  constexpr int R = T / kBuildBlock;
  constexpr uint32_t NB = T / 64;  // 64-position blocks of a tile
  static_assert(T % kBuildBlock == 0 && TLOG > 6, "tile holds whole workgroups of 64-position blocks");
  // block boxes of levels 6..TLOG; level j starts at 2 NB - (2 NB >> (j - 6))
  __shared__ float s_box[2 * NB][6];
  SegDesc sd;
  uint32_t group;
  if (BATCH) {
    const uint2 t = tiles[blockIdx.x];
    sd = segs[t.x];
    group = t.y;
  } else {
    sd = single;
    // consecutive tiles to one XCD: workgroups are handed to the 8 XCDs
    // round-robin, and neighbouring tiles share face and vertex lines (a
    // bijection for any grid)
    const uint32_t G = gridDim.x, q = G / 8, r = G % 8, x = blockIdx.x % 8;
    group = x * q + min(x, r) + blockIdx.x / 8;
  }
  const uint32_t n = sd.n, N1 = n - 1;
  bvht_BVHNode *__restrict__ nodes = nodesAll + sd.nodeBase;
  const bvht_vec3 *__restrict__ vert = vertAll + sd.vertBase;
  const uint32_t u0 = group << TLOG;
  if (u0 >= n) return;  // block-uniform: a group past the last position owns nothing
  const uint32_t cnt = min(T, n - u0);

  // Every sorted index, face and vertex gather of the thread's leaves is
  // issued before any is used: one chain of dependent loads per thread
  // instead of one per leaf.
  uint32_t f[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const uint32_t k = threadIdx.x + i * kBuildBlock;
    f[i] = k < cnt ? sorted[sd.faceBase + u0 + k] - sd.faceBase : 0u;
    if (f[i] >= n) f[i] = 0;  // never read outside the mesh, even from a corrupt permutation
  }
  bvht_AABB lbs[R];  // every leaf box first: the loads die before the barriers below
  if constexpr (REC) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const float2 *r = rec + 3 * (size_t)(sd.faceBase + f[i]);
      const float2 a = r[0], b = r[1], c = r[2];
      lbs[i].minMax[0] = {a.x, a.y, b.x};
      lbs[i].minMax[1] = {b.y, c.x, c.y};
    }
  } else {
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
    for (int i = 0; i < R; ++i) lbs[i] = triangleBox(va[i], vb[i], vc[i]);
  }
  // leaves first: the gathered vertices are dead before the reductions start.
  // A wave's 64 leaves (positions k - lane .. k - lane + 63) are written
  // through the LDS staging buffer.
  __shared__ uint4 s_stage[kBuildBlock / 64][64 * 3];
  __shared__ uint8_t s_stageOk[kBuildBlock / 64][64];
  uint4 *stage = s_stage[threadIdx.x >> 6];
  uint8_t *stageOk = s_stageOk[threadIdx.x >> 6];
  const uint32_t lane = threadIdx.x & 63;
  float leaf[R][6];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const uint32_t k = threadIdx.x + i * kBuildBlock, p = u0 + k;
    const bvht_AABB lb = lbs[i];
    uint32_t parent = 0, isLeft = 0;
    if (k < cnt) parentOf(p, p, N1, parent, isLeft);
    // kGenLeafNodes (LBVH_CUDA.cu:273-292): left = face, right = 1, level = UINT8_MAX
    storeNodesStaged(stage, stageOk, nodes + N1 + (p - lane), packNode(lb, parent, f[i], 1u, 1u, isLeft, 255u, surfaceArea(lb)),
                     k < cnt);
    leaf[i][0] = lb.minMax[0].x;
    leaf[i][1] = lb.minMax[0].y;
    leaf[i][2] = lb.minMax[0].z;
    leaf[i][3] = lb.minMax[1].x;
    leaf[i][4] = lb.minMax[1].y;
    leaf[i][5] = lb.minMax[1].z;
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const uint32_t k = threadIdx.x + i * kBuildBlock, p = u0 + k;
    // the inner node with index p, if it lies in the tile
    uint32_t l = 0, r = 0, jn = 0;
    if (k < cnt && p < N1) {
      nodeAtIndex(p, N1, l, r);
      if (l >= u0 && r < u0 + cnt) jn = 32 - __builtin_clz(l ^ r);
    }
    float *b = leaf[i];
    float pick[6];
#pragma unroll
    for (uint32_t j = 1; j <= 6; ++j) {
      boxStep(b, k, 1u << (j - 1), cnt);
      if (jn == j) {
#pragma unroll
        for (int c = 0; c < 6; ++c) pick[c] = b[c];
      }
    }
    if ((k & 63) == 0) {
#pragma unroll
      for (int c = 0; c < 6; ++c) s_box[k >> 6][c] = b[c];
    }
    // the inner node with index p is the lane's: staged like the leaves
    storeNodesStaged(stage, stageOk, nodes + (p - lane), innerWords(l, r, N1, pick), jn >= 1 && jn <= 6);
  }
  __syncthreads();
  // levels 7..TLOG: block a of level j merges blocks 2a and 2a + 1 of level j - 1
#pragma unroll
  for (uint32_t j = 7; j <= (uint32_t)TLOG; ++j) {
    const uint32_t src = 2 * NB - ((2 * NB) >> (j - 7)), dst = 2 * NB - ((2 * NB) >> (j - 6));
    const uint32_t a = threadIdx.x;
    if (a < (NB >> (j - 6))) {
      float b[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) b[c] = s_box[src + 2 * a][c];
      if ((2 * a + 1) << (j - 1) < cnt) {
#pragma unroll
        for (int c = 0; c < 3; ++c) b[c] = mergeMin(b[c], s_box[src + 2 * a + 1][c]);
#pragma unroll
        for (int c = 3; c < 6; ++c) b[c] = mergeMax(b[c], s_box[src + 2 * a + 1][c]);
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) s_box[dst + a][c] = b[c];
    }
    __syncthreads();
  }
  // inner nodes of blocks larger than a wave's 64 positions (a few per tile)
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const uint32_t k = threadIdx.x + i * kBuildBlock, p = u0 + k;
    if (k >= cnt || p >= N1) continue;
    uint32_t l, r;
    nodeAtIndex(p, N1, l, r);
    if (l < u0 || r >= u0 + cnt) continue;
    const uint32_t jn = 32 - __builtin_clz(l ^ r);
    if (jn <= 6) continue;
    const uint32_t at = 2 * NB - ((2 * NB) >> (jn - 6)) + (k >> jn);
    float b[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) b[c] = s_box[at][c];
    storeInner(nodes, l, r, N1, b);
  }
}

// One upper level of the tile hierarchy.  Units are aligned runs of
// 2^unitLog leaf positions of one segment (unitLog a multiple of TLOG, at
// least TLOG); a workgroup owns 2^TLOG consecutive units and writes every
// node whose leaf range is a union of 2..2^TLOG of its units, reading the
// units' boxes from their nodes.  Boxes are reduced level by level in LDS:
// no global atomics.
template <int TLOG, bool BATCH>
__global__ __launch_bounds__(kBuildBlock) void k_buildLevelUp(bvht_BVHNode *__restrict__ nodesAll,
                                                              const SegDesc single, const SegDesc *__restrict__ segs,
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
  const uint64_t U = (((uint64_t)n) + (1ull << unitLog) - 1) >> unitLog;
  const uint64_t u0 = (uint64_t)group << TLOG;
  if (u0 >= U) return;  // block-uniform: a group past the last unit owns nothing
  const uint32_t cnt = (uint32_t)min((uint64_t)T, U - u0);

  for (uint32_t k = threadIdx.x; k < cnt; k += kBuildBlock) {
    const uint64_t u = u0 + k;
    const uint32_t l = (uint32_t)(u << unitLog);
    const uint32_t r = (uint32_t)min((u + 1) << unitLog, (uint64_t)n) - 1;
    const float *src = reinterpret_cast<const float *>(nodes + rangeNodeIndex(l, r, N1));
#pragma unroll
    for (int c = 0; c < 6; ++c) s_box[c][k] = src[c];
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
#pragma unroll
      for (int c = 0; c < 3; ++c) res[q][c] = mergeMin(res[q][c], s_box[c][c1]);
#pragma unroll
      for (int c = 3; c < 6; ++c) res[q][c] = mergeMax(res[q][c], s_box[c][c1]);
      const uint32_t k0 = a << j;
      const uint32_t l = (uint32_t)((u0 + k0) << unitLog);
      const uint32_t r = (uint32_t)min((u0 + k0 + (1u << j)) << unitLog, (uint64_t)n) - 1;
      storeInner(nodes, l, r, N1, res[q]);
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

// Batched variant: tilesPerLevel[L] lists (segment, group) pairs of level L;
// rec (if not null) holds every triangle's leaf box (k_buildLevel0's REC).
bool launchTreeLevelsBatch(int tlog, bvht_BVHNode *nodes, const uint32_t *sorted, const bvht_Triangle *faces,
                           const bvht_vec3 *vert, const float2 *rec, const SegDesc *segs,
                           const uint2 *const *tilesPerLevel, const uint32_t *tilesCount, int levels);

}  // namespace bvht
