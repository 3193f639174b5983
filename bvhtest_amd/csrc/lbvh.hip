// LBVH builder for gfx950 (reference: lib/builder/LBVH_CUDA.cu).
//
// Pipeline of one build (all on one stream, no host round trip):
//   k_sceneBox   faces + vertices -> scene box (union of the +-FLT_EPSILON
//                triangle boxes and the origin), offset, scale; also zeroes
//                the sort state.  Last-block finalisation, one launch.
//   k_morton     centroid -> 30-bit (reference) or 63-bit Morton key, plus
//                the digit histograms of every sort pass (LDS, then one
//                global atomic per digit per block).
//   onesweep     stable radix sort of (key, face index); the last pass keeps
//                only the face indices when the topology does not need keys.
//   k_buildLevel0 tiles of 2^kTileLog positions: gathers each sorted
//                triangle, writes the leaf, reduces the tile's subtree
//                (wave shuffles, then LDS) and writes each inner node once
//                (box, topology, SA) from the lane of its index;
//   k_buildLevelUp further launches reduce the tile roots in LDS.
// The reference topology is a closed form of n (see types.hpp), so no key is
// read to build it and no global atomic is needed for the refit.
#include "lbvh.hpp"
#include "lbvh_kernels.hpp"
#include "msd_sort.hpp"
#include "radix_sort.hpp"
#include "split_sort.hpp"
#include "refit.hpp"
#include "wave.hpp"

#include <cstring>
#include <new>

namespace bvht {

namespace {

constexpr int kBlock = 256;
// default: 256 sorted positions per build workgroup (at 4.24M triangles the
// build takes 0.350 ms with 2^8, 0.361 with 2^9, 0.363 with 2^10, 0.490 with 2^11)
constexpr int kTileLog = 8;
// Scene box: 1024-thread workgroups (16 waves hide the face -> vertex load
// chain), two per CU.
constexpr int kBoxBlock = 1024;
constexpr uint32_t kBoxBlocksMax = 1024;

__device__ __forceinline__ bvht_vec3 ldVec3(const bvht_vec3 *p) { return *p; }

// Block reduction of per-thread boxes into partials[blockIdx.x]; returns,
// to every thread of the block, whether this block finished last (every
// other block's partial is then visible).  `done` is reset by the caller.
__device__ bool boxPartials(const float mn[3], const float mx[3], float *__restrict__ partials, uint32_t *done) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ float s_red[6][kBoxBlock / 64];
  __shared__ bool s_last;
  for (int a = 0; a < 3; ++a) {
    float x = waveMin(mn[a]), y = waveMax(mx[a]);
    if (lane == 0) {
      s_red[a][wave] = x;
      s_red[3 + a][wave] = y;
    }
  }
  __syncthreads();
  if (tid == 0) {
    for (int a = 0; a < 3; ++a) {
      float x = s_red[a][0], y = s_red[3 + a][0];
      for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
        x = mergeMin(x, s_red[a][w]);
        y = mergeMax(y, s_red[3 + a][w]);
      }
      // device-coherent stores, complete before the count: no L2-wide
      // write-back (a release fence here wrote back every dirty line, e.g.
      // the keys of the Morton pass, once per block)
      storeRelaxedAgent(partials + blockIdx.x * 6 + a, x);
      storeRelaxedAgent(partials + blockIdx.x * 6 + 3 + a, y);
    }
    waitVectorMemory();
    s_last = atomicAdd(done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  return s_last;
}

// The last block's union of every block's partial, starting from the
// all-zero box (LBVH_CUDA.cu:104-106); the result is in thread 0's out.
__device__ void boxFinal(const float *__restrict__ partials, float out[6]) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ float s_red[6][kBoxBlock / 64];
  float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
  for (uint32_t b = tid; b < gridDim.x; b += blockDim.x)  // device-coherent loads after the count
    for (int a = 0; a < 3; ++a) {
      mn[a] = mergeMin(mn[a], loadRelaxedAgent(partials + b * 6 + a));
      mx[a] = mergeMax(mx[a], loadRelaxedAgent(partials + b * 6 + 3 + a));
    }
  __syncthreads();
  for (int a = 0; a < 3; ++a) {
    float x = waveMin(mn[a]), y = waveMax(mx[a]);
    if (lane == 0) {
      s_red[a][wave] = x;
      s_red[3 + a][wave] = y;
    }
  }
  __syncthreads();
  if (tid == 0) {
    for (int a = 0; a < 3; ++a) {
      float x = s_red[a][0], y = s_red[3 + a][0];
      for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
        x = mergeMin(x, s_red[a][w]);
        y = mergeMax(y, s_red[3 + a][w]);
      }
      out[a] = x;
      out[3 + a] = y;
    }
  }
}

// LBVH_calcMortonCodes host part (LBVH_CUDA.cu:171-180), f32, IEEE division.
__device__ void setSceneBox(BuildParams *prm, const float box[6]) {
  for (int a = 0; a < 3; ++a) {
    prm->sceneMin[a] = box[a];
    prm->sceneMax[a] = box[3 + a];
    const float off = 0.0f - box[a];
    prm->offset[a] = off;
    prm->scale[a] = __fdiv_rn(1.0f, box[3 + a] + off);
  }
}

__global__ __launch_bounds__(kBoxBlock) void k_sceneBox(const bvht_vec3 *__restrict__ vert,
                                                        const bvht_Triangle *__restrict__ faces, uint32_t n,
                                                        BuildParams *__restrict__ prm, float *__restrict__ partials,
                                                        uint32_t *__restrict__ zeroPtr, size_t zeroWords) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (size_t i = (size_t)blockIdx.x * kBoxBlock + tid; i < zeroWords; i += (size_t)gridDim.x * kBoxBlock) zeroPtr[i] = 0;
  // the reduction starts from the all-zero box (LBVH_CUDA.cu:104-106)
  float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
  // a wave takes 64 consecutive triangles per trip (coalesced face loads)
  for (uint32_t base = (blockIdx.x * (kBoxBlock / 64) + wave) * 64; base < n; base += gridDim.x * kBoxBlock) {
    const uint32_t i = base + lane;
    // a repeat of triangle `base` for a lane past the end changes no min/max
    const bvht_Triangle t = faces[i < n ? i : base];
    const bvht_AABB bx = triangleBox(ldVec3(vert + t.v1), ldVec3(vert + t.v2), ldVec3(vert + t.v3));
    mn[0] = mergeMin(mn[0], bx.minMax[0].x);
    mn[1] = mergeMin(mn[1], bx.minMax[0].y);
    mn[2] = mergeMin(mn[2], bx.minMax[0].z);
    mx[0] = mergeMax(mx[0], bx.minMax[1].x);
    mx[1] = mergeMax(mx[1], bx.minMax[1].y);
    mx[2] = mergeMax(mx[2], bx.minMax[1].z);
  }
  if (!boxPartials(mn, mx, partials, &prm->done)) return;
  float box[6];
  boxFinal(partials, box);
  if (tid == 0) {
    setSceneBox(prm, box);
    prm->done = 0;
  }
}

// Scene box from every vertex instead of every triangle (the full build's
// first pass).  A triangle box is its vertices' min - FLT_EPSILON /
// max + FLT_EPSILON; f32 rounding is monotone, so the union over triangles
// equals the min / max over the referenced vertices shifted by FLT_EPSILON
// (a zero's sign is lost in the shift), and over all vertices when every
// vertex is referenced.  Unreferenced vertices outside the triangles' box
// make this box larger: k_morton<..., true> also reduces the triangles' box
// and k_mortonFix compares the two and repairs the keys if they differ.
//
// Neither pass ends in a last-block reduction (one global atomic per block
// on one counter, then a serial tail): each block writes its partial box,
// and every block of the next kernel reduces all partials itself.
__device__ __forceinline__ void shiftedBox(float mn[3], float mx[3]) {
  for (int a = 0; a < 3; ++a) {  // -+FLT_EPSILON, then the origin (LBVH_CUDA.cu:104-106)
    mn[a] = mergeMin(0.0f, mn[a] - kFltEpsilon);
    mx[a] = mergeMax(0.0f, mx[a] + kFltEpsilon);
  }
}

// Block reduction of per-thread boxes; thread 0 writes partials[blockIdx.x].
__device__ void writeBoxPartial(const float mn[3], const float mx[3], float *__restrict__ partials) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ float s_red[6][kBoxBlock / 64];
  for (int a = 0; a < 3; ++a) {
    const float x = waveMin(mn[a]), y = waveMax(mx[a]);
    if (lane == 0) {
      s_red[a][wave] = x;
      s_red[3 + a][wave] = y;
    }
  }
  __syncthreads();
  if (tid < 6) {
    float x = s_red[tid][0];
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) x = tid < 3 ? mergeMin(x, s_red[tid][w]) : mergeMax(x, s_red[tid][w]);
    partials[blockIdx.x * 6 + tid] = x;
  }
}

// Union of `count` partial boxes (written by an earlier kernel), starting
// from the all-zero box; every thread of the block receives it.
__device__ void readBoxPartials(const float *__restrict__ partials, uint32_t count, float box[6]) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ float s_red[6][kBoxBlock / 64];
  __shared__ float s_box[6];
  float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
  for (uint32_t b = tid; b < count; b += blockDim.x)
    for (int a = 0; a < 3; ++a) {
      mn[a] = mergeMin(mn[a], partials[b * 6 + a]);
      mx[a] = mergeMax(mx[a], partials[b * 6 + 3 + a]);
    }
  for (int a = 0; a < 3; ++a) {
    const float x = waveMin(mn[a]), y = waveMax(mx[a]);
    if (lane == 0) {
      s_red[a][wave] = x;
      s_red[3 + a][wave] = y;
    }
  }
  __syncthreads();
  if (tid < 6) {
    float x = s_red[tid][0];
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) x = tid < 3 ? mergeMin(x, s_red[tid][w]) : mergeMax(x, s_red[tid][w]);
    s_box[tid] = x;
  }
  __syncthreads();
  for (int a = 0; a < 6; ++a) box[a] = s_box[a];
}

// LBVH_calcMortonCodes host part (LBVH_CUDA.cu:171-180), f32, IEEE division.
__device__ __forceinline__ void mortonMapping(const float box[6], bvht_vec3 &off, bvht_vec3 &sc) {
  off = {0.0f - box[0], 0.0f - box[1], 0.0f - box[2]};
  sc = {__fdiv_rn(1.0f, box[3] + off.x), __fdiv_rn(1.0f, box[4] + off.y), __fdiv_rn(1.0f, box[5] + off.z)};
}

// Box of all vertices.  The vertex array is read as 16-byte words (float4
// q holds floats 4q .. 4q + 3, so its component k is axis (q + k) mod 3)
// when it is 16-byte aligned: the grid has a multiple of 3 threads, so every
// word a thread reads has the same phase q mod 3 and the thread keeps three
// min/max slots (components 0 and 3 in slot 0), turned into axes at the end.
// The last 3 nv mod 4 floats are read one by one.  An unaligned array is
// read one vertex at a time.
__global__ __launch_bounds__(kBoxBlock) void k_vertBox(const bvht_vec3 *__restrict__ vert, uint32_t nv,
                                                       float *__restrict__ partials, uint32_t *__restrict__ zeroPtr,
                                                       size_t zeroWords) {
  const uint32_t tid = threadIdx.x;
  for (size_t i = (size_t)blockIdx.x * kBoxBlock + tid; i < zeroWords; i += (size_t)gridDim.x * kBoxBlock) zeroPtr[i] = 0;
  float mn[3] = {HUGE_VALF, HUGE_VALF, HUGE_VALF}, mx[3] = {-HUGE_VALF, -HUGE_VALF, -HUGE_VALF};
  const uint32_t g = blockIdx.x * kBoxBlock + tid, T = gridDim.x * kBoxBlock;
  if ((reinterpret_cast<uintptr_t>(vert) & 15u) == 0 && T % 3u == 0) {
    const float4 *q4 = reinterpret_cast<const float4 *>(vert);
    const uint64_t nf = 3ull * nv, nq = nf / 4;
    float sn[3] = {HUGE_VALF, HUGE_VALF, HUGE_VALF}, sx[3] = {-HUGE_VALF, -HUGE_VALF, -HUGE_VALF};
#pragma unroll 4
    for (uint64_t q = g; q < nq; q += T) {
      const float4 v = q4[q];
      sn[0] = fminf(sn[0], fminf(v.x, v.w));
      sx[0] = fmaxf(sx[0], fmaxf(v.x, v.w));
      sn[1] = fminf(sn[1], v.y);
      sx[1] = fmaxf(sx[1], v.y);
      sn[2] = fminf(sn[2], v.z);
      sx[2] = fmaxf(sx[2], v.z);
    }
    const uint32_t ph = g % 3u;  // slot s holds axis (ph + s) mod 3
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const uint32_t sl = ((uint32_t)a + 3u - ph) % 3u;
      mn[a] = sl == 0 ? sn[0] : (sl == 1 ? sn[1] : sn[2]);
      mx[a] = sl == 0 ? sx[0] : (sl == 1 ? sx[1] : sx[2]);
    }
    if (g < nf - 4 * nq) {  // at most 3 floats past the last whole word
      const uint64_t fi = 4 * nq + g;
      const float x = reinterpret_cast<const float *>(vert)[fi];
      const uint32_t a = (uint32_t)(fi % 3);
#pragma unroll
      for (int b = 0; b < 3; ++b)
        if ((uint32_t)b == a) {
          mn[b] = fminf(mn[b], x);
          mx[b] = fmaxf(mx[b], x);
        }
    }
  } else {
#pragma unroll 4
    for (uint32_t i = g; i < nv; i += T) {
      const bvht_vec3 p = vert[i];
      mn[0] = fminf(mn[0], p.x);
      mn[1] = fminf(mn[1], p.y);
      mn[2] = fminf(mn[2], p.z);
      mx[0] = fmaxf(mx[0], p.x);
      mx[1] = fmaxf(mx[1], p.y);
      mx[2] = fmaxf(mx[2], p.z);
    }
  }
  shiftedBox(mn, mx);
  writeBoxPartial(mn, mx, partials);
}

// Morton keys plus the histograms of the digit passes FIRST..PASSES-1 (the
// passes the sort runs).  Workgroups of kMortonBlock threads: 16 waves per
// workgroup hide the face -> vertex load chain while the grid stays at two
// workgroups per CU (one global flush of the histograms per workgroup).
// VERTEX: the mapping comes from k_vertBox's partials (reduced by every
// block; block 0 records the box in prm) and the pass also writes its
// blocks' partial triangle boxes (triParts) for k_mortonFix.  Otherwise the
// mapping is prm's (the staged calls).
// low != null (63-bit keys): each key is written packed, its low 24 bits
// replaced by the triangle index, and the low bits go to low[index].
constexpr int kMortonBlock = 1024;
constexpr uint64_t kLowMask = (1ull << kSplitLoBits) - 1;
template <typename K>
__device__ __forceinline__ void storeKey(K *keys, uint32_t *low, uint32_t i, K code) {
  if constexpr (sizeof(K) == 8) {
    if (low) {
      keys[i] = (code & ~kLowMask) | i;
      low[i] = (uint32_t)(code & kLowMask);
      return;
    }
  }
  keys[i] = code;
}
static_assert(kMortonBlock == kBoxBlock, "the encode passes share the box reductions");
template <typename K, int PASSES, int FIRST, bool VERTEX>
__global__ __launch_bounds__(kMortonBlock) void k_morton(const bvht_vec3 *__restrict__ vert,
                                                         const bvht_Triangle *__restrict__ faces, uint32_t n,
                                                         BuildParams *__restrict__ prm, K *__restrict__ keys,
                                                         uint32_t *__restrict__ low, uint32_t *__restrict__ hist,
                                                         const float *__restrict__ vertParts, uint32_t numVertParts,
                                                         float *__restrict__ triParts) {
  __shared__ uint32_t s_hist[PASSES][kSortRadix];
  for (uint32_t j = threadIdx.x; j < PASSES * kSortRadix; j += kMortonBlock) (&s_hist[0][0])[j] = 0;
  bvht_vec3 off, sc;
  if (VERTEX) {
    float box[6];
    readBoxPartials(vertParts, numVertParts, box);
    mortonMapping(box, off, sc);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      for (int a = 0; a < 3; ++a) {
        prm->sceneMin[a] = box[a];
        prm->sceneMax[a] = box[3 + a];
        prm->offset[a] = (&off.x)[a];
        prm->scale[a] = (&sc.x)[a];
      }
    }
  } else {
    __syncthreads();
    off = {prm->offset[0], prm->offset[1], prm->offset[2]};
    sc = {prm->scale[0], prm->scale[1], prm->scale[2]};
  }
  // VERTEX: min / max of the referenced vertices, shifted at the end like
  // k_vertBox's (the triangles' box, by the same monotonicity)
  float mn[3] = {HUGE_VALF, HUGE_VALF, HUGE_VALF}, mx[3] = {-HUGE_VALF, -HUGE_VALF, -HUGE_VALF};
  // wave-uniform trips (the histogram adds use ballots)
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t base = (blockIdx.x * (kMortonBlock / 64) + wave) * 64; base < n; base += gridDim.x * kMortonBlock) {
    const uint32_t i = base + lane;
    const bool valid = i < n;
    const bvht_Triangle t = faces[valid ? i : base];
    const bvht_vec3 va = ldVec3(vert + t.v1), vb = ldVec3(vert + t.v2), vc = ldVec3(vert + t.v3);
    const bvht_vec3 c = centroid(va, vb, vc);
    if (VERTEX) {  // a repeat of triangle `base` for a lane past the end changes no min/max
      mn[0] = fminf(mn[0], fminf(fminf(va.x, vb.x), vc.x));
      mn[1] = fminf(mn[1], fminf(fminf(va.y, vb.y), vc.y));
      mn[2] = fminf(mn[2], fminf(fminf(va.z, vb.z), vc.z));
      mx[0] = fmaxf(mx[0], fmaxf(fmaxf(va.x, vb.x), vc.x));
      mx[1] = fmaxf(mx[1], fmaxf(fmaxf(va.y, vb.y), vc.y));
      mx[2] = fmaxf(mx[2], fmaxf(fmaxf(va.z, vb.z), vc.z));
    }
    K code;
    if constexpr (sizeof(K) == 8)
      code = morton63(c, off, sc);
    else
      code = morton30(c, off, sc);
    if (valid) storeKey(keys, low, i, code);
#pragma unroll
    for (int p = FIRST; p < PASSES; ++p)
      waveHistAdd(s_hist[p], (uint32_t)(code >> (p * kSortRadixBits)) & (kSortRadix - 1), valid);
  }
  __syncthreads();
  for (uint32_t j = FIRST * kSortRadix + threadIdx.x; j < PASSES * kSortRadix; j += kMortonBlock) {
    const uint32_t c = (&s_hist[0][0])[j];
    if (c) atomicAdd(&hist[j], c);
  }
  if (!VERTEX) return;
  shiftedBox(mn, mx);
  writeBoxPartial(mn, mx, triParts);
}

// Compares the triangles' box (k_morton's partials) with the vertex box the
// keys were made from (prm, from k_morton's block 0); when they differ
// (unreferenced vertices outside the triangles' box) every key is
// recomputed with the triangles' mapping and the histograms are corrected
// by removing the old key's digits and adding the new one's.  Otherwise it
// exits after the comparison.  prm is only read here (every block compares
// against it).  The digits the sort uses are the same in packed keys.
template <typename K, int PASSES, int FIRST>
__global__ __launch_bounds__(kMortonBlock) void k_mortonFix(const bvht_vec3 *__restrict__ vert,
                                                            const bvht_Triangle *__restrict__ faces, uint32_t n,
                                                            BuildParams *__restrict__ prm, K *__restrict__ keys,
                                                            uint32_t *__restrict__ low, uint32_t *__restrict__ hist,
                                                            const float *__restrict__ triParts, uint32_t numTriParts) {
  __shared__ uint32_t s_hist[PASSES][kSortRadix];  // net change per bin (mod 2^32)
  float box[6];
  readBoxPartials(triParts, numTriParts, box);
  bool same = true;
  for (int a = 0; a < 3; ++a)
    same = same && __float_as_uint(box[a]) == __float_as_uint(prm->sceneMin[a]) &&
           __float_as_uint(box[3 + a]) == __float_as_uint(prm->sceneMax[a]);
  if (same) return;  // block-uniform
  bvht_vec3 off, sc;
  mortonMapping(box, off, sc);
  for (uint32_t j = threadIdx.x; j < PASSES * kSortRadix; j += kMortonBlock) (&s_hist[0][0])[j] = 0;
  __syncthreads();  // every block has compared before block 0 records the box below
  for (uint32_t i = blockIdx.x * kMortonBlock + threadIdx.x; i < n; i += gridDim.x * kMortonBlock) {
    const bvht_Triangle t = faces[i];
    const bvht_vec3 c = centroid(ldVec3(vert + t.v1), ldVec3(vert + t.v2), ldVec3(vert + t.v3));
    K code;
    if constexpr (sizeof(K) == 8)
      code = morton63(c, off, sc);
    else
      code = morton30(c, off, sc);
    const K old = keys[i];
    storeKey(keys, low, i, code);
#pragma unroll
    for (int p = FIRST; p < PASSES; ++p) {
      atomicAdd(&s_hist[p][(uint32_t)(old >> (p * kSortRadixBits)) & (kSortRadix - 1)], 0xFFFFFFFFu);
      atomicAdd(&s_hist[p][(uint32_t)(code >> (p * kSortRadixBits)) & (kSortRadix - 1)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t j = FIRST * kSortRadix + threadIdx.x; j < PASSES * kSortRadix; j += kMortonBlock) {
    const uint32_t c = (&s_hist[0][0])[j];
    if (c) atomicAdd(&hist[j], c);
  }
}

// Morton keys plus the top-digit histogram of the MSD sort.  A wave takes
// kMortonRuns groups of 64 consecutive triangles and issues every face and
// vertex load of them before computing, so each wave has 4x the loads in
// flight of a one-triangle loop.
constexpr int kMortonRuns = 4;
template <typename K>
__global__ __launch_bounds__(kBlock) void k_mortonMsd(const bvht_vec3 *__restrict__ vert,
                                                      const bvht_Triangle *__restrict__ faces, uint32_t n,
                                                      const BuildParams *__restrict__ prm, K *__restrict__ keys,
                                                      uint32_t *__restrict__ hist, int shift) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bvht_vec3 off = {prm->offset[0], prm->offset[1], prm->offset[2]};
  const bvht_vec3 sc = {prm->scale[0], prm->scale[1], prm->scale[2]};
  const uint32_t stride = gridDim.x * kBlock * kMortonRuns;
  for (uint32_t base = (blockIdx.x * (kBlock / 64) + wave) * 64 * kMortonRuns; base < n; base += stride) {
    bvht_Triangle t[kMortonRuns];
#pragma unroll
    for (int u = 0; u < kMortonRuns; ++u) {
      const uint32_t i = base + u * 64 + lane;
      t[u] = i < n ? faces[i] : bvht_Triangle{0, 0, 0};
    }
    bvht_vec3 a[kMortonRuns], b[kMortonRuns], c[kMortonRuns];
#pragma unroll
    for (int u = 0; u < kMortonRuns; ++u) {
      a[u] = ldVec3(vert + t[u].v1);
      b[u] = ldVec3(vert + t[u].v2);
      c[u] = ldVec3(vert + t[u].v3);
    }
#pragma unroll
    for (int u = 0; u < kMortonRuns; ++u) {
      const uint32_t i = base + u * 64 + lane;
      const bool valid = i < n;
      const bvht_vec3 ce = centroid(a[u], b[u], c[u]);
      K code;
      if constexpr (sizeof(K) == 8)
        code = morton63(ce, off, sc);
      else
        code = morton30(ce, off, sc);
      if (valid) keys[i] = code;
      if (base + u * 64 < n) msdCountRun(hist, (uint32_t)(code >> shift) & (kMsdBuckets - 1), valid);
    }
  }
}

// Karras 2012 over the whole key (morton, index) -- extension (topology 1).
// delta(i, j) = common prefix length of (key_i, i) and (key_j, j).
template <typename K>
__device__ __forceinline__ int keyDelta(const K *__restrict__ keys, uint32_t n, uint32_t i, uint32_t j) {
  if (j >= n) return -1;
  K a = keys[i], b = keys[j];
  constexpr int W = sizeof(K) * 8;
  if (a == b) return W + __clz(i ^ j);
  if constexpr (sizeof(K) == 8)
    return __clzll((long long)(a ^ b));
  else
    return __clz((int)(a ^ b));
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_karras(bvht_BVHNode *__restrict__ nodes, const K *__restrict__ keys,
                                                   uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i + 1 >= n) return;
  int d = keyDelta(keys, n, i, i + 1) - keyDelta(keys, n, i, i - 1);
  d = d > 0 ? 1 : -1;
  const int dmin = keyDelta(keys, n, i, i - (uint32_t)d);
  uint32_t lMax = 2;
  while (keyDelta(keys, n, i, i + lMax * (uint32_t)d) > dmin) lMax *= 2;
  uint32_t l = 0;
  do {
    lMax >>= 1;
    if (keyDelta(keys, n, i, i + (l + lMax) * (uint32_t)d) > dmin) l += lMax;
  } while (lMax > 1);
  const uint32_t j = i + l * (uint32_t)d;
  const uint32_t first = min(i, j), last = max(i, j);
  const int common = keyDelta(keys, n, first, last);
  uint32_t split = first, step = last - first;
  do {
    step = (step + 1) >> 1;
    const uint32_t ns = split + step;
    if (ns < last && keyDelta(keys, n, first, ns) > common) split = ns;
  } while (step > 1);
  uint32_t a = split, b = split + 1;
  if (a == first) a += n - 1;
  if (b == last) b += n - 1;
  nodes[i].left = a;
  nodes[i].right = b;
  nodes[i].isLeafFlag = 0;
  nodes[i].level = 0;
  nodes[i]._pad = 0;
  nodes[a].parent = i;
  nodes[a].isLeft = 1;
  nodes[b].parent = i;
  nodes[b].isLeft = 0;
  if (i == 0) {
    nodes[0].parent = 0;
    nodes[0].isLeft = 0;
  }
}

__global__ __launch_bounds__(kBlock) void k_leaves(bvht_BVHNode *__restrict__ nodes,
                                                   const uint32_t *__restrict__ sortedFaces,
                                                   const bvht_Triangle *__restrict__ faces,
                                                   const bvht_vec3 *__restrict__ vert, uint32_t n) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const uint32_t f = sortedFaces[p];
  const bvht_Triangle t = faces[f];
  bvht_AABB b = triangleBox(vert[t.v1], vert[t.v2], vert[t.v3]);
  // parent / isLeft are written by k_karras (kGenLeafNodes writes UINT32_MAX)
  storeNode(nodes + n - 1 + p, b, n == 1 ? 0u : UINT32_MAX, f, 1u, 1u, 0u, 255u, surfaceArea(b));
}

struct Internal {
  SortPlan plan;    // onesweep (digitBits 8, the default; 30-bit keys)
  SplitPlan split;  // digitBits 8 with 63-bit keys: top 30 bits, then the groups
  MsdPlan msd;      // MSD sort (digitBits 16)
  bool useMsd = false, useSplit = false;
  bool packedKeys = false;  // keys[0] holds packed 63-bit keys (split_sort.hpp)
  float *partials = nullptr;
  uint32_t boxBlocks = 0;
  int resultBuf = 0;
  bvht_MeshRaw mesh = {};  // mesh of the current build (staged API)
  bool haveMesh = false;
  BuildParams hostParams = {};
};

Internal *internalOf(bvht_LBVH_WorkingMemory *m) { return reinterpret_cast<Internal *>(m->internal); }

struct Timing {
  static constexpr int kMax = 32;
  bool enabled = false;
  bool groupPasses = false;  // the sort's digit passes as one entry (no events between them)
  int lastPass = -1;         // index of the sort's last digit pass
  hipEvent_t ev[kMax] = {};
  const char *name[kMax] = {};
  bool created = false;
  int recorded = 0;
  void start() { recorded = 0; }
  void record(const char *label) {
    if (!enabled || recorded >= kMax) return;
    if (!created) {
      for (auto &e : ev) (void)hipEventCreate(&e);
      created = true;
    }
    (void)hipEventRecord(ev[recorded], stream());
    name[recorded] = label;
    recorded++;
  }
  void recordPass(int pass) {
    if (groupPasses && pass != lastPass) return;
    record("k_onesweep");
  }
};
Timing &timing() {
  static Timing t;
  return t;
}

bool validateMesh(const bvht_MeshRaw *mesh, const char *what) {
  if (!mesh || mesh->numFaces == 0 || !mesh->faces || !mesh->vert) {
    setError("%s: empty or missing mesh", what);
    return false;
  }
  if (mesh->numFaces > (1u << 30)) {
    setError("%s: %u faces exceed the 2^30 limit", what, mesh->numFaces);
    return false;
  }
  return true;
}

uint32_t boxBlocks(uint32_t n) {
  uint32_t b = divUp(n, kBoxBlock * 4);
  uint32_t cap = std::min<uint32_t>(kBoxBlocksMax, (uint32_t)numCUs() * 2);
  return b < 1 ? 1 : (b > cap ? cap : b);
}

// The vertex-box pass: up to two workgroups per CU, a multiple of 3 of them
// (k_vertBox's 16-byte reads keep one phase per thread).
uint32_t vertBlocks(uint32_t nv) {
  uint32_t b = divUp((uint64_t)nv + 1, kBoxBlock);
  const uint32_t cap = std::min<uint32_t>(kBoxBlocksMax, (uint32_t)numCUs() * 2) / 3 * 3;
  b = divUp(b, 3) * 3;
  return b > cap ? cap : b;
}

// Every Morton workgroup ends with one global atomic per non-empty bin of
// every pass into the same 8 x 256 counters, so the grid is kept to two
// workgroups per CU (each then covers ~8K triangles at 4.24M triangles).
uint32_t mortonBlocks(uint32_t n) {
  uint32_t b = divUp(n, kMortonBlock * 4);
  uint32_t cap = (uint32_t)numCUs() * 2;
  return b < 1 ? 1 : (b > cap ? cap : b);
}

// Words of the sort scratch that must be zero before the key producer runs.
size_t sortZeroWords(const Internal *in) {
  if (in->useMsd) return in->msd.zeroWords;
  if (in->useSplit) return in->split.ctrlOff + 16;
  return in->plan.totalWords;
}

// check: the keys are made from k_vertBox's box; verify it against the
// triangles' box and repair the keys if it differs (k_mortonFix).
bool launchMorton(bvht_LBVH_WorkingMemory *mem, const bvht_MeshRaw *mesh, bool check) {
  Internal *in = internalOf(mem);
  uint32_t n = mesh->numFaces;
  if (in->useMsd) {
    uint32_t *hist = msdHist(in->msd, (uint32_t *)mem->sortScratch);
    uint32_t blocks = divUp(n, kBlock * kMortonRuns);
    blocks = std::max<uint32_t>(1, std::min<uint32_t>(blocks, (uint32_t)numCUs() * 8));
    const int shift = (int)mem->opts.mortonBits - kMsdTopBits;
    if (mem->opts.mortonBits == 63)
      k_mortonMsd<uint64_t><<<blocks, kBlock, 0, stream()>>>(mesh->vert, mesh->faces, n, (const BuildParams *)mem->params,
                                                              (uint64_t *)mem->keys[0], hist, shift);
    else
      k_mortonMsd<uint32_t><<<blocks, kBlock, 0, stream()>>>(mesh->vert, mesh->faces, n, (const BuildParams *)mem->params,
                                                              (uint32_t *)mem->keys[0], hist, shift);
    return checkLaunch("k_mortonMsd");
  }
  uint32_t *hist = in->useSplit ? splitHist(in->split, (uint32_t *)mem->sortScratch)
                                : sortHist(in->plan, (uint32_t *)mem->sortScratch);
  BuildParams *prm = (BuildParams *)mem->params;
  const uint32_t blocks = mortonBlocks(n);
  // vertex-box path: k_vertBox's partials in the first half of in->partials,
  // the triangles' partials in the second
  const float *vparts = in->partials;
  const uint32_t nvparts = vertBlocks(mesh->numVert);
  float *tparts = in->partials + (size_t)kBoxBlocksMax * 6;
  // the full build packs its 63-bit keys (the staged calls hand real keys
  // to the caller)
  in->packedKeys = check && in->useSplit && in->split.canPack;
  uint32_t *low = in->packedKeys ? splitLow(in->split, (uint32_t *)mem->sortScratch) : nullptr;
  if (mem->opts.mortonBits == 63 && check)
    k_morton<uint64_t, 8, kSplitFirstPass, true><<<blocks, kMortonBlock, 0, stream()>>>(
        mesh->vert, mesh->faces, n, prm, (uint64_t *)mem->keys[0], low, hist, vparts, nvparts, tparts);
  else if (mem->opts.mortonBits == 63)
    k_morton<uint64_t, 8, kSplitFirstPass, false><<<blocks, kMortonBlock, 0, stream()>>>(
        mesh->vert, mesh->faces, n, prm, (uint64_t *)mem->keys[0], nullptr, hist, nullptr, 0u, nullptr);
  else if (check)
    k_morton<uint32_t, 4, 0, true><<<blocks, kMortonBlock, 0, stream()>>>(
        mesh->vert, mesh->faces, n, prm, (uint32_t *)mem->keys[0], nullptr, hist, vparts, nvparts, tparts);
  else
    k_morton<uint32_t, 4, 0, false><<<blocks, kMortonBlock, 0, stream()>>>(
        mesh->vert, mesh->faces, n, prm, (uint32_t *)mem->keys[0], nullptr, hist, nullptr, 0u, nullptr);
  if (!checkLaunch("k_morton")) return false;
  if (!check) return true;
  timing().record("k_morton");
  const uint32_t fixBlocks = std::min<uint32_t>(blocks, (uint32_t)numCUs());
  if (mem->opts.mortonBits == 63)
    k_mortonFix<uint64_t, 8, kSplitFirstPass><<<fixBlocks, kMortonBlock, 0, stream()>>>(
        mesh->vert, mesh->faces, n, prm, (uint64_t *)mem->keys[0], low, hist, tparts, blocks);
  else
    k_mortonFix<uint32_t, 4, 0><<<fixBlocks, kMortonBlock, 0, stream()>>>(
        mesh->vert, mesh->faces, n, prm, (uint32_t *)mem->keys[0], nullptr, hist, tparts, blocks);
  if (!checkLaunch("k_mortonFix")) return false;
  timing().record("k_mortonFix");
  return true;
}

bool runSort(bvht_LBVH_WorkingMemory *mem) {
  Internal *in = internalOf(mem);
  const bool keepKeys = mem->opts.topology == 1;
  Timing &tm = timing();
  if (in->useMsd) {
    // pairs end in keys[0] / values[1]; keys[0] doubles as split scratch
    if (!msdRun(in->msd, mem->keys[0], mem->keys[1], mem->values[0], mem->keys[0], mem->values[1], keepKeys,
                (uint32_t *)mem->sortScratch, tm.enabled ? [](const char *s) { timing().record(s); } : nullptr))
      return false;
    in->resultBuf = 1;
    mem->sortedValues = mem->values[1];
    mem->sortedKeys = keepKeys ? mem->keys[0] : nullptr;
    return true;
  }
  tm.lastPass = (in->useSplit ? in->split.plan.passes : in->plan.passes) - 1;
  if (in->useSplit) {
    uint64_t *k2[2] = {(uint64_t *)mem->keys[0], (uint64_t *)mem->keys[1]};
    const int r = splitRun(in->split, k2, mem->values, (uint32_t *)mem->sortScratch, keepKeys, in->packedKeys,
                           tm.enabled ? [](int pass) { timing().recordPass(pass); } : nullptr,
                           tm.enabled ? [](const char *s) { timing().record(s); } : nullptr);
    if (r < 0) return false;
    in->resultBuf = r;
    mem->sortedValues = mem->values[r];
    mem->sortedKeys = keepKeys ? mem->keys[r] : nullptr;
    return true;
  }
  int r = sortRun(in->plan, mem->keys, mem->values, true, keepKeys, (uint32_t *)mem->sortScratch,
                  tm.enabled ? [](int pass) { timing().recordPass(pass); } : nullptr);
  if (r < 0) return false;
  in->resultBuf = r;
  mem->sortedValues = mem->values[r];
  mem->sortedKeys = keepKeys ? mem->keys[r] : nullptr;
  return true;
}

bool runTree(bvht_LBVH_WorkingMemory *mem, const bvht_MeshRaw *mesh, bvht_CUDAMemoryBVHPointer *bvh) {
  const uint32_t n = mesh->numFaces;
  if (mem->opts.topology == 1) {
    k_leaves<<<divUp(n, kBlock), kBlock, 0, stream()>>>(bvh->nodes, mem->sortedValues, mesh->faces, mesh->vert, n);
    if (!checkLaunch("k_leaves")) return false;
    if (n > 1) {
      if (mem->opts.mortonBits == 63)
        k_karras<uint64_t><<<divUp(n, kBlock), kBlock, 0, stream()>>>(bvh->nodes, (const uint64_t *)mem->sortedKeys, n);
      else
        k_karras<uint32_t><<<divUp(n, kBlock), kBlock, 0, stream()>>>(bvh->nodes, (const uint32_t *)mem->sortedKeys, n);
      if (!checkLaunch("k_karras")) return false;
      return refitFromLeaves(bvh->nodes, n - 1, n, nullptr, mem->atomicLocks, mem->numLocks);
    }
    return true;
  }
  return launchTreeLevels((int)mem->opts.tileLog2, bvh->nodes, mem->sortedValues, mesh->faces, mesh->vert, n,
                          timing().enabled ? [](bool l0) { timing().record(l0 ? "k_buildLevel0" : "k_buildLevelUp"); }
                                           : nullptr);
}

// The whole build on the stream: scene box, keys, sort, tree.
bool buildAll(bvht_LBVH_WorkingMemory *mem, const bvht_MeshRaw *mesh, bvht_CUDAMemoryBVHPointer *bvh) {
  Internal *in = internalOf(mem);
  in->mesh = *mesh;
  in->haveMesh = true;
  Timing &tm = timing();
  tm.start();
  tm.record("begin");
  // the vertex box (one streaming pass) when the keys come from k_morton,
  // which checks it; the MSD key pass takes the triangles' box
  const bool vertexBox = !in->useMsd;
  if (vertexBox) {
    k_vertBox<<<vertBlocks(mesh->numVert), kBoxBlock, 0, stream()>>>(mesh->vert, mesh->numVert, in->partials,
                                                                    (uint32_t *)mem->sortScratch, sortZeroWords(in));
    if (!checkLaunch("k_vertBox")) return false;
    tm.record("k_vertBox");
  } else {
    k_sceneBox<<<in->boxBlocks, kBoxBlock, 0, stream()>>>(mesh->vert, mesh->faces, mesh->numFaces,
                                                       (BuildParams *)mem->params, in->partials,
                                                       (uint32_t *)mem->sortScratch, sortZeroWords(in));
    if (!checkLaunch("k_sceneBox")) return false;
    tm.record("k_sceneBox");
  }
  if (!launchMorton(mem, mesh, vertexBox)) return false;
  if (!vertexBox) tm.record("k_morton");
  if (!runSort(mem)) return false;
  return runTree(mem, mesh, bvh);
}

}  // namespace

template <int TLOG>
static bool treeLevelsT(bvht_BVHNode *nodes, const uint32_t *sorted, const bvht_Triangle *faces, const bvht_vec3 *vert,
                        uint32_t n, void (*afterLevel)(bool)) {
  const SegDesc sd = {n, 0, 0, 0};
  uint32_t unitLog = 0;
  bool first = true;
  while (true) {
    const uint64_t U = (((uint64_t)n) + (1ull << unitLog) - 1) >> unitLog;
    const uint32_t blocks = (uint32_t)((U + (1u << TLOG) - 1) >> TLOG);
    if (first)
      k_buildLevel0<TLOG, false, false><<<blocks, kBuildBlock, 0, stream()>>>(nodes, sorted, faces, vert, nullptr, sd,
                                                                              nullptr, nullptr);
    else
      k_buildLevelUp<TLOG, false><<<blocks, kBuildBlock, 0, stream()>>>(nodes, sd, nullptr, nullptr, unitLog);
    if (!checkLaunch(first ? "k_buildLevel0" : "k_buildLevelUp")) return false;
    if (afterLevel) afterLevel(first);
    first = false;
    if (blocks <= 1) break;
    unitLog += TLOG;
  }
  return true;
}

bool launchTreeLevels(int tlog, bvht_BVHNode *nodes, const uint32_t *sorted, const bvht_Triangle *faces,
                      const bvht_vec3 *vert, uint32_t n, void (*afterLevel)(bool)) {
  switch (tlog) {
    case 8: return treeLevelsT<8>(nodes, sorted, faces, vert, n, afterLevel);
    case 9: return treeLevelsT<9>(nodes, sorted, faces, vert, n, afterLevel);
    case 10: return treeLevelsT<10>(nodes, sorted, faces, vert, n, afterLevel);
    case 11: return treeLevelsT<11>(nodes, sorted, faces, vert, n, afterLevel);
  }
  setError("launchTreeLevels: unsupported tile log2 %d", tlog);
  return false;
}

template <int TLOG>
static bool treeLevelsBatchT(bvht_BVHNode *nodes, const uint32_t *sorted, const bvht_Triangle *faces,
                             const bvht_vec3 *vert, const float2 *rec, const SegDesc *segs, const uint2 *const *tiles,
                             const uint32_t *counts, int levels) {
  const SegDesc none = {0, 0, 0, 0};
  for (int L = 0; L < levels; ++L) {
    if (counts[L] == 0) continue;
    if (L == 0 && rec)
      k_buildLevel0<TLOG, true, true><<<counts[L], kBuildBlock, 0, stream()>>>(nodes, sorted, faces, vert, rec, none,
                                                                                 segs, tiles[L]);
    else if (L == 0)
      k_buildLevel0<TLOG, true, false><<<counts[L], kBuildBlock, 0, stream()>>>(nodes, sorted, faces, vert, nullptr,
                                                                                  none, segs, tiles[L]);
    else
      k_buildLevelUp<TLOG, true><<<counts[L], kBuildBlock, 0, stream()>>>(nodes, none, segs, tiles[L],
                                                                            (uint32_t)(L * TLOG));
    if (!checkLaunch(L == 0 ? "k_buildLevel0 batch" : "k_buildLevelUp batch")) return false;
  }
  return true;
}

bool launchTreeLevelsBatch(int tlog, bvht_BVHNode *nodes, const uint32_t *sorted, const bvht_Triangle *faces,
                           const bvht_vec3 *vert, const float2 *rec, const SegDesc *segs, const uint2 *const *tiles,
                           const uint32_t *counts, int levels) {
  switch (tlog) {
    case 8: return treeLevelsBatchT<8>(nodes, sorted, faces, vert, rec, segs, tiles, counts, levels);
    case 9: return treeLevelsBatchT<9>(nodes, sorted, faces, vert, rec, segs, tiles, counts, levels);
    case 10: return treeLevelsBatchT<10>(nodes, sorted, faces, vert, rec, segs, tiles, counts, levels);
    case 11: return treeLevelsBatchT<11>(nodes, sorted, faces, vert, rec, segs, tiles, counts, levels);
  }
  setError("launchTreeLevelsBatch: unsupported tile log2 %d", tlog);
  return false;
}

}  // namespace bvht

using namespace bvht;

extern "C" {

bvht_LBVH_Options bvht_LBVH_defaultOptions(void) {
  bvht_LBVH_Options o;
  o.mortonBits = 30;
  o.topology = 0;
  o.tileLog2 = kTileLog;
  o.digitBits = kSortRadixBits;
  return o;
}

bool LBVH_allocateBVH(bvht_CUDAMemoryBVHPointer *bvh, const bvht_MeshRaw *mesh) {
  if (!bvh || !validateMesh(mesh, "LBVH_allocateBVH")) return false;
  if (bvh->bvh != nullptr || bvh->nodes != nullptr) {  // LBVH_CUDA.cu:400
    setError("LBVH_allocateBVH: BVH memory already allocated");
    return false;
  }
  bvh->numNodes = mesh->numFaces * 2 - 1;
  bvht_BVH hdr;
  std::memset(&hdr, 0, sizeof(hdr));
  if (!devAlloc(&bvh->bvh, 1, "LBVH_allocateBVH header") || !devAlloc(&bvh->nodes, bvh->numNodes, "LBVH_allocateBVH nodes"))
    goto error;
  hdr.bvh = bvh->nodes;
  hdr.vSize = bvh->numNodes;
  hdr.vCapacity = bvh->numNodes;
  hdr.vRootIndex = 0;
  if (!check(hipMemcpyAsync(bvh->bvh, &hdr, sizeof(hdr), hipMemcpyHostToDevice, stream()), "LBVH_allocateBVH header") ||
      !check(hipStreamSynchronize(stream()), "LBVH_allocateBVH sync"))
    goto error;
  return true;
error:
  devFree(bvh->nodes);
  devFree(bvh->bvh);
  bvh->nodes = nullptr;
  bvh->bvh = nullptr;
  bvh->numNodes = 0;
  return false;
}

bvht_LBVH_WorkingMemory bvht_LBVH_allocateWorkingMemoryEx(const bvht_MeshRaw *mesh, const bvht_LBVH_Options *opts) {
  bvht_LBVH_WorkingMemory m;
  std::memset(&m, 0, sizeof(m));
  m.opts = opts ? *opts : bvht_LBVH_defaultOptions();
  if (m.opts.mortonBits != 30 && m.opts.mortonBits != 63) {
    setError("LBVH_allocateWorkingMemory: mortonBits must be 30 or 63 (got %u)", m.opts.mortonBits);
    return m;
  }
  if (m.opts.topology > 1) {
    setError("LBVH_allocateWorkingMemory: topology must be 0 or 1 (got %u)", m.opts.topology);
    return m;
  }
  if (m.opts.tileLog2 == 0) m.opts.tileLog2 = kTileLog;
  if (m.opts.tileLog2 < 8 || m.opts.tileLog2 > 11) {
    setError("LBVH_allocateWorkingMemory: tileLog2 must be in [8, 11] (got %u)", m.opts.tileLog2);
    return m;
  }
  if (m.opts.digitBits == 0) m.opts.digitBits = kSortRadixBits;
  if (m.opts.digitBits != kMsdTopBits && m.opts.digitBits != kSortRadixBits) {
    setError("LBVH_allocateWorkingMemory: digitBits must be %d (MSD sort) or %d (onesweep) (got %u)", kMsdTopBits,
             kSortRadixBits, m.opts.digitBits);
    return m;
  }
  if (!validateMesh(mesh, "LBVH_allocateWorkingMemory")) return m;
  const uint32_t n = mesh->numFaces;
  const int keyBytes = m.opts.mortonBits == 63 ? 8 : 4;
  Internal *in = new (std::nothrow) Internal();
  if (!in) {
    setError("LBVH_allocateWorkingMemory: host allocation failed");
    return m;
  }
  m.internal = in;
  m.numFaces = n;
  in->useMsd = m.opts.digitBits == kMsdTopBits;
  in->useSplit = !in->useMsd && m.opts.mortonBits == 63;
  if (in->useMsd) {
    in->msd = makeMsdPlan(n, keyBytes, (int)m.opts.mortonBits);
    m.sortScratchBytes = msdScratchBytes(in->msd);
  } else if (in->useSplit) {
    in->split = makeSplitPlan(n);
    m.sortScratchBytes = splitScratchBytes(in->split);
  } else {
    in->plan = makeSortPlan(n, keyBytes, (int)m.opts.mortonBits);
    m.sortScratchBytes = sortScratchBytes(in->plan);
  }
  in->boxBlocks = boxBlocks(n);
  m.numLocks = 2 * n - 1;
  if (!devMalloc(&m.keys[0], (size_t)keyBytes * n, "LBVH keys") || !devMalloc(&m.keys[1], (size_t)keyBytes * n, "LBVH keys") ||
      !devAlloc(&m.values[0], n, "LBVH values") || !devAlloc(&m.values[1], n, "LBVH values") ||
      !devMalloc(&m.sortScratch, m.sortScratchBytes, "LBVH sort scratch") ||
      !devMalloc(&m.params, sizeof(BuildParams), "LBVH params") ||
      !devAlloc(&in->partials, (size_t)kBoxBlocksMax * 12, "LBVH partials") ||
      !devAlloc(&m.atomicLocks, m.numLocks, "LBVH locks")) {
    LBVH_freeWorkingMemory(&m);
    return m;
  }
  if (!check(hipMemsetAsync(m.params, 0, sizeof(BuildParams), stream()), "LBVH params init") ||
      !check(hipMemsetAsync(m.atomicLocks, 0, sizeof(uint32_t) * m.numLocks, stream()), "LBVH locks init") ||
      // the sort's counters and error words read 0 before the first build
      !check(hipMemsetAsync(m.sortScratch, 0, sizeof(uint32_t) * sortZeroWords(in), stream()), "LBVH sort init") ||
      !check(hipStreamSynchronize(stream()), "LBVH init sync")) {
    LBVH_freeWorkingMemory(&m);
    return m;
  }
  m.lRes = true;
  return m;
}

bvht_LBVH_WorkingMemory LBVH_allocateWorkingMemory(const bvht_MeshRaw *mesh) {
  return bvht_LBVH_allocateWorkingMemoryEx(mesh, nullptr);
}

void LBVH_freeWorkingMemory(bvht_LBVH_WorkingMemory *mem) {
  if (!mem) return;
  Internal *in = internalOf(mem);
  if (in) devFree(in->partials);
  delete in;
  devFree(mem->keys[0]);
  devFree(mem->keys[1]);
  devFree(mem->values[0]);
  devFree(mem->values[1]);
  devFree(mem->sortScratch);
  devFree(mem->params);
  devFree(mem->atomicLocks);
  bvht_LBVH_Options o = mem->opts;
  std::memset(mem, 0, sizeof(*mem));
  mem->opts = o;
}

static bool checkMem(bvht_LBVH_WorkingMemory *mem, const bvht_MeshRaw *mesh, const char *what) {
  if (!mem || !mem->lRes || !mem->internal) {
    setError("%s: working memory not allocated", what);
    return false;
  }
  if (!validateMesh(mesh, what)) return false;
  if (mesh->numFaces != mem->numFaces) {
    setError("%s: mesh has %u faces, working memory was sized for %u", what, mesh->numFaces, mem->numFaces);
    return false;
  }
  return true;
}

bvht_AABB LBVH_initTriData(bvht_LBVH_WorkingMemory *mem, const bvht_MeshRaw *mesh) {
  bvht_AABB box;
  std::memset(&box, 0, sizeof(box));
  if (!checkMem(mem, mesh, "LBVH_initTriData")) return box;
  Internal *in = internalOf(mem);
  in->mesh = *mesh;
  in->haveMesh = true;
  k_sceneBox<<<in->boxBlocks, kBoxBlock, 0, stream()>>>(mesh->vert, mesh->faces, mesh->numFaces, (BuildParams *)mem->params,
                                                     in->partials, (uint32_t *)mem->sortScratch, sortZeroWords(in));
  if (!checkLaunch("k_sceneBox")) return box;
  BuildParams p;
  if (check(hipMemcpyAsync(&p, mem->params, sizeof(p), hipMemcpyDeviceToHost, stream()), "LBVH_initTriData box") &&
      check(hipStreamSynchronize(stream()), "LBVH_initTriData sync")) {
    box.minMax[0] = {p.sceneMin[0], p.sceneMin[1], p.sceneMin[2]};
    box.minMax[1] = {p.sceneMax[0], p.sceneMax[1], p.sceneMax[2]};
  }
  return box;
}

void LBVH_calcMortonCodes(bvht_LBVH_WorkingMemory *mem, bvht_AABB sceneAABB) {
  if (!mem || !mem->lRes || !mem->internal || !internalOf(mem)->haveMesh) {
    setError("LBVH_calcMortonCodes: call LBVH_initTriData first");
    return;
  }
  Internal *in = internalOf(mem);
  // offset / scale on the host exactly as LBVH_CUDA.cu:171-180 (f32, IEEE)
  BuildParams p;
  std::memset(&p, 0, sizeof(p));
  const float mn[3] = {sceneAABB.minMax[0].x, sceneAABB.minMax[0].y, sceneAABB.minMax[0].z};
  const float mx[3] = {sceneAABB.minMax[1].x, sceneAABB.minMax[1].y, sceneAABB.minMax[1].z};
  for (int a = 0; a < 3; ++a) {
    p.sceneMin[a] = mn[a];
    p.sceneMax[a] = mx[a];
    p.offset[a] = 0.0f - mn[a];
    volatile float m = mx[a] + p.offset[a];
    p.scale[a] = 1.0f / m;
  }
  in->hostParams = p;
  if (!check(hipMemcpyAsync(mem->params, &in->hostParams, sizeof(p), hipMemcpyHostToDevice, stream()),
             "LBVH_calcMortonCodes params"))
    return;
  if (!fillU32((uint32_t *)mem->sortScratch, 0u, sortZeroWords(in))) return;
  launchMorton(mem, &in->mesh, false);
}

void LBVH_sortMortonCodes(bvht_LBVH_WorkingMemory *mem) {
  if (!mem || !mem->lRes || !mem->internal) {
    setError("LBVH_sortMortonCodes: working memory not allocated");
    return;
  }
  runSort(mem);
}

void LBVH_buildBVHTree(bvht_LBVH_WorkingMemory *mem, bvht_CUDAMemoryBVHPointer *bvh) {
  if (!mem || !mem->lRes || !mem->internal || !internalOf(mem)->haveMesh || !mem->sortedValues) {
    setError("LBVH_buildBVHTree: run LBVH_initTriData, LBVH_calcMortonCodes and LBVH_sortMortonCodes first");
    return;
  }
  if (!bvh || !bvh->nodes || bvh->numNodes != 2 * mem->numFaces - 1) {
    setError("LBVH_buildBVHTree: BVH not allocated for %u faces", mem->numFaces);
    return;
  }
  runTree(mem, &internalOf(mem)->mesh, bvh);
}

void LBVH_fixAABB(bvht_LBVH_WorkingMemory *mem, bvht_CUDAMemoryBVHPointer *bvh) {
  // Full bottom-up refit from the leaves (kFixAABBTree semantics).  The
  // build pass above already produced the same boxes; this recomputes them
  // from the current leaves, e.g. after leaves were edited.
  if (!mem || !mem->lRes || !bvh || !bvh->nodes) {
    setError("LBVH_fixAABB: working memory or BVH not allocated");
    return;
  }
  const uint32_t n = mem->numFaces;
  if (n > 1) refitFromLeaves(bvh->nodes, n - 1, n, nullptr, mem->atomicLocks, mem->numLocks);
}

void LBVH_doCUDASyc(void) { check(hipStreamSynchronize(stream()), "LBVH_doCUDASyc"); }

bool bvht_LBVH_build(bvht_LBVH_WorkingMemory *mem, const bvht_MeshRaw *mesh, bvht_CUDAMemoryBVHPointer *bvh) {
  if (!checkMem(mem, mesh, "bvht_LBVH_build")) return false;
  if (!bvh || !bvh->nodes || bvh->numNodes != 2 * mesh->numFaces - 1) {
    setError("bvht_LBVH_build: BVH not allocated for %u faces", mesh->numFaces);
    return false;
  }
  return buildAll(mem, mesh, bvh);
}

uint32_t bvht_LBVH_sortStatus(bvht_LBVH_WorkingMemory *mem) {
  if (!mem || !mem->lRes || !mem->internal) {
    setError("bvht_LBVH_sortStatus: working memory not allocated");
    return 0xFFFFFFFFu;
  }
  const Internal *in = internalOf(mem);
  const uint32_t *sc = (const uint32_t *)mem->sortScratch;
  const size_t word = in->useMsd ? in->msd.ctrlOff + 2 : (in->useSplit ? in->split.plan.errOff : in->plan.errOff);
  uint32_t v = 0xFFFFFFFFu, w = 0;
  if (!check(hipMemcpyAsync(&v, sc + word, 4, hipMemcpyDeviceToHost, stream()), "bvht_LBVH_sortStatus") ||
      (in->useSplit &&
       !check(hipMemcpyAsync(&w, sc + in->split.ctrlOff + 2, 4, hipMemcpyDeviceToHost, stream()), "bvht_LBVH_sortStatus")) ||
      !check(hipStreamSynchronize(stream()), "bvht_LBVH_sortStatus sync"))
    return 0xFFFFFFFFu;
  return v | (w << 3);
}

void bvht_LBVH_setTiming(bool enable) { timing().enabled = enable; }
void bvht_LBVH_setTimingPassGroups(bool group) { timing().groupPasses = group; }

int bvht_LBVH_stageTimes(float *ms, int maxStages) {
  Timing &tm = timing();
  if (!tm.enabled || tm.recorded < 2) return 0;
  if (!check(hipEventSynchronize(tm.ev[tm.recorded - 1]), "bvht_LBVH_stageTimes")) return 0;
  int k = 0;
  for (int i = 0; i + 1 < tm.recorded && k < maxStages; ++i, ++k) {
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, tm.ev[i], tm.ev[i + 1]);
    ms[k] = t;
  }
  return k;
}

const char *bvht_LBVH_stageName(int i) {
  Timing &tm = timing();
  if (i < 0 || i + 1 >= tm.recorded) return "";
  return tm.name[i + 1];
}

}  // extern "C"

// ---- a build that allocates nothing ----------------------------------------
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): the working memory of bvht_LBVH_build lives
// in `scratch`, a file-scope device array that every GPU the library runs on
// holds until the process ends.  Meshes of up to kNoAllocMaxFaces faces fit
// (30- or 63-bit keys, default options otherwise).  Every call on one GPU
// uses that GPU's array, so calls that overlap on one GPU race; the array
// keeps the last call's contents, and the parts that must start at zero are
// cleared by the call itself (fill kernels, no memset nodes).

constexpr uint32_t kNoAllocMaxFaces = 1u << 20;
constexpr size_t kNoAllocBytes = 32u << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace bvht {
namespace {

constexpr int kMaxDevices = 64;

// The device address of `scratch` on the current device (looked up once per
// device with that device current; each GPU holds its own copy).
uint8_t *scratchAddress() {
  static uint8_t *addr[kMaxDevices] = {};
  int dev = -1;
  if (!check(hipGetDevice(&dev), "bvht_LBVH_buildNoAlloc device")) return nullptr;
  if (dev < 0 || dev >= kMaxDevices) {
    setError("bvht_LBVH_buildNoAlloc: device %d out of range", dev);
    return nullptr;
  }
  if (!addr[dev]) {
    void *p = nullptr;
    if (!check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "bvht_LBVH_buildNoAlloc scratch address")) return nullptr;
    addr[dev] = static_cast<uint8_t *>(p);
  }
  return addr[dev];
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Carves the working memory of an n-face build out of `base`; returns the
// bytes used (the plans in `in` must be set).
size_t carve(uint8_t *base, uint32_t n, int keyBytes, bvht_LBVH_WorkingMemory &m, Internal &in) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    uint8_t *p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  m.keys[0] = take((size_t)keyBytes * n);
  m.keys[1] = take((size_t)keyBytes * n);
  m.values[0] = reinterpret_cast<uint32_t *>(take(4ull * n));
  m.values[1] = reinterpret_cast<uint32_t *>(take(4ull * n));
  m.sortScratch = take(m.sortScratchBytes);
  m.params = take(sizeof(BuildParams));
  in.partials = reinterpret_cast<float *>(take(sizeof(float) * kBoxBlocksMax * 12));
  return off;
}

}  // namespace
}  // namespace bvht

extern "C" bool bvht_LBVH_buildNoAlloc(const bvht_MeshRaw *mesh, bvht_CUDAMemoryBVHPointer *bvh, uint32_t mortonBits) {
  if (!validateMesh(mesh, "bvht_LBVH_buildNoAlloc")) return false;
  if (mortonBits != 30 && mortonBits != 63) {
    setError("bvht_LBVH_buildNoAlloc: mortonBits must be 30 or 63 (got %u)", mortonBits);
    return false;
  }
  const uint32_t n = mesh->numFaces;
  if (n > kNoAllocMaxFaces) {
    setError("bvht_LBVH_buildNoAlloc: %u faces exceed the maximum of %u for the build without allocation", n,
             kNoAllocMaxFaces);
    return false;
  }
  if (!bvh || !bvh->nodes || bvh->numNodes != 2 * n - 1) {
    setError("bvht_LBVH_buildNoAlloc: BVH not allocated for %u faces", n);
    return false;
  }
  Internal in;
  bvht_LBVH_WorkingMemory m;
  std::memset(&m, 0, sizeof(m));
  m.opts = bvht_LBVH_defaultOptions();
  m.opts.mortonBits = mortonBits;
  m.numFaces = n;
  m.internal = &in;
  m.lRes = true;
  const int keyBytes = mortonBits == 63 ? 8 : 4;
  in.useSplit = mortonBits == 63;
  if (in.useSplit) {
    in.split = makeSplitPlan(n);
    m.sortScratchBytes = splitScratchBytes(in.split);
  } else {
    in.plan = makeSortPlan(n, keyBytes, (int)mortonBits);
    m.sortScratchBytes = sortScratchBytes(in.plan);
  }
  in.boxBlocks = boxBlocks(n);
  if (carve(nullptr, n, keyBytes, m, in) > kNoAllocBytes) {  // cannot happen for n <= kNoAllocMaxFaces
    setError("bvht_LBVH_buildNoAlloc: %u faces need more than %zu bytes of scratch", n, kNoAllocBytes);
    return false;
  }
  uint8_t *base = scratchAddress();
  if (!base) return false;
  carve(base, n, keyBytes, m, in);
  // the scene-box reduction's completion counter starts at zero
  if (!fillU32(reinterpret_cast<uint32_t *>(m.params), 0u, sizeof(BuildParams) / 4)) return false;
  return buildAll(&m, mesh, bvh);
}
