// Batched (segmented) LBVH build: many meshes of uneven size in one launch
// sequence, optionally captured once into a hipGraph and replayed.
//
// Every segment's output equals the standalone build of that mesh (own scene
// box, local face and node indices): the reference build of
// lib/builder/LBVH_CUDA.cu applied per mesh.  Stages:
//   k_segInit    zeroes the sort scratch, initialises the mesh boxes;
//   k_segVertBox per run of one mesh's vertices: block reduce, 6 atomic
//                min/max (order-preserving u32 encoding) into the mesh's
//                vertex box (the reference's triangle box once shifted by
//                -+FLT_EPSILON, unless an unreferenced vertex lies outside);
//   k_segMorton  per chunk of one mesh: the mesh's offset/scale (as
//                LBVH_CUDA.cu:171-180), the 30-bit Morton code, the
//                triangle's leaf box, the mesh's digit histograms of every
//                sort pass and the box of the referenced vertices;
//   k_segFix     per chunk: the two mesh boxes compared; a mesh whose
//                boxes differ gets its keys and histograms recomputed;
//   k_segPrefix  per (mesh, pass): exclusive scan of the histogram;
//   k_segOnesweep one segmented onesweep pass: sort tiles never cross a
//                mesh, the look-back stops at the mesh's first tile, and
//                digit offsets start at the mesh's prefix; u32 keys, 8- or
//                10-bit digits (4 or 3 passes, chosen by the autotuner);
//   k_segOnesweepWide the 8-bit pass over tiles of 1024 threads x 11 keys,
//                look-back in rounds of 16 predecessors of the mesh;
//   k_buildLevel0 / k_buildLevelUp the tile hierarchy per mesh (tiles
//                never cross meshes).
#include "lbvh_kernels.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace bvht {

namespace {

// Per-mesh boxes, kBoxWords order-preserving words per mesh: [0, 6) the min
// and max of all the mesh's vertices (k_segVertBox), [6, 12) those of the
// vertices its triangles reference (k_segMorton).  The reference's scene
// box (triangle boxes, each its vertices' min - FLT_EPSILON and max +
// FLT_EPSILON, with the origin included: LBVH_CUDA.cu:60-119) is the
// second pair shifted (f32 rounding is monotone); the keys are made from
// the first, which equals it unless an unreferenced vertex lies outside
// the triangles' box, and k_segFix repairs such meshes.
constexpr int kBoxWords = 12;
constexpr uint32_t kVertChunk = 4096;  // vertices per k_segVertBox workgroup

__device__ __forceinline__ uint32_t boxInitWord(uint32_t w) { return (w % 6) < 3 ? f2o(HUGE_VALF) : f2o(-HUGE_VALF); }

// The reference's box from a raw vertex min / max pair (6 words at w):
// -+FLT_EPSILON, then the origin.
__device__ __forceinline__ void shiftedMeshBox(const uint32_t *w, float m[6]) {
  for (int a = 0; a < 3; ++a) {
    m[a] = mergeMin(0.0f, o2f(w[a]) - kFltEpsilon);
    m[3 + a] = mergeMax(0.0f, o2f(w[3 + a]) + kFltEpsilon);
  }
}

// Block reduction of per-thread min / max; threads 0..5 add the result to
// the 6 order-preserving words at dst with atomics.
__device__ void blockBoxAtomics(const float mn[3], const float mx[3], uint32_t *dst) {
  __shared__ float s_red[6][kBuildBlock / 64];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int a = 0; a < 3; ++a) {
    const float x = waveMin(mn[a]), y = waveMax(mx[a]);
    if (lane == 0) {
      s_red[a][wave] = x;
      s_red[3 + a][wave] = y;
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    float v = s_red[a][0];
    for (int w = 1; w < kBuildBlock / 64; ++w) v = a < 3 ? fminf(v, s_red[a][w]) : fmaxf(v, s_red[a][w]);
    if (a < 3)
      atomicMin(dst + a, f2o(v));
    else
      atomicMax(dst + a, f2o(v));
  }
}

// Zeroes the sort scratch and initialises the per-mesh boxes (one launch).
__global__ void k_segInit(uint32_t *__restrict__ scratch, size_t scratchWords, uint32_t *__restrict__ boxOrd,
                          size_t boxWords) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < scratchWords + boxWords;
       i += (size_t)gridDim.x * blockDim.x) {
    if (i < scratchWords)
      scratch[i] = 0;
    else
      boxOrd[i - scratchWords] = boxInitWord((uint32_t)((i - scratchWords) % kBoxWords));
  }
}

// Min / max of a run of up to kVertChunk vertices of one mesh (vchunks:
// mesh, first vertex, count) into the mesh's first box.
__global__ __launch_bounds__(kBuildBlock) void k_segVertBox(const bvht_vec3 *__restrict__ vertAll,
                                                            const uint4 *__restrict__ vchunks,
                                                            uint32_t *__restrict__ boxOrd) {
  const uint4 c = vchunks[blockIdx.x];
  float mn[3] = {HUGE_VALF, HUGE_VALF, HUGE_VALF}, mx[3] = {-HUGE_VALF, -HUGE_VALF, -HUGE_VALF};
#pragma unroll 4
  for (uint32_t i = threadIdx.x; i < c.z; i += kBuildBlock) {
    const bvht_vec3 p = vertAll[c.y + i];
    mn[0] = fminf(mn[0], p.x);
    mn[1] = fminf(mn[1], p.y);
    mn[2] = fminf(mn[2], p.z);
    mx[0] = fmaxf(mx[0], p.x);
    mx[1] = fmaxf(mx[1], p.y);
    mx[2] = fmaxf(mx[2], p.z);
  }
  blockBoxAtomics(mn, mx, boxOrd + (size_t)kBoxWords * c.x);
}

constexpr int kSegItems = 15;                               // keys per thread of a sort tile
constexpr uint32_t kSegTile = kSortBlock * kSegItems;       // pairs per sort tile
constexpr int kSegWideItems = 11;                           // keys per thread of a wide sort tile
constexpr uint32_t kSegWideTile = kWideBlock * kSegWideItems;
// triangles per key-pass workgroup: 2^11, two trips of four per thread (512
// meshes of 38.4M triangles in all: 2^15 left most of the GPU idle in the
// last round of workgroups, 2.48 ms per build against 2.34)
constexpr int kChunkLog = 11;
constexpr int kSegMaxHist = 3 * 1024;                       // passes x radix, both digit widths
constexpr uint32_t kSegFlagAgg = 1u << 30, kSegFlagInc = 2u << 30, kSegCountMask = (1u << 30) - 1;

__host__ __device__ constexpr int segPasses(int radixBits) { return (30 + radixBits - 1) / radixBits; }

// Morton keys of one chunk (up to 2^kChunkLog triangles of one mesh) and
// the mesh's histograms of every pass (LDS, then one atomic per non-empty
// bin).  Four triangles per thread per trip, loads issued before use.  The
// triangle's leaf box (the tile build's, from the same vertices) goes to
// rec, three float2 per triangle, so that the tile build gathers one
// record per leaf instead of a face and three vertices.  A wave's 64
// records are contiguous: they pass through 1.5 KB of LDS and are written
// as 8-byte words by consecutive lanes (512 contiguous bytes per store
// instruction instead of 8 bytes every 24).  Trips are wave-uniform, so the
// exchange is inside the wave: its LDS operations complete in order, and
// compiler barriers keep the stores and loads where they are.
template <int RB>
__global__ __launch_bounds__(kBuildBlock) void k_segMorton(const bvht_vec3 *__restrict__ vertAll,
                                                           const bvht_Triangle *__restrict__ faces,
                                                           const SegDesc *__restrict__ segs,
                                                           const uint2 *__restrict__ chunks, uint32_t *boxOrd,
                                                           uint32_t *__restrict__ keys, float2 *__restrict__ rec,
                                                           uint32_t *__restrict__ segHist) {
  constexpr int R = 1 << RB, P = segPasses(RB);
  static_assert(P * R <= kSegMaxHist, "histogram fits the LDS array");
  __shared__ uint32_t s_hist[P * R];
  for (int i = threadIdx.x; i < P * R; i += kBuildBlock) s_hist[i] = 0;
  const uint2 c = chunks[blockIdx.x];
  const SegDesc sd = segs[c.x];
  bvht_vec3 off, sc;
  {
    float m[6];
    shiftedMeshBox(boxOrd + (size_t)kBoxWords * c.x, m);
    // LBVH_calcMortonCodes host part (LBVH_CUDA.cu:171-180), f32, IEEE division
    off = {0.0f - m[0], 0.0f - m[1], 0.0f - m[2]};
    sc = {__fdiv_rn(1.0f, m[3] + off.x), __fdiv_rn(1.0f, m[4] + off.y), __fdiv_rn(1.0f, m[5] + off.z)};
  }
  __syncthreads();
  // min / max of the referenced vertices (a lane past the chunk repeats triangle p0)
  float rmn[3] = {HUGE_VALF, HUGE_VALF, HUGE_VALF}, rmx[3] = {-HUGE_VALF, -HUGE_VALF, -HUGE_VALF};
  const bvht_vec3 *vert = vertAll + sd.vertBase;
  const bvht_Triangle *fc = faces + sd.faceBase;
  const uint32_t p0 = c.y << kChunkLog, p1 = min(p0 + (1u << kChunkLog), sd.n);
  constexpr int U = 4;
  __shared__ float2 s_rec[kBuildBlock / 64][64 * 3];
  float2 *wrec = s_rec[threadIdx.x >> 6];
  const uint32_t lane = threadIdx.x & 63;
  // trips are wave-uniform (the histogram adds use ballots)
  for (uint32_t p = p0 + threadIdx.x; p - (threadIdx.x & 63) < p1; p += U * kBuildBlock) {
    bvht_Triangle tri[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t q = p + u * kBuildBlock;
      tri[u] = fc[q < p1 ? q : p0];  // p0 < p1: a chunk is never empty
    }
    bvht_vec3 va[U], vb[U], vc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      va[u] = vert[tri[u].v1];
      vb[u] = vert[tri[u].v2];
      vc[u] = vert[tri[u].v3];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t q = p + u * kBuildBlock;
      const bool valid = q < p1;
      rmn[0] = fminf(rmn[0], fminf(fminf(va[u].x, vb[u].x), vc[u].x));
      rmn[1] = fminf(rmn[1], fminf(fminf(va[u].y, vb[u].y), vc[u].y));
      rmn[2] = fminf(rmn[2], fminf(fminf(va[u].z, vb[u].z), vc[u].z));
      rmx[0] = fmaxf(rmx[0], fmaxf(fmaxf(va[u].x, vb[u].x), vc[u].x));
      rmx[1] = fmaxf(rmx[1], fmaxf(fmaxf(va[u].y, vb[u].y), vc[u].y));
      rmx[2] = fmaxf(rmx[2], fmaxf(fmaxf(va[u].z, vb[u].z), vc[u].z));
      const uint32_t key = morton30(centroid(va[u], vb[u], vc[u]), off, sc);
      if (valid) keys[sd.faceBase + q] = key;
      {
        const bvht_AABB b = triangleBox(va[u], vb[u], vc[u]);
        wrec[3 * lane] = make_float2(b.minMax[0].x, b.minMax[0].y);
        wrec[3 * lane + 1] = make_float2(b.minMax[0].z, b.minMax[1].x);
        wrec[3 * lane + 2] = make_float2(b.minMax[1].y, b.minMax[1].z);
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        // the wave's valid triangles are its first m lanes (q grows with the lane)
        const uint32_t m = (uint32_t)__popcll(__ballot(valid));
        float2 *dst = rec + 3 * (size_t)(sd.faceBase + q - lane);
#pragma unroll
        for (uint32_t j = lane; j < 3 * 64; j += 64)
          if (j < 3 * m) dst[j] = wrec[j];
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
      }
#pragma unroll
      for (int pass = 0; pass < P; ++pass) waveHistAdd(s_hist + pass * R, (key >> (pass * RB)) & (R - 1), valid);
    }
  }
  __syncthreads();
  uint32_t *h = segHist + (size_t)c.x * (P * R);
  for (int i = threadIdx.x; i < P * R; i += kBuildBlock) {
    const uint32_t n = s_hist[i];
    if (n) atomicAdd(h + i, n);
  }
  blockBoxAtomics(rmn, rmx, boxOrd + (size_t)kBoxWords * c.x + 6);
}

// Compares a chunk's mesh's two boxes (shifted: the boxes the keys were
// made from and the reference's); when they differ (an unreferenced vertex
// outside the triangles' box), recomputes the chunk's keys with the
// triangles' box and corrects the mesh's histograms (old digits out, new
// in).  Otherwise it exits after the comparison.  The leaf-box records do
// not depend on the box and stay.
template <int RB>
__global__ __launch_bounds__(kBuildBlock) void k_segFix(const bvht_vec3 *__restrict__ vertAll,
                                                        const bvht_Triangle *__restrict__ faces,
                                                        const SegDesc *__restrict__ segs,
                                                        const uint2 *__restrict__ chunks,
                                                        const uint32_t *__restrict__ boxOrd,
                                                        uint32_t *__restrict__ keys, uint32_t *__restrict__ segHist) {
  constexpr int R = 1 << RB, P = segPasses(RB);
  const uint2 c = chunks[blockIdx.x];
  float vb[6], tb[6];
  shiftedMeshBox(boxOrd + (size_t)kBoxWords * c.x, vb);
  shiftedMeshBox(boxOrd + (size_t)kBoxWords * c.x + 6, tb);
  bool same = true;
  for (int a = 0; a < 6; ++a) same = same && __float_as_uint(vb[a]) == __float_as_uint(tb[a]);
  if (same) return;  // block-uniform
  __shared__ uint32_t s_hist[P * R];  // net change per bin (mod 2^32)
  for (int i = threadIdx.x; i < P * R; i += kBuildBlock) s_hist[i] = 0;
  __syncthreads();
  const SegDesc sd = segs[c.x];
  const bvht_vec3 off = {0.0f - tb[0], 0.0f - tb[1], 0.0f - tb[2]};
  const bvht_vec3 sc = {__fdiv_rn(1.0f, tb[3] + off.x), __fdiv_rn(1.0f, tb[4] + off.y), __fdiv_rn(1.0f, tb[5] + off.z)};
  const bvht_vec3 *vert = vertAll + sd.vertBase;
  const uint32_t p0 = c.y << kChunkLog, p1 = min(p0 + (1u << kChunkLog), sd.n);
  for (uint32_t q = p0 + threadIdx.x; q < p1; q += kBuildBlock) {
    const bvht_Triangle t = faces[sd.faceBase + q];
    const uint32_t key = morton30(centroid(vert[t.v1], vert[t.v2], vert[t.v3]), off, sc);
    const uint32_t old = keys[sd.faceBase + q];
    keys[sd.faceBase + q] = key;
#pragma unroll
    for (int pass = 0; pass < P; ++pass) {
      atomicAdd(&s_hist[pass * R + ((old >> (pass * RB)) & (R - 1))], 0xFFFFFFFFu);
      atomicAdd(&s_hist[pass * R + ((key >> (pass * RB)) & (R - 1))], 1u);
    }
  }
  __syncthreads();
  uint32_t *h = segHist + (size_t)c.x * (P * R);
  for (int i = threadIdx.x; i < P * R; i += kBuildBlock) {
    const uint32_t d = s_hist[i];
    if (d) atomicAdd(h + i, d);
  }
}

// Exclusive scan of one (mesh, pass) histogram row, in place.
template <int RB>
__global__ __launch_bounds__(kSortBlock) void k_segPrefix(uint32_t *__restrict__ segHist) {
  constexpr int R = 1 << RB, D = R / kSortBlock;
  __shared__ uint32_t s_scan[kSortBlock / 64];
  uint32_t *row = segHist + (size_t)blockIdx.x * R + threadIdx.x * D;
  uint32_t v[D], sum = 0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    v[j] = row[j];
    sum += v[j];
  }
  uint32_t run = blockExclusiveScan<kSortBlock>(sum, s_scan);
#pragma unroll
  for (int j = 0; j < D; ++j) {
    row[j] = run;
    run += v[j];
  }
}

// One pass of the segmented onesweep: stable by (mesh, digit).  Sort tiles
// (tileTab: mesh, tile within the mesh) are taken in order from `counter`;
// a tile ranks its keys with RB wave ballots, publishes per-digit counts
// (flag | count in one u32), looks back only over earlier tiles of its own
// mesh, and scatters through LDS.  Thread t owns digits [t D, t D + D).
template <int RB>
__global__ __launch_bounds__(kSortBlock) void k_segOnesweep(const uint32_t *__restrict__ keysIn,
                                                            const uint32_t *__restrict__ valsIn,
                                                            uint32_t *__restrict__ keysOut, uint32_t *__restrict__ valsOut,
                                                            const uint2 *__restrict__ tileTab,
                                                            const SegDesc *__restrict__ segs, uint32_t shift,
                                                            const uint32_t *__restrict__ prefix, int passes,
                                                            uint32_t *status, uint32_t *counter,
                                                            uint32_t *err) {
  constexpr int R = 1 << RB, D = R / kSortBlock, WAVES = kSortBlock / 64, TILE = (int)kSegTile;
  __shared__ uint32_t s_wave[WAVES][R];
  __shared__ uint32_t s_bstart[R], s_gstart[R];
  __shared__ uint32_t s_scan[WAVES];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_keys[TILE], s_vals[TILE];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (tid == 0) s_tile = atomicAdd(counter, 1u);
  for (int i = tid; i < WAVES * R; i += kSortBlock) (&s_wave[0][0])[i] = 0;
  __syncthreads();
  const uint32_t g = s_tile;
  const uint2 tt = tileTab[g];
  const SegDesc sd = segs[tt.x];
  const uint32_t lt = tt.y, segStart = sd.faceBase, segN = sd.n;
  const uint32_t base = lt * TILE + wave * 64 * kSegItems;

  uint32_t key[kSegItems], val[kSegItems], rank[kSegItems];
#pragma unroll
  for (int i = 0; i < kSegItems; ++i) {
    const uint32_t idx = base + i * 64 + lane;
    const bool valid = idx < segN;
    key[i] = valid ? keysIn[segStart + idx] : 0u;
    val[i] = valid ? (valsIn ? valsIn[segStart + idx] : segStart + idx) : 0u;
  }
  // stable in-wave ranking (see k_onesweep): item-major order inside a
  // wave's segment is tile order
  uint32_t *cnt = &s_wave[wave][0];
  const uint64_t lt64 = __lanemask_lt();
#pragma unroll
  for (int i = 0; i < kSegItems; ++i) {
    const bool valid = base + i * 64 + lane < segN;
    const uint32_t d = (key[i] >> shift) & (R - 1);
    const uint64_t peers = matchDigit<RB>(d, __ballot(valid));
    const uint32_t lower = (uint32_t)__popcll(peers & lt64);
    const uint32_t c = __hip_atomic_load(cnt + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    rank[i] = c + lower;
    if (valid && lower == 0)
      __hip_atomic_store(cnt + d, c + (uint32_t)__popcll(peers), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();

  // per owned digit: per-wave exclusive offsets and the tile total
  uint32_t tot[D], sum = 0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const uint32_t d = tid * D + j;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
      const uint32_t c = s_wave[w][d];
      s_wave[w][d] = run;
      run += c;
    }
    tot[j] = run;
    sum += run;
  }
  uint32_t *st = status + (size_t)g * R + tid * D;
#pragma unroll
  for (int j = 0; j < D; ++j) storeRelaxedAgent(st + j, (lt == 0 ? kSegFlagInc : kSegFlagAgg) | tot[j]);
  const uint32_t bex = blockExclusiveScan<kSortBlock>(sum, s_scan);
  uint32_t excl[D];
#pragma unroll
  for (int j = 0; j < D; ++j) excl[j] = 0;
  if (lt > 0) {
    // walk back over the mesh's earlier tiles; its first tile publishes an
    // inclusive count at once, so every walk ends at or before it
    uint32_t t = g - 1, spins = 0, doneMask = 0;
    while (doneMask != (1u << D) - 1) {
      uint32_t w[D];
      bool ready = true;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        w[j] = loadRelaxedAgent(status + (size_t)t * R + tid * D + j);
        if (!(doneMask & (1u << j)) && (w[j] & ~kSegCountMask) == 0) ready = false;
      }
      if (!ready) {
        if (++spins > (1u << 24)) {  // a predecessor that never publishes: error flag instead of a hang
          atomicOr(err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
#pragma unroll
      for (int j = 0; j < D; ++j) {
        if (doneMask & (1u << j)) continue;
        excl[j] += w[j] & kSegCountMask;
        if ((w[j] & ~kSegCountMask) == kSegFlagInc) doneMask |= 1u << j;
      }
      --t;
    }
#pragma unroll
    for (int j = 0; j < D; ++j) storeRelaxedAgent(st + j, kSegFlagInc | (excl[j] + tot[j]));
  }
  const uint32_t *pre = prefix + (size_t)tt.x * passes * R + tid * D;
  uint32_t run = bex;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    s_bstart[tid * D + j] = run;
    s_gstart[tid * D + j] = segStart + pre[j] + excl[j];
    run += tot[j];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSegItems; ++i) {
    if (base + i * 64 + lane < segN) {
      const uint32_t d = (key[i] >> shift) & (R - 1);
      const uint32_t pos = s_bstart[d] + s_wave[wave][d] + rank[i];
      s_keys[pos] = key[i];
      s_vals[pos] = val[i];
    }
  }
  __syncthreads();
  const uint32_t tileCount = min((uint32_t)TILE, segN - lt * TILE);
  const uint32_t segEnd = segStart + segN;
  for (uint32_t j = tid; j < tileCount; j += kSortBlock) {
    const uint32_t k = s_keys[j];
    const uint32_t d = (k >> shift) & (R - 1);
    const uint32_t out = s_gstart[d] + (j - s_bstart[d]);
    if (out >= segEnd) continue;  // only reachable after a bounded-wait error
    if (keysOut) keysOut[out] = k;
    valsOut[out] = s_vals[j];
  }
}

// One pass of the segmented onesweep with 8-bit digits over wide tiles
// (kSegWideTile keys; two workgroups per CU: 8 waves per SIMD, at most 64
// VGPRs, 63 KB of LDS each), for batches whose key array
// stays below 2^31 bytes (buffer accesses with 32-bit offsets; a load past
// the end reads 0).  Same order as k_segOnesweep: stable by (mesh, digit).
// Look-back in rounds of 16 predecessors of the same mesh: thread (g, q)
// loads digits 4q .. 4q + 3 of the predecessor at distance r + g + 1 (one
// 16-byte coherent load), waits only until that word is published
// (aggregate or inclusive), and each digit's owner adds the counts back to
// the nearest inclusive one.  No tile waits for another's look-back, so the
// walk is one round unless a digit's nearest inclusive count lies more than
// 16 tiles back; a mesh's first tile publishes an inclusive count at once.
__global__ __launch_bounds__(kWideBlock, 8) void k_segOnesweepWide(
    const uint32_t *__restrict__ keysIn, const uint32_t *__restrict__ valsIn, uint32_t *__restrict__ keysOut,
    uint32_t *__restrict__ valsOut, const uint2 *__restrict__ tileTab, const SegDesc *__restrict__ segs,
    uint32_t total, uint32_t shift, const uint32_t *__restrict__ prefix, int passes, uint32_t *status,
    uint32_t *counter, uint32_t *err) {
  constexpr int WAVES = kWideBlock / 64, ITEMS = kSegWideItems, R = kSortRadix;
  constexpr uint32_t TILE = kSegWideTile, QUADS = R / 4, NONE = 0x80000000u;
  static_assert(kWideBlock / QUADS == WAVES, "one look-back group per wave");
  __shared__ uint32_t s_wave[WAVES][R];  // per-wave digit counters, then look-back words
  __shared__ uint32_t s_bstart[R], s_gstart[R];
  __shared__ uint32_t s_scan[WAVES];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_buf[TILE];  // the staged keys, then the values

  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (tid == 0) s_tile = atomicAdd(counter, 1u);
  for (uint32_t i = tid; i < (uint32_t)WAVES * R; i += kWideBlock) (&s_wave[0][0])[i] = 0;
  __syncthreads();
  const uint32_t g = s_tile;
  const uint2 tt = tileTab[g];
  const SegDesc sd = segs[tt.x];
  const uint32_t lt = tt.y, segStart = sd.faceBase, segN = sd.n;
  const uint32_t first = lt * TILE + wave * 64 * ITEMS + lane;  // mesh-local index of item 0
  const uint32_t meshPrefix = tid < R ? prefix[(size_t)tt.x * passes * R + tid] : 0u;  // issued now, used last
  const BufferRsrc keys = bufferRsrc(keysIn, total * 4u);
  uint32_t kw[ITEMS], rank[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t idx = first + i * 64;
    kw[i] = bufferLoadU32(keys, idx < segN ? (segStart + idx) * 4u : NONE);
  }
  uint32_t *cnt = &s_wave[wave][0];
  const uint64_t lt64 = __lanemask_lt();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const bool valid = first + i * 64 < segN;
    const uint32_t d = (kw[i] >> shift) & (R - 1);
    const uint64_t peers = matchDigit<kSortRadixBits>(d, __ballot(valid));
    const uint32_t lower = (uint32_t)__popcll(peers & lt64);
    const uint32_t c = __hip_atomic_load(cnt + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    rank[i] = c + lower;
    if (valid && lower == 0)
      __hip_atomic_store(cnt + d, c + (uint32_t)__popcll(peers), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();

  // digit owners (tid < R): per-wave exclusive offsets and the tile's count,
  // published at once (inclusive for the mesh's first tile)
  uint32_t count = 0;
  if (tid < R) {
    for (int w = 0; w < WAVES; ++w) {
      const uint32_t c = s_wave[w][tid];
      s_wave[w][tid] = count;
      count += c;
    }
    storeRelaxedAgent(status + (size_t)g * R + tid, (lt == 0 ? kSegFlagInc : kSegFlagAgg) | count);
  }
  const uint32_t bstart = blockExclusiveScan<kWideBlock>(count, s_scan);
  if (tid < R) s_bstart[tid] = bstart;
  __syncthreads();

  // keys to their tile positions (digits and item indices recomputed)
  uint32_t first2 = first;
  recomputeAfter(first2);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) recomputeAfter(kw[i]);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (first2 + i * 64 < segN) {
      const uint32_t d = (kw[i] >> shift) & (R - 1);
      rank[i] += s_bstart[d] + s_wave[wave][d];
      s_buf[rank[i]] = kw[i];
    }
  }
  __syncthreads();  // s_wave is free from here: it holds the look-back words

  uint32_t excl = 0;
  if (lt > 0) {  // block-uniform
    const uint32_t q = tid % QUADS, grp = tid / QUADS;
    bool done = false;
    uint32_t spins = 0;
    for (uint32_t r0 = 0;; r0 += WAVES) {
      const uint32_t dist = r0 + grp + 1;
      BufferU32x4 v = BufferU32x4{kSegFlagInc, kSegFlagInc, kSegFlagInc, kSegFlagInc};  // before the mesh: 0
      if (dist <= lt) {
        const BufferRsrc st = bufferRsrc(status + (size_t)(g - dist) * R, R * 4u);
        v = bufferLoadU32x4Coherent(st, 16u * q);
        // a predecessor publishes right after its ranking; the wait is bounded
        while ((v[0] & ~kSegCountMask) == 0 || (v[1] & ~kSegCountMask) == 0 || (v[2] & ~kSegCountMask) == 0 ||
               (v[3] & ~kSegCountMask) == 0) {
          if (++spins > (1u << 24)) {
            atomicOr(err, 1u);
            v = BufferU32x4{kSegFlagInc, kSegFlagInc, kSegFlagInc, kSegFlagInc};
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          memoryChanged();
          v = bufferLoadU32x4Coherent(st, 16u * q);
        }
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) s_wave[grp][4 * q + c] = v[c];
      __syncthreads();
      if (tid < R && !done) {
        for (int w = 0; w < WAVES; ++w) {
          const uint32_t x = s_wave[w][tid];
          excl += x & kSegCountMask;
          if ((x & ~kSegCountMask) == kSegFlagInc) {
            done = true;
            break;
          }
        }
      }
      // the barrier also keeps the next round's words from overwriting these
      if (!__syncthreads_or(tid < R && !done)) break;
    }
  }
  if (tid < R) {
    if (lt > 0) storeRelaxedAgent(status + (size_t)g * R + tid, kSegFlagInc | (excl + count));
    s_gstart[tid] = segStart + meshPrefix + excl;
  }
  __syncthreads();

  // the destination of tile position j is s_gstart[d] + j - s_bstart[d];
  // the digits of a thread's positions are kept four to a register
  const uint32_t tileCount = min(TILE, segN - lt * TILE), segEnd = segStart + segN;
  const BufferRsrc kout = bufferRsrc(keysOut, keysOut ? total * 4u : 0u);
  const BufferRsrc vout = bufferRsrc(valsOut, total * 4u);
  auto dest = [&](uint32_t j, uint32_t d) {
    const uint32_t o = s_gstart[d] + (j - s_bstart[d]);
    return j < tileCount && o < segEnd ? o * 4u : NONE;  // past the mesh only after a bounded-wait error
  };
  uint32_t dig[(ITEMS + 3) / 4] = {};
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t j = tid + k * kWideBlock;
    const uint32_t kk = s_buf[j < tileCount ? j : 0];
    const uint32_t d = (kk >> shift) & (R - 1);
    dig[k / 4] |= d << (8 * (k % 4));
    if (keysOut) bufferStore(kk, kout, dest(j, d));
  }
  // values: loaded only now (the other workgroup on the CU covers the
  // latency), so that they are not live beside the ranks during the keys'
  // look-back and scatter
  uint32_t first3 = first;
  recomputeAfter(first3);
  uint32_t val[ITEMS];
  if (valsIn) {
    const BufferRsrc vals = bufferRsrc(valsIn, total * 4u);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t idx = first3 + i * 64;
      val[i] = bufferLoadU32(vals, idx < segN ? (segStart + idx) * 4u : NONE);
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) val[i] = segStart + first3 + i * 64;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if (first3 + i * 64 < segN) s_buf[rank[i]] = val[i];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t j = tid + k * kWideBlock;
    bufferStore(s_buf[j < tileCount ? j : 0], vout, dest(j, (dig[k / 4] >> (8 * (k % 4))) & (R - 1)));
  }
}

struct BatchInternal {
  uint32_t M = 0, totalFaces = 0, totalNodes = 0, totalVert = 0;
  int tlog = 10;
  std::vector<SegDesc> segsHost;
  SegDesc *segs = nullptr;
  std::vector<uint2 *> tiles;
  std::vector<uint32_t> tileCounts;
  uint32_t *boxOrd = nullptr;
  uint32_t *keys[2] = {nullptr, nullptr};
  uint32_t *vals[2] = {nullptr, nullptr};
  float2 *rec = nullptr;  // leaf box per triangle (three float2), written by k_segMorton
  int radixBits = 8;
  uint2 *chunks = nullptr;     // (mesh, chunk) of the Morton pass
  uint32_t numChunks = 0;
  uint4 *vchunks = nullptr;    // (mesh, first vertex, count) of the vertex box pass
  uint32_t numVChunks = 0;
  uint2 *sortTiles = nullptr;  // (mesh, tile) of the sort passes
  uint32_t numSortTiles = 0;
  uint2 *wideTiles = nullptr;  // (mesh, tile) of the wide 8-bit passes
  uint32_t numWideTiles = 0;
  // sort scratch: [0, 16) pass counters + error word, then the mesh
  // histograms / prefixes (M x passes x radix), then the look-back status
  // (passes x sort tiles x radix); all of it is zeroed before every build
  uint32_t *sortScratch = nullptr;
  size_t histWords = 0, statusWords = 0;
  hipStream_t capStream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *capVert = nullptr, *capFaces = nullptr, *capNodes = nullptr;
};

// Wide 8-bit passes apply while the key array stays below 2^31 bytes.
bool useWideSegSort(const BatchInternal *in) { return in->radixBits == 8 && in->totalFaces < (1u << 29); }

template <int RB>
bool segSort(BatchInternal *in, int &result) {
  constexpr int R = 1 << RB, P = segPasses(RB);
  uint32_t *counters = in->sortScratch, *hist = in->sortScratch + 16;
  uint32_t *status = hist + (size_t)in->M * P * R;
  k_segPrefix<RB><<<in->M * P, kSortBlock, 0, stream()>>>(hist);
  if (!checkLaunch("k_segPrefix")) return false;
  int cur = 0;
  for (int pass = 0; pass < P; ++pass) {
    const bool last = pass == P - 1;
    if (RB == 8 && useWideSegSort(in)) {
      k_segOnesweepWide<<<in->numWideTiles, kWideBlock, 0, stream()>>>(
          in->keys[cur], pass == 0 ? nullptr : in->vals[cur], last ? nullptr : in->keys[cur ^ 1], in->vals[cur ^ 1],
          in->wideTiles, in->segs, in->totalFaces, (uint32_t)(pass * RB), hist + (size_t)pass * R, P,
          status + (size_t)pass * in->numWideTiles * R, counters + pass, counters + 15);
      if (!checkLaunch("k_segOnesweepWide")) return false;
      cur ^= 1;
      continue;
    }
    k_segOnesweep<RB><<<in->numSortTiles, kSortBlock, 0, stream()>>>(
        in->keys[cur], pass == 0 ? nullptr : in->vals[cur], last ? nullptr : in->keys[cur ^ 1], in->vals[cur ^ 1],
        in->sortTiles, in->segs, (uint32_t)(pass * RB), hist + (size_t)pass * R, P,
        status + (size_t)pass * in->numSortTiles * R, counters + pass, counters + 15);
    if (!checkLaunch("k_segOnesweep")) return false;
    cur ^= 1;
  }
  result = cur;
  return true;
}

template <int TLOG>
bool batchStagesT(BatchInternal *in, const bvht_vec3 *vert, const bvht_Triangle *faces, bvht_BVHNode *nodes) {
  const int P = segPasses(in->radixBits), R = 1 << in->radixBits;
  const uint32_t sortTiles = useWideSegSort(in) ? in->numWideTiles : in->numSortTiles;
  const size_t scratchWords = 16 + (size_t)in->M * P * R + (size_t)P * sortTiles * R;
  const size_t boxWords = (size_t)kBoxWords * in->M;
  const uint32_t initBlocks = (uint32_t)std::min<size_t>(divUp(scratchWords + boxWords, 256), (size_t)numCUs() * 8);
  k_segInit<<<initBlocks, 256, 0, stream()>>>(in->sortScratch, scratchWords, in->boxOrd, boxWords);
  if (!checkLaunch("k_segInit")) return false;
  if (in->numVChunks) {
    k_segVertBox<<<in->numVChunks, kBuildBlock, 0, stream()>>>(vert, in->vchunks, in->boxOrd);
    if (!checkLaunch("k_segVertBox")) return false;
  }
  int r = 0;
  if (in->radixBits == 10) {
    k_segMorton<10><<<in->numChunks, kBuildBlock, 0, stream()>>>(vert, faces, in->segs, in->chunks, in->boxOrd,
                                                                 in->keys[0], in->rec, in->sortScratch + 16);
    if (!checkLaunch("k_segMorton")) return false;
    k_segFix<10><<<in->numChunks, kBuildBlock, 0, stream()>>>(vert, faces, in->segs, in->chunks, in->boxOrd,
                                                              in->keys[0], in->sortScratch + 16);
    if (!checkLaunch("k_segFix") || !segSort<10>(in, r)) return false;
  } else {
    k_segMorton<8><<<in->numChunks, kBuildBlock, 0, stream()>>>(vert, faces, in->segs, in->chunks, in->boxOrd,
                                                                in->keys[0], in->rec, in->sortScratch + 16);
    if (!checkLaunch("k_segMorton")) return false;
    k_segFix<8><<<in->numChunks, kBuildBlock, 0, stream()>>>(vert, faces, in->segs, in->chunks, in->boxOrd,
                                                             in->keys[0], in->sortScratch + 16);
    if (!checkLaunch("k_segFix") || !segSort<8>(in, r)) return false;
  }
  return launchTreeLevelsBatch(TLOG, nodes, in->vals[r], faces, vert, in->rec, in->segs, in->tiles.data(),
                               in->tileCounts.data(), (int)in->tiles.size());
}

bool batchStages(BatchInternal *in, const bvht_vec3 *vert, const bvht_Triangle *faces, bvht_BVHNode *nodes) {
  switch (in->tlog) {
    case 8: return batchStagesT<8>(in, vert, faces, nodes);
    case 9: return batchStagesT<9>(in, vert, faces, nodes);
    case 10: return batchStagesT<10>(in, vert, faces, nodes);
    case 11: return batchStagesT<11>(in, vert, faces, nodes);
  }
  setError("bvht_LBVH_buildBatch: unsupported tile log2 %d", in->tlog);
  return false;
}

void freeTiles(BatchInternal *in) {
  for (auto *t : in->tiles) devFree(t);
  in->tiles.clear();
  in->tileCounts.clear();
}

// Tile tables per level for tile log2 `tlog`: level 0 lists every mesh,
// level L >= 1 only the meshes whose previous level left more than one tile
// root.  Every (mesh, group) pair satisfies group << tlog < units of the level.
bool buildTiles(BatchInternal *in, int tlog) {
  freeTiles(in);
  for (int L = 0;; ++L) {
    std::vector<uint2> t;
    const uint32_t unitLog = (uint32_t)(L * tlog);
    bool more = false;
    for (uint32_t i = 0; i < in->M; ++i) {
      const uint64_t n = in->segsHost[i].n;
      const uint64_t U = unitLog >= 40 ? 1 : (n + (1ull << unitLog) - 1) >> unitLog;
      if (L > 0 && U <= 1) continue;
      const uint64_t G = (U + (1ull << tlog) - 1) >> tlog;
      for (uint64_t g = 0; g < G; ++g) t.push_back(make_uint2(i, (uint32_t)g));
      if (G > 1) more = true;
    }
    uint2 *d = nullptr;
    if (!t.empty()) {
      if (!devAlloc(&d, t.size(), "batch tiles") ||
          !check(hipMemcpy(d, t.data(), sizeof(uint2) * t.size(), hipMemcpyHostToDevice), "batch tiles")) {
        devFree(d);
        return false;
      }
    }
    in->tiles.push_back(d);
    in->tileCounts.push_back((uint32_t)t.size());
    if (!more) break;
  }
  in->tlog = tlog;
  return true;
}

void dropGraph(BatchInternal *in) {
  if (in->exec) (void)hipGraphExecDestroy(in->exec);
  if (in->graph) (void)hipGraphDestroy(in->graph);
  in->exec = nullptr;
  in->graph = nullptr;
}

}  // namespace

}  // namespace bvht

using namespace bvht;

extern "C" {

bvht_LBVH_BatchWorkingMemory bvht_LBVH_allocateBatch(uint32_t numMeshes, const uint32_t *faceCounts,
                                                     const uint32_t *vertCounts, uint32_t tileLog2) {
  bvht_LBVH_BatchWorkingMemory m;
  std::memset(&m, 0, sizeof(m));
  if (numMeshes == 0 || !faceCounts || !vertCounts) {
    setError("bvht_LBVH_allocateBatch: no meshes");
    return m;
  }
  if (numMeshes > (1u << 24)) {
    setError("bvht_LBVH_allocateBatch: at most 2^24 meshes");
    return m;
  }
  if (tileLog2 == 0) tileLog2 = 10;
  if (tileLog2 < 8 || tileLog2 > 11) {
    setError("bvht_LBVH_allocateBatch: tileLog2 must be in [8, 11]");
    return m;
  }
  BatchInternal *in = new (std::nothrow) BatchInternal();
  if (!in) {
    setError("bvht_LBVH_allocateBatch: host allocation failed");
    return m;
  }
  m.internal = in;
  in->M = numMeshes;
  in->tlog = (int)tileLog2;
  uint64_t fo = 0, vo = 0, no = 0;
  for (uint32_t i = 0; i < numMeshes; ++i) {
    if (faceCounts[i] == 0 || faceCounts[i] >= (1u << 30)) {  // sort counts carry 30 bits
      setError("bvht_LBVH_allocateBatch: mesh %u has %u faces", i, faceCounts[i]);
      bvht_LBVH_freeBatch(&m);
      return m;
    }
    in->segsHost.push_back(SegDesc{faceCounts[i], (uint32_t)fo, (uint32_t)no, (uint32_t)vo});
    fo += faceCounts[i];
    vo += vertCounts[i];
    no += 2ull * faceCounts[i] - 1;
  }
  if (no >= (1ull << 32) || vo >= (1ull << 32)) {
    setError("bvht_LBVH_allocateBatch: batch exceeds 32-bit indices");
    bvht_LBVH_freeBatch(&m);
    return m;
  }
  in->totalFaces = (uint32_t)fo;
  in->totalVert = (uint32_t)vo;
  in->totalNodes = (uint32_t)no;
  if (!buildTiles(in, in->tlog)) {
    bvht_LBVH_freeBatch(&m);
    return m;
  }
  // Morton chunks and sort tiles: per mesh, never crossing one
  std::vector<uint2> chunks, sortTiles, wideTiles;
  std::vector<uint4> vchunks;
  for (uint32_t i = 0; i < numMeshes; ++i) {
    const uint32_t nv = vertCounts[i], vb = in->segsHost[i].vertBase;
    for (uint32_t c = 0; (uint64_t)c * kVertChunk < nv; ++c)
      vchunks.push_back(make_uint4(i, vb + c * kVertChunk, std::min(kVertChunk, nv - c * kVertChunk), 0u));
  }
  for (uint32_t i = 0; i < numMeshes; ++i) {
    const uint32_t n = faceCounts[i];
    for (uint32_t c = 0; (uint64_t)c << kChunkLog < n; ++c) chunks.push_back(make_uint2(i, c));
    for (uint32_t t = 0; (uint64_t)t * kSegTile < n; ++t) sortTiles.push_back(make_uint2(i, t));
    for (uint32_t t = 0; (uint64_t)t * kSegWideTile < n; ++t) wideTiles.push_back(make_uint2(i, t));
  }
  in->numChunks = (uint32_t)chunks.size();
  in->numVChunks = (uint32_t)vchunks.size();
  in->numSortTiles = (uint32_t)sortTiles.size();
  in->numWideTiles = (uint32_t)wideTiles.size();
  // scratch for the larger of the two digit widths (10 bits: 3 passes x 1024)
  in->histWords = (size_t)numMeshes * kSegMaxHist;
  in->statusWords = (size_t)3 * in->numSortTiles * 1024;
  static_assert(segPasses(8) * 256 <= kSegMaxHist && segPasses(10) * 1024 <= kSegMaxHist, "histogram rows");
  if (!devAlloc(&in->chunks, chunks.size(), "batch chunks") ||
      !check(hipMemcpy(in->chunks, chunks.data(), sizeof(uint2) * chunks.size(), hipMemcpyHostToDevice), "batch chunks") ||
      !devAlloc(&in->sortTiles, sortTiles.size(), "batch sort tiles") ||
      !check(hipMemcpy(in->sortTiles, sortTiles.data(), sizeof(uint2) * sortTiles.size(), hipMemcpyHostToDevice),
             "batch sort tiles") ||
      !devAlloc(&in->wideTiles, wideTiles.size(), "batch wide sort tiles") ||
      !devAlloc(&in->vchunks, std::max<size_t>(vchunks.size(), 1), "batch vertex chunks") ||
      !check(hipMemcpy(in->vchunks, vchunks.data(), sizeof(uint4) * vchunks.size(), hipMemcpyHostToDevice),
             "batch vertex chunks") ||
      !check(hipMemcpy(in->wideTiles, wideTiles.data(), sizeof(uint2) * wideTiles.size(), hipMemcpyHostToDevice),
             "batch wide sort tiles") ||
      !devAlloc(&in->segs, numMeshes, "batch segments") ||
      !check(hipMemcpy(in->segs, in->segsHost.data(), sizeof(SegDesc) * numMeshes, hipMemcpyHostToDevice),
             "batch segments") ||
      !devAlloc(&in->boxOrd, (size_t)kBoxWords * numMeshes, "batch boxes") || !devAlloc(&in->keys[0], in->totalFaces, "batch keys") ||
      !devAlloc(&in->keys[1], in->totalFaces, "batch keys") || !devAlloc(&in->vals[0], in->totalFaces, "batch values") ||
      !devAlloc(&in->vals[1], in->totalFaces, "batch values") ||
      !devAlloc(&in->rec, 3ull * in->totalFaces, "batch leaf boxes") ||
      !devAlloc(&in->sortScratch, 16 + in->histWords + in->statusWords, "batch sort") ||
      !fillU32(in->sortScratch, 0u, 16) ||  // the error word reads 0 before the first build
      !check(hipStreamCreateWithFlags(&in->capStream, hipStreamNonBlocking), "batch capture stream")) {
    bvht_LBVH_freeBatch(&m);
    return m;
  }
  m.numMeshes = numMeshes;
  m.totalFaces = in->totalFaces;
  m.totalVert = in->totalVert;
  m.totalNodes = in->totalNodes;
  m.tileLog2 = (uint32_t)in->tlog;
  m.radixBits = (uint32_t)in->radixBits;
  m.lRes = true;
  return m;
}

bool bvht_LBVH_buildBatch(bvht_LBVH_BatchWorkingMemory *mem, const bvht_vec3 *vert, const bvht_Triangle *faces,
                          bvht_BVHNode *nodes, bool useGraph) {
  if (!mem || !mem->lRes || !mem->internal) {
    setError("bvht_LBVH_buildBatch: working memory not allocated");
    return false;
  }
  if (!vert || !faces || !nodes) {
    setError("bvht_LBVH_buildBatch: missing vertices, faces or nodes");
    return false;
  }
  BatchInternal *in = reinterpret_cast<BatchInternal *>(mem->internal);
  if (mem->tileLog2 != (uint32_t)in->tlog) {  // changed by the caller / autotuner
    if (mem->tileLog2 < 8 || mem->tileLog2 > 11) {
      setError("bvht_LBVH_buildBatch: tileLog2 must be in [8, 11]");
      return false;
    }
    dropGraph(in);
    if (!buildTiles(in, (int)mem->tileLog2)) return false;  // tables are per tile size
  }
  if (mem->radixBits != (uint32_t)in->radixBits) {
    if (mem->radixBits != 8 && mem->radixBits != 10) {
      setError("bvht_LBVH_buildBatch: radixBits must be 8 or 10");
      return false;
    }
    dropGraph(in);
    in->radixBits = (int)mem->radixBits;
  }
  if (!useGraph) return batchStages(in, vert, faces, nodes);
  if (in->exec && (in->capVert != vert || in->capFaces != faces || in->capNodes != nodes)) dropGraph(in);
  if (!in->exec) {
    hipStream_t user = rt().stream;
    rt().stream = in->capStream;
    bool ok = check(hipStreamBeginCapture(in->capStream, hipStreamCaptureModeThreadLocal), "batch capture");
    if (ok) ok = batchStages(in, vert, faces, nodes);
    hipGraph_t g = nullptr;
    const bool ended = check(hipStreamEndCapture(in->capStream, &g), "batch capture end");
    rt().stream = user;
    if (!ok || !ended) {
      if (g) (void)hipGraphDestroy(g);
      return false;
    }
    in->graph = g;
    if (!check(hipGraphInstantiate(&in->exec, g, nullptr, nullptr, 0), "batch graph instantiate")) {
      dropGraph(in);
      return false;
    }
    in->capVert = vert;
    in->capFaces = faces;
    in->capNodes = nodes;
  }
  return check(hipGraphLaunch(in->exec, rt().stream), "batch graph launch");
}

void bvht_LBVH_freeBatch(bvht_LBVH_BatchWorkingMemory *mem) {
  if (!mem) return;
  BatchInternal *in = reinterpret_cast<BatchInternal *>(mem->internal);
  if (in) {
    dropGraph(in);
    if (in->capStream) (void)hipStreamDestroy(in->capStream);
    freeTiles(in);
    devFree(in->chunks);
    devFree(in->sortTiles);
    devFree(in->wideTiles);
    devFree(in->vchunks);
    devFree(in->segs);
    devFree(in->boxOrd);
    devFree(in->keys[0]);
    devFree(in->keys[1]);
    devFree(in->vals[0]);
    devFree(in->vals[1]);
    devFree(in->rec);
    devFree(in->sortScratch);
    delete in;
  }
  std::memset(mem, 0, sizeof(*mem));
}

// Error word of the last batched sort: bit 0 = a look-back wait gave up.
uint32_t bvht_LBVH_batchStatus(const bvht_LBVH_BatchWorkingMemory *mem) {
  if (!mem || !mem->internal) {
    setError("bvht_LBVH_batchStatus: working memory not allocated");
    return 0xFFFFFFFFu;
  }
  const BatchInternal *in = reinterpret_cast<const BatchInternal *>(mem->internal);
  uint32_t v = 0xFFFFFFFFu;
  if (!check(hipMemcpyAsync(&v, in->sortScratch + 15, 4, hipMemcpyDeviceToHost, stream()), "bvht_LBVH_batchStatus") ||
      !check(hipStreamSynchronize(stream()), "bvht_LBVH_batchStatus sync"))
    return 0xFFFFFFFFu;
  return v;
}

// Offsets of mesh i inside the concatenated arrays.
bool bvht_LBVH_batchOffsets(const bvht_LBVH_BatchWorkingMemory *mem, uint32_t i, uint32_t *faceOffset,
                            uint32_t *vertOffset, uint32_t *nodeOffset) {
  if (!mem || !mem->internal) return false;
  const BatchInternal *in = reinterpret_cast<const BatchInternal *>(mem->internal);
  if (i >= in->M) return false;
  if (faceOffset) *faceOffset = in->segsHost[i].faceBase;
  if (vertOffset) *vertOffset = in->segsHost[i].vertBase;
  if (nodeOffset) *nodeOffset = in->segsHost[i].nodeBase;
  return true;
}

}  // extern "C"

namespace bvht {
namespace {

// Mean device time (ms) of `reps` runs of fn on the library stream.
template <typename F>
float timeRuns(F fn, int reps) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return -1.0f;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return -1.0f;
  }
  fn();  // warm-up
  (void)hipEventRecord(a, stream());
  bool ok = true;
  for (int i = 0; i < reps; ++i) ok = ok && fn();
  (void)hipEventRecord(b, stream());
  float ms = -1.0f;
  if (ok && hipEventSynchronize(b) == hipSuccess) {
    (void)hipEventElapsedTime(&ms, a, b);
    ms /= reps;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

}  // namespace
}  // namespace bvht

extern "C" {

// Autotuner of the build tile size (results are identical for every
// choice): times tile log2 8..11 on the caller's own input and keeps the
// fastest in mem->opts.tileLog2.  msOut[4] receives the mean times.
int bvht_LBVH_autotune(bvht_LBVH_WorkingMemory *mem, const bvht_MeshRaw *mesh, bvht_CUDAMemoryBVHPointer *bvh,
                       float *msOut) {
  if (!mem || !mem->lRes) {
    setError("bvht_LBVH_autotune: working memory not allocated");
    return 0;
  }
  int best = (int)mem->opts.tileLog2;
  float bestMs = 1e30f;
  for (int t = 8; t <= 11; ++t) {
    mem->opts.tileLog2 = (uint32_t)t;
    const float ms = timeRuns([&] { return bvht_LBVH_build(mem, mesh, bvh); }, 3);
    if (msOut) msOut[t - 8] = ms;
    if (ms > 0 && ms < bestMs) {
      bestMs = ms;
      best = t;
    }
  }
  mem->opts.tileLog2 = (uint32_t)best;
  return best;
}

int bvht_LBVH_autotuneBatch(bvht_LBVH_BatchWorkingMemory *mem, const bvht_vec3 *vert, const bvht_Triangle *faces,
                            bvht_BVHNode *nodes, float *msOut) {
  if (!mem || !mem->lRes) {
    setError("bvht_LBVH_autotuneBatch: working memory not allocated");
    return 0;
  }
  int best = (int)mem->tileLog2, bestBits = (int)mem->radixBits;
  float bestMs = 1e30f;
  for (int bits : {8, 10}) {
    for (int t = 8; t <= 11; ++t) {
      mem->tileLog2 = (uint32_t)t;
      mem->radixBits = (uint32_t)bits;
      const float ms = timeRuns([&] { return bvht_LBVH_buildBatch(mem, vert, faces, nodes, false); }, 2);
      if (msOut) msOut[(bits == 10 ? 4 : 0) + t - 8] = ms;
      if (ms > 0 && ms < bestMs) {
        bestMs = ms;
        best = t;
        bestBits = bits;
      }
    }
  }
  mem->tileLog2 = (uint32_t)best;
  mem->radixBits = (uint32_t)bestBits;
  return best;
}

}  // extern "C"
