// Batched (segmented) LBVH build: many meshes of uneven size in one launch
// sequence, optionally captured once into a hipGraph and replayed.
//
// Every segment's output equals the standalone build of that mesh (own scene
// box, local face and node indices): the reference build of
// lib/builder/LBVH_CUDA.cu applied per mesh.  Stages:
//   k_segBox     per tile of one mesh: triangle boxes, block reduce, 6 atomic
//                min/max (order-preserving u32 encoding) into the mesh box;
//   k_segMorton  per tile: the mesh's offset/scale (as LBVH_CUDA.cu:171-180),
//                key = mesh id << 30 | 30-bit Morton code, sort histograms;
//   onesweep     one stable sort over all meshes (segment id in the high
//                bits keeps segments apart, stability keeps face order);
//   k_buildLevel the tile hierarchy per mesh (tiles never cross meshes).
#include "lbvh_kernels.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace bvht {

namespace {

constexpr uint32_t kOrdZero = 0x80000000u;  // f2o(+0.0f)

template <int TLOG>
__global__ __launch_bounds__(kBuildBlock) void k_segBox(const bvht_vec3 *__restrict__ vertAll,
                                                        const bvht_Triangle *__restrict__ faces,
                                                        const SegDesc *__restrict__ segs, const uint2 *__restrict__ tiles,
                                                        uint32_t *__restrict__ boxOrd) {
  const uint2 t = tiles[blockIdx.x];
  const SegDesc sd = segs[t.x];
  const bvht_vec3 *vert = vertAll + sd.vertBase;
  const uint32_t p0 = t.y << TLOG, p1 = min(p0 + (1u << TLOG), sd.n);
  float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};  // the zero box of LBVH_CUDA.cu:104-106
  for (uint32_t p = p0 + threadIdx.x; p < p1; p += kBuildBlock) {
    const bvht_Triangle tri = faces[sd.faceBase + p];
    const bvht_AABB b = triangleBox(vert[tri.v1], vert[tri.v2], vert[tri.v3]);
    mn[0] = mergeMin(mn[0], b.minMax[0].x);
    mn[1] = mergeMin(mn[1], b.minMax[0].y);
    mn[2] = mergeMin(mn[2], b.minMax[0].z);
    mx[0] = mergeMax(mx[0], b.minMax[1].x);
    mx[1] = mergeMax(mx[1], b.minMax[1].y);
    mx[2] = mergeMax(mx[2], b.minMax[1].z);
  }
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
    for (int w = 1; w < kBuildBlock / 64; ++w) v = a < 3 ? mergeMin(v, s_red[a][w]) : mergeMax(v, s_red[a][w]);
    if (a < 3)
      atomicMin(boxOrd + 6 * t.x + a, f2o(v));
    else
      atomicMax(boxOrd + 6 * t.x + a, f2o(v));
  }
}

template <int TLOG>
__global__ __launch_bounds__(kBuildBlock) void k_segMorton(const bvht_vec3 *__restrict__ vertAll,
                                                           const bvht_Triangle *__restrict__ faces,
                                                           const SegDesc *__restrict__ segs,
                                                           const uint2 *__restrict__ tiles,
                                                           const uint32_t *__restrict__ boxOrd,
                                                           uint64_t *__restrict__ keys, uint32_t *__restrict__ hist,
                                                           int passes) {
  __shared__ uint32_t s_hist[8][kSortRadix];
  for (int p = 0; p < passes; ++p) s_hist[p][threadIdx.x] = 0;
  const uint2 t = tiles[blockIdx.x];
  const SegDesc sd = segs[t.x];
  bvht_vec3 off, sc;
  {
    float m[6];
    for (int a = 0; a < 6; ++a) m[a] = o2f(boxOrd[6 * t.x + a]);
    // LBVH_calcMortonCodes host part (LBVH_CUDA.cu:171-180), f32, IEEE division
    off = {0.0f - m[0], 0.0f - m[1], 0.0f - m[2]};
    sc = {__fdiv_rn(1.0f, m[3] + off.x), __fdiv_rn(1.0f, m[4] + off.y), __fdiv_rn(1.0f, m[5] + off.z)};
  }
  __syncthreads();
  const bvht_vec3 *vert = vertAll + sd.vertBase;
  const uint32_t p0 = t.y << TLOG, p1 = min(p0 + (1u << TLOG), sd.n);
  for (uint32_t p = p0 + threadIdx.x; p < p1; p += kBuildBlock) {
    const bvht_Triangle tri = faces[sd.faceBase + p];
    const bvht_vec3 c = centroid(vert[tri.v1], vert[tri.v2], vert[tri.v3]);
    const uint64_t key = ((uint64_t)t.x << 30) | (uint64_t)morton30(c, off, sc);
    keys[sd.faceBase + p] = key;
    for (int q = 0; q < passes; ++q) atomicAdd(&s_hist[q][(uint32_t)(key >> (q * kSortRadixBits)) & (kSortRadix - 1)], 1u);
  }
  __syncthreads();
  for (int q = 0; q < passes; ++q) {
    const uint32_t c = s_hist[q][threadIdx.x];
    if (c) atomicAdd(hist + q * kSortRadix + threadIdx.x, c);
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
  uint64_t *keys[2] = {nullptr, nullptr};
  uint32_t *vals[2] = {nullptr, nullptr};
  uint32_t *sortScratch = nullptr;
  SortPlan plan;
  hipStream_t capStream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *capVert = nullptr, *capFaces = nullptr, *capNodes = nullptr;
};

template <int TLOG>
bool batchStagesT(BatchInternal *in, const bvht_vec3 *vert, const bvht_Triangle *faces, bvht_BVHNode *nodes) {
  if (!fillU32(in->boxOrd, kOrdZero, 6ull * in->M)) return false;
  if (!sortReset(in->plan, in->sortScratch)) return false;
  k_segBox<TLOG><<<in->tileCounts[0], kBuildBlock, 0, stream()>>>(vert, faces, in->segs, in->tiles[0], in->boxOrd);
  if (!checkLaunch("k_segBox")) return false;
  k_segMorton<TLOG><<<in->tileCounts[0], kBuildBlock, 0, stream()>>>(vert, faces, in->segs, in->tiles[0], in->boxOrd,
                                                                     in->keys[0], sortHist(in->plan, in->sortScratch),
                                                                     in->plan.passes);
  if (!checkLaunch("k_segMorton")) return false;
  void *keys[2] = {in->keys[0], in->keys[1]};
  const int r = sortRun(in->plan, keys, in->vals, true, false, in->sortScratch);
  if (r < 0) return false;
  return launchTreeLevelsBatch(TLOG, nodes, in->vals[r], faces, vert, in->segs, in->tiles.data(),
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
    if (faceCounts[i] == 0 || faceCounts[i] > (1u << 30)) {
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
  int segBits = 1;
  while ((1u << segBits) < numMeshes) segBits++;
  in->plan = makeSortPlan(in->totalFaces, 8, 30 + segBits);
  if (!devAlloc(&in->segs, numMeshes, "batch segments") ||
      !check(hipMemcpy(in->segs, in->segsHost.data(), sizeof(SegDesc) * numMeshes, hipMemcpyHostToDevice),
             "batch segments") ||
      !devAlloc(&in->boxOrd, 6ull * numMeshes, "batch boxes") || !devAlloc(&in->keys[0], in->totalFaces, "batch keys") ||
      !devAlloc(&in->keys[1], in->totalFaces, "batch keys") || !devAlloc(&in->vals[0], in->totalFaces, "batch values") ||
      !devAlloc(&in->vals[1], in->totalFaces, "batch values") ||
      !devMalloc(reinterpret_cast<void **>(&in->sortScratch), sortScratchBytes(in->plan), "batch sort") ||
      !check(hipStreamCreateWithFlags(&in->capStream, hipStreamNonBlocking), "batch capture stream")) {
    bvht_LBVH_freeBatch(&m);
    return m;
  }
  m.numMeshes = numMeshes;
  m.totalFaces = in->totalFaces;
  m.totalVert = in->totalVert;
  m.totalNodes = in->totalNodes;
  m.tileLog2 = (uint32_t)in->tlog;
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
    devFree(in->segs);
    devFree(in->boxOrd);
    devFree(in->keys[0]);
    devFree(in->keys[1]);
    devFree(in->vals[0]);
    devFree(in->vals[1]);
    devFree(in->sortScratch);
    delete in;
  }
  std::memset(mem, 0, sizeof(*mem));
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
  int best = (int)mem->tileLog2;
  float bestMs = 1e30f;
  for (int t = 8; t <= 11; ++t) {
    mem->tileLog2 = (uint32_t)t;
    const float ms = timeRuns([&] { return bvht_LBVH_buildBatch(mem, vert, faces, nodes, false); }, 2);
    if (msOut) msOut[t - 8] = ms;
    if (ms > 0 && ms < bestMs) {
      bestMs = ms;
      best = t;
    }
  }
  mem->tileLog2 = (uint32_t)best;
  return best;
}

}  // extern "C"
