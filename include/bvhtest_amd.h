/*
 * bvhtest_amd.h -- C ABI of the MI355X (gfx950) BVH pipeline.
 *
 * Every entry point below is `extern "C"`, takes plain pointers, sizes and
 * plain-old-data structs, and has no C++ or torch types in its signature.
 * Each declaration cites the reference interface it replaces
 * (mensinda/BVHTest, paths relative to the reference repository root).
 *
 * Layouts: the POD structs reproduce the reference's memory layout byte for
 * byte (BVHNode 48 B, AABB 24 B, Triangle 12 B, Ray 48 B, BVH header 32 B),
 * so buffers written by this library can be handed to reference code and
 * vice versa.  Static asserts in bvhtest_amd/csrc/types.hpp pin the offsets.
 *
 * Error convention (reference: `CUDA_RUN` macros, lib/builder/LBVH_CUDA.cu:29-34):
 * functions return what the reference returns (bool / struct with a result
 * flag / void).  Instead of printing, the library records the last failure;
 * read it with bvht_getLastError().  Kernel launches are checked.
 *
 * Threading: like the reference (single host thread, one stream), calls are
 * not re-entrant.  All device work is issued on the stream set with
 * bvht_setStream() (default: the null stream).
 */
#ifndef BVHTEST_AMD_H
#define BVHTEST_AMD_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Every declaration below is exported; the library hides everything else. */
#pragma GCC visibility push(default)

/* ------------------------------------------------------------------------ */
/* Data model (reference: lib/base/BVH.hpp:38-238, lib/base/Ray.hpp:38-80)   */
/* ------------------------------------------------------------------------ */

/* glm::vec3 (3 x f32, align 4) */
typedef struct bvht_vec3 {
  float x, y, z;
} bvht_vec3;

/* AABB {vec3 minMax[2]} -- lib/base/BVH.hpp:59-60 */
typedef struct bvht_AABB {
  bvht_vec3 minMax[2]; /* 0: min, 1: max */
} bvht_AABB;

/* Triangle {u32 v1, v2, v3} -- lib/base/BVH.hpp:38-42 */
typedef struct bvht_Triangle {
  uint32_t v1, v2, v3;
} bvht_Triangle;

/* BVHNode, alignas(16), 48 B -- lib/base/BVH.hpp:113-123.
 * Leaves: left = first face index, right = face count (1 for LBVH). */
typedef struct __attribute__((aligned(16))) bvht_BVHNode {
  bvht_AABB bbox;       /* @0  */
  uint32_t parent;      /* @24 */
  uint32_t left;        /* @28 */
  uint32_t right;       /* @32 */
  uint16_t isLeafFlag;  /* @36 */
  uint16_t isLeft;      /* @38 */
  uint16_t level;       /* @40 */
  uint16_t _pad;        /* @42 (padding in the reference) */
  float surfaceArea;    /* @44 */
} bvht_BVHNode;

/* Device-side BVH header, 32 B -- lib/base/BVH.hpp:138-144 (class BVH data
 * members: node pointer, size, capacity, root index, max level). */
typedef struct bvht_BVH {
  bvht_BVHNode *bvh;
  size_t vSize;
  size_t vCapacity;
  uint32_t vRootIndex;
  uint16_t vMaxLevel;
  uint16_t _pad;
} bvht_BVH;

/* MeshRaw -- lib/base/BVH.hpp:50-57 (device pointers + counts) */
typedef struct bvht_MeshRaw {
  bvht_vec3 *vert;
  bvht_vec3 *norm;
  bvht_Triangle *faces;
  uint32_t numVert;
  uint32_t numNorm;
  uint32_t numFaces;
} bvht_MeshRaw;

/* CUDAMemoryBVHPointer -- lib/base/BVH.hpp:234-238 */
typedef struct bvht_CUDAMemoryBVHPointer {
  bvht_BVH *bvh;        /* device header */
  bvht_BVHNode *nodes;  /* device nodes  */
  uint32_t numNodes;
} bvht_CUDAMemoryBVHPointer;

/* Ray, alignas(16), 48 B -- lib/base/Ray.hpp:38-53 */
typedef struct __attribute__((aligned(16))) bvht_Ray {
  bvht_vec3 pos;     /* @0  */
  bvht_vec3 dir;     /* @12 */
  bvht_vec3 invDir;  /* @24 */
  uint8_t sign[3];   /* @36 (struct Sign, align 4) */
  uint8_t _pad;
  uint32_t px;       /* @40 */
  uint32_t py;       /* @44 */
} bvht_Ray;

/* CUDAPixel {u8 hit, u8 diffuse, u16 intCount} -- lib/tracer/CUDAKernels.hpp:33-37 */
typedef struct bvht_CUDAPixel {
  uint8_t hit;
  uint8_t diffuse;
  uint16_t intCount;
} bvht_CUDAPixel;

/* Hit record (extension: the reference writes no triangle id or distance,
 * lib/tracer/CUDAKernels.cu:208-220).  triId = UINT32_MAX on a miss. */
typedef struct bvht_Hit {
  uint32_t triId;
  float t;
} bvht_Hit;

/* ------------------------------------------------------------------------ */
/* Library control (no reference counterpart: the reference uses the legacy  */
/* default stream and prints errors)                                        */
/* ------------------------------------------------------------------------ */

/* Stream every later call issues its work on (hipStream_t; NULL = default). */
void bvht_setStream(void *stream);
void *bvht_getStream(void);
/* Message of the last failure ("" if none) and reset of it. */
const char *bvht_getLastError(void);
void bvht_clearError(void);
/* Host-test hook: make the n-th device allocation from now fail (n >= 1);
 * 0 disables.  Exercises the CUDA_ERROR paths, e.g. lib/builder/LBVH.cpp:29-35. */
void bvht_setAllocFailAfter(int n);
/* Library version string and the gfx target it was built for. */
const char *bvht_version(void);
/* Number of visible HIP devices (0 when none; never fails). */
int bvht_deviceCount(void);

/* ------------------------------------------------------------------------ */
/* Memory and copies -- lib/cuda/cudaFN.hpp:27-42 (extern "C" in reference)  */
/* ------------------------------------------------------------------------ */

/* runMalloc/runMemcpy/runFree -- lib/cuda/cudaFN.hpp:37-40.
 * runMemcpy kinds: 0 host->host, 1 host->device, 2 device->host,
 * 3 device->device (hipMemcpyKind values). */
bool runMalloc(void **ptr, size_t size);
bool runMemcpy(void *dest, const void *src, size_t size, int kind);
void runFree(void *ptr);

/* copyBVHToGPU -- lib/cuda/cudaFN.cu:32-74.  Host nodes + root/maxLevel in,
 * device header + nodes out.  numNodes == 0 gives an empty device BVH. */
bool copyBVHToGPU(const bvht_BVHNode *hostNodes, uint32_t numNodes, uint32_t root, uint16_t maxLevel,
                  bvht_CUDAMemoryBVHPointer *out);
/* copyMeshToGPU -- lib/cuda/cudaFN.cu:76-109. */
bool copyMeshToGPU(const bvht_vec3 *vert, uint32_t numVert, const bvht_vec3 *norm, uint32_t numNorm,
                   const bvht_Triangle *faces, uint32_t numFaces, bvht_MeshRaw *out);
/* copyBVHToHost -- lib/cuda/cudaFN.cu:113-146.  Reads the device header for
 * root/size/maxLevel, copies nodes into hostNodes (capacity nodes) and frees
 * the device BVH like the reference.  numNodes, root and maxLevel (optional)
 * receive the header fields. */
bool copyBVHToHost(bvht_CUDAMemoryBVHPointer *bvh, bvht_BVHNode *hostNodes, uint32_t capacity,
                   uint32_t *numNodes, uint32_t *root, uint16_t *maxLevel);
/* copyMeshToHost -- lib/cuda/cudaFN.cu:149-182 (copies, then frees the device mesh). */
bool copyMeshToHost(bvht_MeshRaw *mesh, bvht_vec3 *vert, bvht_vec3 *norm, bvht_Triangle *faces);
/* resetCUDA -- lib/cuda/cudaFN.cu:184-187 (here: device sync + release of
 * library-owned scratch; no hipDeviceReset, which would invalidate the
 * caller's other allocations). */
void resetCUDA(void);
/* Free a device BVH / mesh without copying (no reference counterpart; the
 * reference frees only inside copy*ToHost). */
void bvht_freeBVH(bvht_CUDAMemoryBVHPointer *bvh);
void bvht_freeMesh(bvht_MeshRaw *mesh);

/* ------------------------------------------------------------------------ */
/* LBVH builder -- lib/builder/LBVH_CUDA.hpp:27-59 (C++ linkage in reference) */
/* ------------------------------------------------------------------------ */

/* Build options (extension).  Defaults (bvht_LBVH_defaultOptions) reproduce
 * the reference bit for bit: 30-bit Morton keys, index-only topology. */
typedef struct bvht_LBVH_Options {
  uint32_t mortonBits;   /* 30 (reference, lib/builder/LBVH_CUDA.cu:130-166) or 63 (21 bits/axis) */
  uint32_t topology;     /* 0: reference topology (delta = clz32 of the low word, i.e. of the
                            sorted index, lib/builder/LBVH_CUDA.cu:205-212); 1: Karras 2012 over
                            the full 64-bit key (morton<<32 | index) */
  uint32_t tileLog2;     /* log2 of leaves per build workgroup (8..11), 0 = default */
  uint32_t digitBits;    /* first sort digit: 8 (default, 0 = default) = onesweep LSD radix sort,
                            8-bit digits (63-bit keys: over the top 30 bits, then each group of
                            equal top bits ordered by (key, index)); 16 = most-significant-digit
                            split on the top 16 key bits, then in-LDS sorts of the buckets by
                            (key, index).  All give the permutation of the reference's stable
                            sort. */
} bvht_LBVH_Options;

/* LBVH_WorkingMemory -- lib/builder/LBVH_CUDA.hpp:37-47 (scratch handles). */
typedef struct bvht_LBVH_WorkingMemory {
  bool lRes;
  bvht_LBVH_Options opts;
  uint32_t numFaces;
  void *keys[2];          /* ping-pong sort keys (u32 or u64) */
  uint32_t *values[2];    /* ping-pong face indices */
  uint32_t *sortedValues; /* -> face index of leaf i after LBVH_sortMortonCodes */
  void *sortedKeys;       /* -> Morton key of leaf i (NULL if the sort dropped keys) */
  void *sortScratch;      /* histograms, look-back state, tile counters */
  size_t sortScratchBytes;
  void *params;           /* device: scene box, offset, scale, counters */
  uint32_t *atomicLocks;  /* refit arrival counters */
  uint32_t numLocks;
  void *internal;         /* host-side bookkeeping */
} bvht_LBVH_WorkingMemory;

bvht_LBVH_Options bvht_LBVH_defaultOptions(void);

/* LBVH_allocateBVH -- lib/builder/LBVH_CUDA.cu:396-424.  Fails if *bvh already
 * holds memory.  Writes the device header (root 0, size 2n-1). */
bool LBVH_allocateBVH(bvht_CUDAMemoryBVHPointer *bvh, const bvht_MeshRaw *mesh);
/* LBVH_allocateWorkingMemory -- lib/builder/LBVH_CUDA.cu:426-475 (reference options). */
bvht_LBVH_WorkingMemory LBVH_allocateWorkingMemory(const bvht_MeshRaw *mesh);
/* Same with explicit options (extension). */
bvht_LBVH_WorkingMemory bvht_LBVH_allocateWorkingMemoryEx(const bvht_MeshRaw *mesh, const bvht_LBVH_Options *opts);
/* LBVH_initTriData -- lib/builder/LBVH_CUDA.cu:98-119: scene box = union of the
 * per-triangle boxes (+-FLT_EPSILON) and the origin.  Synchronises and returns
 * the box like the reference (the fused build below does not). */
bvht_AABB LBVH_initTriData(bvht_LBVH_WorkingMemory *mem, const bvht_MeshRaw *mesh);
/* LBVH_calcMortonCodes -- lib/builder/LBVH_CUDA.cu:168-183 (offset/scale from the box). */
void LBVH_calcMortonCodes(bvht_LBVH_WorkingMemory *mem, bvht_AABB sceneAABB);
/* LBVH_sortMortonCodes -- lib/builder/LBVH_CUDA.cu:185-193 (stable LSD radix sort). */
void LBVH_sortMortonCodes(bvht_LBVH_WorkingMemory *mem);
/* LBVH_buildBVHTree -- lib/builder/LBVH_CUDA.cu:326-330: leaves + Karras
 * topology.  Here the same pass also refits every node below the top tiles. */
void LBVH_buildBVHTree(bvht_LBVH_WorkingMemory *mem, bvht_CUDAMemoryBVHPointer *bvh);
/* LBVH_fixAABB -- lib/builder/LBVH_CUDA.cu:376-385: completes the bottom-up refit. */
void LBVH_fixAABB(bvht_LBVH_WorkingMemory *mem, bvht_CUDAMemoryBVHPointer *bvh);
/* LBVH_freeWorkingMemory -- lib/builder/LBVH_CUDA.cu:477-488. */
void LBVH_freeWorkingMemory(bvht_LBVH_WorkingMemory *mem);
/* LBVH_doCUDASyc -- lib/builder/LBVH_CUDA.cu:490 (stream sync). */
void LBVH_doCUDASyc(void);
/* Whole build on the stream without host round trips (extension: the scene
 * box stays on the device).  Same result as the five staged calls above.
 * The scene box comes from one pass over all mesh->numVert vertices (which
 * must all be readable), checked against the triangles' box by the key
 * pass, which repairs the keys when unreferenced vertices lie outside it
 * (digitBits 16 uses the triangles' box directly). */
bool bvht_LBVH_build(bvht_LBVH_WorkingMemory *mem, const bvht_MeshRaw *mesh, bvht_CUDAMemoryBVHPointer *bvh);
/* Error bits of the last sort (extension, synchronises): 0 = the sort ran
 * to completion (also before the first build).  MSD sort: 1 = histogram
 * does not match the keys, 2 = a bucket ran out of key bits, 4 = a piece
 * exceeded the in-LDS limit;
 * onesweep: 1 = a look-back wait gave up; 63-bit keys also 8 = the list of
 * long groups of equal top bits overflowed.
 * 0xFFFFFFFF on a bad argument or a failed copy. */
uint32_t bvht_LBVH_sortStatus(bvht_LBVH_WorkingMemory *mem);
/* LBVH build that allocates nothing (no counterpart in the reference).  For
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc: the same
 * build as bvht_LBVH_build with default options (mortonBits 30 or 63), its
 * working memory taken from a library-owned device array of 32 MiB per GPU
 * (held until the process ends).  Meshes of at most 2^20 faces; larger
 * ones are refused with an error.  The caller passes the device mesh and a
 * BVH allocated for it (LBVH_allocateBVH).  Calls that overlap on one GPU
 * race (they share its array).  No hipMalloc or hipFree is called. */
bool bvht_LBVH_buildNoAlloc(const bvht_MeshRaw *mesh, bvht_CUDAMemoryBVHPointer *bvh, uint32_t mortonBits);
/* Sorts the n (key, index) pairs of device keys[0..n) (keyBits 30: uint32
 * keys, 63: uint64 keys, each below 2^keyBits) by key then index, i.e. the
 * permutation of the reference's stable radix sort with index values
 * (lib/builder/LBVH_CUDA.cu:185-193), with the LBVH build's sort (digitBits
 * 8: onesweep, 16: MSD, 0: default = 8).  valsOut (device, n) receives the
 * indices, keysOut (device, optional) the sorted keys.  Extension for
 * testing the sort alone; allocates its own scratch and synchronises. */
bool bvht_sortKeyIndexPairs(const void *keys, uint32_t n, uint32_t keyBits, uint32_t digitBits, void *keysOut,
                            uint32_t *valsOut);
/* Per-kernel device times (ms, HIP events on the library stream) of the last
 * bvht_LBVH_build when timing is on; bvht_LBVH_stageName(i) names kernel i
 * (k_sceneBox, k_morton or k_mortonMsd, the sort's kernels, k_buildLevel0,
 * k_buildLevelUp).
 * Returns the number of entries written. */
void bvht_LBVH_setTiming(bool enable);
/* With group on, the onesweep digit passes of a timed build are one entry,
 * k_onesweep, covering all of them (no events between the passes, so no
 * event overhead inside; divide by the pass count for a mean pass). */
void bvht_LBVH_setTimingPassGroups(bool group);
int bvht_LBVH_stageTimes(float *ms, int maxStages);
const char *bvht_LBVH_stageName(int i);

/* Autotuner (extension): times tile sizes 2^8..2^11 of the build on the
 * caller's input and keeps the fastest in mem->opts.tileLog2 (every choice
 * gives the same tree).  msOut (4 floats, optional) receives the mean times.
 * Returns the chosen log2. */
int bvht_LBVH_autotune(bvht_LBVH_WorkingMemory *mem, const bvht_MeshRaw *mesh, bvht_CUDAMemoryBVHPointer *bvh,
                       float *msOut);

/* Batched (segmented) build of many meshes (extension; the reference builds
 * one mesh per LBVH command).  Meshes are concatenated: mesh i owns faces
 * [faceOffset_i, +faceCounts[i]) whose indices are local to its vertices
 * [vertOffset_i, +vertCounts[i]), and writes its 2n-1 nodes at nodeOffset_i
 * with local node and face indices -- exactly the standalone build of that
 * mesh (30-bit keys, reference topology). */
typedef struct bvht_LBVH_BatchWorkingMemory {
  bool lRes;
  uint32_t numMeshes;
  uint32_t totalFaces;
  uint32_t totalVert;
  uint32_t totalNodes;
  uint32_t tileLog2;
  uint32_t radixBits; /* digit width of the segmented sort: 8 (4 passes) or 10 (3 passes) */
  void *internal;
} bvht_LBVH_BatchWorkingMemory;

bvht_LBVH_BatchWorkingMemory bvht_LBVH_allocateBatch(uint32_t numMeshes, const uint32_t *faceCounts,
                                                     const uint32_t *vertCounts, uint32_t tileLog2);
/* useGraph: capture the launch sequence into a hipGraph on the first call and
 * replay it afterwards (re-captured when the pointers change). */
bool bvht_LBVH_buildBatch(bvht_LBVH_BatchWorkingMemory *mem, const bvht_vec3 *vert, const bvht_Triangle *faces,
                          bvht_BVHNode *nodes, bool useGraph);
void bvht_LBVH_freeBatch(bvht_LBVH_BatchWorkingMemory *mem);
bool bvht_LBVH_batchOffsets(const bvht_LBVH_BatchWorkingMemory *mem, uint32_t i, uint32_t *faceOffset,
                            uint32_t *vertOffset, uint32_t *nodeOffset);
/* Error bits of the last batched sort (extension, no counterpart in the
 * reference; synchronises): 0 = the sort ran to completion (also before the
 * first build), 1 = a look-back wait gave up (the trees of that build are
 * not valid).  0xFFFFFFFF on a bad argument or a failed copy. */
uint32_t bvht_LBVH_batchStatus(const bvht_LBVH_BatchWorkingMemory *mem);
/* Autotuner of the batched build: times every tile size 2^8..2^11 with both
 * digit widths (8 and 10 bits) on the caller's input and keeps the fastest
 * pair in mem->tileLog2 / mem->radixBits (every choice gives the same trees).
 * msOut (8 floats, optional) receives the mean times, 8-bit digits first.
 * Returns the chosen tile log2. */
int bvht_LBVH_autotuneBatch(bvht_LBVH_BatchWorkingMemory *mem, const bvht_vec3 *vert, const bvht_Triangle *faces,
                            bvht_BVHNode *nodes, float *msOut);

/* ------------------------------------------------------------------------ */
/* Bittner13 GPU optimizer -- lib/builder/Bittner13CUDA.hpp:56-83           */
/* ------------------------------------------------------------------------ */

/* AlgoCFG -- lib/builder/Bittner13CUDA.hpp:56-66 (+ extension fields).
 *   offsetAccess  chunk k takes sorted entry k * numChunks + c instead of
 *                 c * chunkSize + k (Bittner13CUDA.cu:516).
 *   altFindNode   findNode2 (Bittner13CUDA.cu:194-264): the 16-slot
 *                 approximate search, reproduced exactly; otherwise the
 *                 exact branch-and-bound search of findNode1.
 *   altSort       threshold selection (Bittner13CUDA.cu:682-699): the nodes
 *                 whose priority is at least the sorted priority at index
 *                 numChunks * chunkSize, in inner-list order; otherwise the
 *                 stable descending sort.  Heuristic priority only.
 *   sort          must be true (the reference's alternative is disabled:
 *                 Bittner13CUDA.cu:701-722); false is an error.
 *   altFixTree    the reference refits the touched paths (true) or the whole
 *                 tree (false) after each chunk; boxes are exact min/max,
 *                 so both give the same bytes and the touched-path refit
 *                 always runs here.
 *   localPatchCPY the reference keeps a patch in registers until it is
 *                 complete (true) or builds it in global memory, where a
 *                 failed patch is later applied half done (false, a defect);
 *                 this library always keeps it local.
 *   blockSize     the reference's CUDA thread-block size; the kernels here
 *                 pick their own launch shapes.  Results do not depend on it. */
typedef struct bvht_AlgoCFG {
  uint32_t numChunks;
  uint32_t chunkSize;
  uint32_t blockSize;
  bool offsetAccess;
  bool altFindNode;
  bool altFixTree;
  bool altSort;
  bool sort;
  bool localPatchCPY;
  /* extension */
  uint32_t priority;       /* 0: reference heuristic (kCalcCost), 1: learned MLP */
  uint32_t lanesPerSearch; /* reserved, 0 */
} bvht_AlgoCFG;

/* GPUWorkingMemory -- lib/builder/Bittner13CUDA.hpp:39-54 (opaque here). */
typedef struct bvht_GPUWorkingMemory {
  bool result;
  uint32_t numLeafNodes;
  uint32_t numInnerNodes;
  uint32_t numPatches;
  uint32_t numSkipped;
  uint32_t numNodesToFix;
  uint32_t numFlags;
  void *internal;
} bvht_GPUWorkingMemory;

/* allocateMemory -- lib/builder/Bittner13CUDA.cu:836-919 */
bvht_GPUWorkingMemory allocateMemory(bvht_CUDAMemoryBVHPointer *bvh, uint32_t batchSize, uint32_t numFaces);
/* freeMemory -- lib/builder/Bittner13CUDA.cu:921-936 */
void freeMemory(bvht_GPUWorkingMemory *data);
/* initData -- lib/builder/Bittner13CUDA.cu:939-987 */
void initData(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, uint32_t blockSize);
/* fixTree1 -- lib/builder/Bittner13CUDA.cu:641-648 (full bottom-up refit) */
void fixTree1(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, uint32_t blockSize);
/* fixTree3 -- lib/builder/Bittner13CUDA.cu:650-656 (refit of the touched paths) */
void fixTree3(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, uint32_t blockSize);
/* bn13_selectNodes -- lib/builder/Bittner13CUDA.cu:659-726 */
void bn13_selectNodes(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, bvht_AlgoCFG cfg);
/* bn13_rmAndReinsChunk -- lib/builder/Bittner13CUDA.cu:729-766 */
void bn13_rmAndReinsChunk(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, bvht_AlgoCFG cfg,
                          uint32_t chunk);
/* bn13_doAlgorithmStep -- lib/builder/Bittner13CUDA.cu:768-772 */
void bn13_doAlgorithmStep(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, bvht_AlgoCFG cfg);
/* selectDynamicLeafNodes -- lib/tracer/CUDAKernels.cu:456-484: the leaves
 * (device ids, leaf-list order) whose first face is >= beginDynamicTriangle
 * into out[]; their number into *numOut (host). */
void selectDynamicLeafNodes(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, uint32_t *out,
                            uint32_t *numOut, uint32_t beginDynamicTriangle);
/* refitDynamicTris -- lib/tracer/CUDAKernels.cu:492-530: leaf boxes of
 * nodesToUpdate[0..numNodes) from the current vertices, widened by
 * FLT_EPSILON; as in the reference, the leaves' surfaceArea is not
 * rewritten. */
void refitDynamicTris(bvht_CUDAMemoryBVHPointer *bvh, bvht_MeshRaw mesh, uint32_t *nodesToUpdate, uint32_t numNodes,
                      uint32_t blockSize);
/* fixTree3 from explicit start nodes (extension): the reference's viewer
 * points GPUWorkingMemory.nodesToFix at the dynamic leaves and calls fixTree3
 * (lib/viewer/LiveTracer.cpp:284-294); with the opaque working memory here
 * the starts are passed directly.  Leaf starts are allowed, and a start may
 * appear more than once. */
bool bvht_fixTreeFrom(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, const uint32_t *devStarts,
                      uint32_t num);
/* CUDAcalcSAH -- lib/builder/Bittner13CUDA.cu:776-803: (sum SA - SA(root)) / SA(root) */
float CUDAcalcSAH(bvht_CUDAMemoryBVHPointer *bvh);
/* doCudaDevSync -- lib/builder/Bittner13CUDA.cu:824 */
void doCudaDevSync(void);
/* calcNumSkipped -- lib/builder/Bittner13CUDA.cu:806-821 */
uint32_t calcNumSkipped(bvht_GPUWorkingMemory *data);
/* Optimizer counters since initData (extension; no reference counterpart):
 * [0] lock conflicts, [1] patches dropped (removal impossible or reinsertion
 * at the root), [2] patches dropped because their insertion point lies in a
 * subtree another patch moves, [3] search frontier overflows, [4] applied
 * patches, [5] patches the conflict resolution left undecided (0 unless the
 * device breaks memory ordering; they are dropped), [6] searches stopped at
 * the step bound (0 for a tree; the patch is dropped), [7] steps whose
 * priority sort gave up a bounded wait (0 unless the device breaks memory
 * ordering; that step's order is then wrong).  Returns the number of words
 * written (at most 8). */
int bvht_bn13_stats(bvht_GPUWorkingMemory *data, uint32_t *out, int maxWords);
/* Weighted SAH of the device tree (lib/base/BVH.cpp:23-37 semantics:
 * sum over non-root nodes of SA * (leaf ? cLeaf : cInner) / SA(root)),
 * accumulated in f64 (extension). */
double bvht_calcSAHWeighted(bvht_CUDAMemoryBVHPointer *bvh, float cInner, float cLeaf);
/* Validate on the device (lib/misc/Validate.cpp:42-263 without the level
 * checks, which need the host-side fixLevels): errOut[6] receives the counts
 * of tree, bbox, strict-bbox, surface-area, triangle and orphan-triangle
 * errors (the classes are defined in bvhtest_amd/csrc/validate.hip).
 * Returns their sum, or UINT32_MAX on a launch or argument error. */
uint32_t bvht_validate(const bvht_CUDAMemoryBVHPointer *bvh, uint32_t root, bvht_MeshRaw mesh, uint32_t *errOut);

/* ---- learned removal priority (extension; the reference ranks nodes by the
 * fixed heuristic of Bittner13CUDA.cu:480-500).  A 12-64-64-1 MLP with ReLU,
 * BF16 weights on MFMA and FP32 accumulation; the 12 node features are
 * listed in bvhtest_amd/csrc/mlp.hip.  Parameters are kNumParams fp32 values:
 * w1[12][64] (in, out), b1[64], w2[64][64] (in, out), b2[64], w3[64], b3. */
typedef struct bvht_PriorityModel bvht_PriorityModel;
uint32_t bvht_mlpNumParams(void);
/* He-uniform weights from `seed`, zero biases; NULL on failure. */
bvht_PriorityModel *bvht_mlpCreate(uint32_t seed);
void bvht_mlpFree(bvht_PriorityModel *model);
bool bvht_mlpSetParams(bvht_PriorityModel *model, const float *hostParams);
bool bvht_mlpGetParams(bvht_PriorityModel *model, float *hostParams);
/* devOut[i] = model(devFeat[i][0..11]) for i < num (device buffers). */
bool bvht_mlpPredict(bvht_PriorityModel *model, const float *devFeat, uint32_t num, float *devOut);
/* Adam on the mean squared error of model(devFeat[i]) against devTarget[i]:
 * `epochs` passes over a seeded shuffle in batches of `batch` (partial last
 * batch dropped).  Returns the mean loss of the last epoch, -1 on error. */
float bvht_mlpTrain(bvht_PriorityModel *model, const float *devFeat, const float *devTarget, uint32_t num,
                    uint32_t epochs, uint32_t batch, float lr, uint32_t seed);
/* Model used by bn13_selectNodes when AlgoCFG.priority == 1 (NULL clears). */
bool bvht_bn13_setModel(bvht_GPUWorkingMemory *data, bvht_PriorityModel *model);
/* Features of every inner node, in inner-list order (ascending node id),
 * into devFeat[numInnerNodes][12]. */
bool bvht_bn13_features(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, float *devFeat);
/* SAH-gain records: while enabled (maxRecords > 0), every patch a step
 * evaluates appends (features of its node at selection time, gain), where
 * gain = (SA(node) + SA(parent) + shrink of the two path boxes above the
 * removal - both insertion costs) / SA(root).  Re-enabling clears. */
bool bvht_bn13_setRecording(bvht_GPUWorkingMemory *data, uint32_t maxRecords);
/* Number of records held (<= maxRecords); device pointers to them. */
uint32_t bvht_bn13_records(bvht_GPUWorkingMemory *data, float **devFeat, float **devGain);
/* Records the patches of the inner nodes at devPositions[0..num) (inner-list
 * positions) on the current tree without changing it. */
bool bvht_bn13_probe(bvht_GPUWorkingMemory *data, bvht_CUDAMemoryBVHPointer *bvh, bvht_AlgoCFG cfg,
                     const uint32_t *devPositions, uint32_t num);

/* ------------------------------------------------------------------------ */
/* Ray tracer -- lib/tracer/CUDAKernels.hpp:39-78 (extern "C" in reference)  */
/* ------------------------------------------------------------------------ */

bool allocateRays(bvht_Ray **rays, uint32_t num);                 /* CUDAKernels.hpp:39 */
bool allocateImage(bvht_CUDAPixel **img, uint32_t w, uint32_t h); /* CUDAKernels.hpp:40 */
void freeRays(bvht_Ray **rays);                                   /* CUDAKernels.hpp:41 */
void freeImage(bvht_CUDAPixel *img);                              /* CUDAKernels.hpp:42 */
/* generateRays -- lib/tracer/CUDAKernels.cu:386-397 (pinhole camera). */
void generateRays(bvht_Ray *rays, uint32_t w, uint32_t h, bvht_vec3 pos, bvht_vec3 lookAt, bvht_vec3 up, float fov);
/* tracerImage -- lib/tracer/CUDAKernels.cu:399-417.  bundle selects the
 * wave-coherent traversal (reference kTraceRayBundle). */
void tracerImage(bvht_Ray *rays, bvht_CUDAPixel *img, bvht_BVHNode *nodes, uint32_t rootNode, bvht_MeshRaw mesh,
                 bvht_vec3 light, uint32_t w, uint32_t h, bool bundle);
/* copyImageToHost -- CUDAKernels.hpp:67 */
void copyImageToHost(bvht_CUDAPixel *hostPixel, bvht_CUDAPixel *cudaImg, uint32_t w, uint32_t h);
/* tracerDoCudaSync -- CUDAKernels.hpp:77 */
void tracerDoCudaSync(void);

/* Extension: trace n rays with the same traversal and write per-ray hits
 * (and the reference pixel when img != NULL).  mode 0 = reference triangle
 * test (no t-min, CUDAKernels.cu:88-112); mode 1 = t > tMin and the triangle
 * `skipTri[i]` (if skipTri != NULL) is ignored, for secondary rays.
 * Mode-1 rays are traced by resident waves that refill idle lanes from a
 * shared counter; with bvht_setRayOrdering(true) they are first sorted by
 * (direction cell, origin cell).  Every ray's result is the same either way,
 * only the scheduling and memory locality differ. */
bool bvht_traceRays(const bvht_Ray *rays, uint32_t num, bvht_BVHNode *nodes, uint32_t rootNode, bvht_MeshRaw mesh,
                    bvht_vec3 light, int mode, float tMin, const uint32_t *skipTri, bvht_Hit *hits,
                    bvht_CUDAPixel *img);
void bvht_setRayOrdering(bool enable);
/* Extension: one cosine-weighted diffuse bounce per primary hit: builds
 * secondary rays (origin hit + 1e-4 * n, n facing the incoming ray, direction
 * from a per-pixel hash of (seed, pixel)) into outRays; rays of missed pixels
 * get dir = 0 and are skipped by bvht_traceRays (hit = miss). */
bool bvht_makeBounceRays(const bvht_Ray *primary, const bvht_Hit *hits, uint32_t num, bvht_MeshRaw mesh,
                         uint32_t seed, bvht_Ray *outRays, uint32_t *outSkipTri);

/* ------------------------------------------------------------------------ */
/* Synthetic scenes (host only; no reference counterpart -- the reference    */
/* loads meshes from disk with lib/io/Load.cpp)                              */
/* ------------------------------------------------------------------------ */

/* Heightfield gx x gz cells, 2 triangles per cell, heights from seeded
 * 4-octave value noise in [0, amplitude].  Returns counts when buffers are NULL. */
void bvht_sceneHeightfield(uint32_t gx, uint32_t gz, uint32_t seed, float amplitude, bvht_vec3 *vert,
                           bvht_Triangle *faces, uint32_t *numVert, uint32_t *numFaces);
/* Icosphere clusters + triangle soup (the 4M-triangle scene).  Counts first
 * (buffers NULL), then fill. */
void bvht_sceneSpheresSoup(uint32_t numSpheres, uint32_t numSoup, uint32_t seed, bvht_vec3 *vert,
                           bvht_Triangle *faces, uint32_t *numVert, uint32_t *numFaces);
/* Batch of meshes: mesh i (seed + i) has about exp(U(ln minTris, ln maxTris))
 * triangles; kind i mod 3 = heightfield, icosphere cluster, triangle soup.
 * Pass NULL buffers to get the counts (faceCounts/vertCounts per mesh and the
 * totals), then call again with buffers of those sizes. */
void bvht_sceneBatch(uint32_t numMeshes, uint32_t seed, uint32_t minTris, uint32_t maxTris, uint32_t *faceCounts,
                     uint32_t *vertCounts, bvht_vec3 *vert, bvht_Triangle *faces, uint64_t *totalVert,
                     uint64_t *totalFaces);
/* Uniform triangle soup in the unit cube. */
void bvht_sceneSoup(uint32_t numTris, uint32_t seed, float size, bvht_vec3 *vert, bvht_Triangle *faces,
                    uint32_t *numVert, uint32_t *numFaces);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif
#endif /* BVHTEST_AMD_H */
