// Most-significant-digit sort of (Morton key, input index) pairs for gfx950.
//
// Replaces the CUB DeviceRadixSort call of the LBVH build
// (lib/builder/LBVH_CUDA.cu:185-193) when the values are the input indices.
// A stable sort of keys with index values orders the pairs by (key, index),
// and so does any sort by that composite: the result is the same
// permutation, so the sort does not need to be stable on its way there.
//
//   1. histogram of the top 16 key bits, taken by the key producer
//      (msdCountRun, run-length aggregated over a wave);
//   2. k_msdScan: bucket starts, and for every window of kMsdWindow
//      positions the bucket containing its first position;
//   3. k_msdScatter: each pair to its bucket (one cursor atomic per run of
//      equal buckets in a wave; order inside a bucket is arbitrary);
//   4. k_msdWindows: one workgroup per window sorts the window's segment of
//      whole buckets (< 2 windows long) in LDS by (key, index); buckets
//      longer than kMsdLocal go to a list;
//   5. k_msdSplit: workgroups take listed buckets, split them by the next
//      8 composite bits (LDS histogram and cursors), list the pieces that
//      fit in LDS and split the rest again (a per-workgroup stack);
//   6. k_msdJobs: one workgroup per listed piece sorts it in LDS.
// Every loop over global memory issues a thread's 4-16 loads before using
// any: at ~1-2 us per HBM round trip, a dependent load per item would
// dominate.
// Every piece is written once to the output; pieces are whole buckets, so
// sorting a piece by the full composite orders it correctly in place.
#pragma once

#include "runtime.hpp"
#include "wave.hpp"

namespace bvht {

constexpr int kMsdTopBits = 16;
constexpr uint32_t kMsdBuckets = 1u << kMsdTopBits;
constexpr uint32_t kMsdLocal = 4096;             // most pairs one workgroup sorts in LDS
constexpr uint32_t kMsdWindow = kMsdLocal / 2;  // window of the segment partition
constexpr int kMsdCtrlWords = 16;  // [0] big count, [1] big cursor, [2] error bits, [3] job count

struct MsdPlan {
  uint32_t n = 0;
  int keyBytes = 8;
  int keyBits = 63;
  uint32_t windows = 0;  // ceil(n / kMsdWindow)
  uint32_t slots = 0;    // split-stack slots: ceil(n / kMsdLocal) + 1
  uint32_t jobCap = 0;   // pieces the split kernel can hand to the job kernel
  // u32 word offsets into the scratch buffer; [0, zeroWords) must be zero
  // before the key producer runs
  size_t histOff = 0, ctrlOff = 0, zeroWords = 0;
  size_t totalsOff = 0, startOff = 0, cursorOff = 0, winOff = 0, bigOff = 0, jobOff = 0, stackOff = 0, totalWords = 0;
};

MsdPlan makeMsdPlan(uint32_t n, int keyBytes, int keyBits);
inline size_t msdScratchBytes(const MsdPlan &p) { return p.totalWords * 4; }
inline uint32_t *msdHist(const MsdPlan &p, uint32_t *scratch) { return scratch + p.histOff; }

// Adds a wave's keys to the top-digit histogram: lanes [0, popc(valid)) are
// valid and hold consecutive input positions; every lane of the wave calls.
__device__ __forceinline__ void msdCountRun(uint32_t *hist, uint32_t bucket, bool valid) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t prev = __shfl_up(bucket, 1, 64);
  const bool head = valid && (lane == 0 || prev != bucket);
  const uint64_t heads = __ballot(head);
  const uint64_t validMask = __ballot(valid);
  if (head) {
    const uint64_t above = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
    const uint32_t end = above ? (uint32_t)__builtin_ctzll(above) : (uint32_t)__popcll(validMask);
    atomicAdd(hist + bucket, end - lane);
  }
}

// Runs steps 2-5 (the histogram must be complete).  keysIn holds the keys
// of pairs (key, i); keysTmp/valsTmp receive the bucket scatter; the sorted
// pairs end in keysOut/valsOut (keysOut is also used as split scratch and
// receives the sorted keys only when keepKeys).  keysIn and keysOut may be
// the same buffer.
bool msdRun(const MsdPlan &p, const void *keysIn, void *keysTmp, uint32_t *valsTmp, void *keysOut, uint32_t *valsOut,
            bool keepKeys, uint32_t *scratch, void (*afterStage)(const char *) = nullptr);

}  // namespace bvht
