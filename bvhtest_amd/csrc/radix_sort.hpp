// Onesweep LSD radix sort of (key, u32 value) pairs for gfx950.
//
// Replaces the CUB DeviceRadixSort calls of the reference
// (lib/builder/LBVH_CUDA.cu:185-193, lib/builder/Bittner13CUDA.cu:674-680).
// Stable, ascending, 8-bit digits.  One histogram of every pass is gathered
// up front (fused into the producer of the keys where possible), then one
// kernel per pass ranks a tile with wave64 ballots, publishes the tile's
// per-digit counts, resolves its global offsets (a full prefix over the
// published counts when the tiles fit the CUs in one round, decoupled
// look-back otherwise) and scatters through LDS so global writes are
// contiguous runs.
#pragma once

#include "runtime.hpp"

namespace bvht {

constexpr int kSortRadixBits = 8;
constexpr int kSortRadix = 1 << kSortRadixBits;
constexpr int kSortBlock = 256;
constexpr int kSortItems32 = 20;  // keys per thread, 32-bit keys
constexpr int kSortItems64 = 22;  // keys per thread, 64-bit keys: 753 tiles at 4.24M keys, one wave of workgroups
// Wide tiles (k_onesweepWide): 1024 threads x 17 keys, 244 tiles at 4.24M keys.
constexpr int kWideBlock = 1024;
constexpr int kWideItems = 17;
constexpr uint32_t kWideTile = kWideBlock * kWideItems;

struct SortPlan {
  uint32_t n = 0;
  int keyBytes = 4;  // 4 or 8
  int keyBits = 32;  // significant low bits of the key
  int passes = 0;
  uint32_t tile = 0;
  uint32_t numTiles = 0;
  bool wide = false;  // k_onesweepWide (tiles of kWideTile keys)
  // u32 word offsets into the scratch buffer
  size_t histWords = 0, counterWords = 0, statusWords = 0;
  size_t histOff = 0, counterOff = 0, statusOff = 0, totalWords = 0;
  size_t errOff = 0;  // the error word: nonzero after a bounded wait gave up (any pass)
};

SortPlan makeSortPlan(uint32_t n, int keyBytes, int keyBits);
inline size_t sortScratchBytes(const SortPlan &p) { return p.totalWords * 4; }
inline uint32_t *sortHist(const SortPlan &p, uint32_t *scratch) { return scratch + p.histOff; }

// Zeroes histograms, tile counters and look-back state (one fill kernel).
bool sortReset(const SortPlan &p, uint32_t *scratch);
// Histogram pass for keys produced elsewhere (the LBVH fuses it into the
// Morton kernel instead).
bool sortHistogram(const SortPlan &p, const void *keys, uint32_t *scratch);
// Runs every pass.  keys[0]/vals[0] hold the input (vals[0] may be null:
// values are then the input index; vals itself null: a keys-only sort, for
// wide plans of 64-bit keys with keepKeys).  Returns the buffer index (0/1)
// holding the result.  With keepKeys == false the last pass writes values
// only.
// With firstPass > 0 the lower digits are left unsorted: the result is
// ordered stably by key >> (8 firstPass).
int sortRun(const SortPlan &p, void *keys[2], uint32_t *vals[2], bool identityValues, bool keepKeys,
            uint32_t *scratch, void (*afterPass)(int) = nullptr, int firstPass = 0);

}  // namespace bvht
