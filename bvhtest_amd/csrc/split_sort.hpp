// Sort of (63-bit key, input index) pairs in two steps, for the LBVH build's
// 63-bit keys (replaces the CUB DeviceRadixSort call of
// lib/builder/LBVH_CUDA.cu:185-193 for that key width).
//
//   1. the onesweep runs only its passes over bits 24..62 (5 of the 8), so
//      the pairs come out ordered stably by key >> 24, full keys alongside;
//   2. k_splitRank finds each pair's group of equal top bits (galloping
//      search, O(log m)) and places the pair at group start + the number of
//      group members below it by (key, index): O(m) compares per pair, for
//      groups of at most kSplitGroupMax pairs (at the 4.24M-triangle scene
//      the 39 top bits leave 3.57M groups of at most 41 pairs, pair-weighted
//      mean size 2.5);
//   3. k_splitLong sorts longer groups (listed by their first pair) with a
//      stable per-workgroup LSD over the low 24 bits, so any input is
//      sorted in bounded time.
// The result is the permutation of a stable sort by the full key.
//
// Packed keys (the fused build, n < 2^24, wide sort tiles): the key
// producer writes each key with its low 24 bits replaced by the input
// index, and the low 24 bits to splitLow().  The onesweep then sorts 8-byte
// words without values (16 bytes per pair and pass instead of 24), and the
// group step gathers the low bits of the pairs it compares.
#pragma once

#include "radix_sort.hpp"

namespace bvht {

constexpr int kSplitFirstPass = 3;                            // digit passes left to the group step
constexpr int kSplitLoBits = kSplitFirstPass * kSortRadixBits;  // 24
constexpr uint32_t kSplitGroupMax = 4096;

struct SplitPlan {
  uint32_t n = 0;
  SortPlan plan;  // 8-pass 63-bit onesweep plan; passes from kSplitFirstPass on are run
  bool canPack = false;  // packed keys possible (wide tiles, n < 2^24)
  uint32_t listCap = 0;
  // u32 word offsets into the scratch buffer
  size_t ctrlOff = 0, listOff = 0, lowOff = 0, totalWords = 0;
};

SplitPlan makeSplitPlan(uint32_t n);
inline size_t splitScratchBytes(const SplitPlan &p) { return p.totalWords * 4; }
// Histograms of the 8 digit passes, filled by the key producer (the first
// ones go unused).
inline uint32_t *splitHist(const SplitPlan &p, uint32_t *scratch) { return sortHist(p.plan, scratch); }
// The low 24 bits of every key, by input index (packed keys only).
inline uint32_t *splitLow(const SplitPlan &p, uint32_t *scratch) { return scratch + p.lowOff; }
// Zeroes the onesweep state and the long-group list.
bool splitReset(const SplitPlan &p, uint32_t *scratch);

// keys[0] / vals: the n keys in input order (vals unused on entry: the
// input index is the value), packed ones if `packed`.  Returns the buffer
// index (0/1) of keys and vals holding the sorted pairs (-1 on a launch
// error); with keepKeys false only the values there are complete, and with
// keepKeys the keys there are full keys in either mode.
int splitRun(const SplitPlan &p, uint64_t *keys[2], uint32_t *vals[2], uint32_t *scratch, bool keepKeys, bool packed,
             void (*afterPass)(int) = nullptr, void (*afterStage)(const char *) = nullptr);

}  // namespace bvht
