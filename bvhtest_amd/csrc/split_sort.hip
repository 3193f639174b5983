// Two-step sort of (63-bit key, input index) pairs (see split_sort.hpp).
#include "split_sort.hpp"
#include "wave.hpp"


namespace bvht {

namespace {

enum : uint32_t { kErrList = 1 };
constexpr int kLongItems = 16;  // pairs per thread per chunk of the long-group sort
constexpr uint32_t kLongChunk = kSortBlock * kLongItems;

__device__ __forceinline__ uint64_t hiOf(uint64_t k) { return k >> kSplitLoBits; }

// First and one-past-last position of the run of top bits `hi` around p
// (the array is sorted by top bits): galloping out from p, then a binary
// search, so a pair of a group of m pairs costs O(log m) loads.
__device__ __forceinline__ uint32_t groupStart(const uint64_t *kf, uint32_t p, uint64_t hi) {
  uint32_t q = p, step = 1;
  while (q >= step && hiOf(kf[q - step]) == hi) {  // kf[q] is in the group
    q -= step;
    step <<= 1;
  }
  uint32_t lo = q >= step ? q - step + 1 : 0, h = q;  // the start lies in [lo, h]
  while (lo < h) {
    const uint32_t mid = lo + (h - lo) / 2;
    if (hiOf(kf[mid]) == hi)
      h = mid;
    else
      lo = mid + 1;
  }
  return lo;
}
__device__ __forceinline__ uint32_t groupEnd(const uint64_t *kf, uint32_t n, uint32_t p, uint64_t hi) {
  uint32_t q = p, step = 1;
  while (q + step < n && hiOf(kf[q + step]) == hi) {
    q += step;
    step <<= 1;
  }
  uint32_t lo = q, h = q + step < n ? q + step : n;  // the last member lies in [lo, h)
  while (h - lo > 1) {
    const uint32_t mid = lo + (h - lo) / 2;
    if (hiOf(kf[mid]) == hi)
      lo = mid;
    else
      h = mid;
  }
  return lo + 1;
}

// Packed words (the fused build's keys, n < 2^24): a 63-bit key with its
// low 24 bits replaced by the input index.  The top bits, which the
// onesweep passes sort, are the key's; the low bits are kept in `low`,
// indexed by input index.
constexpr uint64_t kPackMask = (1ull << kSplitLoBits) - 1;
__device__ __forceinline__ uint32_t packedIndex(uint64_t w) { return (uint32_t)(w & kPackMask); }
__device__ __forceinline__ uint64_t packedKey(uint64_t w, const uint32_t *low) {
  return (w & ~kPackMask) | low[packedIndex(w)];
}

// Rank of pair (key, v) inside its group [gs, ge) of the sorted-by-top-bits
// array, read from global memory (the general path).
template <bool PACKED>
__device__ __forceinline__ uint32_t rankGlobal(const uint64_t *kf, const uint32_t *vals, const uint32_t *low,
                                               uint32_t gs, uint32_t ge, uint64_t key, uint32_t v) {
  uint32_t rank = 0;
#pragma unroll 4
  for (uint32_t q = gs; q < ge; ++q) {
    const uint64_t kq = PACKED ? packedKey(kf[q], low) : kf[q];
    const uint32_t vq = PACKED ? packedIndex(kf[q]) : vals[q];
    rank += (kq < key || (kq == key && vq < v)) ? 1u : 0u;
  }
  return rank;
}

__device__ __forceinline__ uint64_t readLane64(uint64_t x, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int waveMaxInt(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, 64));
  return x;
}

// Each pair of a group of equal top bits (m <= kSplitGroupMax) goes to
// group start + its rank by (key, index) inside the group.  A wave owns 64
// consecutive positions and also loads the 64 before and the 64 after
// (coalesced; neighbouring waves read them too, from cache).  Group
// boundaries in those 192 positions are three ballots, so a group that
// starts and ends inside them is found with bit scans and ranked against
// its members in an LDS copy of the 192 positions, with no dependent global
// loads.  A pair whose group reaches past them takes the
// general path: galloping search for the bounds, members read from global
// memory (O(m) compares).  A group longer than kSplitGroupMax is listed once,
// by its first pair, for k_splitLong.
// PACKED: kf holds packed words and vals is unused; the full key of every
// pair outside a one-pair group is gathered from `low` (a one-pair group
// needs no compare).
template <bool PACKED>
__global__ __launch_bounds__(256) void k_splitRank(const uint64_t *__restrict__ kf, const uint32_t *__restrict__ vals,
                                                   const uint32_t *__restrict__ low,
                                                   uint32_t n, uint32_t *__restrict__ valsOut,
                                                   uint64_t *__restrict__ keysOut, uint32_t *__restrict__ ctrl,
                                                   uint2 *__restrict__ list, uint32_t listCap) {
  __shared__ uint64_t s_k[4][192];  // per wave: keys of positions P0 - 64 .. P0 + 127
  __shared__ uint32_t s_v[4][192];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t P0 = blockIdx.x * 256 + (threadIdx.x & ~63u);  // the wave's first position
  if (P0 >= n) return;                                          // wave-uniform
  // top bits given to positions below 0 and from n on (never a key's: keys have 63 bits)
  constexpr uint64_t kBelow = 1ull << 62, kAbove = 1ull << 61;
  uint64_t k[3], h[3];
  uint32_t v[3];
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    const int64_t pos = (int64_t)P0 + 64 * (w - 1) + lane;
    const bool in = pos >= 0 && pos < (int64_t)n;
    k[w] = in ? kf[pos] : 0ull;
    v[w] = PACKED ? packedIndex(k[w]) : (in ? vals[pos] : 0u);
    h[w] = in ? hiOf(k[w]) : (pos < 0 ? kBelow : kAbove);
  }
  // bit l of b[w]: a group starts at position P0 + 64 (w - 1) + l.  Whether
  // one starts at P0 - 64 is left open (0) unless that is position 0.
  uint64_t b[3];
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    // every lane takes part in the shuffle (no cross-lane read inside a
    // branch); lane 0 then takes the last top bits of the window before
    const uint64_t up = __shfl_up(h[w], 1, 64);
    const uint64_t before = w > 0 ? readLane64(h[w - 1], 63) : (P0 == 64 ? kBelow : h[0]);
    b[w] = __ballot(h[w] != (lane == 0 ? before : up));
  }
  if (PACKED) {
    // Full keys where a compare can read them: the wave's own positions
    // outside one-pair groups, and the members of the groups that reach
    // from the wave's first position back into the window before and from
    // its last position into the window after.
    const bool s0 = (b[1] >> lane) & 1ull;
    const bool s1 = lane < 63 ? ((b[1] >> (lane + 1)) & 1ull) : (b[2] & 1ull) != 0;
    if (P0 + lane < n && !(s0 && s1)) k[1] = packedKey(k[1], low);
    if (!(b[1] & 1ull) && b[0]) {  // the first group starts in the window before, at its last start
      const uint32_t from = 63 - __clzll((long long)b[0]);
      if (lane >= from) k[0] = packedKey(k[0], low);  // positions P0 - 64 + lane >= 0 here
    }
    if (!(b[2] & 1ull)) {  // the last group reaches into the window after, up to its first start
      const uint32_t upto = b[2] ? (uint32_t)__ffsll((long long)b[2]) - 1 : 64u;
      if (lane < upto && (uint64_t)P0 + 64 + lane < n) k[2] = packedKey(k[2], low);
    }
  }
  const bool valid = P0 + lane < n;
  const uint64_t key = k[1];
  const uint32_t val = v[1];
  // own group [gs, ge) in window coordinates (0 = position P0 - 64); -1: not inside
  const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1);  // bits 0..lane
  int gs = -1, ge = -1;
  if (b[1] & le)
    gs = 64 + (63 - __clzll((long long)(b[1] & le)));
  else if (b[0])
    gs = 63 - __clzll((long long)b[0]);
  if (b[1] & ~le)
    ge = 64 + __ffsll((long long)(b[1] & ~le)) - 1;
  else if (b[2])
    ge = 128 + __ffsll((long long)b[2]) - 1;
  else if ((uint64_t)P0 + 128 >= n)
    ge = (int)(n - (P0 - 64));  // the group runs to the end of the array
  const bool local = valid && gs >= 0 && ge >= 0;

  // each lane walks its own group's members in the wave's LDS copy of the
  // 192 positions (lanes of one group read the same words: broadcasts)
  uint64_t *wk = s_k[threadIdx.x >> 6];
  uint32_t *wv = s_v[threadIdx.x >> 6];
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    wk[64 * w + lane] = k[w];
    wv[64 * w + lane] = v[w];
  }
  const int lenMax = __builtin_amdgcn_readfirstlane(waveMaxInt(local ? ge - gs : 0));
  uint32_t rank = 0;
  for (int j = 0; j < lenMax; ++j) {
    const int q = gs + j;
    if (local && q < ge) {
      const uint64_t kq = wk[q];
      const uint32_t vq = wv[q];
      rank += (kq < key || (kq == key && vq < val)) ? 1u : 0u;
    }
  }
  if (!valid) return;
  const uint32_t p = P0 + lane;
  uint32_t start;
  if (local) {
    start = P0 - 64 + (uint32_t)gs;
  } else {
    const uint64_t hi = hiOf(key);
    const uint32_t g0 = groupStart(kf, p, hi), g1 = groupEnd(kf, n, p, hi);
    if (g1 - g0 > kSplitGroupMax) {
      if (p == g0) {
        const uint32_t slot = atomicAdd(ctrl + 0, 1u);
        if (slot < listCap)
          list[slot] = make_uint2(g0, g1 - g0);
        else
          atomicOr(ctrl + 2, kErrList);
      }
      return;
    }
    start = g0;
    rank = rankGlobal<PACKED>(kf, vals, low, g0, g1, key, val);
  }
  valsOut[start + rank] = val;
  // PACKED: `key` is already the full key unless the group has one pair
  if (keysOut) keysOut[start + rank] = PACKED ? ((key & ~kPackMask) | low[val]) : key;
}

// Long groups: one workgroup per group, a stable LSD over the low 24 bits
// (3 passes of 8 bits) through global memory.  The group is in input order
// on entry (the top-bit sort is stable), so the result is ordered by
// (key, index).  Pass 1 reads (kA, vA) and writes (kB, vB), pass 2 the
// reverse, pass 3 again (kA, vA) -> (kB, vB).  Chunks of
// kLongChunk pairs are ranked stably with wave ballots (item-major order in
// a wave, waves in order) and placed after the digit's running offset.
// PACKED: pass 1 reads packed words from kA (full keys from `low`).
template <bool PACKED>
__global__ __launch_bounds__(kSortBlock) void k_splitLong(uint64_t *kA, uint32_t *vA, uint64_t *kB, uint32_t *vB,
                                                          const uint32_t *__restrict__ low,
                                                          const uint2 *__restrict__ list, uint32_t *__restrict__ ctrl,
                                                          uint32_t listCap) {
  constexpr int WAVES = kSortBlock / 64;
  __shared__ uint32_t s_hist[kSortRadix], s_off[kSortRadix], s_wave[WAVES][kSortRadix], s_scan[WAVES];
  __shared__ uint32_t s_item;
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t num = min(ctrl[0], listCap);
  if (num == 0) return;  // the usual case: no long group, no work-queue atomic
  const uint64_t lt = __lanemask_lt();
  while (true) {
    if (tid == 0) s_item = atomicAdd(ctrl + 1, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    __syncthreads();  // s_item is rewritten only after this barrier
    if (item >= num) break;
    const uint2 g = list[item];
    const uint32_t s = g.x, m = g.y;
    for (int pass = 0; pass < kSplitFirstPass; ++pass) {
      const uint32_t shift = 8 * pass;
      const uint64_t *ki = (pass & 1) ? kB : kA;
      const uint32_t *vi = (pass & 1) ? vB : vA;
      uint64_t *ko = (pass & 1) ? kA : kB;
      uint32_t *vo = (pass & 1) ? vA : vB;
      s_hist[tid] = 0;
      __syncthreads();
      for (uint32_t c0 = 0; c0 < m; c0 += kLongChunk) {
        uint32_t d[kLongItems];
#pragma unroll
        for (int i = 0; i < kLongItems; ++i) {
          const uint32_t j = c0 + tid + i * kSortBlock;
          const uint64_t kj = j < m ? ki[s + j] : 0ull;
          d[i] = j < m ? (uint32_t)((PACKED && pass == 0 ? packedKey(kj, low) : kj) >> shift) & (kSortRadix - 1)
                       : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int i = 0; i < kLongItems; ++i)
          if (d[i] != 0xFFFFFFFFu) atomicAdd(&s_hist[d[i]], 1u);
      }
      __syncthreads();
      const uint32_t c = s_hist[tid];
      s_off[tid] = blockExclusiveScan<kSortBlock>(c, s_scan);
      __syncthreads();
      for (uint32_t c0 = 0; c0 < m; c0 += kLongChunk) {
        const uint32_t base = c0 + wave * 64 * kLongItems;
        uint64_t key[kLongItems];
        uint32_t val[kLongItems], rank[kLongItems];
#pragma unroll
        for (int i = 0; i < kLongItems; ++i) {
          const uint32_t j = base + i * 64 + lane;
          key[i] = j < m ? ki[s + j] : 0ull;
          if (PACKED && pass == 0) {
            val[i] = packedIndex(key[i]);
            if (j < m) key[i] = packedKey(key[i], low);
          } else {
            val[i] = j < m ? vi[s + j] : 0u;
          }
        }
        for (int w = 0; w < WAVES; ++w) s_wave[w][tid] = 0;
        __syncthreads();
        uint32_t *cnt = &s_wave[wave][0];
#pragma unroll
        for (int i = 0; i < kLongItems; ++i) {
          const bool valid = base + i * 64 + lane < m;
          const uint32_t dg = (uint32_t)(key[i] >> shift) & (kSortRadix - 1);
          const uint64_t peers = matchDigit<kSortRadixBits>(dg, __ballot(valid));
          const uint32_t lower = (uint32_t)__popcll(peers & lt);
          const uint32_t cc = __hip_atomic_load(cnt + dg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          rank[i] = cc + lower;
          if (valid && lower == 0)
            __hip_atomic_store(cnt + dg, cc + (uint32_t)__popcll(peers), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        __syncthreads();
        // per digit: exclusive offsets of the waves inside the chunk, chunk total
        uint32_t run = 0;
        for (int w = 0; w < WAVES; ++w) {
          const uint32_t x = s_wave[w][tid];
          s_wave[w][tid] = run;
          run += x;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kLongItems; ++i) {
          if (base + i * 64 + lane < m) {
            const uint32_t dg = (uint32_t)(key[i] >> shift) & (kSortRadix - 1);
            const uint32_t pos = s_off[dg] + s_wave[wave][dg] + rank[i];
            ko[s + pos] = key[i];
            vo[s + pos] = val[i];
          }
        }
        __syncthreads();
        s_off[tid] += run;  // the chunk's pairs of digit tid now precede the next chunk's
        __syncthreads();
      }
      __threadfence_block();
      __syncthreads();
    }
  }
}

}  // namespace

SplitPlan makeSplitPlan(uint32_t n) {
  SplitPlan p;
  p.n = n;
  p.plan = makeSortPlan(n, 8, 63);
  p.canPack = p.plan.wide && n < (1u << kSplitLoBits);
  p.listCap = n / (kSplitGroupMax + 1) + 1;
  p.ctrlOff = p.plan.totalWords;
  p.listOff = p.ctrlOff + 16;
  p.lowOff = p.listOff + 2 * (size_t)p.listCap;
  p.totalWords = p.lowOff + (p.canPack ? n : 0);
  return p;
}

bool splitReset(const SplitPlan &p, uint32_t *scratch) {
  return sortReset(p.plan, scratch) && fillU32(scratch + p.ctrlOff, 0u, 16);
}

int splitRun(const SplitPlan &p, uint64_t *keys[2], uint32_t *vals[2], uint32_t *scratch, bool keepKeys, bool packed,
             void (*afterPass)(int), void (*afterStage)(const char *)) {
  if (p.n == 0) return 0;
  if (packed && !p.canPack) {
    setError("packed 63-bit keys need wide sort tiles and fewer than 2^24 keys");
    return -1;
  }
  void *kv[2] = {keys[0], keys[1]};
  // packed words sort without values and carry the index in their low bits;
  // either way the group step reads the sorted keys
  const int r = sortRun(p.plan, kv, packed ? nullptr : vals, true, true, scratch, afterPass, kSplitFirstPass);
  if (r < 0) return -1;
  uint32_t *ctrl = scratch + p.ctrlOff;
  uint2 *list = reinterpret_cast<uint2 *>(scratch + p.listOff);
  const uint32_t *low = packed ? splitLow(p, scratch) : nullptr;
  uint64_t *keysOut = keepKeys ? keys[r ^ 1] : nullptr;
  if (packed)
    k_splitRank<true><<<divUp(p.n, 256), 256, 0, stream()>>>(keys[r], nullptr, low, p.n, vals[r ^ 1], keysOut, ctrl,
                                                             list, p.listCap);
  else
    k_splitRank<false><<<divUp(p.n, 256), 256, 0, stream()>>>(keys[r], vals[r], nullptr, p.n, vals[r ^ 1], keysOut,
                                                              ctrl, list, p.listCap);
  if (!checkLaunch("k_splitRank")) return -1;
  if (afterStage) afterStage("k_splitRank");
  // long groups are rare (none at the 4.24M-triangle scene): a small grid
  // keeps the launch cheap when the list is empty
  if (packed)
    k_splitLong<true><<<32, kSortBlock, 0, stream()>>>(keys[r], vals[r], keys[r ^ 1], vals[r ^ 1], low, list, ctrl,
                                                       p.listCap);
  else
    k_splitLong<false><<<32, kSortBlock, 0, stream()>>>(keys[r], vals[r], keys[r ^ 1], vals[r ^ 1], nullptr, list,
                                                        ctrl, p.listCap);
  if (!checkLaunch("k_splitLong")) return -1;
  if (afterStage) afterStage("k_splitLong");
  return r ^ 1;
}

}  // namespace bvht
