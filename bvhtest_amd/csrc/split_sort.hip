// Two-step sort of (63-bit key, input index) pairs (see split_sort.hpp).
#include "split_sort.hpp"
#include "wave.hpp"


namespace bvht {

namespace {

enum : uint32_t { kErrList = 1 };
constexpr int kLongItems = 16;  // pairs per thread per chunk of the long-group sort
constexpr uint32_t kLongChunk = kSortBlock * kLongItems;
// 64-position chunks per wave in k_splitRank (4.24M 63-bit keys: 1 chunk
// 31.6 us, 2 chunks 29.0, 4 chunks 35.1, 8 chunks 54.2)
constexpr int kSplitChunks = 2;

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
// group start + its rank by (key, index) inside the group.  A wave owns C
// chunks of 64 consecutive positions and also loads the 64 positions before
// its first chunk and the 64 after its last (coalesced; neighbouring waves
// read those too, from cache), so a chunk sees the 64 positions on either
// side of it.  Group boundaries are one ballot per 64 positions, so a group
// that starts and ends inside a chunk's 192 positions is found with bit
// scans and ranked against its members in an LDS copy of them, with no
// dependent global loads.  A pair whose group reaches past them takes the
// general path: galloping search for the bounds, members read from global
// memory (O(m) compares).  A group longer than kSplitGroupMax is listed once,
// by its first pair, for k_splitLong.  All of a wave's loads are issued
// before any of its chunks is ranked.
// PACKED: kf holds packed words and vals is unused; the full key of every
// pair outside a one-pair group is gathered from `low` (a one-pair group
// needs no compare), each word at most once per wave.
template <bool PACKED, int C>
__global__ __launch_bounds__(256) void k_splitRank(const uint64_t *__restrict__ kf, const uint32_t *__restrict__ vals,
                                                   const uint32_t *__restrict__ low,
                                                   uint32_t n, uint32_t *__restrict__ valsOut,
                                                   uint64_t *__restrict__ keysOut, uint32_t *__restrict__ ctrl,
                                                   uint2 *__restrict__ list, uint32_t listCap) {
  constexpr int W = C + 2;  // windows of 64 positions: one before, the wave's C chunks, one after
  __shared__ uint64_t s_k[4][64 * W];  // per wave: keys of positions P0 - 64 .. P0 + 64 (C + 1) - 1
  __shared__ uint32_t s_v[4][64 * W];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t P0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64u * C;  // the wave's first position
  if (P0 >= n) return;                                                   // wave-uniform
  // top bits given to positions below 0 and from n on (never a key's: keys have 63 bits)
  constexpr uint64_t kBelow = 1ull << 62, kAbove = 1ull << 61;
  uint64_t k[W], h[W];
  uint32_t v[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int64_t pos = (int64_t)P0 + 64 * (w - 1) + lane;
    const bool in = pos >= 0 && pos < (int64_t)n;
    k[w] = in ? kf[pos] : 0ull;
    v[w] = PACKED ? packedIndex(k[w]) : (in ? vals[pos] : 0u);
    h[w] = in ? hiOf(k[w]) : (pos < 0 ? kBelow : kAbove);
  }
  // bit l of b[w]: a group starts at position P0 + 64 (w - 1) + l.  Whether
  // one starts at P0 - 64 is left open (0) unless that is position 0.
  uint64_t b[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    // every lane takes part in the shuffle (no cross-lane read inside a
    // branch); lane 0 then takes the last top bits of the window before
    const uint64_t up = __shfl_up(h[w], 1, 64);
    const uint64_t before = w > 0 ? readLane64(h[w - 1], 63) : (P0 == 64 ? kBelow : h[0]);
    b[w] = __ballot(h[w] != (lane == 0 ? before : up));
  }
  if (PACKED) {
    // Full keys where a compare can read them.  For chunk c (window c + 1):
    // its positions outside one-pair groups, and the members of the groups
    // that reach from its first position back into window c and from its
    // last position into window c + 2.  The masks of all chunks are joined
    // first, so each word is gathered once.
    uint64_t need[W];
#pragma unroll
    for (int w = 0; w < W; ++w) need[w] = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const uint64_t bp = b[c], bo = b[c + 1], bn = b[c + 2];
      need[c + 1] |= ~(bo & ((bo >> 1) | (bn << 63)));  // not a start followed by a start
      if (!(bo & 1ull) && bp) need[c] |= ~0ull << (63 - __clzll((long long)bp));
      if (!(bn & 1ull)) need[c + 2] |= bn ? (1ull << (__ffsll((long long)bn) - 1)) - 1 : ~0ull;
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const int64_t pos = (int64_t)P0 + 64 * (w - 1) + lane;
      if (pos >= 0 && pos < (int64_t)n && ((need[w] >> lane) & 1ull)) k[w] = packedKey(k[w], low);
    }
  }
  uint64_t *wk = s_k[threadIdx.x >> 6];
  uint32_t *wv = s_v[threadIdx.x >> 6];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    wk[64 * w + lane] = k[w];
    wv[64 * w + lane] = v[w];
  }
  const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1);  // bits 0..lane
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const uint64_t bp = b[c], bo = b[c + 1], bn = b[c + 2];
    const uint32_t Pc = P0 + 64u * c;  // the chunk's first position
    const bool valid = Pc + lane < n;
    const uint64_t key = k[c + 1];
    const uint32_t val = v[c + 1];
    // own group [gs, ge) in chunk coordinates (0 = position Pc - 64); -1: not inside
    int gs = -1, ge = -1;
    if (bo & le)
      gs = 64 + (63 - __clzll((long long)(bo & le)));
    else if (bp)
      gs = 63 - __clzll((long long)bp);
    if (bo & ~le)
      ge = 64 + __ffsll((long long)(bo & ~le)) - 1;
    else if (bn)
      ge = 128 + __ffsll((long long)bn) - 1;
    else if ((uint64_t)Pc + 128 >= n)
      ge = (int)(n - (Pc - 64));  // the group runs to the end of the array
    const bool local = valid && gs >= 0 && ge >= 0;
    // each lane walks its own group's members (lanes of one group read the
    // same words: broadcasts)
    const int lenMax = __builtin_amdgcn_readfirstlane(waveMaxInt(local ? ge - gs : 0));
    uint32_t rank = 0;
    for (int j = 0; j < lenMax; ++j) {
      const int q = gs + j;
      if (local && q < ge) {
        const uint64_t kq = wk[64 * c + q];
        const uint32_t vq = wv[64 * c + q];
        rank += (kq < key || (kq == key && vq < val)) ? 1u : 0u;
      }
    }
    if (valid) {
      const uint32_t p = Pc + lane;
      uint32_t start = 0;
      bool place = true;
      if (local) {
        start = Pc - 64 + (uint32_t)gs;
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
          place = false;
        } else {
          start = g0;
          rank = rankGlobal<PACKED>(kf, vals, low, g0, g1, key, val);
        }
      }
      if (place) {
        valsOut[start + rank] = val;
        // PACKED: `key` is already the full key unless the group has one pair
        if (keysOut) keysOut[start + rank] = PACKED ? ((key & ~kPackMask) | low[val]) : key;
      }
    }
  }
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
  const uint32_t rankBlocks = divUp(p.n, 256 * kSplitChunks);
  if (packed)
    k_splitRank<true, kSplitChunks><<<rankBlocks, 256, 0, stream()>>>(keys[r], nullptr, low, p.n, vals[r ^ 1],
                                                                      keysOut, ctrl, list, p.listCap);
  else
    k_splitRank<false, kSplitChunks><<<rankBlocks, 256, 0, stream()>>>(keys[r], vals[r], nullptr, p.n, vals[r ^ 1],
                                                                       keysOut, ctrl, list, p.listCap);
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
