// Onesweep radix sort kernels (see radix_sort.hpp).
#include "radix_sort.hpp"
#include "wave.hpp"

namespace bvht {

namespace {

constexpr uint32_t kFlagAgg = 1u << 30;  // tile aggregate published
constexpr uint32_t kFlagInc = 2u << 30;  // inclusive prefix published
constexpr uint32_t kCountMask = (1u << 30) - 1;
// The wide pass's status words: two digits' 15-bit tile counts per word,
// bit 15 and bit 31 set when published.
constexpr uint32_t kHalfFlags = 0x80008000u, kHalfCount = 0x7FFFu;
constexpr uint32_t kWideStatusWords = kSortRadix / 2;  // per tile
static_assert(kWideTile <= kHalfCount, "a tile's count of one digit fits 15 bits");
constexpr uint32_t kLookWindow = 8;  // look-back statuses loaded per round trip

// Histogram of all passes: LDS counters per block, then one global atomic
// per (pass, digit) per block.
template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_histogram(const K *__restrict__ keys, uint32_t n, int passes,
                                                          uint32_t *__restrict__ hist) {
  __shared__ uint32_t s_hist[8][kSortRadix];
  for (int p = 0; p < passes; ++p) s_hist[p][threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * kSortBlock + threadIdx.x; i < n; i += gridDim.x * kSortBlock) {
    K k = keys[i];
    for (int p = 0; p < passes; ++p) atomicAdd(&s_hist[p][(uint32_t)(k >> (p * kSortRadixBits)) & (kSortRadix - 1)], 1u);
  }
  __syncthreads();
  for (int p = 0; p < passes; ++p) {
    uint32_t c = s_hist[p][threadIdx.x];
    if (c) atomicAdd(&hist[p * kSortRadix + threadIdx.x], c);
  }
}

// One onesweep pass.  Each block takes the next tile id from `counter`
// (tiles start in id order, so every tile a block waits for has started),
// ranks its keys stably with wave64 ballots, publishes per-digit counts,
// looks back over earlier tiles for its global offsets and writes the tile
// out digit-run by digit-run from LDS.
// Three workgroups per CU (168 VGPRs; 64-bit tiles spill 13 keys to scratch
// during the ranking, measured faster than fewer, smaller tiles: 41 vs 45 us
// per 4.24M-key pass; 32-bit keys at 20 per thread 40 vs 43 us at 15).
template <typename K, int ITEMS>
__global__ __launch_bounds__(kSortBlock, 3) void k_onesweep(const K *__restrict__ keysIn, const uint32_t *__restrict__ valsIn,
                                                         K *__restrict__ keysOut, uint32_t *__restrict__ valsOut,
                                                         uint32_t n, uint32_t shift, const uint32_t *__restrict__ hist,
                                                         uint32_t *status, uint32_t *counter, uint32_t *err) {
  constexpr int WAVES = kSortBlock / 64;
  constexpr int TILE = kSortBlock * ITEMS;
  static_assert(kSortBlock == kSortRadix, "one digit per thread");
  __shared__ uint32_t s_wave[WAVES][kSortRadix];
  __shared__ uint32_t s_bstart[kSortRadix];
  __shared__ uint32_t s_gstart[kSortRadix];
  __shared__ uint32_t s_scan[WAVES];
  __shared__ uint32_t s_tile;
  // staging for the scatter: the keys, then the values in the same words,
  // so a 64-bit tile takes 32 KiB of LDS and four workgroups fit on a CU
  __shared__ K s_buf[TILE];

  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (tid == 0) s_tile = atomicAdd(counter, 1u);
  for (int w = 0; w < WAVES; ++w) s_wave[w][tid] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t base = tile * TILE + wave * 64 * ITEMS;

  // the keys now; the values only once the keys are staged (they are not
  // needed before, and their loads then overlap the look-back)
  // (a tile inside the array loads without per-item bounds branches)
  const bool full = (tile + 1) * TILE <= n;
  K key[ITEMS];
  uint32_t rank[ITEMS];
  if (full) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) key[i] = keysIn[base + i * 64 + lane];
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t idx = base + i * 64 + lane;
      key[i] = idx < n ? keysIn[idx] : (K)0;
    }
  }

  // Stable in-wave ranking: item-major order inside a wave's segment equals
  // tile order.  Peers (same digit) are found with 8 ballots (matchDigit); the lowest
  // peer bumps the wave's counter for the digit.  The counters are read and
  // written with relaxed workgroup-scope atomics: plain LDS instructions the
  // compiler may not cache or reorder (a volatile pointer would become a
  // generic one and compile to FLAT accesses that also wait on global loads).
  uint32_t *cnt = &s_wave[wave][0];
  const uint64_t lt = __lanemask_lt();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    bool valid = base + i * 64 + lane < n;
    uint32_t d = (uint32_t)(key[i] >> shift) & (kSortRadix - 1);
    const uint64_t peers = matchDigit<kSortRadixBits>(d, __ballot(valid));
    uint32_t lower = (uint32_t)__popcll(peers & lt);
    uint32_t c = __hip_atomic_load(cnt + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    rank[i] = c + lower;
    if (valid && lower == 0)
      __hip_atomic_store(cnt + d, c + (uint32_t)__popcll(peers), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();

  // Thread tid owns digit tid: per-wave exclusive offsets and tile total.
  uint32_t total = 0;
  for (int w = 0; w < WAVES; ++w) {
    uint32_t c = s_wave[w][tid];
    s_wave[w][tid] = total;
    total += c;
  }
  uint32_t *st = status + (size_t)tile * kSortRadix + tid;
  storeRelaxedAgent(st, (tile == 0 ? kFlagInc : kFlagAgg) | total);
  const uint32_t hbase = blockExclusiveScan<kSortBlock>(hist[tid], s_scan);
  const uint32_t bstart = blockExclusiveScan<kSortBlock>(total, s_scan);
  s_bstart[tid] = bstart;
  __syncthreads();

  // Keys to their tile positions before the look-back (their registers are
  // free while it waits); rank[i] becomes the item's position in the tile.
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (base + i * 64 + lane < n) {
      const uint32_t d = (uint32_t)(key[i] >> shift) & (kSortRadix - 1);
      rank[i] += s_bstart[d] + s_wave[wave][d];
      s_buf[rank[i]] = key[i];
    }
  }
  uint32_t val[ITEMS];
  if (!valsIn) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) val[i] = base + i * 64 + lane;
  } else if (full) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) val[i] = valsIn[base + i * 64 + lane];
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t idx = base + i * 64 + lane;
      val[i] = idx < n ? valsIn[idx] : 0u;
    }
  }

  uint32_t excl = 0;
  if (tile > 0) {
    // Look back over a window of kLookWindow predecessors per round trip:
    // their words are loaded together (newest first) and consumed in order
    // up to the first one not yet published, where the next round resumes.
    constexpr uint32_t W = kLookWindow;
    uint32_t next = tile, spins = 0;
    bool done = false;
    while (!done) {
      uint32_t sv[W];
#pragma unroll
      for (uint32_t w = 0; w < W; ++w)  // tile 0 publishes an inclusive count at once: nothing lies before it
        sv[w] = next >= w + 1 ? loadRelaxedAgent(status + (size_t)(next - 1 - w) * kSortRadix + tid) : kFlagInc;
      uint32_t used = 0;
#pragma unroll
      for (uint32_t w = 0; w < W; ++w) {
        if (done || used != w) continue;
        const uint32_t f = sv[w] & ~kCountMask;
        if (f == 0) continue;  // not yet published: resume here
        excl += sv[w] & kCountMask;
        used = w + 1;
        if (f == kFlagInc) done = true;
      }
      next -= used;
      if (!done && used < W) {
        // bounded wait: a predecessor that never publishes (impossible when
        // the tile counter starts at zero) ends the pass with an error flag
        // instead of a hang
        if (++spins > (1u << 24)) {
          atomicOr(err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    storeRelaxedAgent(st, kFlagInc | (excl + total));
  }
  s_gstart[tid] = hbase + excl;
  __syncthreads();

  // each staged key gives its position's output index (kept in registers
  // for the values)
  const uint32_t tileCount = min((uint32_t)TILE, n - tile * TILE);
  uint32_t out[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t j = tid + k * kSortBlock;
    out[k] = n;  // n marks "nothing to write"
    if (j < tileCount) {
      const K kk = s_buf[j];
      const uint32_t d = (uint32_t)(kk >> shift) & (kSortRadix - 1);
      out[k] = s_gstart[d] + (j - s_bstart[d]);
      if (out[k] >= n) out[k] = n;  // only reachable after a bounded-wait error
      if (keysOut && out[k] < n) keysOut[out[k]] = kk;
    }
  }
  __syncthreads();
  uint32_t *s_v = reinterpret_cast<uint32_t *>(s_buf);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if (base + i * 64 + lane < n) s_v[rank[i]] = val[i];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < ITEMS; ++k)
    if (out[k] < n) valsOut[out[k]] = s_v[tid + k * kSortBlock];
}

// Wide tiles: workgroups of 1024 threads x 17 keys (17,408 keys per tile),
// used when the tiles of a pass fit the CUs in one round (244 tiles at 4.24M
// keys: a dispatch ramp of 244 workgroups instead of 753).  Ranking and
// scatter are those of k_onesweep (16 waves, 256 digit owners).
//  - Global accesses are buffer loads and stores with 32-bit byte offsets
//    (the wide path only runs below 2^23 keys): one address register per
//    access instead of a 64-bit pointer per item, which keeps the kernel
//    within 128 VGPRs without spilling.  Loads past the end of an array read
//    0 and stores past it are dropped by the buffer's range check.
//  - A 64-bit key is loaded whole and ranked on the 32-bit word that holds
//    the pass's digit (a digit never straddles the words).
//  - VALS: kValsLoaded (values move with the keys), kValsIdentity (the
//    first pass: values are the input index, made when they are staged) or
//    kValsNone (a keys-only sort).
//  - The tile's global offsets are the sums of every earlier tile's
//    published per-digit counts, loaded by all 1024 threads at once instead
//    of a chained look-back.  The counts are published two digits to a word
//    (15 bits each), so a 16-byte load brings eight digits and, with 8 loads
//    in flight per thread, one round trip covers 256 predecessors.
//    Tile ids come from `counter`, so every tile a workgroup waits for has
//    started; the wait for a count is bounded and raises the error flag.
enum { kValsLoaded = 0, kValsIdentity = 1, kValsNone = 2 };
template <typename K, int VALS>
__global__ __launch_bounds__(kWideBlock, 1) void k_onesweepWide(const K *__restrict__ keysIn,
                                                             const uint32_t *__restrict__ valsIn,
                                                             K *__restrict__ keysOut, uint32_t *__restrict__ valsOut,
                                                             uint32_t n, uint32_t shift,
                                                             const uint32_t *__restrict__ hist,
                                                             uint32_t *status,
                                                             uint32_t *counter, uint32_t *err) {
  constexpr int WAVES = kWideBlock / 64;
  constexpr int ITEMS = kWideItems;
  constexpr uint32_t TILE = kWideTile;
  constexpr bool WIDE = sizeof(K) == 8;
  __shared__ uint32_t s_wave[WAVES][kSortRadix];  // per-wave digit counters, then prefix partials
  __shared__ uint32_t s_bstart[kSortRadix];
  __shared__ uint32_t s_gstart[kSortRadix];
  __shared__ uint32_t s_scan[WAVES];
  __shared__ uint32_t s_tile;
  __shared__ K s_buf[TILE];  // the staged keys, then the values in the same words

  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (tid == 0) s_tile = atomicAdd(counter, 1u);
  for (uint32_t i = tid; i < (uint32_t)WAVES * kSortRadix; i += kWideBlock) (&s_wave[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t first = tile * TILE + wave * 64 * ITEMS + lane;  // item i is key first + 64 i
  const uint32_t histDigit = tid < kSortRadix ? hist[tid] : 0u;  // issued now, used after the ranking

  const BufferRsrc keys = bufferRsrc(keysIn, n * (uint32_t)sizeof(K));
  const uint32_t dw = (WIDE && shift >= 32) ? 1u : 0u;  // the digit's word
  const uint32_t dshift = WIDE ? (shift & 31u) : shift;
  uint32_t kw[ITEMS], rank[ITEMS], other[WIDE ? ITEMS : 1];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (WIDE) {
      const BufferU32x2 k2 = __builtin_amdgcn_raw_buffer_load_b64(keys, (first + i * 64) * 8u, 0, 0);
      kw[i] = dw ? k2[1] : k2[0];
      other[WIDE ? i : 0] = dw ? k2[0] : k2[1];
    } else {
      kw[i] = bufferLoadU32(keys, (first + i * 64) * 4u);
    }
  }

  // stable in-wave ranking, as in k_onesweep
  uint32_t *cnt = &s_wave[wave][0];
  const uint64_t lt = __lanemask_lt();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const bool valid = first + i * 64 < n;
    const uint32_t d = (kw[i] >> dshift) & (kSortRadix - 1);
    const uint64_t peers = matchDigit<kSortRadixBits>(d, __ballot(valid));
    const uint32_t lower = (uint32_t)__popcll(peers & lt);
    const uint32_t c = __hip_atomic_load(cnt + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    rank[i] = c + lower;
    if (valid && lower == 0)
      __hip_atomic_store(cnt + d, c + (uint32_t)__popcll(peers), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();

  // digit owners (tid < 256): per-wave exclusive offsets, the tile's count,
  // published at once
  uint32_t total = 0;
  if (tid < kSortRadix) {
    for (int w = 0; w < WAVES; ++w) {
      const uint32_t c = s_wave[w][tid];
      s_wave[w][tid] = total;
      total += c;
    }
    // two digits per word: 15-bit counts (a tile holds at most kWideTile
    // keys), each half flagged when published
    const uint32_t odd = __shfl_xor(total, 1, 64);  // waves 0..3, all lanes
    if ((tid & 1u) == 0)
      storeRelaxedAgent(status + (size_t)tile * kWideStatusWords + tid / 2, kHalfFlags | total | (odd << 16));
  }
  const uint32_t hbase = blockExclusiveScan<kWideBlock>(histDigit, s_scan);
  const uint32_t bstart = blockExclusiveScan<kWideBlock>(total, s_scan);
  if (tid < kSortRadix) s_bstart[tid] = bstart;
  __syncthreads();

  // keys to their tile positions; rank[i] becomes the item's tile position
  // (digits and item indices are recomputed here, not kept from the ranking)
  uint32_t first2 = first;
  recomputeAfter(first2);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) recomputeAfter(kw[i]);
  if (WIDE) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if (first2 + i * 64 < n) {
        const uint32_t d = (kw[i] >> dshift) & (kSortRadix - 1);
        rank[i] += s_bstart[d] + s_wave[wave][d];
        s_buf[rank[i]] = (K)(dw ? (((uint64_t)kw[i] << 32) | other[WIDE ? i : 0]) : (((uint64_t)other[WIDE ? i : 0] << 32) | kw[i]));
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if (first2 + i * 64 < n) {
        const uint32_t d = (kw[i] >> dshift) & (kSortRadix - 1);
        rank[i] += s_bstart[d] + s_wave[wave][d];
        s_buf[rank[i]] = (K)kw[i];
      }
    }
  }
  uint32_t val[VALS == kValsLoaded ? ITEMS : 1];
  if (VALS == kValsLoaded) {
    const BufferRsrc vals = bufferRsrc(valsIn, n * 4u);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) val[VALS == kValsLoaded ? i : 0] = bufferLoadU32(vals, (first2 + i * 64) * 4u);
  }
  __syncthreads();  // s_wave is free from here: it holds the prefix partials

  // offsets: the counts of every earlier tile.  Thread (g, q) sums digits
  // 8q .. 8q + 7 over predecessors g, g + 32, ... with 16-byte coherent
  // loads, 8 in flight (one round trip for up to 256 predecessors); the 32
  // partial sums of a digit are then added in two steps.
  {
    constexpr uint32_t OCTS = kSortRadix / 8, QG = kWideBlock / OCTS, QB = 8;
    static_assert(QG == 2 * WAVES, "two look-back groups per row of s_wave");
    const uint32_t q = tid % OCTS, g = tid / OCTS;
    const BufferRsrc st = bufferRsrc(status, tile * kWideStatusWords * 4u);  // the earlier tiles' words
    uint32_t sum[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u}, spins = 0;
    for (uint32_t j0 = g; j0 < tile; j0 += QG * QB) {
      BufferU32x4 sv[QB];
#pragma unroll
      for (uint32_t b = 0; b < QB; ++b) {
        const uint32_t j = j0 + b * QG;
        sv[b] = j < tile ? bufferLoadU32x4Coherent(st, (j * kWideStatusWords + 4 * q) * 4u)
                         : BufferU32x4{kHalfFlags, kHalfFlags, kHalfFlags, kHalfFlags};
      }
#pragma unroll
      for (uint32_t b = 0; b < QB; ++b) {
        const uint32_t j = j0 + b * QG;
        BufferU32x4 v = sv[b];
        // not yet published (j < tile: tile 0 never waits); the wait is bounded
        while (((v[0] & v[1] & v[2] & v[3]) & kHalfFlags) != kHalfFlags) {
          if (++spins > (1u << 24)) {
            atomicOr(err, 1u);
            v = BufferU32x4{kHalfFlags, kHalfFlags, kHalfFlags, kHalfFlags};
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          memoryChanged();
          v = bufferLoadU32x4Coherent(st, (j * kWideStatusWords + 4 * q) * 4u);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          sum[2 * c] += v[c] & kHalfCount;
          sum[2 * c + 1] += (v[c] >> 16) & kHalfCount;
        }
      }
    }
    if (g >= WAVES) {
#pragma unroll
      for (int c = 0; c < 8; ++c) s_wave[g - WAVES][8 * q + c] = sum[c];
    }
    __syncthreads();
    if (g < WAVES) {
#pragma unroll
      for (int c = 0; c < 8; ++c) s_wave[g][8 * q + c] += sum[c];
    }
  }
  __syncthreads();
  if (tid < kSortRadix) {
    uint32_t excl = 0;
#pragma unroll
    for (uint32_t g = 0; g < WAVES; ++g) excl += s_wave[g][tid];
    s_gstart[tid] = hbase + excl;
  }
  __syncthreads();

  const uint32_t tileCount = min(TILE, n - tile * TILE);
  const BufferRsrc kout = bufferRsrc(keysOut, keysOut ? n * (uint32_t)sizeof(K) : 0u);
  const BufferRsrc vout = bufferRsrc(valsOut, n * 4u);
  uint32_t out[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t j = tid + k * kWideBlock;
    out[k] = n;
    if (j < tileCount) {
      const K kk = s_buf[j];
      const uint32_t d = (uint32_t)(kk >> shift) & (kSortRadix - 1);
      out[k] = s_gstart[d] + (j - s_bstart[d]);
      if (out[k] >= n) out[k] = n;  // only reachable after a bounded-wait error
      if (keysOut && out[k] < n) bufferStore(kk, kout, out[k] * (uint32_t)sizeof(K));
    }
  }
  if (VALS == kValsNone) return;
  __syncthreads();
  uint32_t *s_v = reinterpret_cast<uint32_t *>(s_buf);
  uint32_t first3 = first;
  recomputeAfter(first3);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if (first3 + i * 64 < n) s_v[rank[i]] = VALS == kValsIdentity ? first3 + i * 64 : val[VALS == kValsLoaded ? i : 0];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < ITEMS; ++k)
    if (out[k] < n) bufferStore(s_v[tid + k * kWideBlock], vout, out[k] * 4u);
}

}  // namespace

SortPlan makeSortPlan(uint32_t n, int keyBytes, int keyBits) {
  SortPlan p;
  p.n = n;
  p.keyBytes = keyBytes;
  p.keyBits = keyBits;
  p.passes = (keyBits + kSortRadixBits - 1) / kSortRadixBits;
  // wide tiles when they fit the CUs in one round of workgroups (their
  // 32-bit byte offsets need n * 8 < 2^31: true for any device below 30,000 CUs)
  const uint32_t wideTiles = n == 0 ? 0 : divUp(n, kWideTile);
  p.wide = n > 0 && wideTiles <= (uint32_t)numCUs() && (uint64_t)n * 8 < (1ull << 31);
  p.tile = p.wide ? kWideTile : kSortBlock * (keyBytes == 8 ? kSortItems64 : kSortItems32);
  p.numTiles = n == 0 ? 0 : divUp(n, p.tile);
  p.histWords = (size_t)p.passes * kSortRadix;
  p.counterWords = 16;  // a tile counter per pass (at most 8), then the error word
  p.statusWords = (size_t)p.passes * p.numTiles * kSortRadix;
  p.histOff = 0;
  p.counterOff = p.histOff + p.histWords;
  p.statusOff = p.counterOff + p.counterWords;
  p.errOff = p.counterOff + 8;
  p.totalWords = p.statusOff + p.statusWords;
  return p;
}

bool sortReset(const SortPlan &p, uint32_t *scratch) { return fillU32(scratch, 0u, p.totalWords); }

bool sortHistogram(const SortPlan &p, const void *keys, uint32_t *scratch) {
  if (p.n == 0) return true;
  uint32_t blocks = divUp(p.n, kSortBlock * 16);
  if (blocks > (uint32_t)numCUs() * 4) blocks = numCUs() * 4;
  if (p.keyBytes == 8)
    k_histogram<uint64_t><<<blocks, kSortBlock, 0, stream()>>>((const uint64_t *)keys, p.n, p.passes, scratch + p.histOff);
  else
    k_histogram<uint32_t><<<blocks, kSortBlock, 0, stream()>>>((const uint32_t *)keys, p.n, p.passes, scratch + p.histOff);
  return checkLaunch("sortHistogram");
}

int sortRun(const SortPlan &p, void *keys[2], uint32_t *vals[2], bool identityValues, bool keepKeys, uint32_t *scratch,
            void (*afterPass)(int), int firstPass) {
  if (p.n == 0) return 0;
  if (((uintptr_t)(scratch + p.statusOff) & 15) != 0) {  // the wide pass reads statuses 16 bytes at a time
    setError("sort scratch must be 16-byte aligned");
    return -1;
  }
  const bool keysOnly = vals == nullptr;
  if (keysOnly && !(p.wide && p.keyBytes == 8 && keepKeys)) {
    setError("keys-only sort: wide tiles and 64-bit keys only");
    return -1;
  }
  int cur = 0;
  for (int pass = firstPass; pass < p.passes; ++pass) {
    const uint32_t *hist = scratch + p.histOff + (size_t)pass * kSortRadix;
    uint32_t *counter = scratch + p.counterOff + pass;
    uint32_t *err = scratch + p.errOff;  // one word for every pass
    uint32_t *status = scratch + p.statusOff + (size_t)pass * p.numTiles * kSortRadix;
    const uint32_t *vin = (keysOnly || (pass == firstPass && identityValues)) ? nullptr : vals[cur];
    bool last = pass == p.passes - 1;
    void *kout = (last && !keepKeys) ? nullptr : keys[cur ^ 1];
    uint32_t shift = pass * kSortRadixBits;
    if (keysOnly)
      k_onesweepWide<uint64_t, kValsNone><<<p.numTiles, kWideBlock, 0, stream()>>>(
          (const uint64_t *)keys[cur], nullptr, (uint64_t *)kout, nullptr, p.n, shift, hist, status, counter, err);
    else if (p.wide && p.keyBytes == 8 && !vin)
      k_onesweepWide<uint64_t, kValsIdentity><<<p.numTiles, kWideBlock, 0, stream()>>>(
          (const uint64_t *)keys[cur], vin, (uint64_t *)kout, vals[cur ^ 1], p.n, shift, hist, status, counter, err);
    else if (p.wide && p.keyBytes == 8)
      k_onesweepWide<uint64_t, kValsLoaded><<<p.numTiles, kWideBlock, 0, stream()>>>(
          (const uint64_t *)keys[cur], vin, (uint64_t *)kout, vals[cur ^ 1], p.n, shift, hist, status, counter, err);
    else if (p.wide && !vin)
      k_onesweepWide<uint32_t, kValsIdentity><<<p.numTiles, kWideBlock, 0, stream()>>>(
          (const uint32_t *)keys[cur], vin, (uint32_t *)kout, vals[cur ^ 1], p.n, shift, hist, status, counter, err);
    else if (p.wide)
      k_onesweepWide<uint32_t, kValsLoaded><<<p.numTiles, kWideBlock, 0, stream()>>>(
          (const uint32_t *)keys[cur], vin, (uint32_t *)kout, vals[cur ^ 1], p.n, shift, hist, status, counter, err);
    else if (p.keyBytes == 8)
      k_onesweep<uint64_t, kSortItems64><<<p.numTiles, kSortBlock, 0, stream()>>>(
          (const uint64_t *)keys[cur], vin, (uint64_t *)kout, vals[cur ^ 1], p.n, shift, hist, status, counter, err);
    else
      k_onesweep<uint32_t, kSortItems32><<<p.numTiles, kSortBlock, 0, stream()>>>(
          (const uint32_t *)keys[cur], vin, (uint32_t *)kout, vals[cur ^ 1], p.n, shift, hist, status, counter, err);
    if (!checkLaunch("onesweep pass")) return -1;
    if (afterPass) afterPass(pass);
    cur ^= 1;
  }
  return cur;
}

}  // namespace bvht
