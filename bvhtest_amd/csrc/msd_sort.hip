// Most-significant-digit sort of (key, input index) pairs (see msd_sort.hpp).
#include "msd_sort.hpp"
#include "radix_sort.hpp"
#include "split_sort.hpp"

#include <algorithm>

namespace bvht {

namespace {

constexpr int kSortThreads = 256;  // workgroup of the LDS sorts
constexpr int kRunItems = 16;      // pairs per thread in the LDS sort
static_assert(kSortThreads * kRunItems == (int)kMsdLocal, "one run of 16 per thread covers a piece");
constexpr uint32_t kPadded = kMsdLocal + kMsdLocal / 16;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kScanThreads = 1024;
constexpr int kScanBlocks = kMsdBuckets / kScanThreads;
static_assert(kScanBlocks <= 64, "one wave sums the block totals");
constexpr int kScatterGroups = 4;

enum : uint32_t { kErrCount = 1, kErrDepth = 2, kErrPiece = 4 };

// One padding slot after every 16 pairs: the 16 pairs a thread owns start
// 17 slots apart, so the lanes of a wave hit different banks.
__device__ __forceinline__ uint32_t padIdx(uint32_t i) { return i + (i >> 4); }

template <typename K>
__device__ __forceinline__ bool pairLess(K ka, uint32_t ia, K kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

template <typename K>
__device__ __forceinline__ void orderPair(K &ka, uint32_t &ia, K &kb, uint32_t &ib) {
  if (pairLess(kb, ib, ka, ia)) {
    const K tk = ka;
    ka = kb;
    kb = tk;
    const uint32_t ti = ia;
    ia = ib;
    ib = ti;
  }
}

// Bitonic network over a thread's 16 registers, ascending.
template <typename K>
__device__ __forceinline__ void sortRegisters(K (&k)[kRunItems], uint32_t (&x)[kRunItems]) {
#pragma unroll
  for (int size = 2; size <= kRunItems; size <<= 1)
#pragma unroll
    for (int j = size >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < kRunItems; ++i) {
        const int l = i ^ j;
        if (l > i) {
          if ((i & size) == 0)
            orderPair(k[i], x[i], k[l], x[l]);
          else
            orderPair(k[l], x[l], k[i], x[i]);
        }
      }
}

template <typename K>
struct SortLds {
  K key[kPadded];
  uint32_t idx[kPadded];
};

// Sorts the m <= kMsdLocal pairs src[start, start + m) by (key, index) into
// out[start, start + m) (src may be out).  Workgroup-uniform: barriers.
// Runs of 16 are sorted in registers, then merged pairwise in LDS (merge
// path: each thread finds where its 16 outputs start in the two runs and
// merges them serially).
template <typename K>
__device__ void blockSort(SortLds<K> &s, const K *srcK, const uint32_t *srcV, K *outK, uint32_t *outV, bool writeKeys,
                          uint32_t start, uint32_t m, uint32_t *ctrl) {
  const uint32_t tid = threadIdx.x;
  if (m > kMsdLocal) {  // excluded by the piece construction; never overrun LDS
    if (tid == 0) atomicOr(ctrl + 2, kErrPiece);
    return;
  }
  uint32_t P = kRunItems;
  while (P < m) P <<= 1;
  {
    // every global load of the piece is issued before the first is used
    K lk[kRunItems];
    uint32_t lv[kRunItems];
#pragma unroll
    for (int j = 0; j < kRunItems; ++j) {
      const uint32_t i = tid + j * kSortThreads;
      const bool v = i < m;
      lk[j] = v ? srcK[start + i] : ~(K)0;  // padding sorts last: real keys have a clear top bit
      lv[j] = v ? srcV[start + i] : kNone;
    }
#pragma unroll
    for (int j = 0; j < kRunItems; ++j) {
      const uint32_t i = tid + j * kSortThreads;
      if (i < P) {
        s.key[padIdx(i)] = lk[j];
        s.idx[padIdx(i)] = lv[j];
      }
    }
  }
  __syncthreads();
  const uint32_t first = tid * kRunItems;
  const bool active = first < P;
  K k[kRunItems];
  uint32_t x[kRunItems];
  if (active) {
#pragma unroll
    for (int j = 0; j < kRunItems; ++j) {
      k[j] = s.key[padIdx(first + j)];
      x[j] = s.idx[padIdx(first + j)];
    }
    sortRegisters(k, x);
  }
  for (uint32_t L = kRunItems; L < P; L <<= 1) {
    __syncthreads();
    if (active) {
#pragma unroll
      for (int j = 0; j < kRunItems; ++j) {
        s.key[padIdx(first + j)] = k[j];
        s.idx[padIdx(first + j)] = x[j];
      }
    }
    __syncthreads();
    if (active) {
      const uint32_t aBeg = first & ~(2 * L - 1), bBeg = aBeg + L, diag = first - aBeg;
      uint32_t lo = diag > L ? diag - L : 0, hi = diag < L ? diag : L;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t bi = padIdx(bBeg + diag - 1 - mid), ai = padIdx(aBeg + mid);
        if (!pairLess(s.key[bi], s.idx[bi], s.key[ai], s.idx[ai]))
          lo = mid + 1;
        else
          hi = mid;
      }
      uint32_t a = lo, b = diag - lo;
      K ka = 0, kb = 0;
      uint32_t ia = 0, ib = 0;
      if (a < L) {
        ka = s.key[padIdx(aBeg + a)];
        ia = s.idx[padIdx(aBeg + a)];
      }
      if (b < L) {
        kb = s.key[padIdx(bBeg + b)];
        ib = s.idx[padIdx(bBeg + b)];
      }
#pragma unroll
      for (int j = 0; j < kRunItems; ++j) {
        const bool takeA = b >= L || (a < L && !pairLess(kb, ib, ka, ia));
        if (takeA) {
          k[j] = ka;
          x[j] = ia;
          if (++a < L) {
            ka = s.key[padIdx(aBeg + a)];
            ia = s.idx[padIdx(aBeg + a)];
          }
        } else {
          k[j] = kb;
          x[j] = ib;
          if (++b < L) {
            kb = s.key[padIdx(bBeg + b)];
            ib = s.idx[padIdx(bBeg + b)];
          }
        }
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int j = 0; j < kRunItems; ++j) {
      s.key[padIdx(first + j)] = k[j];
      s.idx[padIdx(first + j)] = x[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kRunItems; ++j) {
    const uint32_t i = tid + j * kSortThreads;
    if (i < m) {
      outV[start + i] = s.idx[padIdx(i)];
      if (writeKeys) outK[start + i] = s.key[padIdx(i)];
    }
  }
  __syncthreads();
}

// Totals of the 64 blocks of 1024 top-digit bins.
__global__ __launch_bounds__(kScanThreads) void k_msdScanTotals(const uint32_t *__restrict__ hist,
                                                                uint32_t *__restrict__ totals) {
  __shared__ uint32_t s_tmp[kScanThreads / 64];
  const uint32_t v = hist[blockIdx.x * kScanThreads + threadIdx.x];
  const uint32_t ex = blockExclusiveScan<kScanThreads>(v, s_tmp);
  if (threadIdx.x == kScanThreads - 1) totals[blockIdx.x] = ex + v;
}

// Bucket starts (exclusive scan of the top-digit histogram), the cursors of
// the scatter, and for every window the bucket holding its first position.
__global__ __launch_bounds__(kScanThreads) void k_msdScan(uint32_t n, uint32_t windows,
                                                          const uint32_t *__restrict__ hist,
                                                          const uint32_t *__restrict__ totals,
                                                          uint32_t *__restrict__ starts, uint32_t *__restrict__ cursors,
                                                          uint32_t *__restrict__ winBucket, uint32_t *__restrict__ ctrl) {
  __shared__ uint32_t s_tmp[kScanThreads / 64];
  __shared__ uint32_t s_prefix;
  const uint32_t t = threadIdx.x, blk = blockIdx.x, bin = blk * kScanThreads + t;
  if (t < 64) {
    uint32_t x = t < blk ? totals[t] : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (t == 0) s_prefix = x;
  }
  const uint32_t c = hist[bin];
  const uint32_t ex = blockExclusiveScan<kScanThreads>(c, s_tmp) + s_prefix;  // the scan's barriers order s_prefix
  starts[bin] = ex;
  cursors[bin] = ex;
  if (c) {
    for (uint32_t w = divUpDev(ex, kMsdWindow); w < windows && w * kMsdWindow < ex + c; ++w) winBucket[w] = bin;
  }
  if (bin == kMsdBuckets - 1) {
    starts[kMsdBuckets] = ex + c;
    if (ex + c != n) atomicOr(ctrl + 2, kErrCount);  // the later stages then do nothing
  }
}

// Each pair to its bucket.  A wave takes kScatterGroups groups of 64
// consecutive pairs; consecutive lanes with the same bucket share one cursor
// atomic, and every load and atomic of the wave is issued before the first
// result is used.
template <typename K>
__global__ __launch_bounds__(256) void k_msdScatter(const K *__restrict__ keys, uint32_t n, int shift,
                                                    uint32_t *__restrict__ cursors, K *__restrict__ tmpK,
                                                    uint32_t *__restrict__ tmpV, const uint32_t *__restrict__ ctrl) {
  if (ctrl[2] & kErrCount) return;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t base = (blockIdx.x * 4 + wave) * 64 * kScatterGroups;
  if (base >= n) return;
  K key[kScatterGroups];
#pragma unroll
  for (int u = 0; u < kScatterGroups; ++u) {
    const uint32_t i = base + u * 64 + lane;
    key[u] = i < n ? keys[i] : (K)0;
  }
  uint32_t pos[kScatterGroups], head[kScatterGroups];
#pragma unroll
  for (int u = 0; u < kScatterGroups; ++u) {
    pos[u] = 0;
    head[u] = 0;
    if (base + u * 64 >= n) continue;  // wave-uniform; lane 0 of a started group is valid
    const bool valid = base + u * 64 + lane < n;
    const uint32_t b = (uint32_t)(key[u] >> shift) & (kMsdBuckets - 1);
    const uint32_t prev = __shfl_up(b, 1, 64);
    const bool isHead = valid && (lane == 0 || prev != b);
    const uint64_t heads = __ballot(isHead);
    const uint64_t validMask = __ballot(valid);
    if (isHead) {
      const uint64_t above = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
      const uint32_t end = above ? (uint32_t)__builtin_ctzll(above) : (uint32_t)__popcll(validMask);
      pos[u] = atomicAdd(cursors + b, end - lane);
    }
    const uint64_t upTo = lane == 63 ? heads : heads & ((2ull << lane) - 1);
    head[u] = 63u - (uint32_t)__builtin_clzll(upTo);
  }
#pragma unroll
  for (int u = 0; u < kScatterGroups; ++u) {
    if (base + u * 64 >= n) continue;
    const uint32_t i = base + u * 64 + lane;
    const uint32_t p = __shfl(pos[u], (int)head[u], 64) + (lane - head[u]);
    if (i < n && p < n) {
      tmpK[p] = key[u];
      tmpV[p] = i;
    }
  }
}

// Pieces of the window [w W, (w + 1) W) of a bucket partition with starts
// st(0..) of a range of length len (W = kMsdWindow).  Cut points are the
// starts of the buckets holding the window starts, so every segment
// between cut points is either < 2W long, or one bucket longer than W
// (emitted by the first window it covers) followed by < W positions.
// emit(start, length, split): split = longer than kMsdLocal.
template <typename StartOf, typename Emit>
__device__ __forceinline__ void windowPieces(uint32_t w, uint32_t len, uint32_t b, uint32_t bNext, bool lastWindow,
                                             StartOf startOf, Emit emit) {
  const uint32_t sb = startOf(b), eb = startOf(b + 1);
  const uint32_t segEnd = lastWindow ? len : startOf(bNext);
  if (eb - sb > kMsdWindow) {
    if (w == 0 || sb > (w - 1) * kMsdWindow) emit(sb, eb - sb, eb - sb > kMsdLocal);
    if (segEnd > eb) emit(eb, segEnd - eb, false);
  } else if (segEnd > sb) {
    emit(sb, segEnd - sb, false);
  }
}

template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_msdWindows(const K *__restrict__ tmpK, const uint32_t *__restrict__ tmpV,
                                                             K *__restrict__ outK, uint32_t *__restrict__ outV,
                                                             bool keepKeys, uint32_t n,
                                                             const uint32_t *__restrict__ starts,
                                                             const uint32_t *__restrict__ winBucket,
                                                             uint32_t *__restrict__ ctrl, uint2 *__restrict__ bigList) {
  __shared__ SortLds<K> s;
  if (ctrl[2] & kErrCount) return;
  const uint32_t w = blockIdx.x;
  const bool last = (w + 1) * kMsdWindow >= n;
  const uint32_t b = winBucket[w], bNext = last ? 0u : winBucket[w + 1];
  windowPieces(
      w, n, b, bNext, last, [&](uint32_t x) { return starts[x]; },
      [&](uint32_t st, uint32_t len, bool split) {
        if (split) {
          if (threadIdx.x == 0) bigList[atomicAdd(ctrl + 0, 1u)] = make_uint2(st, len);
        } else {
          blockSort(s, tmpK, tmpV, outK, outV, keepKeys, st, len, ctrl);
        }
      });
}

template <typename K>
__device__ __forceinline__ uint32_t compositeDigit(K key, uint32_t idx, int r, int wd) {
  const uint32_t mask = (1u << wd) - 1;
  if (r > 32) return (uint32_t)(key >> (r - 32 - wd)) & mask;
  return (idx >> (r - wd)) & mask;
}

// Buckets longer than kMsdLocal.  A piece (start, len, r, buffer) has its
// top composite bits above r consumed; splitting it by the next digit moves
// it to the other buffer.  The split's pieces that fit in LDS go to the job
// list (sorted by k_msdJobs, many workgroups at once); longer ones wait on
// a per-workgroup stack kept in `stack`: live pieces are disjoint and
// longer than kMsdLocal, so slot start / kMsdLocal is free for each.
// Loads go in chunks of kMsdLocal pairs, 16 per thread, all issued before
// the first is used.
template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_msdSplit(K *__restrict__ k0, uint32_t *__restrict__ v0,
                                                           K *__restrict__ k1, uint32_t *__restrict__ v1, bool keepKeys,
                                                           const uint2 *__restrict__ bigList, uint32_t *__restrict__ ctrl,
                                                           uint4 *__restrict__ stack, uint2 *__restrict__ jobs,
                                                           uint32_t jobCap, int rootBits) {
  __shared__ SortLds<K> s;  // only for pieces that do not fit the job list
  __shared__ uint32_t s_hist[256], s_cur[256], s_sub[257], s_tmp[kSortThreads / 64];
  __shared__ uint32_t s_jobStart[2 * kSortThreads], s_jobLen[2 * kSortThreads];
  __shared__ uint32_t s_pushStart[kSortThreads], s_pushLen[kSortThreads];
  __shared__ uint4 s_piece;  // start, len, r, buffer (len 0: none)
  __shared__ uint32_t s_top, s_have;
  const uint32_t tid = threadIdx.x;
  if (ctrl[2] & kErrCount) return;
  const uint32_t numBig = ctrl[0];
  while (true) {
    if (tid == 0) {
      const uint32_t j = atomicAdd(ctrl + 1, 1u);
      s_top = kNone;
      s_have = 0;
      if (j < numBig) {
        const uint2 bk = bigList[j];
        const uint32_t slot = bk.x / kMsdLocal;
        stack[slot] = make_uint4(bk.x, bk.y, (uint32_t)rootBits, kNone);
        s_top = slot;
        s_have = 1;
      }
    }
    __syncthreads();
    // s_have is written again only after the barriers of the pieces below
    if (s_have == 0) break;
    while (true) {
      if (tid == 0) {
        if (s_top == kNone) {
          s_piece = make_uint4(0, 0, 0, 0);
        } else {
          const uint4 e = stack[s_top];
          s_top = e.w;
          s_piece = make_uint4(e.x, e.y, e.z & 0xFF, e.z >> 8);
        }
      }
      __syncthreads();
      const uint4 pc = s_piece;
      if (pc.y == 0) break;
      const uint32_t st = pc.x, len = pc.y, buf = pc.w;
      const int r = (int)pc.z;
      if (r == 0) {  // composites are distinct, so a piece runs out of bits only when it is one pair
        if (tid == 0) atomicOr(ctrl + 2, kErrDepth);
        __syncthreads();
        continue;
      }
      const int wd = r > 32 ? min(8, r - 32) : min(8, r);
      const K *sk = buf ? k1 : k0;
      const uint32_t *sv = buf ? v1 : v0;
      K *dk = buf ? k0 : k1;
      uint32_t *dv = buf ? v0 : v1;
      s_hist[tid] = 0;
      __syncthreads();
      for (uint32_t c0 = 0; c0 < len; c0 += kMsdLocal) {
        uint32_t d[kRunItems];
#pragma unroll
        for (int j = 0; j < kRunItems; ++j) {
          const uint32_t i = c0 + tid + j * kSortThreads;
          d[j] = i < len ? compositeDigit(sk[st + i], sv[st + i], r, wd) : kNone;
        }
#pragma unroll
        for (int j = 0; j < kRunItems; ++j)
          if (d[j] != kNone) atomicAdd(&s_hist[d[j]], 1u);
      }
      __syncthreads();
      const uint32_t c = s_hist[tid];
      const uint32_t ex = blockExclusiveScan<kSortThreads>(c, s_tmp);
      s_sub[tid] = ex;
      s_cur[tid] = ex;
      if (tid == kSortThreads - 1) s_sub[256] = ex + c;
      __syncthreads();
      for (uint32_t c0 = 0; c0 < len; c0 += kMsdLocal) {
        K kk[kRunItems];
        uint32_t vv[kRunItems];
#pragma unroll
        for (int j = 0; j < kRunItems; ++j) {
          const uint32_t i = c0 + tid + j * kSortThreads;
          const bool ok = i < len;
          kk[j] = ok ? sk[st + i] : (K)0;
          vv[j] = ok ? sv[st + i] : 0u;
        }
        uint32_t p[kRunItems];
#pragma unroll
        for (int j = 0; j < kRunItems; ++j)
          p[j] = c0 + tid + j * kSortThreads < len ? atomicAdd(&s_cur[compositeDigit(kk[j], vv[j], r, wd)], 1u) : kNone;
#pragma unroll
        for (int j = 0; j < kRunItems; ++j) {
          if (p[j] != kNone) {
            dk[st + p[j]] = kk[j];
            dv[st + p[j]] = vv[j];
          }
        }
      }
      __threadfence_block();
      __syncthreads();
      const uint32_t windows = divUpDev(len, kMsdWindow);
      auto bucketOf = [&](uint32_t q) {
        uint32_t lo = 0, hi = 255;
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (s_sub[mid] <= q)
            lo = mid;
          else
            hi = mid - 1;
        }
        return lo;
      };
      const uint32_t dstBuf = buf ^ 1u;
      for (uint32_t wBase = 0; wBase < windows; wBase += kSortThreads) {
        const uint32_t w = wBase + tid;
        s_jobLen[2 * tid] = 0;
        s_jobLen[2 * tid + 1] = 0;
        s_pushLen[tid] = 0;
        if (w < windows) {
          const bool last = (w + 1) * kMsdWindow >= len;
          const uint32_t b = bucketOf(w * kMsdWindow), bNext = last ? 0u : bucketOf((w + 1) * kMsdWindow);
          uint32_t slot = 2 * tid;
          windowPieces(
              w, len, b, bNext, last, [&](uint32_t x) { return s_sub[x]; },
              [&](uint32_t ps, uint32_t pl, bool split) {
                if (split) {
                  s_pushStart[tid] = ps;
                  s_pushLen[tid] = pl;
                } else {
                  const uint32_t jslot = atomicAdd(ctrl + 3, 1u);
                  if (jslot < jobCap) {
                    jobs[jslot] = make_uint2(st + ps, pl | (dstBuf << 16));
                  } else {  // list full: this workgroup sorts the piece itself
                    s_jobStart[slot] = ps;
                    s_jobLen[slot] = pl;
                    ++slot;
                  }
                }
              });
        }
        __syncthreads();
        if (tid == 0) {
          for (uint32_t t = 0; t < kSortThreads; ++t) {
            if (s_pushLen[t] == 0) continue;
            const uint32_t ps = st + s_pushStart[t], slot = ps / kMsdLocal;
            stack[slot] = make_uint4(ps, s_pushLen[t], (uint32_t)(r - wd) | (dstBuf << 8), s_top);
            s_top = slot;
          }
        }
        for (uint32_t j = 0; j < 2 * kSortThreads; ++j) {
          const uint32_t pl = s_jobLen[j];
          if (pl) blockSort(s, dk, dv, k1, v1, keepKeys, st + s_jobStart[j], pl, ctrl);
        }
        __syncthreads();
      }
    }
  }
}

// Sorts the listed pieces into the output, one workgroup per piece.
template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_msdJobs(const K *__restrict__ k0, const uint32_t *__restrict__ v0,
                                                          K *__restrict__ k1, uint32_t *__restrict__ v1, bool keepKeys,
                                                          const uint2 *__restrict__ jobs, uint32_t jobCap,
                                                          uint32_t *__restrict__ ctrl) {
  __shared__ SortLds<K> s;
  if (ctrl[2] & kErrCount) return;
  const uint32_t num = min(ctrl[3], jobCap);
  for (uint32_t j = blockIdx.x; j < num; j += gridDim.x) {
    const uint2 job = jobs[j];
    const uint32_t len = job.y & 0xFFFFu;
    if (job.y >> 16)
      blockSort(s, (const K *)k1, (const uint32_t *)v1, k1, v1, keepKeys, job.x, len, ctrl);
    else
      blockSort(s, k0, v0, k1, v1, keepKeys, job.x, len, ctrl);
  }
}

// Top-digit histogram of keys that are already in memory.
template <typename K>
__global__ __launch_bounds__(256) void k_msdHistogram(const K *__restrict__ keys, uint32_t n, int shift,
                                                      uint32_t *__restrict__ hist) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t base = blockIdx.x * 256 + wave * 64; base < n; base += gridDim.x * 256) {
    const uint32_t i = base + lane;
    const bool valid = i < n;
    const K key = valid ? keys[i] : (K)0;
    msdCountRun(hist, (uint32_t)(key >> shift) & (kMsdBuckets - 1), valid);
  }
}

}  // namespace

MsdPlan makeMsdPlan(uint32_t n, int keyBytes, int keyBits) {
  MsdPlan p;
  p.n = n;
  p.keyBytes = keyBytes;
  p.keyBits = keyBits;
  p.windows = n == 0 ? 0 : divUp(n, kMsdWindow);
  p.slots = divUp(n, kMsdLocal) + 1;
  // pieces one split level emits: at most 2 per window of its length, plus
  // one per piece; 12 levels at most (95 composite bits, 8 per level)
  p.jobCap = 40 * (p.windows + 1);
  p.histOff = 0;
  p.ctrlOff = p.histOff + kMsdBuckets;
  p.zeroWords = p.ctrlOff + kMsdCtrlWords;
  p.totalsOff = p.zeroWords;
  p.startOff = p.totalsOff + 64;
  p.cursorOff = p.startOff + kMsdBuckets + 4;  // starts has kMsdBuckets + 1 entries; keep 16-byte alignment
  p.winOff = p.cursorOff + kMsdBuckets;
  p.bigOff = p.winOff + ((p.windows + 3) & ~3u);
  p.jobOff = p.bigOff + 2 * (size_t)((p.windows + 1) & ~1u);
  p.stackOff = p.jobOff + 2 * (size_t)((p.jobCap + 1) & ~1u);
  p.totalWords = p.stackOff + 4 * (size_t)p.slots;
  return p;
}

bool msdRun(const MsdPlan &p, const void *keysIn, void *keysTmp, uint32_t *valsTmp, void *keysOut, uint32_t *valsOut,
            bool keepKeys, uint32_t *scratch, void (*afterStage)(const char *)) {
  if (p.n == 0) return true;
  uint32_t *hist = scratch + p.histOff, *ctrl = scratch + p.ctrlOff, *starts = scratch + p.startOff;
  uint32_t *cursors = scratch + p.cursorOff, *win = scratch + p.winOff, *totals = scratch + p.totalsOff;
  uint2 *big = reinterpret_cast<uint2 *>(scratch + p.bigOff);
  uint2 *jobs = reinterpret_cast<uint2 *>(scratch + p.jobOff);
  uint4 *stack = reinterpret_cast<uint4 *>(scratch + p.stackOff);
  k_msdScanTotals<<<kScanBlocks, kScanThreads, 0, stream()>>>(hist, totals);
  if (!checkLaunch("k_msdScanTotals")) return false;
  k_msdScan<<<kScanBlocks, kScanThreads, 0, stream()>>>(p.n, p.windows, hist, totals, starts, cursors, win, ctrl);
  if (!checkLaunch("k_msdScan")) return false;
  if (afterStage) afterStage("k_msdScan");
  const int shift = p.keyBits - kMsdTopBits;
  const int rootBits = p.keyBits - kMsdTopBits + 32;
  const uint32_t scatterBlocks = divUp(p.n, 256 * kScatterGroups);
  const uint32_t splitBlocks = (uint32_t)numCUs();
  const uint32_t jobBlocks = std::max<uint32_t>(1, std::min<uint32_t>(p.windows, (uint32_t)numCUs() * 3));
  if (p.keyBytes == 8) {
    using K = uint64_t;
    k_msdScatter<K><<<scatterBlocks, 256, 0, stream()>>>((const K *)keysIn, p.n, shift, cursors, (K *)keysTmp, valsTmp,
                                                         ctrl);
    if (!checkLaunch("k_msdScatter")) return false;
    if (afterStage) afterStage("k_msdScatter");
    k_msdWindows<K><<<p.windows, kSortThreads, 0, stream()>>>((const K *)keysTmp, valsTmp, (K *)keysOut, valsOut, keepKeys,
                                                             p.n, starts, win, ctrl, big);
    if (!checkLaunch("k_msdWindows")) return false;
    if (afterStage) afterStage("k_msdWindows");
    k_msdSplit<K><<<splitBlocks, kSortThreads, 0, stream()>>>((K *)keysTmp, valsTmp, (K *)keysOut, valsOut, keepKeys, big,
                                                              ctrl, stack, jobs, p.jobCap, rootBits);
    if (!checkLaunch("k_msdSplit")) return false;
    if (afterStage) afterStage("k_msdSplit");
    k_msdJobs<K><<<jobBlocks, kSortThreads, 0, stream()>>>((const K *)keysTmp, valsTmp, (K *)keysOut, valsOut, keepKeys,
                                                           jobs, p.jobCap, ctrl);
  } else {
    using K = uint32_t;
    k_msdScatter<K><<<scatterBlocks, 256, 0, stream()>>>((const K *)keysIn, p.n, shift, cursors, (K *)keysTmp, valsTmp,
                                                         ctrl);
    if (!checkLaunch("k_msdScatter")) return false;
    if (afterStage) afterStage("k_msdScatter");
    k_msdWindows<K><<<p.windows, kSortThreads, 0, stream()>>>((const K *)keysTmp, valsTmp, (K *)keysOut, valsOut, keepKeys,
                                                             p.n, starts, win, ctrl, big);
    if (!checkLaunch("k_msdWindows")) return false;
    if (afterStage) afterStage("k_msdWindows");
    k_msdSplit<K><<<splitBlocks, kSortThreads, 0, stream()>>>((K *)keysTmp, valsTmp, (K *)keysOut, valsOut, keepKeys, big,
                                                              ctrl, stack, jobs, p.jobCap, rootBits);
    if (!checkLaunch("k_msdSplit")) return false;
    if (afterStage) afterStage("k_msdSplit");
    k_msdJobs<K><<<jobBlocks, kSortThreads, 0, stream()>>>((const K *)keysTmp, valsTmp, (K *)keysOut, valsOut, keepKeys,
                                                           jobs, p.jobCap, ctrl);
  }
  if (!checkLaunch("k_msdJobs")) return false;
  if (afterStage) afterStage("k_msdJobs");
  return true;
}

}  // namespace bvht

using namespace bvht;

extern "C" {

bool bvht_sortKeyIndexPairs(const void *keys, uint32_t n, uint32_t keyBits, uint32_t digitBits, void *keysOut,
                            uint32_t *valsOut) {
  if (keyBits != 30 && keyBits != 63) {
    setError("bvht_sortKeyIndexPairs: keyBits must be 30 or 63 (got %u)", keyBits);
    return false;
  }
  if (digitBits == 0) digitBits = kSortRadixBits;
  if (digitBits != (uint32_t)kMsdTopBits && digitBits != (uint32_t)kSortRadixBits) {
    setError("bvht_sortKeyIndexPairs: digitBits must be %d or %d (got %u)", kMsdTopBits, kSortRadixBits, digitBits);
    return false;
  }
  if (n == 0) return true;
  if (!keys || !valsOut || n > (1u << 30)) {
    setError("bvht_sortKeyIndexPairs: null buffer or more than 2^30 pairs");
    return false;
  }
  const int keyBytes = keyBits == 63 ? 8 : 4;
  const size_t kb = (size_t)keyBytes * n;
  void *k0 = nullptr, *k1 = nullptr;
  uint32_t *v0 = nullptr, *v1 = nullptr, *scratch = nullptr;
  bool ok = devMalloc(&k0, kb, "sort keys") && devMalloc(&k1, kb, "sort keys") && devAlloc(&v0, n, "sort values") &&
            devAlloc(&v1, n, "sort values") &&
            check(hipMemcpyAsync(k0, keys, kb, hipMemcpyDeviceToDevice, stream()), "sort key copy");
  const uint32_t blocks = std::min<uint32_t>(divUp(n, 1024), (uint32_t)numCUs() * 8);
  if (ok && digitBits == (uint32_t)kMsdTopBits) {
    const MsdPlan p = makeMsdPlan(n, keyBytes, (int)keyBits);
    ok = devMalloc((void **)&scratch, msdScratchBytes(p), "sort scratch") && fillU32(scratch, 0u, p.zeroWords);
    if (ok) {
      const int shift = (int)keyBits - kMsdTopBits;
      if (keyBytes == 8)
        k_msdHistogram<uint64_t><<<blocks, 256, 0, stream()>>>((const uint64_t *)k0, n, shift, msdHist(p, scratch));
      else
        k_msdHistogram<uint32_t><<<blocks, 256, 0, stream()>>>((const uint32_t *)k0, n, shift, msdHist(p, scratch));
      ok = checkLaunch("k_msdHistogram") && msdRun(p, k0, k1, v0, k0, v1, true, scratch);
    }
    uint32_t err = 0;
    ok = ok && check(hipMemcpyAsync(&err, scratch + p.ctrlOff + 2, 4, hipMemcpyDeviceToHost, stream()), "sort status") &&
         check(hipStreamSynchronize(stream()), "sort sync");
    if (ok && err) {
      setError("bvht_sortKeyIndexPairs: MSD sort error bits 0x%x", err);
      ok = false;
    }
    ok = ok && check(hipMemcpyAsync(valsOut, v1, 4 * (size_t)n, hipMemcpyDeviceToDevice, stream()), "sort value copy") &&
         (!keysOut || check(hipMemcpyAsync(keysOut, k0, kb, hipMemcpyDeviceToDevice, stream()), "sort key copy"));
  } else if (ok && keyBits == 63) {
    const SplitPlan p = makeSplitPlan(n);
    uint64_t *kk[2] = {(uint64_t *)k0, (uint64_t *)k1};
    uint32_t *vv[2] = {v0, v1};
    ok = devMalloc((void **)&scratch, splitScratchBytes(p), "sort scratch") && splitReset(p, scratch) &&
         sortHistogram(p.plan, k0, scratch);
    const int r = ok ? splitRun(p, kk, vv, scratch, true, false) : -1;
    uint32_t words[2] = {0, 0};
    ok = r >= 0 &&
         check(hipMemcpyAsync(&words[0], scratch + p.plan.errOff, 4, hipMemcpyDeviceToHost, stream()),
               "sort status") &&
         check(hipMemcpyAsync(&words[1], scratch + p.ctrlOff + 2, 4, hipMemcpyDeviceToHost, stream()), "sort status") &&
         check(hipStreamSynchronize(stream()), "sort sync");
    if (ok && (words[0] | words[1])) {
      setError("bvht_sortKeyIndexPairs: sort error bits 0x%x", words[0] | (words[1] << 3));
      ok = false;
    }
    ok = ok && check(hipMemcpyAsync(valsOut, vv[r], 4 * (size_t)n, hipMemcpyDeviceToDevice, stream()), "sort value copy") &&
         (!keysOut || check(hipMemcpyAsync(keysOut, kk[r], kb, hipMemcpyDeviceToDevice, stream()), "sort key copy"));
  } else if (ok) {
    const SortPlan p = makeSortPlan(n, keyBytes, (int)keyBits);
    void *kk[2] = {k0, k1};
    uint32_t *vv[2] = {v0, v1};
    ok = devMalloc((void **)&scratch, sortScratchBytes(p), "sort scratch") && sortReset(p, scratch) &&
         sortHistogram(p, k0, scratch);
    const int r = ok ? sortRun(p, kk, vv, true, true, scratch) : -1;
    ok = r >= 0 &&
         check(hipMemcpyAsync(valsOut, vv[r], 4 * (size_t)n, hipMemcpyDeviceToDevice, stream()), "sort value copy") &&
         (!keysOut || check(hipMemcpyAsync(keysOut, kk[r], kb, hipMemcpyDeviceToDevice, stream()), "sort key copy"));
  }
  ok = check(hipStreamSynchronize(stream()), "sort sync") && ok;
  devFree(k0);
  devFree(k1);
  devFree(v0);
  devFree(v1);
  devFree(scratch);
  return ok;
}

}  // extern "C"
