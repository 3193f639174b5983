// wave64 / workgroup primitives shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bvht {

// Device-coherent (all XCDs) relaxed accesses.  On gfx950 these compile to
// loads/stores with the sc1 cache-coherence bit and no cache maintenance;
// use them for self-contained words (flag + payload in one u32).
__device__ __forceinline__ void storeRelaxedAgent(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t loadRelaxedAgent(const uint32_t *p) {
  return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float loadRelaxedAgent(const float *p) {
  return __uint_as_float(loadRelaxedAgent(reinterpret_cast<const uint32_t *>(p)));
}
__device__ __forceinline__ void storeRelaxedAgent(float *p, float v) {
  storeRelaxedAgent(reinterpret_cast<uint32_t *>(p), __float_as_uint(v));
}
// Waits until this wave's outstanding vector memory operations are done
// (s_waitcnt vmcnt(0)).  After device-coherent stores (the relaxed
// agent-scope stores above write through the XCD's L2) this orders them
// before a later atomic as a release does, without the L2 write-back of
// every dirty line that a release fence performs.
__device__ __forceinline__ void waitVectorMemory() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Hides a register's value from the optimiser (no instruction), so values
// derived from it are recomputed after this point instead of being kept
// live (or spilled) from before it.
__device__ __forceinline__ void recomputeAfter(uint32_t &x) { asm volatile("" : "+v"(x)); }

// Raw buffer accesses (gfx9 buffer resource: base, `bytes` records, 32-bit
// data format in word 3).  Offsets are 32-bit byte offsets in a VGPR (plus
// the immediate the compiler folds in); both take part in the range check,
// so a load at or past `bytes` returns 0 and such a store is dropped.
// `bytes` must stay below 2^31.
typedef __amdgpu_buffer_rsrc_t BufferRsrc;
typedef unsigned int BufferU32x2 __attribute__((vector_size(8)));
__device__ __forceinline__ BufferRsrc bufferRsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t bufferLoadU32(BufferRsrc r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
// 16 bytes (16-byte aligned) with the cache policy of a relaxed agent-scope
// atomic load (sc1: past the XCD's L2), for words other workgroups publish
// with storeRelaxedAgent; each dword is read whole.  To the optimiser this
// is a plain read: a reload in a spin loop must follow memoryChanged(),
// or it is merged with the first load and the loop never sees the word
// change.
typedef unsigned int BufferU32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ BufferU32x4 bufferLoadU32x4Coherent(BufferRsrc r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 16);
}
// Tells the optimiser that memory may have changed (other workgroups'
// stores); no instruction.
__device__ __forceinline__ void memoryChanged() { asm volatile("" ::: "memory"); }
__device__ __forceinline__ void bufferStore(uint32_t v, BufferRsrc r, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, 0);
}
__device__ __forceinline__ void bufferStore(uint64_t v, BufferRsrc r, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b64(BufferU32x2{(uint32_t)v, (uint32_t)(v >> 32)}, r, off, 0, 0);
}

// Agent-scope fences for publishing data through a relaxed counter RMW.
// In a loop that stores and then increments a counter on the next trip, an
// acq_rel RMW at the loop head is lowered here to `buffer_wbl2` + the atomic
// without waiting for the stores of the previous trip (vmcnt), so a reader
// that sees the count can still read the old data.  A release fence placed
// right after the stores lowers to `buffer_wbl2 sc1; s_waitcnt vmcnt(0)`.
__device__ __forceinline__ void fenceReleaseAgent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fenceAcquireAgent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

__device__ __forceinline__ uint32_t divUpDev(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// The lanes of the wave whose BITS-bit digit equals this lane's digit d,
// ANDed with `peers`.  One ballot per digit bit; each ballot is folded into
// the 64-bit mask with one 3-input bitwise op per half (gfx950
// v_bitop3_b32 with table 0x90: a & ~(b ^ c)), c being the lane's own bit
// sign-extended to 0 or all ones, so the mask keeps the ballot where the
// lane's bit is 1 and its complement where it is 0.
template <int BITS>
__device__ __forceinline__ uint64_t matchDigit(uint32_t d, uint64_t peers) {
  uint32_t lo = (uint32_t)peers, hi = (uint32_t)(peers >> 32);
  uint32_t t[BITS];
#pragma unroll
  for (int b = 0; b < BITS; ++b) t[b] = (uint32_t)__builtin_amdgcn_sbfe((int32_t)d, b, 1);
  if constexpr (BITS == 8)  // ballots of t itself: one compare per bit instead of a shift and a compare
    asm("" : "+v"(t[0]), "+v"(t[1]), "+v"(t[2]), "+v"(t[3]), "+v"(t[4]), "+v"(t[5]), "+v"(t[6]), "+v"(t[7]));
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const uint64_t m = __ballot(t[b] != 0u);
    lo = __builtin_amdgcn_bitop3_b32(lo, (uint32_t)m, t[b], 0x90);
    hi = __builtin_amdgcn_bitop3_b32(hi, (uint32_t)(m >> 32), t[b], 0x90);
  }
  return ((uint64_t)hi << 32) | lo;
}

// Adds one to bin `bin` of an LDS histogram for every valid lane of the
// wave (every lane calls).  When all valid lanes hold the same bin (common
// for the high digits of spatially coherent keys) one lane adds the count:
// same-address LDS atomics from one instruction are serialised lane by lane.
__device__ __forceinline__ void waveHistAdd(uint32_t *hist, uint32_t bin, bool valid) {
  const uint64_t validMask = __ballot(valid);
  if (validMask == 0) return;
  const int first = __builtin_ctzll(validMask);
  const uint32_t b0 = __builtin_amdgcn_readlane(bin, first);
  const uint64_t differ = __ballot(valid && bin != b0);
  if (differ == 0) {
    if ((int)(threadIdx.x & 63) == first) atomicAdd(hist + b0, (uint32_t)__popcll(validMask));
  } else if (valid) {
    atomicAdd(hist + bin, 1u);
  }
}

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t waveInclusiveScan(uint32_t x) {
  const uint32_t lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  return x;
}

// Exclusive scan across a workgroup of BLOCK threads (BLOCK multiple of 64).
// s_tmp needs BLOCK / 64 words.  Contains barriers: call from uniform code.
template <int BLOCK>
__device__ __forceinline__ uint32_t blockExclusiveScan(uint32_t v, uint32_t *s_tmp) {
  constexpr int WAVES = BLOCK / 64;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = waveInclusiveScan(v);
  if (lane == 63) s_tmp[wave] = inc;
  __syncthreads();
  uint32_t prefix = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w)
    if ((uint32_t)w < wave) prefix += s_tmp[w];
  __syncthreads();
  return prefix + inc - v;
}

__device__ __forceinline__ float waveMin(float x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    float y = __shfl_xor(x, o, 64);
    x = y < x ? y : x;
  }
  return x;
}
__device__ __forceinline__ float waveMax(float x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    float y = __shfl_xor(x, o, 64);
    x = y > x ? y : x;
  }
  return x;
}

}  // namespace bvht
