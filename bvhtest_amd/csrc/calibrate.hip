// Counter calibration kernels: three access patterns of known byte counts
// whose FETCH_SIZE / WRITE_SIZE readings convert the same counters on the
// LBVH kernels into HBM bytes.  Built as its own library
// (libbvht_calibrate.so) because it is a measurement aid, not part of the
// BVH interface.
//
//   bool calib_copy(void *dst, const void *src, size_t bytes)
//       dst[i] = src[i], 16-byte words; reads and writes `bytes`.
//   bool calib_read(const void *src, size_t bytes, void *out)
//       sums `bytes` of 32-bit words (one f32 partial per workgroup written
//       to `out`, which holds kReadBlocks floats); reads `bytes`.
//   bool calib_gather(const void *src, const uint32_t *rows, size_t numRows, void *dst)
//       dst row r = src row rows[r], rows of 128 bytes; reads the rows once
//       each in the order of `rows` plus the 4-byte indices, writes
//       numRows * 128 bytes contiguously.
// `bytes` must be a multiple of 16.  Every call synchronises the null stream
// and returns false (with the HIP error printed) when a call fails.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kReadBlocks = 2048;

__global__ __launch_bounds__(kBlock) void k_calibCopy(uint4 *__restrict__ dst, const uint4 *__restrict__ src,
                                                      size_t n) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_calibRead(const uint4 *__restrict__ src, size_t n,
                                                      float *__restrict__ out) {
  __shared__ float s_part[kBlock / 64];
  float acc = 0.f;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
    const uint4 w = src[i];
    acc += __uint_as_float(w.x) + __uint_as_float(w.y) + __uint_as_float(w.z) + __uint_as_float(w.w);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (uint32_t w = 0; w < kBlock / 64; ++w) t += s_part[w];
    out[blockIdx.x] = t;
  }
}

// 8 lanes per 128-byte row, 16 bytes each
__global__ __launch_bounds__(kBlock) void k_calibGather(const uint4 *__restrict__ src,
                                                        const uint32_t *__restrict__ rows, size_t numRows,
                                                        uint4 *__restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const size_t r = t >> 3;
  if (r >= numRows) return;
  dst[t] = src[(size_t)rows[r] * 8 + (t & 7)];
}

bool done(const char *what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return false;
  }
  return true;
}

uint32_t gridFor(size_t items) {
  const size_t g = (items + kBlock - 1) / kBlock;
  return (uint32_t)(g < 8192 ? (g == 0 ? 1 : g) : 8192);
}

}  // namespace

extern "C" __attribute__((visibility("default"))) bool calib_copy(void *dst, const void *src, size_t bytes) {
  if (!dst || !src || bytes % 16) return false;
  const size_t n = bytes / 16;
  k_calibCopy<<<gridFor(n), kBlock>>>((uint4 *)dst, (const uint4 *)src, n);
  return done("calib_copy");
}

extern "C" __attribute__((visibility("default"))) bool calib_read(const void *src, size_t bytes, void *out) {
  if (!src || !out || bytes % 16) return false;
  k_calibRead<<<kReadBlocks, kBlock>>>((const uint4 *)src, bytes / 16, (float *)out);
  return done("calib_read");
}

extern "C" __attribute__((visibility("default"))) bool calib_gather(const void *src, const uint32_t *rows,
                                                                    size_t numRows, void *dst) {
  if (!src || !rows || !dst) return false;
  const size_t threads = numRows * 8;
  const size_t grid = (threads + kBlock - 1) / kBlock;
  if (grid == 0) return true;
  if (grid > 0x7fffffffu) return false;
  k_calibGather<<<(uint32_t)grid, kBlock>>>((const uint4 *)src, rows, numRows, (uint4 *)dst);
  return done("calib_gather");
}
