// Library runtime: the stream all work goes to, the last-error record, the
// allocation hook used by host tests, and small launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>

#include "types.hpp"

namespace bvht {

struct Runtime {
  hipStream_t stream = nullptr;
  std::string lastError;
  int allocFailAfter = 0;  // 0: off; n: the n-th allocation from now fails
};

Runtime &rt();
inline hipStream_t stream() { return rt().stream; }

// Records a failure message (printf-style) as the library's last error.
void setError(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// Returns true on hipSuccess; otherwise records "<what>: <hip error>".
bool check(hipError_t e, const char *what);
// Checks the last kernel launch.
inline bool checkLaunch(const char *what) { return check(hipGetLastError(), what); }

// Device allocation honouring the failure-injection hook; *p = nullptr on failure.
bool devMalloc(void **p, size_t bytes, const char *what);
template <typename T>
inline bool devAlloc(T **p, size_t count, const char *what) {
  return devMalloc(reinterpret_cast<void **>(p), count * sizeof(T), what);
}
void devFree(void *p);

inline uint32_t divUp(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Number of compute units of the current device (cached).
int numCUs();

// Fills `count` u32 words with `value` on the stream (a kernel, not
// hipMemsetAsync, so the call is safe inside captured graphs).
bool fillU32(uint32_t *p, uint32_t value, size_t count);

}  // namespace bvht
