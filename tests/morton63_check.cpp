// Checks the build's 63-bit Morton assembly from 32-bit words
// (bvht::expandBits21Halves, bvht::morton63) against the 64-bit spread
// bvht::expandBits21 (LBVH_CUDA.cu:132-138 widened to 21 bits per axis): the
// halves for every 21-bit coordinate, and x * 4 + y * 2 + z for every
// coordinate paired with a pseudo-random pair plus the extreme corners.
// Prints the mismatch count; exit status 1 on any.
#include <cstdint>
#include <cstdio>

#include "../bvhtest_amd/csrc/types.hpp"

int main() {
  uint64_t bad = 0;
  auto combine = [](uint32_t x, uint32_t y, uint32_t z) {
    uint32_t xl, xh, yl, yh, zl, zh;
    bvht::expandBits21Halves(x, xl, xh);
    bvht::expandBits21Halves(y, yl, yh);
    bvht::expandBits21Halves(z, zl, zh);
    const uint32_t lo = (xl << 2) | (yl << 1) | zl;
    const uint32_t hi = (xh << 2) | (xl >> 30) | (yh << 1) | (yl >> 31) | zh;
    return ((uint64_t)hi << 32) | lo;
  };
  for (uint32_t x = 0; x < (1u << 21); ++x) {
    uint32_t lo, hi;
    bvht::expandBits21Halves(x, lo, hi);
    if ((((uint64_t)hi << 32) | lo) != bvht::expandBits21(x) && bad++ < 5) std::printf("spread x=%u\n", x);
    const uint32_t y = (x * 2654435761u) & 0x1FFFFFu, z = (x * 40503u + 17u) & 0x1FFFFFu;
    const uint64_t want = bvht::expandBits21(x) * 4 + bvht::expandBits21(y) * 2 + bvht::expandBits21(z);
    if (combine(x, y, z) != want && bad++ < 10) std::printf("combine %u %u %u\n", x, y, z);
  }
  const uint32_t m = 0x1FFFFFu;
  const uint32_t corners[][3] = {{m, m, m}, {m, 0, 0}, {0, m, 0}, {0, 0, m}, {0, 0, 0}};
  for (const auto &c : corners) {
    const uint64_t want = bvht::expandBits21(c[0]) * 4 + bvht::expandBits21(c[1]) * 2 + bvht::expandBits21(c[2]);
    if (combine(c[0], c[1], c[2]) != want && bad++ < 15) std::printf("corner\n");
  }
  std::printf("mismatches %llu\n", (unsigned long long)bad);
  return bad ? 1 : 0;
}
