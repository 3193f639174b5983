// Shared POD types (host and device) and the exact-arithmetic helpers every
// kernel uses.  The structs come from the C ABI header; this file pins their
// layout and restates the reference's AABB semantics
// (lib/base/BVH.hpp:59-74) in a form usable from HIP device code.
#pragma once

#include "../../include/bvhtest_amd.h"

#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#define BVHT_HD __host__ __device__ __forceinline__
#else
#define BVHT_HD inline
#endif

static_assert(sizeof(bvht_vec3) == 12, "vec3");
static_assert(sizeof(bvht_AABB) == 24, "AABB");
static_assert(sizeof(bvht_Triangle) == 12, "Triangle");
static_assert(sizeof(bvht_BVHNode) == 48 && alignof(bvht_BVHNode) == 16, "BVHNode");
static_assert(offsetof(bvht_BVHNode, parent) == 24 && offsetof(bvht_BVHNode, left) == 28 &&
                  offsetof(bvht_BVHNode, right) == 32 && offsetof(bvht_BVHNode, isLeafFlag) == 36 &&
                  offsetof(bvht_BVHNode, isLeft) == 38 && offsetof(bvht_BVHNode, level) == 40 &&
                  offsetof(bvht_BVHNode, surfaceArea) == 44,
              "BVHNode offsets");
static_assert(sizeof(bvht_Ray) == 48 && offsetof(bvht_Ray, sign) == 36 && offsetof(bvht_Ray, px) == 40, "Ray");
static_assert(sizeof(bvht_BVH) == 32, "BVH header");
static_assert(sizeof(bvht_CUDAPixel) == 4, "CUDAPixel");

namespace bvht {

constexpr float kFltEpsilon = 1.1920928955078125e-07f;  // FLT_EPSILON

// AABB::surfaceArea -- lib/base/BVH.hpp:62-65: 2 * (dx*dy + dx*dz + dy*dz),
// evaluated left to right.  Callers compile with -ffp-contract=off so no FMA
// is formed and the value is a pure function of the box.
BVHT_HD float surfaceArea(float mnx, float mny, float mnz, float mxx, float mxy, float mxz) {
  float dx = mxx - mnx, dy = mxy - mny, dz = mxz - mnz;
  return 2.0f * (dx * dy + dx * dz + dy * dz);
}

BVHT_HD float surfaceArea(const bvht_AABB &b) {
  return surfaceArea(b.minMax[0].x, b.minMax[0].y, b.minMax[0].z, b.minMax[1].x, b.minMax[1].y, b.minMax[1].z);
}

// AABB::mergeWith -- lib/base/BVH.hpp:67-74 (`a < b ? a : b` for the minimum,
// `a > b ? a : b` for the maximum; b wins ties, which only matters for +-0).
BVHT_HD float mergeMin(float a, float b) { return a < b ? a : b; }
BVHT_HD float mergeMax(float a, float b) { return a > b ? a : b; }

BVHT_HD void mergeInto(bvht_AABB &a, const bvht_AABB &b) {
  a.minMax[0].x = mergeMin(a.minMax[0].x, b.minMax[0].x);
  a.minMax[0].y = mergeMin(a.minMax[0].y, b.minMax[0].y);
  a.minMax[0].z = mergeMin(a.minMax[0].z, b.minMax[0].z);
  a.minMax[1].x = mergeMax(a.minMax[1].x, b.minMax[1].x);
  a.minMax[1].y = mergeMax(a.minMax[1].y, b.minMax[1].y);
  a.minMax[1].z = mergeMax(a.minMax[1].z, b.minMax[1].z);
}

// Triangle box as kInitTriData builds it (lib/builder/LBVH_CUDA.cu:72-90):
// min starts at v1 and max at v2, then branchy compares, then -+FLT_EPSILON.
BVHT_HD bvht_AABB triangleBox(const bvht_vec3 &v1, const bvht_vec3 &v2, const bvht_vec3 &v3) {
  bvht_AABB b;
  b.minMax[0] = v1;
  b.minMax[1] = v2;
  if (v2.x < b.minMax[0].x) b.minMax[0].x = v2.x;
  if (v2.y < b.minMax[0].y) b.minMax[0].y = v2.y;
  if (v2.z < b.minMax[0].z) b.minMax[0].z = v2.z;
  if (v3.x < b.minMax[0].x) b.minMax[0].x = v3.x;
  if (v3.y < b.minMax[0].y) b.minMax[0].y = v3.y;
  if (v3.z < b.minMax[0].z) b.minMax[0].z = v3.z;
  if (v1.x > b.minMax[1].x) b.minMax[1].x = v1.x;
  if (v1.y > b.minMax[1].y) b.minMax[1].y = v1.y;
  if (v1.z > b.minMax[1].z) b.minMax[1].z = v1.z;
  if (v3.x > b.minMax[1].x) b.minMax[1].x = v3.x;
  if (v3.y > b.minMax[1].y) b.minMax[1].y = v3.y;
  if (v3.z > b.minMax[1].z) b.minMax[1].z = v3.z;
  b.minMax[0].x -= kFltEpsilon;
  b.minMax[0].y -= kFltEpsilon;
  b.minMax[0].z -= kFltEpsilon;
  b.minMax[1].x += kFltEpsilon;
  b.minMax[1].y += kFltEpsilon;
  b.minMax[1].z += kFltEpsilon;
  return b;
}

// x / 3.0f, correctly rounded, in three operations instead of a general
// division: q0 = x * RN(1/3), the remainder r = x - 3 q0 exactly (one FMA),
// q0 + r * RN(1/3) rounded once (Markstein's correction).  An exact
// quotient (r = 0: keeps the sign of a zero) or an infinite x (r is NaN)
// returns q0.  Checked against IEEE division for every f32 bit pattern
// (equal except for NaN payloads; tests/div3_check.c).
BVHT_HD float div3(float x) {
  const float third = 0.3333333432674408f;  // RN(1/3)
  const float q0 = x * third;
  const float r = __builtin_fmaf(-q0, 3.0f, x);
  const float q1 = __builtin_fmaf(r, third, q0);
  return (r == 0.0f || r != r) ? q0 : q1;
}

// Centroid (v1 + v2 + v3) / 3.0f -- lib/builder/LBVH_CUDA.cu:92.
BVHT_HD bvht_vec3 centroid(const bvht_vec3 &v1, const bvht_vec3 &v2, const bvht_vec3 &v3) {
  bvht_vec3 c;
  c.x = div3(v1.x + v2.x + v3.x);
  c.y = div3(v1.y + v2.y + v3.y);
  c.z = div3(v1.z + v2.z + v3.z);
  return c;
}

// expandBits -- lib/builder/LBVH_CUDA.cu:132-138: insert two zero bits after
// each of the low 10 bits.
BVHT_HD uint32_t expandBits10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// 21-bit variant for 63-bit keys (extension): spread bit b of v to bit 3b.
BVHT_HD uint64_t expandBits21(uint64_t v) {
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// Per-axis quantisation of kCalcMortonCodes (lib/builder/LBVH_CUDA.cu:145-162):
// ((c + offset) * scale) * Q, clamped to [0, Q-1], truncated to an integer.
BVHT_HD float quantize(float c, float offset, float scale, float q, float qmax) {
  c += offset;
  c *= scale;
  c *= q;
  if (c < 0.0f) c = 0.0f;
  if (c > qmax) c = qmax;
  return c;
}

BVHT_HD uint32_t morton30(const bvht_vec3 &c, const bvht_vec3 &off, const bvht_vec3 &scale) {
  uint32_t x = expandBits10((uint32_t)quantize(c.x, off.x, scale.x, 1024.0f, 1023.0f));
  uint32_t y = expandBits10((uint32_t)quantize(c.y, off.y, scale.y, 1024.0f, 1023.0f));
  uint32_t z = expandBits10((uint32_t)quantize(c.z, off.z, scale.z, 1024.0f, 1023.0f));
  return x * 4 + y * 2 + z;
}

// expandBits21 as two 32-bit words: the low 11 bits spread to bits 0..30
// (expandBits10's steps place bit 10 at bit 30 as well), the high 10 bits
// to bits 33..60, i.e. bits 1..28 of the high word.  Equal to expandBits21
// for every 21-bit input (checked exhaustively); 32-bit operations only.
BVHT_HD void expandBits21Halves(uint32_t v, uint32_t &lo, uint32_t &hi) {
  lo = expandBits10(v & 0x7FFu);
  hi = expandBits10(v >> 11) << 1;
}

// x * 4 + y * 2 + z of the three spread coordinates (their bits are
// disjoint, so the sum is an OR), assembled from 32-bit words.  The
// quantised value is below 2^21, so its conversion goes through u32.
BVHT_HD uint64_t morton63(const bvht_vec3 &c, const bvht_vec3 &off, const bvht_vec3 &scale) {
  uint32_t xl, xh, yl, yh, zl, zh;
  expandBits21Halves((uint32_t)quantize(c.x, off.x, scale.x, 2097152.0f, 2097151.0f), xl, xh);
  expandBits21Halves((uint32_t)quantize(c.y, off.y, scale.y, 2097152.0f, 2097151.0f), yl, yh);
  expandBits21Halves((uint32_t)quantize(c.z, off.z, scale.z, 2097152.0f, 2097151.0f), zl, zh);
  const uint32_t lo = (xl << 2) | (yl << 1) | zl;
  const uint32_t hi = (xh << 2) | (xl >> 30) | (yh << 1) | (yl >> 31) | zh;
  return ((uint64_t)hi << 32) | lo;
}

// ---------------------------------------------------------------------------
// Closed form of the reference LBVH topology.
//
// kBuildTree (lib/builder/LBVH_CUDA.cu:205-324) passes the 64-bit key
// (morton << 32 | index) to __clz, whose only overload takes a 32-bit int, so
// every delta is clz32(i ^ j): the radix tree is the one over the keys
// 0..n-1.  For those keys a node covering leaf positions [l, r] (l < r)
// splits at m = r with the bits below the highest bit of (l ^ r) cleared;
// the children are [l, m-1] and [m, r].  Every node's range is either an
// aligned power-of-two block or a "spine" range [N1 & ~(2^k - 1), N1]
// (N1 = n - 1).  Karras numbering: a left child is stored at index r, a right
// child at index l, the root at 0, leaf position p at N1 + p.
// ---------------------------------------------------------------------------

BVHT_HD uint32_t ilog2(uint32_t v) { return 31u - (uint32_t)__builtin_clz(v); }

// True iff [l, r] is the left child of its parent (a full aligned block
// [a*2s, a*2s + s - 1] whose right sibling starts inside [0, N1]).
BVHT_HD bool rangeIsLeft(uint32_t l, uint32_t r, uint32_t N1) {
  uint32_t s = r - l + 1;
  if ((s & (s - 1)) != 0) return false;
  if ((l & (2 * s - 1)) != 0) return false;
  return l + s <= N1;
}

// Node index of the node covering [l, r].
BVHT_HD uint32_t rangeNodeIndex(uint32_t l, uint32_t r, uint32_t N1) {
  if (l == r) return N1 + l;
  if (l == 0 && r == N1) return 0;
  return rangeIsLeft(l, r, N1) ? r : l;
}

// Range of the parent of the (non-root) node covering [l, r].
BVHT_HD void rangeParent(uint32_t l, uint32_t r, uint32_t N1, uint32_t &pl, uint32_t &pr) {
  if (rangeIsLeft(l, r, N1)) {
    uint32_t s = r - l + 1;
    pl = l;
    pr = (l + 2 * s - 1 <= N1 && l + 2 * s - 1 >= l) ? l + 2 * s - 1 : N1;
  } else {
    pl = l - (l & (0u - l));
    pr = r;
  }
}

// Split position m of the node [l, r] (children [l, m-1] and [m, r]).
BVHT_HD uint32_t rangeSplit(uint32_t l, uint32_t r) {
  uint32_t h = ilog2(l ^ r);
  return (r >> h) << h;
}

}  // namespace bvht
