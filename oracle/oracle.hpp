// ORACLE -- test infrastructure, not product code (see oracle.cpp).
// Shared structs and the restated AABB arithmetic of the reference
// (lib/base/BVH.hpp:59-136, lib/base/Ray.hpp:38-80).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {

struct V3 {
  float x, y, z;
};
struct Box {
  V3 mn, mx;  // minMax[0], minMax[1]
};
struct alignas(16) Node {
  Box bbox;
  uint32_t parent, left, right;
  uint16_t isLeafFlag, isLeft, level, pad;
  float surfaceArea;
};
static_assert(sizeof(Node) == 48, "BVHNode is 48 bytes");

struct alignas(16) Ray {
  V3 pos, dir, inv;
  uint8_t sign[3];
  uint8_t pad;
  uint32_t px, py;
};
static_assert(sizeof(Ray) == 48, "Ray is 48 bytes");

struct Pixel {
  uint8_t hit, diffuse;
  uint16_t intCount;
};
struct Hit {
  uint32_t tri;
  float t;
};

inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
// glm::cross / glm::dot (dot sums left to right: (x + y) + z)
inline V3 cross(V3 x, V3 y) { return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// glm::normalize = v * inversesqrt(dot(v, v)), inversesqrt = 1 / sqrt
inline V3 normalize(V3 v) { return mul(v, 1.0f / std::sqrt(dot(v, v))); }

// AABB::surfaceArea -- lib/base/BVH.hpp:62-65
inline float boxSA(const Box &b) {
  V3 d = sub(b.mx, b.mn);
  return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}
// AABB::mergeWith -- lib/base/BVH.hpp:67-74
inline void mergeWith(Box &a, const Box &b) {
  a.mn.x = a.mn.x < b.mn.x ? a.mn.x : b.mn.x;
  a.mn.y = a.mn.y < b.mn.y ? a.mn.y : b.mn.y;
  a.mn.z = a.mn.z < b.mn.z ? a.mn.z : b.mn.z;
  a.mx.x = a.mx.x > b.mx.x ? a.mx.x : b.mx.x;
  a.mx.y = a.mx.y > b.mx.y ? a.mx.y : b.mx.y;
  a.mx.z = a.mx.z > b.mx.z ? a.mx.z : b.mx.z;
}

}  // namespace orc
