// ORACLE -- test infrastructure, not product code.
//
// CPU restatement of mensinda/BVHTest's GPU BVH pipeline, used only by
// tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the
// checker.  It shares no code with bvhtest_amd/ (its own structs, its own
// arithmetic), so a mistake in the product is not silently mirrored here.
// Every function cites the reference file:line it follows.
//
// Floating point: compiled with -ffp-contract=off (no FMA), SSE f32; divisions
// are IEEE.  Where the reference relies on unpinned vendor arithmetic (nvcc
// FMA contraction, __fdividef, glm's matrix inverse) the restatement uses
// plain IEEE operations; those spots are marked "unpinned" below.
//
// Parity pinning status of each entry point is listed in oracle/README.md.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <tuple>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

#include "oracle.hpp"

namespace orc {

// ---------------------------------------------------------------------------
// LBVH -- lib/builder/LBVH_CUDA.cu
// ---------------------------------------------------------------------------

struct TriData {  // LBVH_TriData -- lib/builder/LBVH_CUDA.hpp:27-32
  Box bbox;
  V3 centroid;
  uint32_t faceIndex;
};

// kInitTriData -- lib/builder/LBVH_CUDA.cu:60-96
TriData initTri(const V3 *vert, const uint32_t *f, uint32_t i) {
  TriData r;
  r.faceIndex = i;
  V3 v1 = vert[f[3 * i]], v2 = vert[f[3 * i + 1]], v3 = vert[f[3 * i + 2]];
  r.bbox.mn = v1;
  r.bbox.mx = v2;
  if (v2.x < r.bbox.mn.x) r.bbox.mn.x = v2.x;
  if (v2.y < r.bbox.mn.y) r.bbox.mn.y = v2.y;
  if (v2.z < r.bbox.mn.z) r.bbox.mn.z = v2.z;
  if (v3.x < r.bbox.mn.x) r.bbox.mn.x = v3.x;
  if (v3.y < r.bbox.mn.y) r.bbox.mn.y = v3.y;
  if (v3.z < r.bbox.mn.z) r.bbox.mn.z = v3.z;
  if (v1.x > r.bbox.mx.x) r.bbox.mx.x = v1.x;
  if (v1.y > r.bbox.mx.y) r.bbox.mx.y = v1.y;
  if (v1.z > r.bbox.mx.z) r.bbox.mx.z = v1.z;
  if (v3.x > r.bbox.mx.x) r.bbox.mx.x = v3.x;
  if (v3.y > r.bbox.mx.y) r.bbox.mx.y = v3.y;
  if (v3.z > r.bbox.mx.z) r.bbox.mx.z = v3.z;
  const float eps = std::numeric_limits<float>::epsilon();  // FLT_EPSILON
  r.bbox.mn = {r.bbox.mn.x - eps, r.bbox.mn.y - eps, r.bbox.mn.z - eps};
  r.bbox.mx = {r.bbox.mx.x + eps, r.bbox.mx.y + eps, r.bbox.mx.z + eps};
  V3 s = add(add(v1, v2), v3);
  r.centroid = {s.x / 3.0f, s.y / 3.0f, s.z / 3.0f};
  return r;
}

// expandBits -- lib/builder/LBVH_CUDA.cu:132-138, restated as a plain bit
// loop (bit b of v goes to bit 3b): the magic-number form of the reference and
// of the product is checked against this.
inline uint64_t spread3(uint64_t v, int bits) {
  uint64_t r = 0;
  for (int b = 0; b < bits; ++b) r |= ((v >> b) & 1ull) << (3 * b);
  return r;
}

// kCalcMortonCodes -- lib/builder/LBVH_CUDA.cu:140-166 (Q = 1024 for the
// reference's 30-bit keys; Q = 2^21 for the 63-bit extension).
inline uint64_t mortonOf(V3 c, V3 off, V3 scale, int bits) {
  const int per = bits / 3;
  const float q = (float)(1u << per), qmax = (float)((1u << per) - 1u);
  float v[3] = {c.x, c.y, c.z}, o[3] = {off.x, off.y, off.z}, s[3] = {scale.x, scale.y, scale.z};
  uint64_t e[3];
  for (int a = 0; a < 3; ++a) {
    float t = v[a];
    t += o[a];
    t *= s[a];
    t *= q;
    if (t < 0.0f) t = 0.0f;
    if (t > qmax) t = qmax;
    e[a] = spread3((uint64_t)(uint32_t)t, per);
  }
  return e[0] * 4 + e[1] * 2 + e[2];
}

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }
inline int clz64(uint64_t x) { return x == 0 ? 64 : __builtin_clzll(x); }

// delta / findSplit / determineRange -- lib/builder/LBVH_CUDA.cu:205-271.
// topology 0 restates the reference as compiled: CODE(x) is a 64-bit value
// passed to __clz(int), so only its low word (the index) reaches clz.
// topology 1 (extension) uses the whole key (morton, index).
struct Karras {
  const uint64_t *codes;
  uint32_t n;
  int topology, keyBits;
  int delta(uint32_t i, uint32_t j) const {
    if (j >= n) return -1;
    if (topology == 0) return clz32(i ^ j);
    uint64_t a = codes[i], b = codes[j];
    int w = keyBits <= 32 ? 32 : 64;
    if (a == b) return w + clz32(i ^ j);
    return w == 32 ? clz32((uint32_t)(a ^ b)) : clz64(a ^ b);
  }
  uint32_t findSplit(uint32_t first, uint32_t last) const {
    int common = delta(first, last);
    uint32_t split = first, step = last - first;
    do {
      step = (step + 1) >> 1;
      uint32_t ns = split + step;
      if (ns < last) {
        if (delta(first, ns) > common) split = ns;
      }
    } while (step > 1);
    return split;
  }
  void determineRange(uint32_t i, uint32_t &lo, uint32_t &hi) const {
    int32_t d = delta(i, i + 1) - delta(i, i - 1);
    d = d > 0 ? 1 : -1;
    int deltaMin = delta(i, i - (uint32_t)d);
    uint32_t lMax = 2;
    while (delta(i, i + lMax * (uint32_t)d) > deltaMin) lMax *= 2;
    uint32_t l = 0;
    do {
      lMax >>= 1;
      if (delta(i, i + (l + lMax) * (uint32_t)d) > deltaMin) l += lMax;
    } while (lMax > 1);
    uint32_t j = i + l * (uint32_t)d;
    lo = i < j ? i : j;
    hi = i < j ? j : i;
  }
};

}  // namespace orc

using namespace orc;

extern "C" {

// Whole LBVH build: kInitTriData + scene reduce (LBVH_CUDA.cu:60-119), Morton
// codes (:130-183), stable radix sort (:185-193), kGenLeafNodes (:273-292),
// kBuildTree (:294-324), kFixAABBTree (:341-374).
//   scene[12] = box min xyz, box max xyz, offset xyz, scale xyz.
// Fields the reference leaves unwritten are defined as: root.parent = 0
// (self), root.isLeft = 0, inner level = 0, pad = 0.  Returns 0 on success.
int orc_lbvh(const float *vert, uint32_t nv, const uint32_t *faces, uint32_t nf, int mortonBits, int topology,
             uint64_t *keysSorted, uint32_t *sortedFaces, Node *nodes, float *scene) {
  (void)nv;
  if (nf == 0) return 1;
  const V3 *V = reinterpret_cast<const V3 *>(vert);
  std::vector<TriData> tri(nf);
  for (uint32_t i = 0; i < nf; ++i) tri[i] = initTri(V, faces, i);
  // DeviceReduce::Reduce with ReduceRootAABB from an all-zero box (:104-113)
  Box sb = {{0, 0, 0}, {0, 0, 0}};
  for (uint32_t i = 0; i < nf; ++i) mergeWith(sb, tri[i].bbox);
  // LBVH_calcMortonCodes host part (:171-180)
  V3 off = sub(V3{0.0f, 0.0f, 0.0f}, sb.mn);
  V3 mx = add(sb.mx, off);
  V3 scale = {1.0f / mx.x, 1.0f / mx.y, 1.0f / mx.z};
  if (scene) {
    float s[12] = {sb.mn.x, sb.mn.y, sb.mn.z, sb.mx.x, sb.mx.y, sb.mx.z, off.x, off.y, off.z, scale.x, scale.y, scale.z};
    std::memcpy(scene, s, sizeof(s));
  }
  std::vector<uint64_t> code(nf);
  for (uint32_t i = 0; i < nf; ++i) code[i] = mortonOf(tri[i].centroid, off, scale, mortonBits);
  // CUB SortPairs is a stable LSD radix sort
  std::vector<uint32_t> perm(nf);
  for (uint32_t i = 0; i < nf; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
  std::vector<uint64_t> sorted(nf);
  for (uint32_t i = 0; i < nf; ++i) sorted[i] = code[perm[i]];
  if (keysSorted) std::memcpy(keysSorted, sorted.data(), sizeof(uint64_t) * nf);
  if (sortedFaces)
    for (uint32_t i = 0; i < nf; ++i) sortedFaces[i] = tri[perm[i]].faceIndex;

  const uint32_t N = 2 * nf - 1;
  std::memset(nodes, 0, sizeof(Node) * N);
  // kGenLeafNodes
  for (uint32_t i = 0; i < nf; ++i) {
    const TriData &d = tri[perm[i]];
    Node &l = nodes[nf - 1 + i];
    l.bbox = d.bbox;
    l.parent = UINT32_MAX;
    l.isLeafFlag = 1;
    l.left = d.faceIndex;
    l.right = 1;
    l.level = UINT8_MAX;
    l.surfaceArea = boxSA(l.bbox);
  }
  // kBuildTree
  Karras K{sorted.data(), nf, topology, mortonBits};
  for (uint32_t i = 0; i + 1 < nf; ++i) {
    uint32_t first, last;
    K.determineRange(i, first, last);
    uint32_t split = K.findSplit(first, last);
    uint32_t a = split, b = split + 1;
    if (a == first) a += nf - 1;
    if (b == last) b += nf - 1;
    nodes[i].left = a;
    nodes[i].right = b;
    nodes[i].isLeafFlag = 0;
    nodes[a].parent = i;
    nodes[a].isLeft = 1;
    nodes[b].parent = i;
    nodes[b].isLeft = 0;
  }
  nodes[0].parent = 0;
  nodes[0].isLeft = 0;
  if (nf == 1) return 0;
  // kFixAABBTree: each leaf walks up; the first arrival at a node stops, the
  // second merges both children (atomicCAS(lock, 0, 1) semantics).
  std::vector<uint32_t> lock(N, 0);
  for (uint32_t i = nf - 1; i < N; ++i) {
    uint32_t node = nodes[i].parent;
    while (true) {
      if (lock[node] == 0) {
        lock[node] = 1;
        break;
      }
      Box b = nodes[nodes[node].left].bbox;
      mergeWith(b, nodes[nodes[node].right].bbox);
      nodes[node].bbox = b;
      nodes[node].surfaceArea = boxSA(b);
      if (node == nodes[node].parent) break;
      node = nodes[node].parent;
    }
  }
  return 0;
}

// The reference Morton helpers alone (for known-answer checks).
uint64_t orc_spread3(uint64_t v, int bits) { return spread3(v, bits); }
void orc_init_tri(const float *vert, const uint32_t *faces, uint32_t i, float *box6, float *centroid3) {
  TriData t = initTri(reinterpret_cast<const V3 *>(vert), faces, i);
  float b[6] = {t.bbox.mn.x, t.bbox.mn.y, t.bbox.mn.z, t.bbox.mx.x, t.bbox.mx.y, t.bbox.mx.z};
  std::memcpy(box6, b, sizeof(b));
  centroid3[0] = t.centroid.x;
  centroid3[1] = t.centroid.y;
  centroid3[2] = t.centroid.z;
}
float orc_box_sa(const float *box6) {
  Box b = {{box6[0], box6[1], box6[2]}, {box6[3], box6[4], box6[5]}};
  return boxSA(b);
}

// BVH::calcSAH -- lib/base/BVH.cpp:23-37 (f64 accumulation instead of the
// reference's order-dependent f32 OpenMP reduction).
double orc_calc_sah(const Node *nodes, uint32_t N, uint32_t root, float cInner, float cLeaf) {
  double s = 0.0;
  for (uint32_t i = 0; i < N; ++i) {
    if (i == root) continue;
    float c = nodes[i].surfaceArea;
    c *= nodes[i].isLeafFlag != 0 ? cLeaf : cInner;
    s += c;
  }
  return s / nodes[root].surfaceArea;
}

// BVH::fixLevels -- lib/base/BVH.cpp:39-70 (128-bit bit stack DFS).
// Returns the max level.
uint32_t orc_fix_levels(Node *nodes, uint32_t N, uint32_t root) {
  if (N == 0) return 0;
  unsigned __int128 stack = 0;
  uint16_t level = 0, maxLevel = 0;
  Node *node = nodes + root;
  while (true) {
    node->level = level;
    if (node->isLeafFlag == 0) {
      stack = (stack << 1) | 1;
      node = nodes + node->left;
      level++;
      maxLevel = std::max(level, maxLevel);
      continue;
    }
    while ((stack & 1) == 0) {
      if (stack == 0) return maxLevel;
      node = nodes + node->parent;
      stack >>= 1;
      level--;
    }
    node = nodes + nodes[node->parent].right;
    stack ^= 1;
  }
}

// Validate -- lib/misc/Validate.cpp:42-263.  err[0..5] = tree errors
// (checkTree total), bbox errors, strict bbox errors, surface-area errors,
// triangle errors, orphan triangles.  Returns the sum.
uint32_t orc_validate(const Node *nodes, uint32_t N, uint32_t root, uint32_t maxLevel, const float *vert,
                      const uint32_t *faces, uint32_t nf, uint32_t *err) {
  const V3 *V = reinterpret_cast<const V3 *>(vert);
  uint32_t e[6] = {0, 0, 0, 0, 0, 0};
  if (N > 0) {
    // checkTree (:42-124)
    unsigned __int128 stack = 0;
    uint32_t level = 0, mlev = 0, node = root, last = root;
    std::vector<uint8_t> seen(N, 0);
    if (node != nodes[node].parent) e[0]++;
    uint64_t guard = 0;
    while (true) {
      if (++guard > 4ull * N + 16) {
        e[0]++;
        break;
      }
      seen[node] = 1;
      if (nodes[node].level != level) e[0]++;
      if (node != root) {
        if (nodes[node].isLeft != 0 && nodes[nodes[node].parent].left != node) e[0]++;
        if (nodes[node].isLeft == 0 && nodes[nodes[node].parent].right != node) e[0]++;
      }
      if (nodes[node].isLeafFlag == 0) {
        if (nodes[node].left == nodes[node].right) e[0]++;
        if (node == nodes[node].right) e[0]++;
        if (node == nodes[node].left) e[0]++;
        stack = (stack << 1) | 1;
        last = node;
        node = nodes[node].left;
        if (node >= N) {
          e[0]++;
          break;
        }
        if (nodes[node].parent != last) e[0]++;
        level++;
        mlev = std::max(level, mlev);
        continue;
      }
      bool done = false;
      while ((stack & 1) == 0) {
        if (stack == 0) {
          done = true;
          break;
        }
        node = nodes[node].parent;
        stack >>= 1;
        level--;
      }
      if (done) break;
      last = nodes[node].parent;
      node = nodes[last].right;
      if (nodes[node].parent != last) e[0]++;
      stack ^= 1;
    }
    for (uint32_t i = 0; i < N; ++i)
      if (!seen[i]) e[0]++;
    if (mlev != maxLevel) e[0]++;
  }
  std::vector<uint8_t> tseen(nf, 0);
  for (uint32_t i = 0; i < N; ++i) {
    const Node &n = nodes[i];
    if (n.isLeafFlag == 0) {
      const Box &b = n.bbox, &l = nodes[n.left].bbox, &r = nodes[n.right].bbox;
      bool bad = !(b.mn.x <= l.mn.x && b.mn.y <= l.mn.y && b.mn.z <= l.mn.z && b.mx.x >= l.mx.x && b.mx.y >= l.mx.y &&
                   b.mx.z >= l.mx.z && b.mn.x <= r.mn.x && b.mn.y <= r.mn.y && b.mn.z <= r.mn.z &&
                   b.mx.x >= r.mx.x && b.mx.y >= r.mx.y && b.mx.z >= r.mx.z);
      if (bad) e[1]++;
      bool strict = b.mn.x == std::min(l.mn.x, r.mn.x) && b.mn.y == std::min(l.mn.y, r.mn.y) &&
                    b.mn.z == std::min(l.mn.z, r.mn.z) && b.mx.x == std::max(l.mx.x, r.mx.x) &&
                    b.mx.y == std::max(l.mx.y, r.mx.y) && b.mx.z == std::max(l.mx.z, r.mx.z);
      if (!strict) e[2]++;
    } else {
      uint32_t errs = 0;
      for (uint32_t j = n.left; j < n.left + n.right; ++j) {
        if (j >= nf) {
          errs++;
          continue;
        }
        tseen[j] = 1;
        for (int k = 0; k < 3; ++k) {
          V3 v = V[faces[3 * j + k]];
          if (!(n.bbox.mn.x <= v.x && n.bbox.mn.y <= v.y && n.bbox.mn.z <= v.z && n.bbox.mx.x >= v.x &&
                n.bbox.mx.y >= v.y && n.bbox.mx.z >= v.z))
            errs++;
        }
      }
      if (errs) e[4]++;
    }
    if (std::fabs(n.surfaceArea - boxSA(n.bbox)) >= 0.000001f) e[3]++;
  }
  for (uint32_t j = 0; j < nf; ++j)
    if (!tseen[j]) e[5]++;
  uint32_t total = 0;
  for (int k = 0; k < 6; ++k) {
    if (err) err[k] = e[k];
    total += e[k];
  }
  return total;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Rays and traversal -- lib/tracer/CUDAKernels.cu, lib/misc/Camera.cpp
// ---------------------------------------------------------------------------

namespace orc {

// Ray::set -- lib/base/Ray.hpp:65-74
inline void setRay(Ray &r, V3 pos, V3 dir, uint32_t x, uint32_t y) {
  r.pos = pos;
  r.dir = dir;
  r.px = x;
  r.py = y;
  r.inv = {1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z};
  r.sign[0] = r.inv.x < 0 ? 1 : 0;
  r.sign[1] = r.inv.y < 0 ? 1 : 0;
  r.sign[2] = r.inv.z < 0 ? 1 : 0;
  r.pad = 0;
}

inline float boxCoord(const Box &b, int side, int axis) {
  const V3 &v = side ? b.mx : b.mn;
  return axis == 0 ? v.x : (axis == 1 ? v.y : v.z);
}

// intersectRayAABB -- lib/tracer/CUDAKernels.cu:66-85
inline bool hitBox(const Box &b, const Ray &r, float t0, float t1, float &tmin) {
  float tmax, tymin, tymax, tzmin, tzmax;
  tmin = (boxCoord(b, r.sign[0], 0) - r.pos.x) * r.inv.x;
  tmax = (boxCoord(b, 1 - r.sign[0], 0) - r.pos.x) * r.inv.x;
  tymin = (boxCoord(b, r.sign[1], 1) - r.pos.y) * r.inv.y;
  tymax = (boxCoord(b, 1 - r.sign[1], 1) - r.pos.y) * r.inv.y;
  tzmin = (boxCoord(b, r.sign[2], 2) - r.pos.z) * r.inv.z;
  tzmax = (boxCoord(b, 1 - r.sign[2], 2) - r.pos.z) * r.inv.z;
  if (tymin > tmin) tmin = tymin;
  if (tzmin > tmin) tmin = tzmin;
  if (tymax < tmax) tmax = tymax;
  if (tzmax < tmax) tmax = tzmax;
  return (tmin < tmax) && (tmin < t1) && (tmax > t0);
}

// intersectRayTriangle2 -- lib/tracer/CUDAKernels.cu:88-112 (face culling on,
// the meson default).  Unpinned: __fdividef(1, det) is an IEEE division here,
// and nvcc may contract the dot/cross products into FMAs.
inline bool hitTri(V3 orig, V3 dir, V3 v0, V3 v1, V3 v2, float &dist, bool culling) {
  V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
  V3 p = cross(dir, e2);
  float det = dot(e1, p);
  V3 tv = sub(orig, v0);
  V3 q = cross(tv, e1);
  float invDet = 1.0f / det;
  float u = dot(tv, p);
  float v = dot(dir, q);
  dist = dot(e2, q) * invDet;
  bool front = (det > 0.0f) && ((u >= 0.0f && u <= det) && (v >= 0.0f && (u + v) <= det));
  if (culling) return front;
  return front || ((det < 0.0f) && ((u <= 0.0f && u >= det) && (v <= 0.0f && (u + v) >= det)));
}

}  // namespace orc

extern "C" {

// kGenerateRays + generateRays -- lib/tracer/CUDAKernels.cu:43-59, :386-397
// (same as Camera::genRays, lib/misc/Camera.cpp:67-88).  Pixel offsets are
// computed in f64 as in the reference; the camera-to-world matrix is the
// exact inverse of lookAtRH computed in f64 (unpinned: glm's f32 inverse).
void orc_gen_rays(uint32_t w, uint32_t h, const float *pos, const float *lookAt, const float *up, float fov,
                  Ray *rays) {
  double P[3] = {pos[0], pos[1], pos[2]};
  double f[3] = {lookAt[0] - P[0], lookAt[1] - P[1], lookAt[2] - P[2]};
  double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double &c : f) c /= fl;
  double U[3] = {up[0], up[1], up[2]};
  double s[3] = {f[1] * U[2] - f[2] * U[1], f[2] * U[0] - f[0] * U[2], f[0] * U[1] - f[1] * U[0]};
  double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (double &c : s) c /= sl;
  double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  float ratio = (float)w / (float)h;
  float scale = (float)std::tan((0.5 * (double)fov) * (M_PI / 180.0));
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      float px = (float)((2 * ((x + 0.5) / w) - 1) * scale * ratio);
      float py = (float)((1 - 2 * ((y + 0.5) / h)) * scale);
      // inverse(lookAtRH) has columns s, u, -f, pos: dir = px*s + py*u + f
      double d[3];
      for (int k = 0; k < 3; ++k) d[k] = (double)px * s[k] + (double)py * u[k] + f[k];
      double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      V3 dir = {(float)(d[0] / dl), (float)(d[1] / dl), (float)(d[2] / dl)};
      setRay(rays[y * w + x], V3{pos[0], pos[1], pos[2]}, dir, x, y);
    }
}

// Ray::Ray(pos, dir) for externally built rays.
void orc_set_ray(Ray *r, const float *pos, const float *dir, uint32_t x, uint32_t y) {
  setRay(*r, V3{pos[0], pos[1], pos[2]}, V3{dir[0], dir[1], dir[2]}, x, y);
}

// kTraceRay -- lib/tracer/CUDAKernels.cu:114-223 (stackless traversal with a
// 128-bit bit stack, Afra & Szirmay-Kalos 2014), plus the shading of
// :208-218.  mode 0: reference triangle test; mode 1: additionally require
// t > tMin and skip triangle skip[i] (secondary rays).  Rays with dir == 0
// are reported as misses.
void orc_trace(const Ray *rays, uint32_t num, const Node *nodes, uint32_t root, const float *vert,
               const uint32_t *faces, const float *light, int culling, int mode, float tMin, const uint32_t *skip,
               Pixel *pix, Hit *hits) {
  const V3 *V = reinterpret_cast<const V3 *>(vert);
  V3 L = {light[0], light[1], light[2]};
#pragma omp parallel for schedule(dynamic, 512)
  for (int64_t ri = 0; ri < (int64_t)num; ++ri) {
    const Ray &ray = rays[ri];
    Pixel res = {0, 0, 0};
    float nearest = HUGE_VALF;
    uint32_t closest = UINT32_MAX;
    uint32_t skipTri = skip ? skip[ri] : UINT32_MAX;
    bool active = !(ray.dir.x == 0.0f && ray.dir.y == 0.0f && ray.dir.z == 0.0f);
    unsigned __int128 stack = 0;
    uint32_t node = root;
    float dist, minL, minR;
    while (active) {
      const Node &nd = nodes[node];
      if (nd.isLeafFlag == 0) {
        res.intCount++;
        uint32_t l = nd.left, r = nd.right;
        bool hl = hitBox(nodes[l].bbox, ray, 0.01f, nearest + 0.01f, minL);
        bool hr = hitBox(nodes[r].bbox, ray, 0.01f, nearest + 0.01f, minR);
        if (hl || hr) {
          stack <<= 1;
          if (hl && hr) {
            stack |= 1;
            node = minL < minR ? l : r;
          } else {
            node = hl ? l : r;
          }
          continue;
        }
      } else {
        for (uint32_t k = 0; k < nd.right; ++k) {
          uint32_t t = nd.left + k;
          const uint32_t *f = faces + 3 * t;
          bool h = hitTri(ray.pos, ray.dir, V[f[0]], V[f[1]], V[f[2]], dist, culling != 0);
          if (mode == 1) h = h && dist > tMin && t != skipTri;
          if (h && dist < nearest) {
            nearest = dist;
            closest = t;
          }
        }
      }
      bool end = false;
      while ((stack & 1) == 0) {
        if (stack == 0) {
          end = true;
          break;
        }
        node = nodes[node].parent;
        stack >>= 1;
      }
      if (end) break;
      uint32_t p = nodes[node].parent;
      node = nodes[node].isLeft == 0 ? nodes[p].left : nodes[p].right;
      stack ^= 1;
    }
    if (nearest < HUGE_VALF) {
      const uint32_t *f = faces + 3 * closest;
      V3 nrm = normalize(cross(sub(V[f[1]], V[f[0]]), sub(V[f[2]], V[f[0]])));  // glm::triangleNormal
      V3 hp = add(ray.pos, mul(ray.dir, nearest));
      V3 ld = normalize(sub(L, hp));
      float diff = 1.0f + dot(nrm, ld);
      diff = diff > 0.0f ? diff : 0.0f;
      res.hit = 1;
      res.diffuse = (uint8_t)(diff * 127.0f);
    }
    if (pix) pix[ri] = res;
    if (hits) hits[ri] = Hit{closest, nearest};
  }
}

// kTraceRayBundle -- lib/tracer/CUDAKernels.cu:240-383: the rays of an 8x8
// pixel tile (one thread block) walk one shared node sequence.  At an inner
// node every ray tests both child boxes against its own nearest hit; a child
// is entered when any ray hits it; when both are, the trail bit is set and
// the majority prefers the near child (a ray votes +1 for the left child,
// -1 for the right, 0 without preference; a sum > 0 enters the left child,
// otherwise the right).  The other child is entered later through the
// trail (Afra & Szirmay-Kalos backtracking, as kTraceRay) without a second
// box test.  At a leaf every ray tests the triangles and keeps the first one
// at a strictly smaller distance.  intCount counts the inner nodes the
// bundle visits (the same for every ray of a tile).  Lanes of a tile that
// lie outside the image take no part (in the reference their vote entries
// are left unwritten).  Shading as kTraceRay.
void orc_trace_bundle(const Ray *rays, uint32_t w, uint32_t h, const Node *nodes, uint32_t root, const float *vert,
                      const uint32_t *faces, const float *light, Pixel *pix, Hit *hits) {
  const V3 *V = reinterpret_cast<const V3 *>(vert);
  const V3 L = {light[0], light[1], light[2]};
  const uint32_t tx = (w + 7) / 8, ty = (h + 7) / 8;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t tile = 0; tile < (int64_t)tx * ty; ++tile) {
    const uint32_t bx = (uint32_t)(tile % tx) * 8, by = (uint32_t)(tile / tx) * 8;
    bool valid[64];
    uint32_t idx[64], closest[64];
    float nearest[64];
    for (int l = 0; l < 64; ++l) {
      const uint32_t x = bx + (l & 7), y = by + (l >> 3);
      valid[l] = x < w && y < h;
      idx[l] = valid[l] ? y * w + x : 0;
      nearest[l] = HUGE_VALF;
      closest[l] = UINT32_MAX;
    }
    uint16_t intCount = 0;
    unsigned __int128 stack = 0;
    uint32_t node = root;
    float dist;
    while (true) {
      const Node &nd = nodes[node];
      if (nd.isLeafFlag == 0) {
        intCount++;
        const uint32_t lc = nd.left, rc = nd.right;
        bool anyL = false, anyR = false;
        int vote = 0;
        for (int l = 0; l < 64; ++l) {
          if (!valid[l]) continue;
          float minL = 0.0f, minR = 0.0f;
          const Ray &ray = rays[idx[l]];
          const bool hl = hitBox(nodes[lc].bbox, ray, 0.01f, nearest[l] + 0.01f, minL);
          const bool hr = hitBox(nodes[rc].bbox, ray, 0.01f, nearest[l] + 0.01f, minR);
          anyL |= hl;
          anyR |= hr;
          vote += (hl && (minL < minR || !hr)) - (hr && (minR < minL || !hl));
        }
        if (anyL || anyR) {
          stack <<= 1;
          if (anyL && anyR) {
            stack |= 1;
            node = vote > 0 ? lc : rc;
          } else {
            node = anyL ? lc : rc;
          }
          continue;
        }
      } else {
        for (int l = 0; l < 64; ++l) {
          if (!valid[l]) continue;
          const Ray &ray = rays[idx[l]];
          for (uint32_t k = 0; k < nd.right; ++k) {
            const uint32_t t = nd.left + k;
            const uint32_t *f = faces + 3 * t;
            if (hitTri(ray.pos, ray.dir, V[f[0]], V[f[1]], V[f[2]], dist, true) && dist < nearest[l]) {
              nearest[l] = dist;
              closest[l] = t;
            }
          }
        }
      }
      bool end = false;
      while ((stack & 1) == 0) {
        if (stack == 0) {
          end = true;
          break;
        }
        node = nodes[node].parent;
        stack >>= 1;
      }
      if (end) break;
      const uint32_t p = nodes[node].parent;
      node = nodes[node].isLeft == 0 ? nodes[p].left : nodes[p].right;
      stack ^= 1;
    }
    for (int l = 0; l < 64; ++l) {
      if (!valid[l]) continue;
      Pixel res = {0, 0, intCount};
      if (nearest[l] < HUGE_VALF) {
        const Ray &ray = rays[idx[l]];
        const uint32_t *f = faces + 3 * closest[l];
        V3 nrm = normalize(cross(sub(V[f[1]], V[f[0]]), sub(V[f[2]], V[f[0]])));
        V3 hp = add(ray.pos, mul(ray.dir, nearest[l]));
        V3 ld = normalize(sub(L, hp));
        float diff = 1.0f + dot(nrm, ld);
        diff = diff > 0.0f ? diff : 0.0f;
        res.hit = 1;
        res.diffuse = (uint8_t)(diff * 127.0f);
      }
      if (pix) pix[idx[l]] = res;
      if (hits) hits[idx[l]] = Hit{closest[l], nearest[l]};
    }
  }
}

// kRefitLeafBVHTris -- lib/tracer/CUDAKernels.cu:492-525: the box of each
// listed leaf from its (first) triangle's current vertices, widened by
// FLT_EPSILON; the node is written back with its old surfaceArea (the kernel
// does not recompute it).
void orc_refit_leaf_tris(Node *nodes, const float *vert, const uint32_t *faces, const uint32_t *ids, uint32_t num) {
  const V3 *V = reinterpret_cast<const V3 *>(vert);
  for (uint32_t i = 0; i < num; ++i) {
    Node &n = nodes[ids[i]];
    const uint32_t *f = faces + 3 * n.left;
    const V3 a = V[f[0]], b = V[f[1]], c = V[f[2]];
    Box bb = {a, b};
    if (b.x < bb.mn.x) bb.mn.x = b.x;
    if (b.y < bb.mn.y) bb.mn.y = b.y;
    if (b.z < bb.mn.z) bb.mn.z = b.z;
    if (c.x < bb.mn.x) bb.mn.x = c.x;
    if (c.y < bb.mn.y) bb.mn.y = c.y;
    if (c.z < bb.mn.z) bb.mn.z = c.z;
    if (a.x > bb.mx.x) bb.mx.x = a.x;
    if (a.y > bb.mx.y) bb.mx.y = a.y;
    if (a.z > bb.mx.z) bb.mx.z = a.z;
    if (c.x > bb.mx.x) bb.mx.x = c.x;
    if (c.y > bb.mx.y) bb.mx.y = c.y;
    if (c.z > bb.mx.z) bb.mx.z = c.z;
    const float e = std::numeric_limits<float>::epsilon();
    bb.mn = {bb.mn.x - e, bb.mn.y - e, bb.mn.z - e};
    bb.mx = {bb.mx.x + e, bb.mx.y + e, bb.mx.z + e};
    n.bbox = bb;
  }
}

// Brute force over all faces (lib/tracer/BruteForceTracer.cpp:56-120) with
// the GPU triangle test above, restricted to tMin < t: the independent check
// of the traversal on scenes whose hits all lie in front of the camera.
void orc_trace_bruteforce(const Ray *rays, uint32_t num, const float *vert, const uint32_t *faces, uint32_t nf,
                          int culling, float tMin, Hit *hits) {
  const V3 *V = reinterpret_cast<const V3 *>(vert);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t ri = 0; ri < (int64_t)num; ++ri) {
    const Ray &ray = rays[ri];
    float nearest = HUGE_VALF, dist;
    uint32_t closest = UINT32_MAX;
    for (uint32_t t = 0; t < nf; ++t) {
      const uint32_t *f = faces + 3 * t;
      if (hitTri(ray.pos, ray.dir, V[f[0]], V[f[1]], V[f[2]], dist, culling != 0) && dist > tMin && dist < nearest) {
        nearest = dist;
        closest = t;
      }
    }
    hits[ri] = Hit{closest, nearest};
  }
}

}  // extern "C"
