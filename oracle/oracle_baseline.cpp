// ORACLE -- test infrastructure, not product code (see oracle.cpp).
//
// OpenMP CPU LBVH used as the measured CPU baseline of bench.py: the same
// stages and arithmetic as orc_lbvh (lib/builder/LBVH_CUDA.cu:60-374), run
// with OpenMP loops, a parallel stable sort, and an atomic bottom-up refit
// (the kFixAABBTree protocol with acquire/release atomics).  Its output is
// checked against orc_lbvh in tests/test_oracle.py.
#include "oracle.hpp"

#include <atomic>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

namespace orc {
namespace {

inline uint32_t spread10(uint32_t v) {
  uint32_t r = 0;
  for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
  return r;
}
inline uint64_t spread21(uint64_t v) {
  uint64_t r = 0;
  for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (3 * b);
  return r;
}

inline Box triBox(const V3 *V, const uint32_t *f, V3 &c) {
  V3 v1 = V[f[0]], v2 = V[f[1]], v3 = V[f[2]];
  Box b{v1, v2};
  if (v2.x < b.mn.x) b.mn.x = v2.x;
  if (v2.y < b.mn.y) b.mn.y = v2.y;
  if (v2.z < b.mn.z) b.mn.z = v2.z;
  if (v3.x < b.mn.x) b.mn.x = v3.x;
  if (v3.y < b.mn.y) b.mn.y = v3.y;
  if (v3.z < b.mn.z) b.mn.z = v3.z;
  if (v1.x > b.mx.x) b.mx.x = v1.x;
  if (v1.y > b.mx.y) b.mx.y = v1.y;
  if (v1.z > b.mx.z) b.mx.z = v1.z;
  if (v3.x > b.mx.x) b.mx.x = v3.x;
  if (v3.y > b.mx.y) b.mx.y = v3.y;
  if (v3.z > b.mx.z) b.mx.z = v3.z;
  const float e = std::numeric_limits<float>::epsilon();
  b.mn = {b.mn.x - e, b.mn.y - e, b.mn.z - e};
  b.mx = {b.mx.x + e, b.mx.y + e, b.mx.z + e};
  V3 s = add(add(v1, v2), v3);
  c = {s.x / 3.0f, s.y / 3.0f, s.z / 3.0f};
  return b;
}

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }
inline int dlt(uint32_t n, uint32_t i, uint32_t j) { return j >= n ? -1 : clz32(i ^ j); }

}  // namespace
}  // namespace orc

using namespace orc;

extern "C" {

// Returns wall seconds of the build (allocation excluded, like
// LBVH::runImpl timing in lib/base/Command.cpp:52-54).
double orc_lbvh_omp(const float *vert, const uint32_t *faces, uint32_t nf, int mortonBits, Node *nodes,
                    uint32_t *sortedFaces, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  const V3 *V = reinterpret_cast<const V3 *>(vert);
  const uint32_t N = 2 * nf - 1;
  std::vector<Box> box(nf);
  std::vector<V3> cen(nf);
  std::vector<std::pair<uint64_t, uint32_t>> kv(nf);
  std::vector<std::atomic<uint32_t>> lock(N);
  auto t0 = std::chrono::steady_clock::now();
  float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
#pragma omp parallel
  {
    float lmn[3] = {0, 0, 0}, lmx[3] = {0, 0, 0};
#pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)nf; ++i) {
      box[i] = triBox(V, faces + 3 * i, cen[i]);
      const Box &b = box[i];
      lmn[0] = lmn[0] < b.mn.x ? lmn[0] : b.mn.x;
      lmn[1] = lmn[1] < b.mn.y ? lmn[1] : b.mn.y;
      lmn[2] = lmn[2] < b.mn.z ? lmn[2] : b.mn.z;
      lmx[0] = lmx[0] > b.mx.x ? lmx[0] : b.mx.x;
      lmx[1] = lmx[1] > b.mx.y ? lmx[1] : b.mx.y;
      lmx[2] = lmx[2] > b.mx.z ? lmx[2] : b.mx.z;
    }
#pragma omp critical
    for (int a = 0; a < 3; ++a) {
      mn[a] = mn[a] < lmn[a] ? mn[a] : lmn[a];
      mx[a] = mx[a] > lmx[a] ? mx[a] : lmx[a];
    }
  }
  float off[3], scale[3];
  for (int a = 0; a < 3; ++a) {
    off[a] = 0.0f - mn[a];
    scale[a] = 1.0f / (mx[a] + off[a]);
  }
  const int per = mortonBits / 3;
  const float q = (float)(1u << per), qmax = (float)((1u << per) - 1u);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)nf; ++i) {
    float c[3] = {cen[i].x, cen[i].y, cen[i].z};
    uint64_t e[3];
    for (int a = 0; a < 3; ++a) {
      float t = c[a];
      t += off[a];
      t *= scale[a];
      t *= q;
      if (t < 0.0f) t = 0.0f;
      if (t > qmax) t = qmax;
      e[a] = per == 10 ? spread10((uint32_t)t) : spread21((uint64_t)(uint32_t)t);
    }
    kv[i] = {e[0] * 4 + e[1] * 2 + e[2], (uint32_t)i};
  }
  // (key, index) pairs are unique, so any sort equals a stable sort by key.
#ifdef _OPENMP
  __gnu_parallel::sort(kv.begin(), kv.end());
#else
  std::sort(kv.begin(), kv.end());
#endif
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)nf; ++i) {
    Node &l = nodes[nf - 1 + i];
    l.bbox = box[kv[i].second];
    l.parent = UINT32_MAX;
    l.left = kv[i].second;
    l.right = 1;
    l.isLeafFlag = 1;
    l.isLeft = 0;
    l.level = UINT8_MAX;
    l.pad = 0;
    l.surfaceArea = boxSA(l.bbox);
    if (sortedFaces) sortedFaces[i] = kv[i].second;
  }
  // Karras over the reference's index keys (delta = clz32(i ^ j)).
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)nf - 1; ++ii) {
    uint32_t i = (uint32_t)ii;
    int32_t d = dlt(nf, i, i + 1) - dlt(nf, i, i - 1);
    d = d > 0 ? 1 : -1;
    int dmin = dlt(nf, i, i - (uint32_t)d);
    uint32_t lMax = 2;
    while (dlt(nf, i, i + lMax * (uint32_t)d) > dmin) lMax *= 2;
    uint32_t l = 0;
    do {
      lMax >>= 1;
      if (dlt(nf, i, i + (l + lMax) * (uint32_t)d) > dmin) l += lMax;
    } while (lMax > 1);
    uint32_t j = i + l * (uint32_t)d;
    uint32_t first = i < j ? i : j, last = i < j ? j : i;
    int common = clz32(first ^ last);
    uint32_t split = first, step = last - first;
    do {
      step = (step + 1) >> 1;
      uint32_t ns = split + step;
      if (ns < last && clz32(first ^ ns) > common) split = ns;
    } while (step > 1);
    uint32_t a = split, b = split + 1;
    if (a == first) a += nf - 1;
    if (b == last) b += nf - 1;
    nodes[i].left = a;
    nodes[i].right = b;
    nodes[i].isLeafFlag = 0;
    nodes[i].level = 0;
    nodes[i].pad = 0;
    nodes[a].parent = i;
    nodes[a].isLeft = 1;
    nodes[b].parent = i;
    nodes[b].isLeft = 0;
  }
  nodes[0].parent = 0;
  nodes[0].isLeft = 0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)N; ++i) lock[i].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t i = nf - 1; i < (int64_t)N; ++i) {
    uint32_t node = nodes[i].parent;
    while (nf > 1) {
      if (lock[node].fetch_add(1, std::memory_order_acq_rel) == 0) break;
      Box b = nodes[nodes[node].left].bbox;
      mergeWith(b, nodes[nodes[node].right].bbox);
      nodes[node].bbox = b;
      nodes[node].surfaceArea = boxSA(b);
      std::atomic_thread_fence(std::memory_order_release);
      if (node == nodes[node].parent) break;
      node = nodes[node].parent;
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

int orc_omp_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
