// Ray generation and traversal for gfx950
// (reference: lib/tracer/CUDAKernels.cu:43-223, :386-417).
//
// Traversal visits nodes in the reference's order (Afra & Szirmay-Kalos
// 2014, the bit-trail algorithm of kTraceRay): at an inner node both children
// are tested against the current nearest hit, the nearer hit child is entered
// first and the other is remembered by a set trail bit; a remembered child is
// entered later without a second box test.  Two changes make it fast without
// changing that order:
//   * each inner node's two child boxes, ids and leaf triangles are packed
//     into one 64-byte record (k_packPairs), each triangle into (v0, e1, e2)
//     (k_packTris, the same f32 subtractions the test makes), so a visit is
//     one 64-byte load instead of three 48-byte node loads;
//   * the remembered children are cached in a 16-entry per-lane LDS stack, so
//     backtracking pops instead of walking parent links.  The trail stays the
//     authority: when the cache has dropped an entry (deeper than 16 pending
//     children), the walk falls back to the parent links.
// Arithmetic is f32 with no FMA contraction and IEEE division, so the visit
// order, hit triangle and distance are a pure function of the inputs (the
// oracle reproduces them exactly).  The 128-bit trail limits trees to depth
// 128, as in the reference.
//
// Rays are generated in f64 from a camera basis computed in f64 on the host
// (the exact inverse of lookAtRH), then rounded once to f32.
#include "radix_sort.hpp"
#include "runtime.hpp"
#include "types.hpp"

#include <algorithm>
#include <cmath>

namespace bvht {

namespace {

constexpr uint32_t kNoHit = 0xFFFFFFFFu;

struct CamBasis {
  double s[3], u[3], f[3];
  float pos[3];
  float scale, ratio;
};

__device__ __forceinline__ void setRay(bvht_Ray &r, bvht_vec3 pos, bvht_vec3 dir, uint32_t x, uint32_t y) {
  // Ray::set -- lib/base/Ray.hpp:65-74
  r.pos = pos;
  r.dir = dir;
  r.px = x;
  r.py = y;
  r.invDir = {__fdiv_rn(1.0f, dir.x), __fdiv_rn(1.0f, dir.y), __fdiv_rn(1.0f, dir.z)};
  r.sign[0] = r.invDir.x < 0 ? 1 : 0;
  r.sign[1] = r.invDir.y < 0 ? 1 : 0;
  r.sign[2] = r.invDir.z < 0 ? 1 : 0;
  r._pad = 0;
}

// kGenerateRays -- CUDAKernels.cu:43-59
__global__ __launch_bounds__(256) void k_genRays(bvht_Ray *__restrict__ rays, uint32_t w, uint32_t h, CamBasis cam) {
  const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15);
  const uint32_t y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= w || y >= h) return;
  const float px = (float)((2 * ((x + 0.5) / w) - 1) * cam.scale * cam.ratio);
  const float py = (float)((1 - 2 * ((y + 0.5) / h)) * cam.scale);
  double d[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) d[k] = (double)px * cam.s[k] + (double)py * cam.u[k] + cam.f[k];
  const double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  bvht_vec3 dir = {(float)(d[0] / dl), (float)(d[1] / dl), (float)(d[2] / dl)};
  bvht_Ray r;
  setRay(r, bvht_vec3{cam.pos[0], cam.pos[1], cam.pos[2]}, dir, x, y);
  rays[(size_t)y * w + x] = r;
}

__device__ __forceinline__ bvht_vec3 vsub(bvht_vec3 a, bvht_vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ bvht_vec3 vcross(bvht_vec3 x, bvht_vec3 y) {
  return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
__device__ __forceinline__ float vdot(bvht_vec3 a, bvht_vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// glm::normalize = v * (1 / sqrt(dot(v, v))), both operations correctly
// rounded.  sqrtf (not __fsqrt_rn, which maps to the 1-ulp v_sqrt_f32 on this
// toolchain) lowers to v_sqrt_f32 plus an fma residual correction.
__device__ __forceinline__ bvht_vec3 vnormalize(bvht_vec3 v) {
  float s = __fdiv_rn(1.0f, sqrtf(vdot(v, v)));
  return {v.x * s, v.y * s, v.z * s};
}

struct TraceOut {
  uint32_t tri;
  float t;
  uint16_t intCount;
};

// One inner node's children: boxes (min.xyz, max.xyz each), node ids and
// leaf info (kInnerChild, the first face of a one-triangle leaf, or
// kLookupLeaf for a leaf holding several triangles).
struct PackedPair {
  float4 a;  // l.min.x l.min.y l.min.z l.max.x
  float4 b;  // l.max.y l.max.z r.min.x r.min.y
  float4 c;  // r.min.z r.max.x r.max.y r.max.z
  uint4 d;   // l r lInfo rInfo
};
struct PackedTri {
  float4 v0, e1, e2;  // e1 = v1 - v0, e2 = v2 - v0 (w unused)
};
constexpr uint32_t kInnerChild = 0xFFFFFFFFu, kLookupLeaf = 0xFFFFFFFEu, kEmptyLeaf = 0xFFFFFFFDu;
constexpr int kStackCap = 16;  // cached remembered children per lane (power of two)
constexpr uint32_t kTraceWavesPerCU = 20;  // 160 KiB of LDS / 8 KiB of stack cache per wave
constexpr uint32_t kRefillAt = 16;         // idle lanes that make a k_traceRaysDyn wave fetch rays

__device__ __forceinline__ uint32_t leafInfo(const bvht_BVHNode &n) {
  if (n.isLeafFlag == 0) return kInnerChild;
  return n.right == 1 ? n.left : kLookupLeaf;
}

// Packs every inner node below `bound`; a child index at or past `bound`
// becomes an empty leaf and raises *bad (the traversal then never reads past
// the arrays).
__global__ __launch_bounds__(256) void k_packPairs(const bvht_BVHNode *__restrict__ nodes, uint32_t bound,
                                                   PackedPair *__restrict__ pairs, uint32_t *__restrict__ bad) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= bound) return;
  const bvht_BVHNode n = nodes[i];
  if (n.isLeafFlag != 0) return;
  uint32_t l = n.left, r = n.right, li = kEmptyLeaf, ri = kEmptyLeaf;
  bvht_AABB lb = {{{0.f, 0.f, 0.f}, {-1.f, -1.f, -1.f}}}, rb = lb;
  if (l < bound) {
    const bvht_BVHNode c = nodes[l];
    lb = c.bbox;
    li = leafInfo(c);
  } else {
    *bad = 1;
    l = i;
  }
  if (r < bound) {
    const bvht_BVHNode c = nodes[r];
    rb = c.bbox;
    ri = leafInfo(c);
  } else {
    *bad = 1;
    r = i;
  }
  PackedPair p;
  p.a = make_float4(lb.minMax[0].x, lb.minMax[0].y, lb.minMax[0].z, lb.minMax[1].x);
  p.b = make_float4(lb.minMax[1].y, lb.minMax[1].z, rb.minMax[0].x, rb.minMax[0].y);
  p.c = make_float4(rb.minMax[0].z, rb.minMax[1].x, rb.minMax[1].y, rb.minMax[1].z);
  p.d = make_uint4(l, r, li, ri);
  pairs[i] = p;
}

__global__ __launch_bounds__(256) void k_packTris(const bvht_Triangle *__restrict__ faces,
                                                  const bvht_vec3 *__restrict__ vert, uint32_t nf,
                                                  PackedTri *__restrict__ tris) {
  const uint32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nf) return;
  const bvht_Triangle t = faces[f];
  const bvht_vec3 v0 = vert[t.v1], v1 = vert[t.v2], v2 = vert[t.v3];
  const bvht_vec3 e1 = vsub(v1, v0), e2 = vsub(v2, v0);
  tris[f] = PackedTri{make_float4(v0.x, v0.y, v0.z, 0.f), make_float4(e1.x, e1.y, e1.z, 0.f),
                      make_float4(e2.x, e2.y, e2.z, 0.f)};
}

// intersectRayAABB -- CUDAKernels.cu:66-85, on a packed box.
__device__ __forceinline__ bool hitBox6(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                        const bvht_Ray &r, float t0, float t1, float &tmin) {
  float tmax, tymin, tymax, tzmin, tzmax;
  tmin = ((r.sign[0] ? mxx : mnx) - r.pos.x) * r.invDir.x;
  tmax = ((r.sign[0] ? mnx : mxx) - r.pos.x) * r.invDir.x;
  tymin = ((r.sign[1] ? mxy : mny) - r.pos.y) * r.invDir.y;
  tymax = ((r.sign[1] ? mny : mxy) - r.pos.y) * r.invDir.y;
  tzmin = ((r.sign[2] ? mxz : mnz) - r.pos.z) * r.invDir.z;
  tzmax = ((r.sign[2] ? mnz : mxz) - r.pos.z) * r.invDir.z;
  if (tymin > tmin) tmin = tymin;
  if (tzmin > tmin) tmin = tzmin;
  if (tymax < tmax) tmax = tymax;
  if (tzmax < tmax) tmax = tzmax;
  return (tmin < tmax) && (tmin < t1) && (tmax > t0);
}

// intersectRayTriangle2 -- CUDAKernels.cu:88-112, one-sided (the
// cudaFaceCulling default, meson_options.txt:1), with the edges precomputed
// by k_packTris.  IEEE 1/det replaces __fdividef.
__device__ __forceinline__ bool hitTriPacked(const bvht_Ray &r, const PackedTri &t, float &dist) {
  const bvht_vec3 v0 = {t.v0.x, t.v0.y, t.v0.z}, e1 = {t.e1.x, t.e1.y, t.e1.z}, e2 = {t.e2.x, t.e2.y, t.e2.z};
  const bvht_vec3 p = vcross(r.dir, e2);
  const float det = vdot(e1, p);
  const bvht_vec3 tv = vsub(r.pos, v0);
  const bvht_vec3 q = vcross(tv, e1);
  const float invDet = __fdiv_rn(1.0f, det);
  const float u = vdot(tv, p);
  const float v = vdot(r.dir, q);
  dist = vdot(e2, q) * invDet;
  return (det > 0.0f) && ((u >= 0.0f && u <= det) && (v >= 0.0f && (u + v) <= det));
}

struct TraceCtx {
  const PackedPair *pairs;
  const PackedTri *tris;
  const bvht_BVHNode *nodes;  // parent links (fallback) and multi-triangle leaves
  uint32_t bound;             // node count the packing covered
  uint32_t numFaces;
};

// kTraceRay traversal loop -- CUDAKernels.cu:146-204, over packed nodes,
// as a state machine: travInit sets up a walk, each travStep visits one
// node (an inner node's two child boxes, or a leaf's triangles) and moves
// to the next.  MODE 0: reference triangle test.  MODE 1: hits need
// t > tMin and are not the triangle `skip` (secondary rays leaving a
// surface).  `stk` is this lane's column of a [kStackCap][64] LDS array.
struct TravState {
  uint64_t lo, hi;  // bit trail: a 1 marks a level whose second child is still pending
  uint32_t node, info, top, cached, steps;
  float nearest;
  TraceOut o;
};

// Returns false when there is nothing to walk (the root lies outside the tree).
__device__ __forceinline__ bool travInit(TravState &s, const TraceCtx &cx, uint32_t root) {
  s.o = {kNoHit, HUGE_VALF, 0};
  s.nearest = HUGE_VALF;
  s.lo = s.hi = 0;
  s.top = s.cached = s.steps = 0;
  if (root >= cx.bound) return false;
  s.node = root;
  s.info = leafInfo(cx.nodes[root]);
  return true;
}

// One visit; returns true when the walk has ended.
template <int MODE>
__device__ __forceinline__ bool travStep(TravState &s, const bvht_Ray &ray, const TraceCtx &cx, float tMin,
                                         uint32_t skip, uint2 *stk) {
  float dist;
  if (s.info == kInnerChild) {
    s.o.intCount++;
    const PackedPair pp = cx.pairs[s.node];
    float minL, minR;
    const bool hl = hitBox6(pp.a.x, pp.a.y, pp.a.z, pp.a.w, pp.b.x, pp.b.y, ray, 0.01f, s.nearest + 0.01f, minL);
    const bool hr = hitBox6(pp.b.z, pp.b.w, pp.c.x, pp.c.y, pp.c.z, pp.c.w, ray, 0.01f, s.nearest + 0.01f, minR);
    if (hl || hr) {
      s.hi = (s.hi << 1) | (s.lo >> 63);
      s.lo <<= 1;
      if (hl && hr) {
        s.lo |= 1;
        const bool leftFirst = minL < minR;
        stk[s.top * 64] = leftFirst ? make_uint2(pp.d.y, pp.d.w) : make_uint2(pp.d.x, pp.d.z);
        s.top = (s.top + 1) & (kStackCap - 1);
        s.cached = s.cached < kStackCap ? s.cached + 1 : kStackCap;
        s.node = leftFirst ? pp.d.x : pp.d.y;
        s.info = leftFirst ? pp.d.z : pp.d.w;
      } else {
        s.node = hl ? pp.d.x : pp.d.y;
        s.info = hl ? pp.d.z : pp.d.w;
      }
      return false;
    }
  } else if (s.info != kEmptyLeaf) {
    uint32_t first = s.info, count = 1;
    if (s.info == kLookupLeaf) {
      first = cx.nodes[s.node].left;
      count = cx.nodes[s.node].right;
    }
    for (uint32_t k = 0; k < count; ++k) {
      const uint32_t t = first + k;
      if (t >= cx.numFaces) break;
      bool h = hitTriPacked(ray, cx.tris[t], dist);
      if (MODE == 1) h = h && dist > tMin && t != skip;
      if (h && dist < s.nearest) {
        s.nearest = dist;
        s.o.tri = t;
      }
    }
  }
  // backtrack to the deepest remembered child
  if (s.lo == 0 && s.hi == 0) return true;
  uint32_t tz;
  if (s.lo != 0) {
    tz = (uint32_t)__builtin_ctzll(s.lo);
    if (tz) {
      s.lo = (s.lo >> tz) | (s.hi << (64 - tz));
      s.hi >>= tz;
    }
  } else {
    tz = 64 + (uint32_t)__builtin_ctzll(s.hi);
    s.lo = s.hi >> (tz - 64);
    s.hi = 0;
  }
  s.lo ^= 1;
  if (s.cached > 0) {
    s.top = (s.top - 1) & (kStackCap - 1);
    --s.cached;
    const uint2 e = stk[s.top * 64];
    s.node = e.x;
    s.info = e.y;
  } else {  // the entry fell out of the cache: walk the parent links
    uint32_t node = s.node;
    for (uint32_t k = 0; k < tz && node < cx.bound; ++k) node = cx.nodes[node].parent;
    if (node >= cx.bound) return true;
    const uint32_t p = cx.nodes[node].parent;
    if (p >= cx.bound) return true;
    node = cx.nodes[node].isLeft == 0 ? cx.nodes[p].left : cx.nodes[p].right;
    if (node >= cx.bound) return true;
    s.node = node;
    s.info = leafInfo(cx.nodes[node]);
  }
  return false;
}

// A valid tree is walked in at most ~4 visits per node; the step bound only
// ends the walk on a malformed (cyclic) tree.
constexpr uint32_t kMaxSteps = 1u << 26;

template <int MODE>
__device__ TraceOut traversePacked(const bvht_Ray &ray, const TraceCtx &cx, uint32_t root, float tMin, uint32_t skip,
                                   uint2 *stk) {
  TravState s;
  if (travInit(s, cx, root))
    while (!travStep<MODE>(s, ray, cx, tMin, skip, stk) && ++s.steps < kMaxSteps) {
    }
  s.o.t = s.nearest;
  return s.o;
}

// Shading of CUDAKernels.cu:208-218.
__device__ __forceinline__ bvht_CUDAPixel shade(const TraceOut &o, const bvht_Ray &ray, const bvht_vec3 *vert,
                                                const bvht_Triangle *faces, bvht_vec3 light) {
  bvht_CUDAPixel px = {0, 0, o.intCount};
  if (o.t < HUGE_VALF) {
    const bvht_Triangle tri = faces[o.tri];
    const bvht_vec3 v1 = vert[tri.v1], v2 = vert[tri.v2], v3 = vert[tri.v3];
    const bvht_vec3 n = vnormalize(vcross(vsub(v2, v1), vsub(v3, v1)));
    const bvht_vec3 hp = {ray.pos.x + ray.dir.x * o.t, ray.pos.y + ray.dir.y * o.t, ray.pos.z + ray.dir.z * o.t};
    const bvht_vec3 ld = vnormalize(vsub(light, hp));
    float diff = 1.0f + vdot(n, ld);
    diff = diff > 0.0f ? diff : 0.0f;
    px.hit = 1;
    px.diffuse = (uint8_t)(diff * 127.0f);
  }
  return px;
}

// Image path: 8x8 pixel tiles per wave (kTraceRay launch shape, CUDAKernels.cu:409-416).
__global__ __launch_bounds__(64) void k_traceImage(const bvht_Ray *__restrict__ rays, bvht_CUDAPixel *__restrict__ img,
                                                   TraceCtx cx, uint32_t root, const bvht_vec3 *__restrict__ vert,
                                                   const bvht_Triangle *__restrict__ faces, bvht_vec3 light, uint32_t w,
                                                   uint32_t h) {
  __shared__ uint2 s_stk[kStackCap * 64];
  const uint32_t x = blockIdx.x * 8 + (threadIdx.x & 7);
  const uint32_t y = blockIdx.y * 8 + (threadIdx.x >> 3);
  if (x >= w || y >= h) return;
  const size_t i = (size_t)y * w + x;
  const bvht_Ray ray = rays[i];
  const TraceOut o = traversePacked<0>(ray, cx, root, 0.0f, kNoHit, s_stk + threadIdx.x);
  img[i] = shade(o, ray, vert, faces, light);
}

// Ray bundles (kTraceRayBundle, CUDAKernels.cu:240-383): the 64 rays of an
// 8x8 tile -- one wave -- walk one shared node sequence.  At an inner node
// each ray tests both children against its own nearest hit; the bundle
// enters every child some ray hits, the one a majority prefers first (ties:
// right), and backtracks on one shared trail.  Ballots replace the
// reference's shared-memory flags and reduction.  Every ray still finds its
// nearest hit (the shared walk covers its own); among triangles at exactly
// the same distance the first one met may differ from the per-ray order.
__global__ __launch_bounds__(64) void k_traceImageBundle(const bvht_Ray *__restrict__ rays,
                                                         bvht_CUDAPixel *__restrict__ img, TraceCtx cx, uint32_t root,
                                                         const bvht_vec3 *__restrict__ vert,
                                                         const bvht_Triangle *__restrict__ faces, bvht_vec3 light,
                                                         uint32_t w, uint32_t h) {
  __shared__ uint2 s_stk[128];  // pending children: at most one per trail bit
  const uint32_t x = blockIdx.x * 8 + (threadIdx.x & 7);
  const uint32_t y = blockIdx.y * 8 + (threadIdx.x >> 3);
  const bool valid = x < w && y < h;
  bvht_Ray ray;
  if (valid) ray = rays[(size_t)y * w + x];
  TraceOut o = {kNoHit, HUGE_VALF, 0};
  if (root < cx.bound) {
    uint64_t lo = 0, hi = 0;  // wave-uniform from here on
    uint32_t node = root, info = leafInfo(cx.nodes[root]), sp = 0;
    float nearest = HUGE_VALF, dist;
    for (uint32_t steps = 0; steps < (1u << 26); ++steps) {
      if (info == kInnerChild) {
        o.intCount++;
        const PackedPair pp = cx.pairs[node];
        float minL = 0.0f, minR = 0.0f;
        const bool hl =
            valid && hitBox6(pp.a.x, pp.a.y, pp.a.z, pp.a.w, pp.b.x, pp.b.y, ray, 0.01f, nearest + 0.01f, minL);
        const bool hr =
            valid && hitBox6(pp.b.z, pp.b.w, pp.c.x, pp.c.y, pp.c.z, pp.c.w, ray, 0.01f, nearest + 0.01f, minR);
        const bool anyL = __ballot(hl) != 0, anyR = __ballot(hr) != 0;
        if (anyL || anyR) {
          hi = (hi << 1) | (lo >> 63);
          lo <<= 1;
          if (anyL && anyR) {
            lo |= 1;
            const bool wantL = hl && (minL < minR || !hr), wantR = hr && (minR < minL || !hl);
            const int vote = __popcll(__ballot(wantL)) - __popcll(__ballot(wantR));
            const bool leftFirst = vote > 0;
            if (sp < 128) s_stk[sp] = leftFirst ? make_uint2(pp.d.y, pp.d.w) : make_uint2(pp.d.x, pp.d.z);
            ++sp;
            node = leftFirst ? pp.d.x : pp.d.y;
            info = leftFirst ? pp.d.z : pp.d.w;
          } else {
            node = anyL ? pp.d.x : pp.d.y;
            info = anyL ? pp.d.z : pp.d.w;
          }
          continue;
        }
      } else if (info != kEmptyLeaf) {
        uint32_t first = info, count = 1;
        if (info == kLookupLeaf) {
          first = cx.nodes[node].left;
          count = cx.nodes[node].right;
        }
        for (uint32_t k = 0; k < count; ++k) {
          const uint32_t t = first + k;
          if (t >= cx.numFaces) break;
          if (valid && hitTriPacked(ray, cx.tris[t], dist) && dist < nearest) {
            nearest = dist;
            o.tri = t;
          }
        }
      }
      if (lo == 0 && hi == 0) break;
      if (lo != 0) {
        const uint32_t tz = (uint32_t)__builtin_ctzll(lo);
        if (tz) {
          lo = (lo >> tz) | (hi << (64 - tz));
          hi >>= tz;
        }
      } else {
        const uint32_t tz = 64 + (uint32_t)__builtin_ctzll(hi);
        lo = hi >> (tz - 64);
        hi = 0;
      }
      lo ^= 1;
      if (sp == 0 || sp > 128) break;  // a trail deeper than 128 levels (unsupported, as in the reference)
      const uint2 e = s_stk[--sp];
      node = e.x;
      info = e.y;
    }
    o.t = nearest;
  }
  if (valid) img[(size_t)y * w + x] = shade(o, ray, vert, faces, light);
}

template <int MODE>
__global__ __launch_bounds__(64) void k_traceRays(const bvht_Ray *__restrict__ rays, uint32_t num, TraceCtx cx,
                                                  uint32_t root, const bvht_vec3 *__restrict__ vert,
                                                  const bvht_Triangle *__restrict__ faces, bvht_vec3 light, float tMin,
                                                  const uint32_t *__restrict__ skip, bvht_Hit *__restrict__ hits,
                                                  bvht_CUDAPixel *__restrict__ img, const uint32_t *__restrict__ order,
                                                  uint32_t imgW) {
  __shared__ uint2 s_stk[kStackCap * 64];
  uint32_t i;
  if (imgW) {  // the rays of an imgW-wide image, row-major: one 8x8 pixel tile per wave
    const uint32_t tilesX = (imgW + 7) / 8;
    // workgroups go to the 8 XCDs round-robin: runs of 4 consecutive tiles
    // (neighbouring rays, shared nodes) go to one XCD's L2, and the runs of
    // the 8 XCDs interleave, so every XCD keeps sweeping the same band of
    // the image (one stripe of the image per XCD balanced the work badly:
    // 5.0 against 3.8 ms at 1080p); the last partial group keeps its order
    const uint32_t b = blockIdx.x, g0 = b / 32 * 32;
    const uint32_t t = g0 + 32 <= gridDim.x ? g0 + (b % 8) * 4 + (b / 8) % 4 : b;
    const uint32_t x = (t % tilesX) * 8 + (threadIdx.x & 7), y = (t / tilesX) * 8 + (threadIdx.x >> 3);
    if (x >= imgW || (uint64_t)y * imgW + x >= num) return;
    i = y * imgW + x;
  } else {
    const uint32_t slot = blockIdx.x * 64 + threadIdx.x;
    if (slot >= num) return;
    i = order ? order[slot] : slot;
  }
  const bvht_Ray ray = rays[i];
  TraceOut o = {kNoHit, HUGE_VALF, 0};
  if (!(ray.dir.x == 0.0f && ray.dir.y == 0.0f && ray.dir.z == 0.0f))
    o = traversePacked<MODE>(ray, cx, root, tMin, skip ? skip[i] : kNoHit, s_stk + threadIdx.x);
  if (hits) hits[i] = bvht_Hit{o.tri, o.t};
  if (img) img[i] = shade(o, ray, vert, faces, light);
}

// Rays of one call traced by resident waves that refill their idle lanes
// from a shared counter (`next`, zero at launch), so a wave does not idle
// behind its longest walk.  A wave takes at least kRefillAt rays at a time, in
// slot order (neighbouring rays stay together); each ray's walk is the
// same visit sequence as in k_traceRays.  Every wave ends once the counter
// has passed `num` and its lanes are idle.
template <int MODE>
__global__ __launch_bounds__(64) void k_traceRaysDyn(const bvht_Ray *__restrict__ rays, uint32_t num, TraceCtx cx,
                                                     uint32_t root, const bvht_vec3 *__restrict__ vert,
                                                     const bvht_Triangle *__restrict__ faces, bvht_vec3 light,
                                                     float tMin, const uint32_t *__restrict__ skip,
                                                     bvht_Hit *__restrict__ hits, bvht_CUDAPixel *__restrict__ img,
                                                     const uint32_t *__restrict__ order, uint32_t *__restrict__ next) {
  __shared__ uint2 s_stk[kStackCap * 64];
  uint2 *stk = s_stk + threadIdx.x;
  const uint64_t lt = __lanemask_lt();
  bool active = false, drained = false;
  uint32_t i = 0, skipTri = kNoHit;
  bvht_Ray ray;
  TravState s;
  while (true) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nIdle = (uint32_t)__popcll(idle);
    if (!drained && (nIdle >= kRefillAt || nIdle == 64)) {  // wave-uniform
      uint32_t base = 0;
      if (threadIdx.x == 0) base = atomicAdd(next, nIdle);
      base = __shfl(base, 0, 64);
      if (base + nIdle >= num) drained = true;
      if (!active) {
        const uint32_t slot = base + (uint32_t)__popcll(idle & lt);
        if (slot < num) {
          i = order ? order[slot] : slot;
          ray = rays[i];
          skipTri = skip ? skip[i] : kNoHit;
          active = !(ray.dir.x == 0.0f && ray.dir.y == 0.0f && ray.dir.z == 0.0f) && travInit(s, cx, root);
          if (!active) {  // nothing to walk: a miss
            const TraceOut o = {kNoHit, HUGE_VALF, 0};
            if (hits) hits[i] = bvht_Hit{o.tri, o.t};
            if (img) img[i] = shade(o, ray, vert, faces, light);
          }
        }
      }
    }
    if (__ballot(active) == 0) {
      if (drained) break;
      continue;
    }
    if (active && (travStep<MODE>(s, ray, cx, tMin, skipTri, stk) || ++s.steps >= kMaxSteps)) {
      s.o.t = s.nearest;
      if (hits) hits[i] = bvht_Hit{s.o.tri, s.o.t};
      if (img) img[i] = shade(s.o, ray, vert, faces, light);
      active = false;
    }
  }
}

// kRefitLeafBVHTris (CUDAKernels.cu:492-525): leaf boxes from the current
// vertices (triangle box -+ FLT_EPSILON); like the reference, the leaf's
// surfaceArea is not rewritten.
__global__ __launch_bounds__(256) void k_refitLeafTris(const uint32_t *__restrict__ ids, uint32_t num,
                                                       bvht_BVHNode *__restrict__ nodes,
                                                       const bvht_Triangle *__restrict__ faces,
                                                       const bvht_vec3 *__restrict__ vert, uint32_t numFaces) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num) return;
  const uint32_t id = ids[i];
  const uint32_t f = nodes[id].left;
  if (f >= numFaces) return;
  const bvht_Triangle t = faces[f];
  nodes[id].bbox = triangleBox(vert[t.v1], vert[t.v2], vert[t.v3]);  // exact min/max: the reference's box
}

// Counter-based hash RNG (splitmix64 finaliser) -> uniform floats in [0, 1).
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// One cosine-weighted bounce: origin hit + 1e-4 * n (n faces the incoming
// ray), direction sampled around n (Malley's method), orthonormal basis of
// Duff et al. 2017.
__global__ __launch_bounds__(256) void k_bounce(const bvht_Ray *__restrict__ prim, const bvht_Hit *__restrict__ hits,
                                                uint32_t num, const bvht_vec3 *__restrict__ vert,
                                                const bvht_Triangle *__restrict__ faces, uint32_t seed,
                                                bvht_Ray *__restrict__ out, uint32_t *__restrict__ skip) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num) return;
  const bvht_Ray r = prim[i];
  const bvht_Hit hh = hits[i];
  bvht_Ray o;
  if (hh.triId == kNoHit) {
    o.pos = r.pos;
    o.dir = {0.0f, 0.0f, 0.0f};
    o.invDir = {0.0f, 0.0f, 0.0f};
    o.sign[0] = o.sign[1] = o.sign[2] = 0;
    o._pad = 0;
    o.px = r.px;
    o.py = r.py;
    out[i] = o;
    skip[i] = kNoHit;
    return;
  }
  const bvht_Triangle tri = faces[hh.triId];
  const bvht_vec3 v1 = vert[tri.v1], v2 = vert[tri.v2], v3 = vert[tri.v3];
  bvht_vec3 n = vnormalize(vcross(vsub(v2, v1), vsub(v3, v1)));
  if (vdot(n, r.dir) > 0.0f) n = {-n.x, -n.y, -n.z};
  const bvht_vec3 hp = {r.pos.x + r.dir.x * hh.t, r.pos.y + r.dir.y * hh.t, r.pos.z + r.dir.z * hh.t};
  const bvht_vec3 org = {hp.x + 1e-4f * n.x, hp.y + 1e-4f * n.y, hp.z + 1e-4f * n.z};
  const uint64_t h = mix64(((uint64_t)seed << 32) ^ (uint64_t)i);
  const float u1 = (float)(h >> 40) * (1.0f / 16777216.0f);
  const float u2 = (float)((h >> 16) & 0xFFFFFF) * (1.0f / 16777216.0f);
  const float rr = sqrtf(u1), phi = 6.28318530717958647692f * u2;
  const float lx = rr * cosf(phi), ly = rr * sinf(phi), lz = sqrtf(fmaxf(0.0f, 1.0f - u1));
  const float sg = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sg + n.z), b = n.x * n.y * a;
  const bvht_vec3 t1 = {1.0f + sg * n.x * n.x * a, sg * b, -sg * n.x};
  const bvht_vec3 t2 = {b, sg + n.y * n.y * a, -n.y};
  bvht_vec3 d = {lx * t1.x + ly * t2.x + lz * n.x, lx * t1.y + ly * t2.y + lz * n.y, lx * t1.z + ly * t2.z + lz * n.z};
  d = vnormalize(d);
  setRay(o, org, d, r.px, r.py);
  out[i] = o;
  skip[i] = hh.triId;
}

__device__ __forceinline__ uint32_t spread9(uint32_t v) {  // 9 bits -> every third bit of 27
  v &= 511u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ uint32_t cell9(float x, float lo, float inv) {
  const float q = (x - lo) * inv;
  return q <= 0.0f ? 0u : (q >= 511.0f ? 511u : (uint32_t)q);
}

// Ordering key of a secondary ray: its direction cell (4 bins per axis over
// [-1, 1]) above the Morton cell of its origin in the root box (8 bits per
// axis).  Rays that start close together and head the same way walk similar
// nodes, so a wave of them diverges less (measured on the 4.24M-triangle scene
// at 1080p, one-ray-per-lane tracing: bounce trace 5.6 ->
// 4.9 ms with the sort included; octant-only direction cells 4.9, origin-major
// keys 5.2).  A ray with no direction gets the largest key.  Plus the digit
// histograms of the 4 sort passes.
__global__ __launch_bounds__(256) void k_rayKeys(const bvht_Ray *__restrict__ rays, uint32_t num,
                                                 const bvht_BVHNode *__restrict__ nodes, uint32_t root,
                                                 uint32_t *__restrict__ keys, uint32_t *__restrict__ hist) {
  __shared__ uint32_t s_hist[4][kSortRadix];
  for (int p = 0; p < 4; ++p) s_hist[p][threadIdx.x] = 0;
  __syncthreads();
  const bvht_AABB box = nodes[root].bbox;
  const float ix = 512.0f / fmaxf(box.minMax[1].x - box.minMax[0].x, 1e-30f);
  const float iy = 512.0f / fmaxf(box.minMax[1].y - box.minMax[0].y, 1e-30f);
  const float iz = 512.0f / fmaxf(box.minMax[1].z - box.minMax[0].z, 1e-30f);
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < num; i += gridDim.x * 256) {
    const bvht_Ray r = rays[i];
    uint32_t key = (1u << 30) - 1;
    if (!(r.dir.x == 0.0f && r.dir.y == 0.0f && r.dir.z == 0.0f)) {
      // direction: 4 bins per axis over [-1, 1]; origin: 8 bits per axis
      const uint32_t dx = min((uint32_t)fmaxf((r.dir.x + 1.0f) * 2.0f, 0.0f), 3u);
      const uint32_t dy = min((uint32_t)fmaxf((r.dir.y + 1.0f) * 2.0f, 0.0f), 3u);
      const uint32_t dz = min((uint32_t)fmaxf((r.dir.z + 1.0f) * 2.0f, 0.0f), 3u);
      const uint32_t dcell = (dx << 4) | (dy << 2) | dz;
      const uint32_t o = (spread9(cell9(r.pos.x, box.minMax[0].x, ix) >> 1) << 2) |
                         (spread9(cell9(r.pos.y, box.minMax[0].y, iy) >> 1) << 1) |
                         spread9(cell9(r.pos.z, box.minMax[0].z, iz) >> 1);
      key = (dcell << 24) | o;
    }
    keys[i] = key;
#pragma unroll
    for (int p = 0; p < 4; ++p) atomicAdd(&s_hist[p][(key >> (8 * p)) & (kSortRadix - 1)], 1u);
  }
  __syncthreads();
  for (int p = 0; p < 4; ++p) {
    const uint32_t c = s_hist[p][threadIdx.x];
    if (c) atomicAdd(hist + p * kSortRadix + threadIdx.x, c);
  }
}

// Packed copies of the tree and mesh, rebuilt by every trace call (the tree
// may have changed since the last one); the buffers only grow.
struct TraceCache {
  PackedPair *pairs = nullptr;
  size_t pairCap = 0;
  PackedTri *tris = nullptr;
  size_t triCap = 0;
  uint32_t *bad = nullptr;
  uint32_t *next = nullptr;  // ray counter of k_traceRaysDyn
  // the ray buffer generateRays last filled, as a w x h image
  const void *imgRays = nullptr;
  uint32_t imgW = 0, imgH = 0;
  // ray ordering (bvht_traceRays): keys and indices x2, sort scratch
  bool orderRays = false;  // measured: with lane refills the sort costs more than it saves
  uint32_t orderCap = 0;
  uint32_t *okeys[2] = {nullptr, nullptr}, *ovals[2] = {nullptr, nullptr}, *oscratch = nullptr;
};
TraceCache &traceCache() {
  static TraceCache c;
  return c;
}

// Node indices of the traced tree must lie below 2 * numFaces - 1 (every
// binary BVH over the mesh fits); children past that are reported through
// tracerDoCudaSync and traced as empty leaves.
bool prepareTrace(const bvht_BVHNode *nodes, const bvht_MeshRaw &mesh, TraceCtx &cx, const char *what) {
  if (!mesh.faces || !mesh.vert || mesh.numFaces == 0) {
    setError("%s: empty mesh", what);
    return false;
  }
  const uint64_t bound = 2ull * mesh.numFaces - 1;
  if (bound >= kEmptyLeaf) {
    setError("%s: %u faces exceed the 32-bit node range", what, mesh.numFaces);
    return false;
  }
  TraceCache &c = traceCache();
  if (c.pairCap < bound) {
    devFree(c.pairs);
    c.pairs = nullptr;
    c.pairCap = 0;
    if (!devAlloc(&c.pairs, bound, "trace node pairs")) return false;
    c.pairCap = bound;
  }
  if (c.triCap < mesh.numFaces) {
    devFree(c.tris);
    c.tris = nullptr;
    c.triCap = 0;
    if (!devAlloc(&c.tris, mesh.numFaces, "trace triangles")) return false;
    c.triCap = mesh.numFaces;
  }
  if (!c.bad && !devAlloc(&c.bad, 1, "trace flag")) return false;
  if (!fillU32(c.bad, 0u, 1)) return false;
  k_packPairs<<<divUp(bound, 256), 256, 0, stream()>>>(nodes, (uint32_t)bound, c.pairs, c.bad);
  if (!checkLaunch("k_packPairs")) return false;
  k_packTris<<<divUp(mesh.numFaces, 256), 256, 0, stream()>>>(mesh.faces, mesh.vert, mesh.numFaces, c.tris);
  if (!checkLaunch("k_packTris")) return false;
  cx = TraceCtx{c.pairs, c.tris, nodes, (uint32_t)bound, mesh.numFaces};
  return true;
}

// Sorts the rays by k_rayKeys and returns the trace order (nullptr on
// failure, with the error recorded).
const uint32_t *rayOrder(const bvht_Ray *rays, uint32_t num, const bvht_BVHNode *nodes, uint32_t root) {
  TraceCache &c = traceCache();
  const SortPlan plan = makeSortPlan(num, 4, 30);
  if (c.orderCap < num) {
    for (int b = 0; b < 2; ++b) {
      devFree(c.okeys[b]);
      devFree(c.ovals[b]);
      c.okeys[b] = c.ovals[b] = nullptr;
    }
    devFree(c.oscratch);
    c.oscratch = nullptr;
    c.orderCap = 0;
    if (!devAlloc(&c.okeys[0], num, "ray keys") || !devAlloc(&c.okeys[1], num, "ray keys") ||
        !devAlloc(&c.ovals[0], num, "ray order") || !devAlloc(&c.ovals[1], num, "ray order") ||
        !devMalloc(reinterpret_cast<void **>(&c.oscratch), sortScratchBytes(makeSortPlan(num, 4, 30)), "ray sort"))
      return nullptr;
    c.orderCap = num;
  }
  if (!sortReset(plan, c.oscratch)) return nullptr;
  const uint32_t blocks = std::min<uint32_t>(divUp(num, 256 * 4), (uint32_t)numCUs() * 2);
  k_rayKeys<<<blocks, 256, 0, stream()>>>(rays, num, nodes, root, c.okeys[0], sortHist(plan, c.oscratch));
  if (!checkLaunch("k_rayKeys")) return nullptr;
  void *keys[2] = {c.okeys[0], c.okeys[1]};
  const int r = sortRun(plan, keys, c.ovals, true, false, c.oscratch);
  return r < 0 ? nullptr : c.ovals[r];
}

}  // namespace

}  // namespace bvht

using namespace bvht;

extern "C" {

void bvht_setRayOrdering(bool enable) { traceCache().orderRays = enable; }

bool allocateRays(bvht_Ray **rays, uint32_t num) { return devAlloc(rays, num, "allocateRays"); }
bool allocateImage(bvht_CUDAPixel **img, uint32_t w, uint32_t h) {
  return devAlloc(img, (size_t)w * h, "allocateImage");
}
void freeRays(bvht_Ray **rays) {
  if (rays && *rays) {
    devFree(*rays);
    *rays = nullptr;
  }
}
void freeImage(bvht_CUDAPixel *img) { devFree(img); }

void generateRays(bvht_Ray *rays, uint32_t w, uint32_t h, bvht_vec3 pos, bvht_vec3 lookAt, bvht_vec3 up, float fov) {
  if (!rays || w == 0 || h == 0) {
    setError("generateRays: no ray buffer or empty image");
    return;
  }
  // inverse(lookAtRH(pos, lookAt, up)) has columns s, u, -f, pos
  CamBasis c;
  double P[3] = {pos.x, pos.y, pos.z};
  double f[3] = {lookAt.x - P[0], lookAt.y - P[1], lookAt.z - P[2]};
  double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double &v : f) v /= fl;
  double U[3] = {up.x, up.y, up.z};
  double s[3] = {f[1] * U[2] - f[2] * U[1], f[2] * U[0] - f[0] * U[2], f[0] * U[1] - f[1] * U[0]};
  double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (double &v : s) v /= sl;
  double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  for (int k = 0; k < 3; ++k) {
    c.s[k] = s[k];
    c.u[k] = u[k];
    c.f[k] = f[k];
  }
  c.pos[0] = pos.x;
  c.pos[1] = pos.y;
  c.pos[2] = pos.z;
  c.ratio = (float)w / (float)h;
  c.scale = (float)std::tan((0.5 * (double)fov) * (M_PI / 180.0));  // tan(radians(0.5 * fov))
  dim3 grid(divUp(w, 16), divUp(h, 16));
  k_genRays<<<grid, 256, 0, stream()>>>(rays, w, h, c);
  if (!checkLaunch("k_genRays")) return;
  TraceCache &tc = traceCache();
  tc.imgRays = rays;
  tc.imgW = w;
  tc.imgH = h;
}

void tracerImage(bvht_Ray *rays, bvht_CUDAPixel *img, bvht_BVHNode *nodes, uint32_t rootNode, bvht_MeshRaw mesh,
                 bvht_vec3 light, uint32_t w, uint32_t h, bool bundle) {
  if (!rays || !img || !nodes) {
    setError("tracerImage: missing rays, image or nodes");
    return;
  }
  TraceCtx cx;
  if (!prepareTrace(nodes, mesh, cx, "tracerImage")) return;
  dim3 grid(divUp(w, 8), divUp(h, 8));
  if (bundle)
    k_traceImageBundle<<<grid, 64, 0, stream()>>>(rays, img, cx, rootNode, mesh.vert, mesh.faces, light, w, h);
  else
    k_traceImage<<<grid, 64, 0, stream()>>>(rays, img, cx, rootNode, mesh.vert, mesh.faces, light, w, h);
  checkLaunch(bundle ? "k_traceImageBundle" : "k_traceImage");
}

void copyImageToHost(bvht_CUDAPixel *hostPixel, bvht_CUDAPixel *cudaImg, uint32_t w, uint32_t h) {
  if (check(hipMemcpyAsync(hostPixel, cudaImg, sizeof(bvht_CUDAPixel) * (size_t)w * h, hipMemcpyDeviceToHost, stream()),
            "copyImageToHost"))
    check(hipStreamSynchronize(stream()), "copyImageToHost sync");
}

void tracerDoCudaSync(void) {
  if (!check(hipStreamSynchronize(stream()), "tracerDoCudaSync")) return;
  uint32_t bad = 0;
  TraceCache &c = traceCache();
  if (c.bad && check(hipMemcpy(&bad, c.bad, 4, hipMemcpyDeviceToHost), "tracerDoCudaSync flag") && bad)
    setError("tracer: a child index lies outside the 2 * numFaces - 1 nodes of the mesh");
}

bool bvht_traceRays(const bvht_Ray *rays, uint32_t num, bvht_BVHNode *nodes, uint32_t rootNode, bvht_MeshRaw mesh,
                    bvht_vec3 light, int mode, float tMin, const uint32_t *skipTri, bvht_Hit *hits,
                    bvht_CUDAPixel *img) {
  if (!rays || !nodes || (!hits && !img)) {
    setError("bvht_traceRays: missing rays, nodes or outputs");
    return false;
  }
  if (num == 0) return true;
  TraceCtx cx;
  if (!prepareTrace(nodes, mesh, cx, "bvht_traceRays")) return false;
  if (rootNode >= cx.bound) {
    setError("bvht_traceRays: root %u outside the %u nodes of the mesh", rootNode, cx.bound);
    return false;
  }
  // secondary rays only: primary rays of an image are coherent in their own order
  const uint32_t *order = nullptr;  // (measured: ordering primary rays costs more than it saves)
  if (mode == 1 && traceCache().orderRays && !(order = rayOrder(rays, num, nodes, rootNode))) return false;
  if (mode == 0) {  // primary rays: one ray per lane, neighbouring rays in a wave
    // Rays that generateRays wrote as an image are traced in 8x8 pixel tiles
    // (tighter bundles than 64-pixel rows); the order never changes a ray's
    // result, only which rays share a wave.
    const TraceCache &c = traceCache();
    const bool image = rays == c.imgRays && (uint64_t)num == (uint64_t)c.imgW * c.imgH;
    const uint32_t blocks = image ? divUp(c.imgW, 8) * divUp(c.imgH, 8) : divUp(num, 64);
    k_traceRays<0><<<blocks, 64, 0, stream()>>>(rays, num, cx, rootNode, mesh.vert, mesh.faces, light, tMin, skipTri,
                                                hits, img, order, image ? c.imgW : 0u);
    return checkLaunch("k_traceRays");
  }
  // secondary rays: resident waves refill idle lanes (walk lengths vary widely)
  TraceCache &c = traceCache();
  if (!c.next && !devAlloc(&c.next, 1, "trace ray counter")) return false;
  if (!fillU32(c.next, 0u, 1)) return false;
  // resident waves only: kTraceWavesPerCU of them fill a CU (LDS-bound)
  const uint32_t waves = std::min<uint32_t>(divUp(num, 64), (uint32_t)numCUs() * kTraceWavesPerCU);
  k_traceRaysDyn<1><<<waves, 64, 0, stream()>>>(rays, num, cx, rootNode, mesh.vert, mesh.faces, light, tMin, skipTri,
                                                hits, img, order, c.next);
  return checkLaunch("k_traceRaysDyn");
}

void refitDynamicTris(bvht_CUDAMemoryBVHPointer *bvh, bvht_MeshRaw mesh, uint32_t *nodesToUpdate, uint32_t numNodes,
                      uint32_t blockSize) {
  (void)blockSize;
  if (!bvh || !bvh->nodes || (numNodes && !nodesToUpdate) || !mesh.faces || !mesh.vert) {
    setError("refitDynamicTris: missing tree, mesh or node list");
    return;
  }
  if (numNodes == 0) return;
  k_refitLeafTris<<<divUp(numNodes, 256), 256, 0, stream()>>>(nodesToUpdate, numNodes, bvh->nodes, mesh.faces,
                                                              mesh.vert, mesh.numFaces);
  checkLaunch("k_refitLeafTris");
}

bool bvht_makeBounceRays(const bvht_Ray *primary, const bvht_Hit *hits, uint32_t num, bvht_MeshRaw mesh, uint32_t seed,
                         bvht_Ray *outRays, uint32_t *outSkipTri) {
  if (!primary || !hits || !outRays || !outSkipTri) {
    setError("bvht_makeBounceRays: missing buffers");
    return false;
  }
  if (num == 0) return true;
  k_bounce<<<divUp(num, 256), 256, 0, stream()>>>(primary, hits, num, mesh.vert, mesh.faces, seed, outRays, outSkipTri);
  return checkLaunch("k_bounce");
}

}  // extern "C"
