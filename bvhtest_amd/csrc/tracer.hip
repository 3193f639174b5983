// Ray generation and stackless traversal for gfx950
// (reference: lib/tracer/CUDAKernels.cu:43-223, :386-417).
//
// Traversal is the reference's bit-stack algorithm (Afra & Szirmay-Kalos
// 2014): descend to the nearer of the two hit children, remember the other
// with a set bit, backtrack through parent links.  Arithmetic is f32 with no
// FMA contraction and IEEE division, so the visit order, hit triangle and
// distance are a pure function of the inputs (the oracle reproduces them
// exactly).  The 128-bit stack limits trees to depth 128, as in the reference.
//
// Rays are generated in f64 from a camera basis computed in f64 on the host
// (the exact inverse of lookAtRH), then rounded once to f32.
#include "runtime.hpp"

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

__device__ __forceinline__ float boxCoord(const bvht_AABB &b, uint32_t side, int axis) {
  const bvht_vec3 &v = b.minMax[side];
  return axis == 0 ? v.x : (axis == 1 ? v.y : v.z);
}

// intersectRayAABB -- CUDAKernels.cu:66-85
__device__ __forceinline__ bool hitBox(const bvht_AABB &b, const bvht_Ray &r, float t0, float t1, float &tmin) {
  float tmax, tymin, tymax, tzmin, tzmax;
  tmin = (boxCoord(b, r.sign[0], 0) - r.pos.x) * r.invDir.x;
  tmax = (boxCoord(b, 1 - r.sign[0], 0) - r.pos.x) * r.invDir.x;
  tymin = (boxCoord(b, r.sign[1], 1) - r.pos.y) * r.invDir.y;
  tymax = (boxCoord(b, 1 - r.sign[1], 1) - r.pos.y) * r.invDir.y;
  tzmin = (boxCoord(b, r.sign[2], 2) - r.pos.z) * r.invDir.z;
  tzmax = (boxCoord(b, 1 - r.sign[2], 2) - r.pos.z) * r.invDir.z;
  if (tymin > tmin) tmin = tymin;
  if (tzmin > tmin) tmin = tzmin;
  if (tymax < tmax) tmax = tymax;
  if (tzmax < tmax) tmax = tzmax;
  return (tmin < tmax) && (tmin < t1) && (tmax > t0);
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

// intersectRayTriangle2 -- CUDAKernels.cu:88-112, one-sided (the
// cudaFaceCulling default, meson_options.txt:1).  IEEE 1/det replaces
// __fdividef.
__device__ __forceinline__ bool hitTri(const bvht_Ray &r, bvht_vec3 v0, bvht_vec3 v1, bvht_vec3 v2, float &dist) {
  const bvht_vec3 e1 = vsub(v1, v0), e2 = vsub(v2, v0);
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

struct TraceOut {
  uint32_t tri;
  float t;
  uint16_t intCount;
};

// kTraceRay traversal loop -- CUDAKernels.cu:146-204.
// MODE 0: reference triangle test.  MODE 1: hits need t > tMin and are not
// the triangle `skip` (secondary rays leaving a surface).
template <int MODE>
__device__ TraceOut traverse(const bvht_Ray &ray, const bvht_BVHNode *__restrict__ nodes, uint32_t root,
                             const bvht_vec3 *__restrict__ vert, const bvht_Triangle *__restrict__ faces, float tMin,
                             uint32_t skip) {
  TraceOut o = {kNoHit, HUGE_VALF, 0};
  uint64_t lo = 0, hi = 0;
  uint32_t node = root;
  float nearest = HUGE_VALF, dist, minL, minR;
  // A valid tree is walked in at most ~4 visits per node; the bound only
  // ends the loop on a malformed (cyclic) tree.
  for (uint32_t steps = 0; steps < (1u << 26); ++steps) {
    const bvht_BVHNode *nd = nodes + node;
    if (nd->isLeafFlag == 0) {
      o.intCount++;
      const uint32_t l = nd->left, r = nd->right;
      const bool hl = hitBox(nodes[l].bbox, ray, 0.01f, nearest + 0.01f, minL);
      const bool hr = hitBox(nodes[r].bbox, ray, 0.01f, nearest + 0.01f, minR);
      if (hl || hr) {
        hi = (hi << 1) | (lo >> 63);
        lo <<= 1;
        if (hl && hr) {
          lo |= 1;
          node = minL < minR ? l : r;
        } else {
          node = hl ? l : r;
        }
        continue;
      }
    } else {
      const uint32_t first = nd->left, count = nd->right;
      for (uint32_t k = 0; k < count; ++k) {
        const uint32_t t = first + k;
        const bvht_Triangle tri = faces[t];
        bool h = hitTri(ray, vert[tri.v1], vert[tri.v2], vert[tri.v3], dist);
        if (MODE == 1) h = h && dist > tMin && t != skip;
        if (h && dist < nearest) {
          nearest = dist;
          o.tri = t;
        }
      }
    }
    bool end = false;
    while ((lo & 1) == 0) {
      if (lo == 0 && hi == 0) {
        end = true;
        break;
      }
      node = nodes[node].parent;
      lo = (lo >> 1) | (hi << 63);
      hi >>= 1;
    }
    if (end) break;
    const uint32_t p = nodes[node].parent;
    node = nodes[node].isLeft == 0 ? nodes[p].left : nodes[p].right;
    lo ^= 1;
  }
  o.t = nearest;
  return o;
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
                                                   const bvht_BVHNode *__restrict__ nodes, uint32_t root,
                                                   const bvht_vec3 *__restrict__ vert,
                                                   const bvht_Triangle *__restrict__ faces, bvht_vec3 light, uint32_t w,
                                                   uint32_t h) {
  const uint32_t x = blockIdx.x * 8 + (threadIdx.x & 7);
  const uint32_t y = blockIdx.y * 8 + (threadIdx.x >> 3);
  if (x >= w || y >= h) return;
  const size_t i = (size_t)y * w + x;
  const bvht_Ray ray = rays[i];
  const TraceOut o = traverse<0>(ray, nodes, root, vert, faces, 0.0f, kNoHit);
  img[i] = shade(o, ray, vert, faces, light);
}

template <int MODE>
__global__ __launch_bounds__(64) void k_traceRays(const bvht_Ray *__restrict__ rays, uint32_t num,
                                                  const bvht_BVHNode *__restrict__ nodes, uint32_t root,
                                                  const bvht_vec3 *__restrict__ vert,
                                                  const bvht_Triangle *__restrict__ faces, bvht_vec3 light, float tMin,
                                                  const uint32_t *__restrict__ skip, bvht_Hit *__restrict__ hits,
                                                  bvht_CUDAPixel *__restrict__ img) {
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i >= num) return;
  const bvht_Ray ray = rays[i];
  TraceOut o = {kNoHit, HUGE_VALF, 0};
  if (!(ray.dir.x == 0.0f && ray.dir.y == 0.0f && ray.dir.z == 0.0f))
    o = traverse<MODE>(ray, nodes, root, vert, faces, tMin, skip ? skip[i] : kNoHit);
  if (hits) hits[i] = bvht_Hit{o.tri, o.t};
  if (img) img[i] = shade(o, ray, vert, faces, light);
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

}  // namespace

}  // namespace bvht

using namespace bvht;

extern "C" {

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
  checkLaunch("k_genRays");
}

void tracerImage(bvht_Ray *rays, bvht_CUDAPixel *img, bvht_BVHNode *nodes, uint32_t rootNode, bvht_MeshRaw mesh,
                 bvht_vec3 light, uint32_t w, uint32_t h, bool bundle) {
  (void)bundle;  // the wave-coherent variant traces the same hits; see bvht_traceRays
  if (!rays || !img || !nodes) {
    setError("tracerImage: missing rays, image or nodes");
    return;
  }
  dim3 grid(divUp(w, 8), divUp(h, 8));
  k_traceImage<<<grid, 64, 0, stream()>>>(rays, img, nodes, rootNode, mesh.vert, mesh.faces, light, w, h);
  checkLaunch("k_traceImage");
}

void copyImageToHost(bvht_CUDAPixel *hostPixel, bvht_CUDAPixel *cudaImg, uint32_t w, uint32_t h) {
  if (check(hipMemcpyAsync(hostPixel, cudaImg, sizeof(bvht_CUDAPixel) * (size_t)w * h, hipMemcpyDeviceToHost, stream()),
            "copyImageToHost"))
    check(hipStreamSynchronize(stream()), "copyImageToHost sync");
}

void tracerDoCudaSync(void) { check(hipStreamSynchronize(stream()), "tracerDoCudaSync"); }

bool bvht_traceRays(const bvht_Ray *rays, uint32_t num, bvht_BVHNode *nodes, uint32_t rootNode, bvht_MeshRaw mesh,
                    bvht_vec3 light, int mode, float tMin, const uint32_t *skipTri, bvht_Hit *hits,
                    bvht_CUDAPixel *img) {
  if (!rays || !nodes || (!hits && !img)) {
    setError("bvht_traceRays: missing rays, nodes or outputs");
    return false;
  }
  if (num == 0) return true;
  if (mode == 0)
    k_traceRays<0><<<divUp(num, 64), 64, 0, stream()>>>(rays, num, nodes, rootNode, mesh.vert, mesh.faces, light, tMin,
                                                       skipTri, hits, img);
  else
    k_traceRays<1><<<divUp(num, 64), 64, 0, stream()>>>(rays, num, nodes, rootNode, mesh.vert, mesh.faces, light, tMin,
                                                       skipTri, hits, img);
  return checkLaunch("k_traceRays");
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
