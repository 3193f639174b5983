// Library runtime, memory/copy entry points of the C ABI
// (reference: lib/cuda/cudaFN.cu:32-187, lib/cuda/cudaFN.hpp:27-42).
#include "runtime.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

namespace bvht {

Runtime &rt() {
  static Runtime r;
  return r;
}

void setError(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  rt().lastError = buf;
}

bool check(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  setError("%s: %s", what, hipGetErrorString(e));
  return false;
}

bool devMalloc(void **p, size_t bytes, const char *what) {
  *p = nullptr;
  Runtime &r = rt();
  if (r.allocFailAfter > 0) {
    if (--r.allocFailAfter == 0) {
      setError("%s: injected allocation failure (%zu bytes)", what, bytes);
      return false;
    }
  }
  if (bytes == 0) bytes = 16;
  return check(hipMalloc(p, bytes), what);
}

void devFree(void *p) {
  if (p) (void)hipFree(p);
}

int numCUs() {
  static int cached = 0;
  if (cached == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
      cached = cus > 0 ? cus : 256;
    else
      cached = 256;
  }
  return cached;
}

__global__ void k_fillU32(uint32_t *p, uint32_t v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

bool fillU32(uint32_t *p, uint32_t value, size_t count) {
  if (count == 0) return true;
  uint32_t blocks = divUp(count, 256);
  if (blocks > 4096) blocks = 4096;
  k_fillU32<<<blocks, 256, 0, stream()>>>(p, value, count);
  return checkLaunch("fillU32");
}

}  // namespace bvht

using namespace bvht;

extern "C" {

void bvht_setStream(void *s) { rt().stream = reinterpret_cast<hipStream_t>(s); }
void *bvht_getStream(void) { return rt().stream; }
const char *bvht_getLastError(void) { return rt().lastError.c_str(); }
void bvht_clearError(void) { rt().lastError.clear(); }
void bvht_setAllocFailAfter(int n) { rt().allocFailAfter = n < 0 ? 0 : n; }
const char *bvht_version(void) { return "bvhtest_amd 0.1 (gfx950)"; }
int bvht_deviceCount(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

bool runMalloc(void **ptr, size_t size) { return devMalloc(ptr, size, "runMalloc"); }

bool runMemcpy(void *dest, const void *src, size_t size, int kind) {
  if (kind < 0 || kind > 3) {
    setError("runMemcpy: invalid kind %d", kind);
    return false;
  }
  if (!check(hipMemcpyAsync(dest, src, size, (hipMemcpyKind)kind, stream()), "runMemcpy")) return false;
  return check(hipStreamSynchronize(stream()), "runMemcpy sync");
}

void runFree(void *ptr) { devFree(ptr); }

bool copyBVHToGPU(const bvht_BVHNode *hostNodes, uint32_t numNodes, uint32_t root, uint16_t maxLevel,
                  bvht_CUDAMemoryBVHPointer *out) {
  if (!out) {
    setError("copyBVHToGPU: null output");
    return false;
  }
  if (numNodes == 0) {  // cudaFN.cu:35-40
    out->bvh = nullptr;
    out->nodes = nullptr;
    out->numNodes = 0;
    return true;
  }
  if (!hostNodes || root >= numNodes) {
    setError("copyBVHToGPU: invalid host BVH");
    return false;
  }
  bvht_BVH hdr;
  std::memset(&hdr, 0, sizeof(hdr));
  out->numNodes = numNodes;
  if (!devAlloc(&out->nodes, numNodes, "copyBVHToGPU nodes")) goto error;
  if (!devAlloc(&out->bvh, 1, "copyBVHToGPU header")) goto error;
  hdr.bvh = out->nodes;
  hdr.vSize = numNodes;
  hdr.vCapacity = numNodes;
  hdr.vRootIndex = root;
  hdr.vMaxLevel = maxLevel;
  if (!check(hipMemcpyAsync(out->nodes, hostNodes, sizeof(bvht_BVHNode) * numNodes, hipMemcpyHostToDevice, stream()),
             "copyBVHToGPU nodes"))
    goto error;
  if (!check(hipMemcpyAsync(out->bvh, &hdr, sizeof(hdr), hipMemcpyHostToDevice, stream()), "copyBVHToGPU header")) goto error;
  if (!check(hipStreamSynchronize(stream()), "copyBVHToGPU sync")) goto error;
  return true;
error:
  devFree(out->nodes);
  devFree(out->bvh);
  out->nodes = nullptr;
  out->bvh = nullptr;
  out->numNodes = 0;
  return false;
}

bool copyMeshToGPU(const bvht_vec3 *vert, uint32_t numVert, const bvht_vec3 *norm, uint32_t numNorm,
                   const bvht_Triangle *faces, uint32_t numFaces, bvht_MeshRaw *out) {
  if (!out) {
    setError("copyMeshToGPU: null output");
    return false;
  }
  *out = bvht_MeshRaw{nullptr, nullptr, nullptr, 0, 0, 0};
  if (!devAlloc(&out->vert, numVert, "copyMeshToGPU vert")) goto error;
  if (!devAlloc(&out->norm, numNorm, "copyMeshToGPU norm")) goto error;
  if (!devAlloc(&out->faces, numFaces, "copyMeshToGPU faces")) goto error;
  if (numVert && !check(hipMemcpyAsync(out->vert, vert, 12ull * numVert, hipMemcpyHostToDevice, stream()), "copyMeshToGPU vert"))
    goto error;
  if (numNorm && !check(hipMemcpyAsync(out->norm, norm, 12ull * numNorm, hipMemcpyHostToDevice, stream()), "copyMeshToGPU norm"))
    goto error;
  if (numFaces &&
      !check(hipMemcpyAsync(out->faces, faces, 12ull * numFaces, hipMemcpyHostToDevice, stream()), "copyMeshToGPU faces"))
    goto error;
  if (!check(hipStreamSynchronize(stream()), "copyMeshToGPU sync")) goto error;
  out->numVert = numVert;
  out->numNorm = numNorm;
  out->numFaces = numFaces;
  return true;
error:
  devFree(out->vert);
  devFree(out->norm);
  devFree(out->faces);
  *out = bvht_MeshRaw{nullptr, nullptr, nullptr, 0, 0, 0};
  return false;
}

bool copyBVHToHost(bvht_CUDAMemoryBVHPointer *bvh, bvht_BVHNode *hostNodes, uint32_t capacity, uint32_t *numNodes,
                   uint32_t *root, uint16_t *maxLevel) {
  if (!bvh || !bvh->bvh || !bvh->nodes) {  // cudaFN.cu:114
    setError("copyBVHToHost: no device BVH");
    return false;
  }
  bool ok = false;
  bvht_BVH hdr;
  if (check(hipMemcpyAsync(&hdr, bvh->bvh, sizeof(hdr), hipMemcpyDeviceToHost, stream()), "copyBVHToHost header") &&
      check(hipStreamSynchronize(stream()), "copyBVHToHost sync")) {
    if (numNodes) *numNodes = (uint32_t)hdr.vSize;
    if (root) *root = hdr.vRootIndex;
    if (maxLevel) *maxLevel = hdr.vMaxLevel;
    if (hdr.vSize > capacity || !hostNodes) {
      setError("copyBVHToHost: host buffer holds %u nodes, BVH has %zu", capacity, hdr.vSize);
    } else {
      ok = check(hipMemcpyAsync(hostNodes, bvh->nodes, sizeof(bvht_BVHNode) * hdr.vSize, hipMemcpyDeviceToHost, stream()),
                 "copyBVHToHost nodes") &&
           check(hipStreamSynchronize(stream()), "copyBVHToHost sync");
    }
  }
  if (ok) bvht_freeBVH(bvh);  // the reference frees the device BVH (cudaFN.cu:136-143)
  return ok;
}

bool copyMeshToHost(bvht_MeshRaw *mesh, bvht_vec3 *vert, bvht_vec3 *norm, bvht_Triangle *faces) {
  if (!mesh || !mesh->vert || !mesh->faces) {
    setError("copyMeshToHost: no device mesh");
    return false;
  }
  bool ok = (!mesh->numVert ||
             check(hipMemcpyAsync(vert, mesh->vert, 12ull * mesh->numVert, hipMemcpyDeviceToHost, stream()), "copyMeshToHost vert")) &&
            (!mesh->numNorm ||
             check(hipMemcpyAsync(norm, mesh->norm, 12ull * mesh->numNorm, hipMemcpyDeviceToHost, stream()), "copyMeshToHost norm")) &&
            (!mesh->numFaces || check(hipMemcpyAsync(faces, mesh->faces, 12ull * mesh->numFaces, hipMemcpyDeviceToHost, stream()),
                                      "copyMeshToHost faces")) &&
            check(hipStreamSynchronize(stream()), "copyMeshToHost sync");
  if (ok) bvht_freeMesh(mesh);
  return ok;
}

void bvht_freeBVH(bvht_CUDAMemoryBVHPointer *bvh) {
  if (!bvh) return;
  devFree(bvh->nodes);
  devFree(bvh->bvh);
  bvh->nodes = nullptr;
  bvh->bvh = nullptr;
  bvh->numNodes = 0;
}

void bvht_freeMesh(bvht_MeshRaw *mesh) {
  if (!mesh) return;
  devFree(mesh->vert);
  devFree(mesh->norm);
  devFree(mesh->faces);
  *mesh = bvht_MeshRaw{nullptr, nullptr, nullptr, 0, 0, 0};
}

void resetCUDA(void) { (void)hipStreamSynchronize(stream()); }

}  // extern "C"
