// Learned removal priority for Bittner13 (extension; the reference ranks by
// the fixed heuristic of Bittner13CUDA.cu:480-500).
//
//   k_subtreeUp  leaf count and subtree surface-area sum of every node
//                (bottom-up, arrival counters between agent-scope fences; the
//                optimizer runs it once and then keeps both exact in its
//                refit);
//   k_depthInit / k_depthJump  node depths by pointer jumping;
//   k_features   12 features per inner node (below);
//   k_mlpInfer   12-64-64-1 MLP, ReLU, BF16 operands on
//                v_mfma_f32_16x16x16_bf16 with FP32 accumulation, 16 nodes
//                per wave and MFMA group, weights held in registers; writes
//                descending sort keys + digit histograms for the optimizer's
//                radix sort, or raw outputs;
//   k_mlpTrain   forward + backward of a mini-batch on MFMA (the weight
//                gradients are MFMA products over the 16 nodes of a tile,
//                staged through LDS); one gradient partial per block;
//   k_mlpAdam    sums the partials in block order (deterministic) and takes
//                an Adam step on the FP32 master weights;
//   k_mlpPack    re-packs the BF16 MFMA operands.
//
// Features of inner node x (S = SA(x), L/R = SA of its children, P = SA of
// its parent, Sr = SA of the root, boxes B):
//   0 log2(S/Sr)/16          1 L/S                 2 R/S
//   3 S/P (1 at the root)    4 SA(sibling)/P       5 min(L,R)/(L+R)
//   6 log2(c/Sr)/16 with c = 2 S^3 / ((L+R) min(L,R)), the reference priority
//   7 min(depth, 255)/64     8 log2(leaves below x)/24
//   9 vol(B_l n B_r)/vol(B_x)   10 vol(B_x n B_sibling)/vol(B_parent)
//   11 log2(subtree SA sum / (S * leaves))/8
#include "mlp.hpp"
#include "wave.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace bvht {

namespace {

__device__ __forceinline__ short f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }
__device__ __forceinline__ uint32_t orderKey(float f) {  // order-preserving u32
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float safeDiv(float a, float b) { return b > 0.0f ? a / b : 0.0f; }
__device__ __forceinline__ float lg2(float x) { return log2f(fmaxf(x, 1e-30f)); }
__device__ __forceinline__ float boxVol(const bvht_AABB &b) {
  return fmaxf(b.minMax[1].x - b.minMax[0].x, 0.0f) * fmaxf(b.minMax[1].y - b.minMax[0].y, 0.0f) *
         fmaxf(b.minMax[1].z - b.minMax[0].z, 0.0f);
}
__device__ __forceinline__ float overlapVol(const bvht_AABB &a, const bvht_AABB &b) {
  const float dx = fminf(a.minMax[1].x, b.minMax[1].x) - fmaxf(a.minMax[0].x, b.minMax[0].x);
  const float dy = fminf(a.minMax[1].y, b.minMax[1].y) - fmaxf(a.minMax[0].y, b.minMax[0].y);
  const float dz = fminf(a.minMax[1].z, b.minMax[1].z) - fmaxf(a.minMax[0].z, b.minMax[0].z);
  return (dx > 0.0f && dy > 0.0f && dz > 0.0f) ? dx * dy * dz : 0.0f;
}

__global__ __launch_bounds__(256) void k_subtreeUp(const bvht_BVHNode *__restrict__ nodes, uint32_t N,
                                                   const uint32_t *__restrict__ leaves, uint32_t numLeaves,
                                                   uint32_t *leafCnt, float *subSA, uint32_t *arrive) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= numLeaves) return;
  const uint32_t leaf = leaves[i];
  leafCnt[leaf] = 1;
  subSA[leaf] = nodes[leaf].surfaceArea;
  fenceReleaseAgent();
  uint32_t node = nodes[leaf].parent;
  if (node == leaf || node >= N) return;
  for (uint32_t steps = 0; steps < N; ++steps) {
    const uint32_t seen = __hip_atomic_fetch_add(arrive + node, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (seen == 0) return;
    fenceAcquireAgent();
    const uint32_t l = nodes[node].left, r = nodes[node].right;
    leafCnt[node] = leafCnt[l] + leafCnt[r];
    subSA[node] = subSA[l] + subSA[r] + nodes[node].surfaceArea;
    fenceReleaseAgent();
    const uint32_t p = nodes[node].parent;
    if (p == node || p >= N) return;
    node = p;
  }
}

// Pointer jumping: after round r, anc is the ancestor 2^r steps up (or the
// root) and dep the distance to it.
__global__ __launch_bounds__(256) void k_depthInit(const bvht_BVHNode *__restrict__ nodes, uint32_t N, uint32_t root,
                                                   uint32_t *__restrict__ anc, uint32_t *__restrict__ dep) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const uint32_t p = nodes[i].parent;
  const bool top = i == root || p >= N;
  anc[i] = top ? root : p;
  dep[i] = top ? 0u : 1u;
}

__global__ __launch_bounds__(256) void k_depthJump(uint32_t N, uint32_t root, const uint32_t *__restrict__ ancIn,
                                                   const uint32_t *__restrict__ depIn, uint32_t *__restrict__ ancOut,
                                                   uint32_t *__restrict__ depOut) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const uint32_t a = ancIn[i];
  if (a == root) {
    ancOut[i] = a;
    depOut[i] = depIn[i];
  } else {
    ancOut[i] = ancIn[a];
    depOut[i] = depIn[i] + depIn[a];
  }
}

__global__ __launch_bounds__(256) void k_features(const bvht_BVHNode *__restrict__ nodes, uint32_t N, uint32_t root,
                                                  const uint32_t *__restrict__ inner, uint32_t num,
                                                  const uint32_t *__restrict__ leafCnt,
                                                  const float *__restrict__ subSA, const uint32_t *__restrict__ depth,
                                                  float *__restrict__ feat) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num) return;
  const uint32_t x = inner[i];
  const bvht_BVHNode nd = nodes[x];
  const float Sr = nodes[root].surfaceArea;
  const float S = nd.surfaceArea, L = nodes[nd.left].surfaceArea, R = nodes[nd.right].surfaceArea;
  const float mn = fminf(L, R);
  float f[kFeat];
  f[0] = lg2(safeDiv(S, Sr)) * (1.0f / 16.0f);
  f[1] = safeDiv(L, S);
  f[2] = safeDiv(R, S);
  f[5] = safeDiv(mn, L + R);
  f[6] = lg2(safeDiv(safeDiv(S * S * S * 2.0f, (L + R) * mn), Sr)) * (1.0f / 16.0f);
  if (x == root || nd.parent >= N) {
    f[3] = 1.0f;
    f[4] = 0.0f;
    f[10] = 0.0f;
  } else {
    const bvht_BVHNode pn = nodes[nd.parent];
    const uint32_t sib = nd.isLeft ? pn.right : pn.left;
    const bvht_BVHNode sn = nodes[sib < N ? sib : x];
    f[3] = safeDiv(S, pn.surfaceArea);
    f[4] = safeDiv(sn.surfaceArea, pn.surfaceArea);
    f[10] = safeDiv(overlapVol(nd.bbox, sn.bbox), boxVol(pn.bbox));
  }
  f[7] = (float)min(depth[x], 255u) * (1.0f / 64.0f);
  const float lc = (float)leafCnt[x];
  f[8] = lg2(lc) * (1.0f / 24.0f);
  f[9] = safeDiv(overlapVol(nodes[nd.left].bbox, nodes[nd.right].bbox), boxVol(nd.bbox));
  f[11] = lg2(safeDiv(subSA[x], S * lc)) * (1.0f / 8.0f);
  float4 *o = reinterpret_cast<float4 *>(feat + (size_t)i * kFeat);
  o[0] = make_float4(f[0], f[1], f[2], f[3]);
  o[1] = make_float4(f[4], f[5], f[6], f[7]);
  o[2] = make_float4(f[8], f[9], f[10], f[11]);
}

// B operand of X^T for the 16 nodes starting at `base` (lane l: node l % 16,
// features 4 (l / 16) .. +3; features 12..15 are zero).
__device__ __forceinline__ bf16x4 loadFeatB(const float *feat, uint32_t row, bool valid, uint32_t lane) {
  const uint32_t kq = lane >> 4;
  bf16x4 xb = {0, 0, 0, 0};
  if (valid && kq < 3) {
    const float4 v = reinterpret_cast<const float4 *>(feat + (size_t)row * kFeat)[kq];
    xb[0] = f2bf(v.x);
    xb[1] = f2bf(v.y);
    xb[2] = f2bf(v.z);
    xb[3] = f2bf(v.w);
  }
  return xb;
}

__global__ __launch_bounds__(256) void k_mlpInfer(const MLPPacked *__restrict__ pk, const float *__restrict__ feat,
                                                  uint32_t num, uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ hist, float *__restrict__ out) {
  __shared__ uint32_t s_hist[4][kSortRadix];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, kq = lane >> 4;
  if (keys) {
    for (int p = 0; p < 4; ++p) s_hist[p][threadIdx.x] = 0;
    __syncthreads();
  }
  bf16x4 w1t[4], w2t[4][4], w3t[4];
  float b1v[4][4], b2v[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    w1t[mt] = pk->w1t[mt][lane];
    w3t[mt] = pk->w3t[mt][lane];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) w2t[mt][kt] = pk->w2t[mt][kt][lane];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      b1v[mt][j] = pk->b1[mt * 16 + kq * 4 + j];
      b2v[mt][j] = pk->b2[mt * 16 + kq * 4 + j];
    }
  }
  const float b3 = pk->b3;
  const uint32_t tiles = (num + 15) >> 4;
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  for (uint32_t tile = blockIdx.x * 4 + wave; tile < tiles; tile += gridDim.x * 4) {  // wave-uniform
    const uint32_t row = tile * 16 + (lane & 15);
    const bool valid = row < num;
    const bf16x4 xb = loadFeatB(feat, row, valid, lane);
    bf16x4 h1b[4], h2b[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const f32x4 h = mfma(w1t[mt], xb, zero);
#pragma unroll
      for (int j = 0; j < 4; ++j) h1b[mt][j] = f2bf(fmaxf(h[j] + b1v[mt][j], 0.0f));
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 h = zero;
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) h = mfma(w2t[mt][kt], h1b[kt], h);
#pragma unroll
      for (int j = 0; j < 4; ++j) h2b[mt][j] = f2bf(fmaxf(h[j] + b2v[mt][j], 0.0f));
    }
    f32x4 o = zero;
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) o = mfma(w3t[kt], h2b[kt], o);
    if (lane < 16 && valid) {
      const float y = o[0] + b3;
      if (keys) {
        const uint32_t key = ~orderKey(y);  // descending
        keys[row] = key;
#pragma unroll
        for (int p = 0; p < 4; ++p) atomicAdd(&s_hist[p][(key >> (8 * p)) & 255u], 1u);
      } else {
        out[row] = y;
      }
    }
  }
  if (keys) {
    __syncthreads();
    for (int p = 0; p < 4; ++p) {
      const uint32_t c = s_hist[p][threadIdx.x];
      if (c) atomicAdd(hist + p * kSortRadix + threadIdx.x, c);
    }
  }
}

// Sum over the 16 lanes that share l / 16 (the nodes of a tile).
__device__ __forceinline__ float sum16(float v) {
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wave per block; the block's tiles are rows [blockIdx.x * tilesPerBlock
// * 16, ...) of the mini-batch `idx[0..B)`.
__global__ __launch_bounds__(64) void k_mlpTrain(const MLPPacked *__restrict__ pk, const float *__restrict__ feat,
                                                 const float *__restrict__ target, const uint32_t *__restrict__ idx,
                                                 uint32_t B, uint32_t tilesPerBlock, float *__restrict__ partials,
                                                 float *__restrict__ lossOut) {
  __shared__ short sH1[16][kHid], sDH2[16][kHid], sDH1[16][kHid];
  __shared__ float sAcc[3][kHid];  // b1, b2, w3 gradient sums
  const uint32_t lane = threadIdx.x, kq = lane >> 4, c = lane & 15;
  for (int k = lane; k < 3 * kHid; k += 64) (&sAcc[0][0])[k] = 0.0f;
  float accB3 = 0.0f, accLoss = 0.0f;
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  f32x4 gW2[4][4], gW1[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gW1[a] = zero;
#pragma unroll
    for (int b = 0; b < 4; ++b) gW2[a][b] = zero;
  }
  const float invB = 2.0f / (float)B;
  for (uint32_t t = 0; t < tilesPerBlock; ++t) {
    const uint32_t base = (blockIdx.x * tilesPerBlock + t) * 16;
    if (base >= B) break;  // block-uniform
    const bool valid = base + c < B;
    const uint32_t row = valid ? idx[base + c] : 0;
    const bf16x4 xb = loadFeatB(feat, row, valid, lane);
    // forward
    f32x4 h1r[4], h2r[4];
    bf16x4 h1b[4], h2b[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const f32x4 h = mfma(pk->w1t[mt][lane], xb, zero);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        h1r[mt][j] = fmaxf(h[j] + pk->b1[mt * 16 + kq * 4 + j], 0.0f);
        h1b[mt][j] = f2bf(h1r[mt][j]);
      }
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 h = zero;
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) h = mfma(pk->w2t[mt][kt][lane], h1b[kt], h);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        h2r[mt][j] = fmaxf(h[j] + pk->b2[mt * 16 + kq * 4 + j], 0.0f);
        h2b[mt][j] = f2bf(h2r[mt][j]);
      }
    }
    f32x4 o = zero;
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) o = mfma(pk->w3t[kt][lane], h2b[kt], o);
    // loss and dL/dy (lanes 0..15 hold node c's output)
    float dOut = 0.0f;
    if (lane < 16 && valid) {
      const float err = o[0] + pk->b3 - target[row];
      accLoss += err * err;
      dOut = err * invB;
    }
    accB3 += dOut;
    const float dO = __shfl(dOut, (int)c, 64);  // node c's gradient on every lane of column c
    // dH2^T (D layout): dO * w3[k] where the ReLU passed
    bf16x4 dh2b[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = mt * 16 + kq * 4 + j;
        const float g = h2r[mt][j] > 0.0f ? dO * bf2f(pk->w3t[k >> 4][((k & 15) >> 2) << 4][k & 3]) : 0.0f;
        dh2b[mt][j] = f2bf(g);
        const float sW3 = sum16(bf2f(h2b[mt][j]) * dO), sB2 = sum16(g);
        if (c == 0) {
          sAcc[2][k] += sW3;
          sAcc[1][k] += sB2;
        }
      }
    // dH1^T = W2 dH2^T, masked by the first ReLU
    bf16x4 dh1b[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 h = zero;
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) h = mfma(pk->w2[mt][kt][lane], dh2b[kt], h);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float g = h1r[mt][j] > 0.0f ? h[j] : 0.0f;
        dh1b[mt][j] = f2bf(g);
        const float sB1 = sum16(g);
        if (c == 0) sAcc[0][mt * 16 + kq * 4 + j] += sB1;
      }
    }
    // stage [node][unit] for the weight-gradient products
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = mt * 16 + kq * 4 + j;
        sH1[c][k] = h1b[mt][j];
        sDH2[c][k] = dh2b[mt][j];
        sDH1[c][k] = dh1b[mt][j];
      }
    __syncthreads();
    // dW2[in][out] += sum_n H1[n][in] dH2[n][out]; dW1[f][out] += sum_n X[n][f] dH1[n][out]
    bf16x4 aH1[4], bDH2[4], bDH1[4], aX;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        aH1[mt][j] = sH1[kq * 4 + j][mt * 16 + c];
        bDH2[mt][j] = sDH2[kq * 4 + j][mt * 16 + c];
        bDH1[mt][j] = sDH1[kq * 4 + j][mt * 16 + c];
      }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t n = base + kq * 4 + j;
      aX[j] = (c < kFeat && n < B) ? f2bf(feat[(size_t)idx[n] * kFeat + c]) : (short)0;
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ko = 0; ko < 4; ++ko) gW2[mi][ko] = mfma(aH1[mi], bDH2[ko], gW2[mi][ko]);
#pragma unroll
    for (int ko = 0; ko < 4; ++ko) gW1[ko] = mfma(aX, bDH1[ko], gW1[ko]);
    __syncthreads();
  }
  float *P = partials + (size_t)blockIdx.x * kNumParams;
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ko = 0; ko < 4; ++ko)
#pragma unroll
      for (int j = 0; j < 4; ++j) P[kOffW2 + (mi * 16 + kq * 4 + j) * kHid + ko * 16 + c] = gW2[mi][ko][j];
#pragma unroll
  for (int ko = 0; ko < 4; ++ko)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t f = kq * 4 + j;
      if (f < kFeat) P[kOffW1 + f * kHid + ko * 16 + c] = gW1[ko][j];
    }
  __syncthreads();
  for (int k = lane; k < kHid; k += 64) {
    P[kOffB1 + k] = sAcc[0][k];
    P[kOffB2 + k] = sAcc[1][k];
    P[kOffW3 + k] = sAcc[2][k];
  }
  float b3s = accB3, ls = accLoss;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    b3s += __shfl_xor(b3s, o, 64);
    ls += __shfl_xor(ls, o, 64);
  }
  if (lane == 0) {
    P[kOffB3] = b3s;
    lossOut[blockIdx.x] = ls;
  }
}

__global__ __launch_bounds__(256) void k_mlpAdam(float *__restrict__ params, float *__restrict__ m,
                                                 float *__restrict__ v, const float *__restrict__ partials,
                                                 uint32_t blocks, float lr, float c1, float c2) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= (uint32_t)kNumParams) return;
  float g = 0.0f;
  for (uint32_t b = 0; b < blocks; ++b) g += partials[(size_t)b * kNumParams + p];
  constexpr float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const float mm = beta1 * m[p] + (1.0f - beta1) * g;
  const float vv = beta2 * v[p] + (1.0f - beta2) * g * g;
  m[p] = mm;
  v[p] = vv;
  params[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);  // c1 = 1/(1-beta1^t), c2 = 1/(1-beta2^t)
}

__global__ __launch_bounds__(64) void k_mlpPack(const float *__restrict__ P, MLPPacked *__restrict__ pk) {
  const uint32_t l = threadIdx.x, kq = l >> 4, c = l & 15;
  for (int mt = 0; mt < 4; ++mt) {
    bf16x4 a, w3;
    for (int j = 0; j < 4; ++j) {
      const uint32_t k = kq * 4 + j;
      a[j] = k < (uint32_t)kFeat ? f2bf(P[kOffW1 + k * kHid + mt * 16 + c]) : (short)0;
      w3[j] = c == 0 ? f2bf(P[kOffW3 + mt * 16 + k]) : (short)0;
    }
    pk->w1t[mt][l] = a;
    pk->w3t[mt][l] = w3;
    for (int kt = 0; kt < 4; ++kt) {
      bf16x4 t, n;
      for (int j = 0; j < 4; ++j) {
        t[j] = f2bf(P[kOffW2 + (kt * 16 + kq * 4 + j) * kHid + mt * 16 + c]);
        n[j] = f2bf(P[kOffW2 + (mt * 16 + c) * kHid + kt * 16 + kq * 4 + j]);
      }
      pk->w2t[mt][kt][l] = t;
      pk->w2[mt][kt][l] = n;
    }
  }
  pk->b1[l] = P[kOffB1 + l];
  pk->b2[l] = P[kOffB2 + l];
  pk->w3[l] = P[kOffW3 + l];
  if (l == 0) pk->b3 = P[kOffB3];
}

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

bool packModel(PriorityModel &m) {
  k_mlpPack<<<1, 64, 0, stream()>>>(m.params, m.packed);
  return checkLaunch("k_mlpPack");
}

void freeModel(PriorityModel *m) {
  if (!m) return;
  devFree(m->params);
  devFree(m->adamM);
  devFree(m->adamV);
  devFree(m->packed);
  devFree(m->partials);
  devFree(m->loss);
  delete m;
}

}  // namespace

bool computeSubtree(const bvht_BVHNode *nodes, uint32_t N, const uint32_t *leaves, uint32_t numLeaves,
                    uint32_t *leafCnt, float *subSA, uint32_t *arrive) {
  if (numLeaves == 0) return true;
  k_subtreeUp<<<divUp(numLeaves, 256), 256, 0, stream()>>>(nodes, N, leaves, numLeaves, leafCnt, subSA, arrive);
  return checkLaunch("k_subtreeUp") && fillU32(arrive, 0u, N);
}

int computeDepth(const bvht_BVHNode *nodes, uint32_t N, uint32_t root, uint32_t *anc[2], uint32_t *dep[2]) {
  k_depthInit<<<divUp(N, 256), 256, 0, stream()>>>(nodes, N, root, anc[0], dep[0]);
  if (!checkLaunch("k_depthInit")) return -1;
  int cur = 0;
  for (int r = 0; r < kDepthRounds; ++r) {
    k_depthJump<<<divUp(N, 256), 256, 0, stream()>>>(N, root, anc[cur], dep[cur], anc[cur ^ 1], dep[cur ^ 1]);
    if (!checkLaunch("k_depthJump")) return -1;
    cur ^= 1;
  }
  return cur;
}

bool computeFeatures(const bvht_BVHNode *nodes, uint32_t N, uint32_t root, const uint32_t *inner, uint32_t numInner,
                     const uint32_t *leafCnt, const float *subSA, const uint32_t *depth, float *feat) {
  if (numInner == 0) return true;
  k_features<<<divUp(numInner, 256), 256, 0, stream()>>>(nodes, N, root, inner, numInner, leafCnt, subSA, depth, feat);
  return checkLaunch("k_features");
}

bool mlpInfer(const PriorityModel &m, const float *feat, uint32_t num, uint32_t *keys, uint32_t *hist, float *out) {
  if (num == 0) return true;
  uint32_t blocks = divUp(divUp(num, 16), 4);
  const uint32_t cap = (uint32_t)numCUs() * 8;
  if (blocks > cap) blocks = cap;
  k_mlpInfer<<<blocks, 256, 0, stream()>>>(m.packed, feat, num, keys, hist, out);
  return checkLaunch("k_mlpInfer");
}

}  // namespace bvht

using namespace bvht;

extern "C" {

uint32_t bvht_mlpNumParams(void) { return (uint32_t)kNumParams; }

bvht_PriorityModel *bvht_mlpCreate(uint32_t seed) {
  PriorityModel *m = new (std::nothrow) PriorityModel();
  if (!m) {
    setError("bvht_mlpCreate: host allocation failed");
    return nullptr;
  }
  if (!devAlloc(&m->params, kNumParams, "MLP params") || !devAlloc(&m->adamM, kNumParams, "MLP Adam") ||
      !devAlloc(&m->adamV, kNumParams, "MLP Adam") || !devAlloc(&m->packed, 1, "MLP operands")) {
    freeModel(m);
    return nullptr;
  }
  // He-uniform weights (fan-in), zero biases, from a seeded splitmix64
  std::vector<float> h(kNumParams, 0.0f);
  uint64_t s = 0x5EEDull ^ ((uint64_t)seed << 32);
  auto uni = [&](float lim) { return (float)((double)(splitmix(s) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * lim; };
  for (int i = 0; i < kFeat * kHid; ++i) h[kOffW1 + i] = uni(std::sqrt(6.0f / kFeat));
  for (int i = 0; i < kHid * kHid; ++i) h[kOffW2 + i] = uni(std::sqrt(6.0f / kHid));
  for (int i = 0; i < kHid; ++i) h[kOffW3 + i] = uni(std::sqrt(3.0f / kHid));
  if (!bvht_mlpSetParams(reinterpret_cast<bvht_PriorityModel *>(m), h.data())) {
    freeModel(m);
    return nullptr;
  }
  return reinterpret_cast<bvht_PriorityModel *>(m);
}

void bvht_mlpFree(bvht_PriorityModel *model) { freeModel(reinterpret_cast<PriorityModel *>(model)); }

bool bvht_mlpSetParams(bvht_PriorityModel *model, const float *hostParams) {
  PriorityModel *m = reinterpret_cast<PriorityModel *>(model);
  if (!m || !hostParams) {
    setError("bvht_mlpSetParams: no model or parameters");
    return false;
  }
  m->step = 0;
  return check(hipMemcpyAsync(m->params, hostParams, sizeof(float) * kNumParams, hipMemcpyHostToDevice, stream()),
               "bvht_mlpSetParams") &&
         fillU32(reinterpret_cast<uint32_t *>(m->adamM), 0u, kNumParams) &&
         fillU32(reinterpret_cast<uint32_t *>(m->adamV), 0u, kNumParams) && packModel(*m) &&
         check(hipStreamSynchronize(stream()), "bvht_mlpSetParams sync");
}

bool bvht_mlpGetParams(bvht_PriorityModel *model, float *hostParams) {
  PriorityModel *m = reinterpret_cast<PriorityModel *>(model);
  if (!m || !hostParams) {
    setError("bvht_mlpGetParams: no model or output");
    return false;
  }
  return check(hipMemcpyAsync(hostParams, m->params, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, stream()),
               "bvht_mlpGetParams") &&
         check(hipStreamSynchronize(stream()), "bvht_mlpGetParams sync");
}

bool bvht_mlpPredict(bvht_PriorityModel *model, const float *devFeat, uint32_t num, float *devOut) {
  PriorityModel *m = reinterpret_cast<PriorityModel *>(model);
  if (!m || (num && (!devFeat || !devOut))) {
    setError("bvht_mlpPredict: no model, features or output");
    return false;
  }
  return mlpInfer(*m, devFeat, num, nullptr, nullptr, devOut);
}

float bvht_mlpTrain(bvht_PriorityModel *model, const float *devFeat, const float *devTarget, uint32_t num,
                    uint32_t epochs, uint32_t batch, float lr, uint32_t seed) {
  PriorityModel *m = reinterpret_cast<PriorityModel *>(model);
  if (!m || !devFeat || !devTarget || num == 0 || batch == 0) {
    setError("bvht_mlpTrain: no model, records or batch");
    return -1.0f;
  }
  if (batch > num) batch = num;
  constexpr uint32_t kTilesPerBlock = 4;
  const uint32_t blocks = divUp(divUp(batch, 16), kTilesPerBlock);
  if (m->partialBlocks < blocks) {
    devFree(m->partials);
    devFree(m->loss);
    m->partials = nullptr;
    m->loss = nullptr;
    m->partialBlocks = 0;
    if (!devAlloc(&m->partials, (size_t)blocks * kNumParams, "MLP gradients") ||
        !devAlloc(&m->loss, blocks, "MLP loss"))
      return -1.0f;
    m->partialBlocks = blocks;
  }
  uint32_t *perm = nullptr;
  if (!devAlloc(&perm, num, "MLP batch order")) return -1.0f;
  std::vector<uint32_t> order(num);
  std::vector<float> lossHost(blocks);
  uint64_t s = 0xB47Cull ^ ((uint64_t)seed << 32);
  double lastEpochLoss = 0.0;
  bool ok = true;
  const uint32_t steps = num / batch;  // full batches per epoch
  for (uint32_t e = 0; e < epochs && ok; ++e) {
    for (uint32_t i = 0; i < num; ++i) order[i] = i;
    for (uint32_t i = num - 1; i > 0; --i) std::swap(order[i], order[splitmix(s) % (i + 1)]);  // Fisher-Yates
    ok = check(hipMemcpyAsync(perm, order.data(), 4ull * num, hipMemcpyHostToDevice, stream()), "MLP order") &&
         check(hipStreamSynchronize(stream()), "MLP order sync");
    double epochLoss = 0.0;
    for (uint32_t st = 0; st < steps && ok; ++st) {
      k_mlpTrain<<<blocks, 64, 0, stream()>>>(m->packed, devFeat, devTarget, perm + (size_t)st * batch, batch,
                                              kTilesPerBlock, m->partials, m->loss);
      ++m->step;
      const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)m->step));
      const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)m->step));
      k_mlpAdam<<<divUp(kNumParams, 256), 256, 0, stream()>>>(m->params, m->adamM, m->adamV, m->partials, blocks, lr,
                                                             c1, c2);
      ok = checkLaunch("k_mlpTrain / k_mlpAdam") && packModel(*m);
      if (ok && e + 1 == epochs) {
        ok = check(hipMemcpyAsync(lossHost.data(), m->loss, 4ull * blocks, hipMemcpyDeviceToHost, stream()),
                   "MLP loss") &&
             check(hipStreamSynchronize(stream()), "MLP loss sync");
        for (float x : lossHost) epochLoss += x;
      }
    }
    if (e + 1 == epochs) lastEpochLoss = epochLoss / ((double)steps * batch);
  }
  devFree(perm);
  return ok ? (float)lastEpochLoss : -1.0f;
}

}  // extern "C"
