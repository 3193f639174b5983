// Learned node priority for Bittner13: node features and a 12-64-64-1 MLP
// (BF16 weights on MFMA, FP32 accumulation).  Shared by the optimizer
// (bittner.hip) and the model kernels (mlp.hip).
#pragma once

#include "radix_sort.hpp"

namespace bvht {

constexpr int kFeat = 12, kHid = 64;
// fp32 parameter vector: w1[12][64] (in, out), b1[64], w2[64][64] (in, out),
// b2[64], w3[64], b3
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kFeat * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid, kNumParams = kOffB3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// MFMA operands of the weights, one bf16x4 per lane (v_mfma_f32_16x16x16_bf16:
// A lane l holds A[l % 16][4 (l / 16) + j]; B lane l holds B[4 (l / 16) + j][l % 16];
// D lane l holds D[4 (l / 16) + j][l % 16]).  Layers run transposed
// (H^T = W^T X^T), so a layer's D tile is the next layer's B operand.
struct MLPPacked {
  bf16x4 w1t[4][64];     // W1^T tile mt
  bf16x4 w2t[4][4][64];  // W2^T tile (mt over outputs, kt over inputs)
  bf16x4 w2[4][4][64];   // W2 tile (mt over inputs, kt over outputs): backward
  bf16x4 w3t[4][64];     // w3^T (row 0 of a 16-row tile), tile kt
  float b1[kHid], b2[kHid], w3[kHid], b3;
};

struct PriorityModel {
  float *params = nullptr;  // fp32 master, kNumParams
  float *adamM = nullptr, *adamV = nullptr;
  MLPPacked *packed = nullptr;
  float *partials = nullptr;  // training: per-block gradient partials
  uint32_t partialBlocks = 0;
  float *loss = nullptr;  // per-block loss partials
  uint32_t step = 0;
};

// Leaf count and subtree surface-area sum of every node (bottom-up from the
// leaves; `arrive` holds N zeros on entry and on exit).
bool computeSubtree(const bvht_BVHNode *nodes, uint32_t N, const uint32_t *leaves, uint32_t numLeaves,
                    uint32_t *leafCnt, float *subSA, uint32_t *arrive);
// Depth of every node, exact below 2^kDepthRounds (larger depths come out
// as at least that), by pointer jumping over double buffers anc[2][N],
// dep[2][N]; returns the buffer index holding the result (-1 on error).
constexpr int kDepthRounds = 8;
int computeDepth(const bvht_BVHNode *nodes, uint32_t N, uint32_t root, uint32_t *anc[2], uint32_t *dep[2]);
// Features of the inner nodes `inner[0..num)` into feat[num][12] from the
// per-node leaf counts, subtree SA sums and depths.
bool computeFeatures(const bvht_BVHNode *nodes, uint32_t N, uint32_t root, const uint32_t *inner, uint32_t numInner,
                     const uint32_t *leafCnt, const float *subSA, const uint32_t *depth, float *feat);

// Model outputs of feat[num][12] as descending sort keys (~order-preserving
// bits) plus the per-pass digit histograms of a 32-bit sort, or as raw
// values (keys == nullptr, out != nullptr).
bool mlpInfer(const PriorityModel &m, const float *feat, uint32_t num, uint32_t *keys, uint32_t *hist, float *out);

}  // namespace bvht
