// mlp.hpp -- the learned mode selector: a 32 -> 64 -> 64 MLP with ReLU and
// four softmax heads (Golomb-Rice k 0..31, PFD width 1..8, RLE off/on,
// curve Z/Hilbert), FP16 weights with FP32 accumulation.
//
// Parameter blob (floats, in this order): feature mean[32], feature
// 1/std[32], W1[32][64], b1[64], W2[64][64], b2[64], W3[64][44], b3[44];
// W[i][o] multiplies input i into output o.  The kernels run on FP16 copies
// of W1..W3 (products of FP16 operands, summed in FP32); the optimiser keeps
// the FP32 masters.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace xnyz {

constexpr int kFeat = 32, kHid = 64, kOut = 44;
constexpr int kHeadK = 0, kHeadW = 32, kHeadR = 40, kHeadC = 42;  // logit offsets
constexpr int kPMean = 0, kPIstd = 32, kPW1 = 64, kPB1 = kPW1 + kFeat * kHid,
              kPW2 = kPB1 + kHid, kPB2 = kPW2 + kHid * kHid, kPW3 = kPB2 + kHid,
              kPB3 = kPW3 + kHid * kOut, kPTotal = kPB3 + kOut;
constexpr int kTrainable = kPTotal - kPW1;  // everything after the normaliser

struct MlpLabel {  // class indices per head
  int k, w, rle, curve;
};

// Device entry points (mlp.hip); all pointers are device memory.
void mlp_refresh_half(const float* P, __half* Wh, hipStream_t s);
void mlp_train_step(const float* X, const MlpLabel* Y, const u32* idx, u32 nb, float* P,
                    __half* Wh, float* scratch, float* G, float* m, float* v, int t, float lr,
                    float* loss_acc, hipStream_t s);
// scratch: mlp_scratch_floats(nb) floats.
inline size_t mlp_scratch_floats(u32 nb) { return (size_t)nb * (kFeat + 4 * kHid + 2 * kOut); }
constexpr int kHalfWeights = kFeat * kHid + kHid * kHid + kHid * kOut;
// Class of every head and the confidence min_h max softmax_h of n samples.
void mlp_predict(const float* X, u32 n, const float* P, const __half* Wh, MlpLabel* out,
                 float* conf, float* scratch, hipStream_t s);

}  // namespace xnyz
