// mlp.hip -- gfx950 kernels of the learned mode selector (mlp.hpp).
//
// One thread per sample for the layer products (the layers are 32-64 wide:
// a sample's whole forward pass is ~9k multiply-adds, far too small to tile
// onto matrix cores with any benefit), one thread per parameter for the
// weight gradients, Adam on the FP32 masters.  Activations live in a scratch
// buffer laid out [unit][sample] so that every access is coalesced across
// the threads of a wavefront.  FP16 semantics: inputs and activations are
// rounded to FP16 before each product and weights are read from FP16 copies;
// products and sums are FP32.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"
#include "mlp.hpp"

namespace xnyz {
namespace {

constexpr int kH1 = 0, kH2 = kH1 + kFeat * kHid, kH3 = kH2 + kHid * kHid,
              kHTotal = kH3 + kHid * kOut;  // FP16 weight copies

__device__ __forceinline__ float f16r(float x) { return __half2float(__float2half(x)); }

struct Scratch {  // [unit][sample] arrays for nb samples
  float *xn, *h1, *h2, *z, *dz, *dh2, *dh1;
  __device__ __host__ Scratch(float* base, u32 nb)
      : xn(base), h1(xn + kFeat * nb), h2(h1 + kHid * nb), z(h2 + kHid * nb),
        dz(z + kOut * nb), dh2(dz + kOut * nb), dh1(dh2 + kHid * nb) {}
};

__global__ void k_refresh_half(const float* __restrict__ P, __half* __restrict__ Wh) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kHTotal; i += gridDim.x * blockDim.x)
    Wh[i] = __float2half(P[kPW1 + (i < kH2 ? i : i < kH3 ? i - kH2 + (kPW2 - kPW1)
                                                         : i - kH3 + (kPW3 - kPW1))]);
}

// Forward pass of samples idx[0..nb) (or 0..nb); with labels, the softmax
// cross-entropy of the four heads: dz = (softmax - onehot) / nb, loss summed.
__global__ __launch_bounds__(256) void k_fwd(const float* __restrict__ X,
                                             const MlpLabel* __restrict__ Y,
                                             const u32* __restrict__ idx, u32 nb,
                                             const float* __restrict__ P,
                                             const __half* __restrict__ Wh, float* sbase,
                                             float* loss_acc) {
  const u32 s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  const Scratch S(sbase, nb);
  const u32 si = idx ? idx[s] : s;
  for (int i = 0; i < kFeat; ++i)
    S.xn[i * nb + s] = f16r((X[(size_t)si * kFeat + i] - P[kPMean + i]) * P[kPIstd + i]);
  for (int o = 0; o < kHid; ++o) {
    float acc = P[kPB1 + o];
    for (int i = 0; i < kFeat; ++i) acc += S.xn[i * nb + s] * __half2float(Wh[kH1 + i * kHid + o]);
    S.h1[o * nb + s] = f16r(acc > 0.f ? acc : 0.f);
  }
  for (int o = 0; o < kHid; ++o) {
    float acc = P[kPB2 + o];
    for (int i = 0; i < kHid; ++i) acc += S.h1[i * nb + s] * __half2float(Wh[kH2 + i * kHid + o]);
    S.h2[o * nb + s] = f16r(acc > 0.f ? acc : 0.f);
  }
  for (int o = 0; o < kOut; ++o) {
    float acc = P[kPB3 + o];
    for (int i = 0; i < kHid; ++i) acc += S.h2[i * nb + s] * __half2float(Wh[kH3 + i * kOut + o]);
    S.z[o * nb + s] = acc;
  }
  if (!Y) return;
  const MlpLabel y = Y[si];
  const int lab[4] = {y.k, y.w, y.rle, y.curve};
  const int h0[5] = {kHeadK, kHeadW, kHeadR, kHeadC, kOut};
  float loss = 0.f;
  for (int h = 0; h < 4; ++h) {
    float mx = -INFINITY;
    for (int o = h0[h]; o < h0[h + 1]; ++o) mx = fmaxf(mx, S.z[o * nb + s]);
    float den = 0.f;
    for (int o = h0[h]; o < h0[h + 1]; ++o) den += expf(S.z[o * nb + s] - mx);
    for (int o = h0[h]; o < h0[h + 1]; ++o) {
      const float p = expf(S.z[o * nb + s] - mx) / den;
      const bool hit = (o - h0[h]) == lab[h];
      S.dz[o * nb + s] = (p - (hit ? 1.f : 0.f)) / (float)nb;
      if (hit) loss -= logf(fmaxf(p, 1e-30f));
    }
  }
  atomicAdd(loss_acc, loss / (float)nb);
}

// dh2 = relu'(h2) * (dz W3^T), dh1 = relu'(h1) * (dh2 W2^T)
__global__ __launch_bounds__(256) void k_bwd(u32 nb, const __half* __restrict__ Wh,
                                             float* sbase) {
  const u32 s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  const Scratch S(sbase, nb);
  for (int h = 0; h < kHid; ++h) {
    float g = 0.f;
    for (int o = 0; o < kOut; ++o) g += S.dz[o * nb + s] * __half2float(Wh[kH3 + h * kOut + o]);
    S.dh2[h * nb + s] = S.h2[h * nb + s] > 0.f ? g : 0.f;
  }
  for (int h = 0; h < kHid; ++h) {
    float g = 0.f;
    for (int o = 0; o < kHid; ++o) g += S.dh2[o * nb + s] * __half2float(Wh[kH2 + h * kHid + o]);
    S.dh1[h * nb + s] = S.h1[h * nb + s] > 0.f ? g : 0.f;
  }
}

// Gradient of trainable parameter p (offset from kPW1), summed over the batch.
__global__ __launch_bounds__(256) void k_wgrad(u32 nb, float* sbase, float* __restrict__ G) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kTrainable) return;
  const Scratch S(sbase, nb);
  const int q = p + kPW1;
  const float *a = nullptr, *d;
  if (q < kPB1) {
    a = S.xn + ((q - kPW1) / kHid) * nb;
    d = S.dh1 + ((q - kPW1) % kHid) * nb;
  } else if (q < kPW2) {
    d = S.dh1 + (q - kPB1) * nb;
  } else if (q < kPB2) {
    a = S.h1 + ((q - kPW2) / kHid) * nb;
    d = S.dh2 + ((q - kPW2) % kHid) * nb;
  } else if (q < kPW3) {
    d = S.dh2 + (q - kPB2) * nb;
  } else if (q < kPB3) {
    a = S.h2 + ((q - kPW3) / kOut) * nb;
    d = S.dz + ((q - kPW3) % kOut) * nb;
  } else {
    d = S.dz + (q - kPB3) * nb;
  }
  float g = 0.f;
  if (a) {
    for (u32 s = 0; s < nb; ++s) g += a[s] * d[s];
  } else {
    for (u32 s = 0; s < nb; ++s) g += d[s];
  }
  G[p] = g;
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, const float* __restrict__ G,
                                              float* __restrict__ m, float* __restrict__ v,
                                              float lr, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kTrainable) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float g = G[p];
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  P[kPW1 + p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ __launch_bounds__(256) void k_argmax(u32 n, float* sbase, MlpLabel* __restrict__ out,
                                                float* __restrict__ conf) {
  const u32 s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const Scratch S(sbase, n);
  const int h0[5] = {kHeadK, kHeadW, kHeadR, kHeadC, kOut};
  int best[4];
  float cmin = 1.f;
  for (int h = 0; h < 4; ++h) {
    float mx = -INFINITY;
    int am = 0;
    for (int o = h0[h]; o < h0[h + 1]; ++o)
      if (S.z[o * n + s] > mx) {
        mx = S.z[o * n + s];
        am = o - h0[h];
      }
    float den = 0.f;
    for (int o = h0[h]; o < h0[h + 1]; ++o) den += expf(S.z[o * n + s] - mx);
    cmin = fminf(cmin, 1.f / den);
    best[h] = am;
  }
  out[s] = MlpLabel{best[0], best[1], best[2], best[3]};
  if (conf) conf[s] = cmin;
}

inline u32 blocks_for(u32 n) { return (n + 255) / 256; }

}  // namespace

void mlp_refresh_half(const float* P, __half* Wh, hipStream_t s) {
  k_refresh_half<<<blocks_for(kHTotal), 256, 0, s>>>(P, Wh);
  check_hip(hipGetLastError(), "k_refresh_half");
}

void mlp_train_step(const float* X, const MlpLabel* Y, const u32* idx, u32 nb, float* P,
                    __half* Wh, float* scratch, float* G, float* m, float* v, int t, float lr,
                    float* loss_acc, hipStream_t s) {
  k_fwd<<<blocks_for(nb), 256, 0, s>>>(X, Y, idx, nb, P, Wh, scratch, loss_acc);
  check_hip(hipGetLastError(), "k_fwd");
  k_bwd<<<blocks_for(nb), 256, 0, s>>>(nb, Wh, scratch);
  check_hip(hipGetLastError(), "k_bwd");
  k_wgrad<<<blocks_for(kTrainable), 256, 0, s>>>(nb, scratch, G);
  check_hip(hipGetLastError(), "k_wgrad");
  const float c1 = 1.f - std::pow(0.9f, (float)t), c2 = 1.f - std::pow(0.999f, (float)t);
  k_adam<<<blocks_for(kTrainable), 256, 0, s>>>(P, G, m, v, lr, c1, c2);
  check_hip(hipGetLastError(), "k_adam");
  mlp_refresh_half(P, Wh, s);
}

void mlp_predict(const float* X, u32 n, const float* P, const __half* Wh, MlpLabel* out,
                 float* conf, float* scratch, hipStream_t s) {
  k_fwd<<<blocks_for(n), 256, 0, s>>>(X, nullptr, nullptr, n, P, Wh, scratch, nullptr);
  check_hip(hipGetLastError(), "k_fwd");
  k_argmax<<<blocks_for(n), 256, 0, s>>>(n, scratch, out, conf);
  check_hip(hipGetLastError(), "k_argmax");
}

}  // namespace xnyz
