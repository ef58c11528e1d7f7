// common.hpp -- types shared by the host orchestration and the gfx950
// kernels, plus wave64 / 256-thread block primitives.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace xnyz {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = unsigned long long;  // the width HIP's 64-bit atomics take
using i64 = long long;

constexpr int kWave = 64;        // CDNA wavefront
constexpr int kBlock = 256;      // threads per workgroup (4 wavefronts)
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr u32 kSymBins = 256;    // PFD symbol alphabet (repos, quads, sfc deltas)

// One code word for the variable-length packers.  MSB-first streams
// (Huffman, Exp-Golomb): `lead` zero bits, then `len` bits of `val`, most
// significant first.  LSB-first streams (Golomb-Rice): `lead` one bits, then
// `len` bits of `val`, least significant first.
struct Code {
  u64 val;
  u32 len;
  u32 lead;
};

struct HufEntry {  // codebook entry uploaded by the host
  u32 bits;
  u32 len;
};

#if defined(__HIPCC__)

__device__ __forceinline__ u32 lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, kWave);
    v = w < v ? w : v;
  }
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, kWave);
    v = w > v ? w : v;
  }
  return v;
}

// Inclusive scan over the 64 lanes of a wavefront.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const u32 l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T w = __shfl_up(v, o, kWave);
    if (l >= (u32)o) v += w;
  }
  return v;
}

// Exclusive scan over a 256-thread block; returns the thread's exclusive
// prefix and writes the block total.  `sh` holds kWavesPerBlock entries.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* sh, T& total) {
  const u32 l = lane_id(), w = threadIdx.x / kWave;
  T inc = wave_incl_scan(v);
  if (l == kWave - 1) sh[w] = inc;
  __syncthreads();
  T base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kWavesPerBlock; ++i) {
    T s = sh[i];
    if ((u32)i < w) base += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + inc - v;
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* sh) {
  T t;
  (void)block_excl_scan(v, sh, t);
  return t;
}

__device__ __forceinline__ u32 bswap32(u32 v) { return __builtin_bswap32(v); }

__device__ __forceinline__ u32 bitlen64(u64 v) { return v ? 64u - (u32)__clzll((i64)v) : 0u; }

#endif  // __HIPCC__

}  // namespace xnyz
