// common.hpp -- types shared by the host orchestration and the gfx950
// kernels, plus wave64 / 256-thread block primitives and bit writers.
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

// OR `len` (<= 64) bits of `val` into a big-endian-bit-order word array at
// bit `p` (word bit 31 is the first bit of the word).
template <bool kShared>
__device__ __forceinline__ void or_word(u32* w, u32 v, bool be) {
  if (!v) return;
  if (kShared) {
    atomicOr(w, v);
  } else {
    atomicOr(w, be ? bswap32(v) : v);
  }
}

template <bool kShared>
__device__ __forceinline__ void put_msb(u32* words, u64 p, u64 val, u32 len) {
  if (!len) return;
  const u64 t = val << (64 - len);  // left-aligned
  const u32 o = (u32)(p & 31);
  u32* w = words + (p >> 5);
  const u64 hi = t >> o;
  or_word<kShared>(w, (u32)(hi >> 32), true);
  or_word<kShared>(w + 1, (u32)hi, true);
  if (o) or_word<kShared>(w + 2, (u32)((t << (64 - o)) >> 32), true);
}

template <bool kShared>
__device__ __forceinline__ void put_lsb(u32* words, u64 p, u64 val, u32 len) {
  if (!len) return;
  if (len < 64) val &= (1ull << len) - 1;
  const u32 o = (u32)(p & 31);
  u32* w = words + (p >> 5);
  const u64 lo = val << o;
  or_word<kShared>(w, (u32)lo, false);
  or_word<kShared>(w + 1, (u32)(lo >> 32), false);
  if (o) or_word<kShared>(w + 2, (u32)(val >> (64 - o)), false);
}

// A run of `n` one bits at bit `p`, LSB-first.
template <bool kShared>
__device__ __forceinline__ void put_ones_lsb(u32* words, u64 p, u64 n) {
  while (n) {
    const u32 o = (u32)(p & 31);
    const u32 take = (u32)((n < (u64)(32 - o)) ? n : (u64)(32 - o));
    const u32 mask = (take == 32 ? 0xFFFFFFFFu : ((1u << take) - 1u)) << o;
    or_word<kShared>(words + (p >> 5), mask, false);
    p += take;
    n -= take;
  }
}

template <bool kShared, bool kMsb>
__device__ __forceinline__ void put_code(u32* words, u64 p, const Code& c) {
  if (kMsb) {
    put_msb<kShared>(words, p + c.lead, c.val, c.len);
  } else {
    put_ones_lsb<kShared>(words, p, c.lead);
    put_lsb<kShared>(words, p + c.lead, c.val, c.len);
  }
}

__device__ __forceinline__ u32 bitlen64(u64 v) { return v ? 64u - (u32)__clzll((i64)v) : 0u; }

#endif  // __HIPCC__

}  // namespace xnyz
