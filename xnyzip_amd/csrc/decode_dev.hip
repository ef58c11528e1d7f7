// decode_dev.hip -- the decoder on the device (gfx950): assembled stream ->
// points, for BlockDecompressor[RLE]<float>::decompress
// (source/composer/block_decompressor.hpp:31-192,
// block_decompressor_rle.hpp:32-207, recover_from_lcp_meta
// block_quantizer.hpp:153-207).
//
// Entropy-coded sections (Huffman / Exp-Golomb / Golomb-Rice) carry no
// restart points.  They are decoded in parallel by self-synchronisation: the
// section's bits are cut into sub-sequences of kSub bits, one per thread;
// every thread decodes its sub-sequence from its first bit as if a code word
// started there (k_pd_spec), recording where its last code word ends (the
// first code-word start at or past its end, "exit") and how many code words
// it saw.  The true entry of sub-sequence t is the exit of t - 1; rounds of
// k_pd_sync re-decode each sub-sequence from its predecessor's exit until the
// new path meets the old one at a common code-word start (prefix codes fall
// into step within a few code words), after which the rest of the old path,
// its count and its exit stand.  Rounds repeat until no exit changes (each
// round fixes at least the next sub-sequence, so the loop ends).  An
// exclusive scan of the counts gives every sub-sequence its first output
// index, and k_pd_write decodes once more from the settled entries, writing
// the values.  Code words decoded from the final byte's padding fall after
// the section's last index and are not written.
//
// The LCP metadata is then rebuilt by one scan over the elements with a
// composite operator (k_lcp_reduce / k_lcp_tiles / k_lcp_apply): the block
// index (count of block starts), the patch index (count of PFD symbol 0),
// the quad within the block (sum of quad deltas since the block's start) and
// the sub-block key (sum of repos deltas since the last block start or quad
// change), whose patch part comes from a prefix sum of the patch values.  The
// apply pass turns every element into its lattice node, recovers and
// unshifts it, and writes the point (RLE: count copies at its output offset).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "decode_dev.hpp"
#include "exact_div.hpp"

#pragma clang fp contract(off)

namespace xnyz {
namespace {

#define LAUNCH_CHECK(name)                      \
  do {                                          \
    check_hip(hipGetLastError(), name);         \
    if (g_profiler) g_profiler->mark(name, s);  \
  } while (0)

inline u32 blocks_for(u64 n, u32 per_block) {
  const u64 g = (n + per_block - 1) / per_block;
  return (u32)(g ? g : 1);
}

// ---- bit access --------------------------------------------------------------
// The stream buffer is 256-byte aligned and padded with 64 zero bytes, so
// aligned 8-byte loads up to three words past any bit before the stream's
// end stay inside it.

__device__ __forceinline__ u64 ld64(const uint8_t* s, u64 w) {
  return reinterpret_cast<const u64*>(s)[w];
}

// 64 bits from bit g on, the first one in bit 63 (MSB-first streams: Huffman,
// Exp-Golomb; bit i of a byte stream is bit 7 - i%8 of byte i/8).
__device__ __forceinline__ u64 peek_msb(const uint8_t* s, u64 g) {
  const u64 w = g >> 6;
  const u32 o = (u32)(g & 63);
  const u64 a = __builtin_bswap64(ld64(s, w));
  if (!o) return a;
  return (a << o) | (__builtin_bswap64(ld64(s, w + 1)) >> (64 - o));
}

// 64 bits from bit g on, the first one in bit 0 (LSB-first streams:
// Golomb-Rice; bit i is bit i%8 of byte i/8).
__device__ __forceinline__ u64 peek_lsb(const uint8_t* s, u64 g) {
  const u64 w = g >> 6;
  const u32 o = (u32)(g & 63);
  const u64 a = ld64(s, w);
  if (!o) return a;
  return (a >> o) | (ld64(s, w + 1) << (64 - o));
}

// A thread's forward-only view of the stream: three consecutive 64-bit words
// in registers (byte-swapped for the MSB-first codes), so that a walk over a
// sub-sequence reloads once per 64 bits instead of waiting on two loads per
// code word.  Every peek of a walk is at or past the previous one.  The
// buffer's 64 zero bytes of padding keep the two words read ahead inside it.
template <bool kMsb>
struct BitWin {
  const u64* s;
  u64 wi;  // word of w0
  u64 w0, w1, w2;

  __device__ __forceinline__ u64 ld(u64 k) const {
    const u64 x = s[k];
    return kMsb ? __builtin_bswap64(x) : x;
  }
  __device__ __forceinline__ void at(u64 g) {
    wi = g >> 6;
    w0 = ld(wi);
    w1 = ld(wi + 1);
    w2 = ld(wi + 2);
  }
  __device__ __forceinline__ BitWin(const uint8_t* base, u64 g)
      : s(reinterpret_cast<const u64*>(base)) {
    at(g);
  }
  // peek_msb / peek_lsb of bit g (g at or past the last peek)
  __device__ __forceinline__ u64 peek(u64 g) {
    u64 d = g - (wi << 6);
    if (d >= 64) {
      if (d < 128) {
        w0 = w1;
        w1 = w2;
        w2 = ld(wi + 3);
        ++wi;
        d -= 64;
      } else {
        at(g);
        d = g & 63;
      }
    }
    const u32 o = (u32)d;
    if (!o) return w0;
    return kMsb ? (w0 << o) | (w1 >> (64 - o)) : (w0 >> o) | (w1 << (64 - o));
  }
};

template <u32 kKind>
using WinOf = BitWin<kKind != kCodeGR>;

// ---- one code word -----------------------------------------------------------
// Each decodes the code word starting at section bit p (< c.bits) into v and
// returns the bit after it.  A code word that would run past the section's
// end returns c.bits + 1 (no valid stream has one; the caller's walk stops).

// Huffman: a lookup of lb bits, then the tree for longer codes (at most 32
// bits).
__device__ __forceinline__ u64 next_huf_w(BitWin<true>& W, const CodeDesc& c, const u32* lut,
                                          u64 p, u64& v) {
  const u64 w = W.peek(c.base_bit + p);
  const u32 e = lut[w >> (64 - c.lb)];
  const u32 len = e >> 24;
  if (len) {
    v = e & 0xFFFFFFu;
    return p + len;
  }
  u32 node = e & 0xFFFFFFu;
  for (u32 i = c.lb; i < 64; ++i) {
    const u32 ch = c.tree[2 * node + ((w >> (63 - i)) & 1)];
    if (ch == kTreeNull) return c.bits + 1;
    if (ch & kTreeLeaf) {
      v = ch & ~kTreeLeaf;
      return p + i + 1;
    }
    node = ch;
  }
  return c.bits + 1;
}

// huffman_encoder.hpp:365-396: z zeros, then n1 = (n >> k) + 1 in z + 1 bits
// MSB-first, then the k low bits of n LSB-first.
__device__ __forceinline__ u64 next_eg_w(BitWin<true>& W, const CodeDesc& c, u64 p, u64& v) {
  const u64 w = W.peek(c.base_bit + p);
  if (!w) return c.bits + 1;
  const u32 z = (u32)__clzll((i64)w);
  const u64 q = p + z + 1 + z + c.k;
  if (q > c.bits) return c.bits + 1;
  u64 n1;
  if (2 * z + 1 <= 64) {
    n1 = (w << z) >> (63 - z);
  } else {
    n1 = W.peek(c.base_bit + p + z) >> (63 - z);
  }
  u64 low = 0;
  if (c.k) {
    const u64 t = W.peek(c.base_bit + p + 2 * z + 1) >> (64 - c.k);
    low = __builtin_bitreverse64(t) >> (64 - c.k);
  }
  v = ((n1 - 1) << c.k) | low;
  return q;
}

// golomb_rice_encoder.hpp:92-148: q one bits, a zero, then the k low bits
// LSB-first.
__device__ __forceinline__ u64 next_gr_w(BitWin<false>& W, const CodeDesc& c, u64 p, u64& v) {
  u64 q = 0;
  u64 w = W.peek(c.base_bit + p);
  while (w == ~0ull) {
    q += 64;
    if (p + q >= c.bits) return c.bits + 1;
    w = W.peek(c.base_bit + p + q);
  }
  q += (u64)__builtin_ctzll(~w);
  const u64 e = p + q + 1 + c.k;
  if (e > c.bits) return c.bits + 1;
  u64 r = 0;
  if (c.k) {
    r = W.peek(c.base_bit + p + q + 1);
    if (c.k < 64) r &= (1ull << c.k) - 1;
  }
  v = (q << c.k) + r;
  return e;
}

template <u32 kKind>
__device__ __forceinline__ u64 next_code_w(WinOf<kKind>& W, const CodeDesc& c, const u32* lut,
                                           u64 p, u64& v) {
  if constexpr (kKind == kCodeHuf) return next_huf_w(W, c, lut, p, v);
  if constexpr (kKind == kCodeEG) return next_eg_w(W, c, p, v);
  if constexpr (kKind == kCodeGR) return next_gr_w(W, c, p, v);
}

// The Huffman lookup table goes to LDS (at most 2^kLutBits entries).
template <u32 kKind>
__device__ __forceinline__ const u32* stage_lut(const CodeDesc& c, u32* sh) {
  if constexpr (kKind != kCodeHuf) return nullptr;
  const u32 n = 1u << c.lb;
  for (u32 i = threadIdx.x; i < n; i += blockDim.x) sh[i] = c.lut[i];
  __syncthreads();
  return sh;
}

// ---- self-synchronising decode ------------------------------------------------

template <u32 kKind>
__global__ __launch_bounds__(kBlock) void k_pd_spec(const uint8_t* __restrict__ s, CodeDesc c,
                                                    u64 nsub, u64* __restrict__ start,
                                                    u64* __restrict__ exitp,
                                                    u32* __restrict__ cnt) {
  __shared__ u32 lsh[1u << kLutBits];
  const u32* lut = stage_lut<kKind>(c, lsh);
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsub) return;
  constexpr u64 kS = kKind == kCodeGR ? kSubGR : kSub;
  const u64 b = t * kS, end = b + kS < c.bits ? b + kS : c.bits;
  u64 p = b, v;
  u32 n = 0;
  WinOf<kKind> W(s, c.base_bit + p);
  while (p < end) {
    p = next_code_w<kKind>(W, c, lut, p, v);
    ++n;
  }
  start[t] = b;
  exitp[t] = p;
  cnt[t] = n;
}

// One round: sub-sequence t takes the exit of t - 1 from the previous round
// as its entry and walks the new path and its current one in step (always
// advancing the one behind) until they meet at a code-word start inside its
// range; the old path's remaining count and exit then stand.
template <u32 kKind>
__global__ __launch_bounds__(kBlock) void k_pd_sync(const uint8_t* __restrict__ s, CodeDesc c,
                                                    u64 nsub, u64* __restrict__ start,
                                                    const u64* __restrict__ exit_in,
                                                    u64* __restrict__ exit_out,
                                                    u32* __restrict__ cnt,
                                                    u32* __restrict__ changed) {
  __shared__ u32 lsh[1u << kLutBits];
  const u32* lut = stage_lut<kKind>(c, lsh);
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsub) return;
  const u64 old_exit = exit_in[t];
  if (t == 0) {
    exit_out[0] = old_exit;
    return;
  }
  const u64 entry = exit_in[t - 1], cur = start[t];
  if (entry == cur) {
    exit_out[t] = old_exit;
    return;
  }
  constexpr u64 kS = kKind == kCodeGR ? kSubGR : kSub;
  const u64 b = t * kS, end = b + kS < c.bits ? b + kS : c.bits;
  u64 p = entry, q = cur, v;
  u32 np = 0, nq = 0;
  WinOf<kKind> WP(s, c.base_bit + (p < end ? p : 0)), WQ(s, c.base_bit + (q < end ? q : 0));
  while (p != q && p < end && q < end) {
    if (p < q) {
      p = next_code_w<kKind>(WP, c, lut, p, v);
      ++np;
    } else {
      q = next_code_w<kKind>(WQ, c, lut, q, v);
      ++nq;
    }
  }
  u64 ex;
  u32 n;
  if (p == q && p < end) {  // in step: the rest of the old path stands
    ex = old_exit;
    n = np + (cnt[t] - nq);
  } else {
    while (p < end) {
      p = next_code_w<kKind>(WP, c, lut, p, v);
      ++np;
    }
    ex = p;
    n = np;
  }
  start[t] = entry;
  cnt[t] = n;
  exit_out[t] = ex;
  if (ex != old_exit) atomicOr(changed, 1u);
}

template <u32 kKind, typename T>
__global__ __launch_bounds__(kBlock) void k_pd_write(const uint8_t* __restrict__ s, CodeDesc c,
                                                     u64 nsub, const u64* __restrict__ start,
                                                     const u32* __restrict__ base,
                                                     const u32* __restrict__ cnt, u64 count,
                                                     T* __restrict__ out) {
  __shared__ u32 lsh[1u << kLutBits];
  const u32* lut = stage_lut<kKind>(c, lsh);
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsub) return;
  u64 p = start[t], o = base[t];
  const u32 n = o < count ? (u32)(cnt[t] < count - o ? cnt[t] : count - o) : 0u;
  WinOf<kKind> W(s, c.base_bit + (p < c.bits ? p : 0));
  // A thread's values are consecutive in the output: they are decoded kBatch
  // at a time into registers and stored together, so the lines a thread
  // writes are complete while still in L2 (one store per decoded value left
  // every lane's line partly written for a long time, and with ~10^6 threads
  // the partial lines were evicted and written back repeatedly).  Byte
  // outputs: the values up to the first 4-byte boundary are stored byte by
  // byte (that word is shared with the previous thread), then batches start
  // on word boundaries and store whole words, the last partial word byte by
  // byte.
  constexpr u32 kBatch = 16;
  auto decode_one = [&](u64& v) {
    p = next_code_w<kKind>(W, c, lut, p, v);
    if constexpr (kKind == kCodeHuf)
      if (!c.direct_values) v = c.values[v];
    if (c.zigzag) v = (u64)(i64)((v >> 1) ^ (~(v & 1) + 1));
  };
  u32 i0 = 0;
  if constexpr (sizeof(T) == 1) {
    const u32 head = (4u - (u32)(o & 3)) & 3u;
    for (; i0 < head && i0 < n; ++i0) {
      u64 v = 0;
      decode_one(v);
      out[o + i0] = (T)v;
    }
  }
  for (; i0 < n; i0 += kBatch) {
    const u32 m = n - i0 < kBatch ? n - i0 : kBatch;
    u64 v[kBatch];
#pragma unroll
    for (u32 k = 0; k < kBatch; ++k) {
      v[k] = 0;
      if (k < m) decode_one(v[k]);
    }
    T* dst = out + o + i0;  // 4-byte aligned for byte outputs
    if constexpr (sizeof(T) == 1) {
#pragma unroll
      for (u32 q = 0; q < kBatch / 4; ++q) {
        if (4 * q + 3 < m) {
          reinterpret_cast<u32*>(dst)[q] = ((u32)v[4 * q] & 0xFFu) |
                                           (((u32)v[4 * q + 1] & 0xFFu) << 8) |
                                           (((u32)v[4 * q + 2] & 0xFFu) << 16) |
                                           (((u32)v[4 * q + 3] & 0xFFu) << 24);
        } else {
#pragma unroll
          for (u32 b = 0; b < 4; ++b)
            if (4 * q + b < m) dst[4 * q + b] = (T)v[4 * q + b];
        }
      }
    } else {
#pragma unroll
      for (u32 k = 0; k < kBatch; ++k)
        if (k < m) dst[k] = (T)v[k];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_const(T* out, u64 n, T v) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (u64)gridDim.x * blockDim.x)
    out[i] = v;
}

template <u32 kKind, typename T>
void decode_section_t(const uint8_t* src, const CodeDesc& c, u64 count, T* out, DecScratch& w,
                      Arena& ar, hipStream_t s) {
  constexpr u64 kS = kKind == kCodeGR ? kSubGR : kSub;
  const u64 nsub = (c.bits + kS - 1) / kS;
  if (nsub >= 0xFFFFFFFFull) throw Error(-5, "entropy-coded section too long");
  const u32 g = blocks_for(nsub, kBlock);
  u64* start = ar.take<u64>(nsub);
  u64* ex[2] = {ar.take<u64>(nsub), ar.take<u64>(nsub)};
  u32* cnt = ar.take<u32>(nsub);
  u32* base = ar.take<u32>(nsub);
  u32* flag = ar.take<u32>(2);
  k_pd_spec<kKind><<<g, kBlock, 0, s>>>(src, c, nsub, start, ex[0], cnt);
  LAUNCH_CHECK("k_pd_spec");
  int cur = 0;
  // Rounds go in pairs between looks at the flag (the second round of a
  // settled pair costs a launch: every sub-sequence's entry already matches),
  // halving the host round trips; the flag reports the pair's second round.
  for (u64 round = 0;; round += 2) {
    if (round > nsub + 2) throw Error(-7, "decoder: sub-sequences did not settle");
    for (int r = 0; r < 2; ++r) {
      k_fill_const<u32><<<1, 64, 0, s>>>(flag, 1, 0u);
      LAUNCH_CHECK("k_fill_const");
      k_pd_sync<kKind><<<g, kBlock, 0, s>>>(src, c, nsub, start, ex[cur], ex[cur ^ 1], cnt,
                                            flag);
      LAUNCH_CHECK("k_pd_sync");
      cur ^= 1;
    }
    launch_copy_words(flag, w.host_flag, 4, s);  // not the copy engine: uploads may hold it
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (g_profiler) g_profiler->mark(nullptr, s);  // the next kernel starts a new interval
    if (!*w.host_flag) break;
  }
  scan_excl_u32(cnt, base, nsub, flag + 1, ar, s);
  launch_copy_words(flag + 1, w.host_flag, 4, s);
  check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (g_profiler) g_profiler->mark(nullptr, s);
  if (*w.host_flag < count) throw Error(-6, "entropy-coded section holds fewer values than its count");
  k_pd_write<kKind, T><<<g, kBlock, 0, s>>>(src, c, nsub, start, base, cnt, count, out);
  LAUNCH_CHECK("k_pd_write");
}

// ---- fixed-width sections and block starts -----------------------------------

__global__ __launch_bounds__(kBlock) void k_bitunpack(const uint8_t* __restrict__ src, u64 base_bit,
                                                      u32 w, u64 n, u64* __restrict__ out) {
  const u64 mask = w >= 64 ? ~0ull : ((1ull << w) - 1);
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (u64)gridDim.x * blockDim.x)
    out[i] = peek_lsb(src, base_bit + i * w) & mask;
}

__global__ __launch_bounds__(kBlock) void k_add_u64(u64* a, const u64* b, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (u64)gridDim.x * blockDim.x)
    a[i] += b[i];
}

__global__ __launch_bounds__(kBlock) void k_mark_heads(const u64* start, u64 n, uint8_t* flags) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (u64)gridDim.x * blockDim.x)
    flags[start[i]] = 1;
}

// ---- LCP / direct reconstruction ----------------------------------------------

// Prefix state over the coded elements.  Plain counts: nb (block starts), nz
// (patch symbols).  Segmented at block starts: q (sum of quad deltas).
// Segmented at "repos heads" (a block start or a non-zero quad delta,
// block_quantizer.hpp:131-136): a (sum of the non-patch repos deltas), zsr
// (patch symbols before the last repos head).  Direct mode uses nz and a,
// with no heads.
struct LState {
  u64 a;
  u32 nb, nz, q, zsr;
  u32 fb, fr;
};

__device__ __forceinline__ LState lcombine(const LState& x, const LState& y) {
  LState r;
  r.nb = x.nb + y.nb;
  r.nz = x.nz + y.nz;
  r.q = y.fb ? y.q : x.q + y.q;
  r.fb = x.fb | y.fb;
  r.a = y.fr ? y.a : x.a + y.a;
  r.zsr = y.fr ? x.nz + y.zsr : x.zsr;
  r.fr = x.fr | y.fr;
  return r;
}

__device__ __forceinline__ LState lzero() {
  LState r;
  r.a = 0;
  r.nb = r.nz = r.q = r.zsr = r.fb = r.fr = 0;
  return r;
}

__device__ __forceinline__ LState lelem(const LcpArgs& A, u64 j) {
  LState e = lzero();
  const u32 sym = A.sym[j];
  e.nz = sym == 0;
  e.a = sym ? sym - 1 : 0;
  if (!A.direct) {
    const u32 bh = A.head[j];
    const u32 qd = (u32)A.quad[j] - 1u;  // quads PFD symbol v + 1
    e.nb = bh;
    e.q = qd;
    e.fb = bh;
    e.fr = bh | (qd != 0);
  }
  return e;
}

// Ordered block reduction / exclusive scan of one LState per thread (LDS).
__device__ LState block_excl(LState v, LState* sh, LState& total) {
  const u32 t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (u32 o = 1; o < kBlock; o <<= 1) {
    LState x = t >= o ? lcombine(sh[t - o], sh[t]) : sh[t];
    __syncthreads();
    sh[t] = x;
    __syncthreads();
  }
  total = sh[kBlock - 1];
  LState ex = t ? sh[t - 1] : lzero();
  __syncthreads();
  return ex;
}

constexpr u32 kLItems = 16;
constexpr u32 kLTile = kBlock * kLItems;

// A thread's kLItems consecutive elements' symbols, quads and block-start
// flags, one 16-byte load per array (the arrays are padded by 16 bytes past
// the last element; bytes past it are ignored).
struct LBytes {
  u32 sym[4], head[4], quad[4];
};
static_assert(kLItems == 16, "one 16-byte load per array");

__device__ __forceinline__ LBytes load_bytes(const LcpArgs& A, u64 b0) {
  LBytes b;
  const uint4 s4 = *reinterpret_cast<const uint4*>(A.sym + b0);
  b.sym[0] = s4.x, b.sym[1] = s4.y, b.sym[2] = s4.z, b.sym[3] = s4.w;
  if (!A.direct) {
    const uint4 h4 = *reinterpret_cast<const uint4*>(A.head + b0);
    const uint4 q4 = *reinterpret_cast<const uint4*>(A.quad + b0);
    b.head[0] = h4.x, b.head[1] = h4.y, b.head[2] = h4.z, b.head[3] = h4.w;
    b.quad[0] = q4.x, b.quad[1] = q4.y, b.quad[2] = q4.z, b.quad[3] = q4.w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) b.head[k] = b.quad[k] = 0;
  }
  return b;
}

__device__ __forceinline__ u32 byte_at(const u32* w, u32 i) { return (w[i >> 2] >> (8 * (i & 3))) & 0xFFu; }

// lelem() from an item's three bytes.
__device__ __forceinline__ LState lelem_v(const LcpArgs& A, u32 sym, u32 bh, u32 qb) {
  LState e = lzero();
  e.nz = sym == 0;
  e.a = sym ? sym - 1 : 0;
  if (!A.direct) {
    const u32 qd = qb - 1u;
    e.nb = bh;
    e.q = qd;
    e.fb = bh;
    e.fr = bh | (qd != 0);
  }
  return e;
}

// Word w of a four-word array without indexing it (registers, not memory).
__device__ __forceinline__ u32 word_of(const u32* x, u32 w) {
  return w == 0 ? x[0] : w == 1 ? x[1] : w == 2 ? x[2] : x[3];
}

// lelem() of item i from the loaded bytes.
__device__ __forceinline__ LState lelem_b(const LcpArgs& A, const LBytes& b, u32 i) {
  LState e = lzero();
  const u32 sym = byte_at(b.sym, i);
  e.nz = sym == 0;
  e.a = sym ? sym - 1 : 0;
  if (!A.direct) {
    const u32 bh = byte_at(b.head, i);
    const u32 qd = byte_at(b.quad, i) - 1u;
    e.nb = bh;
    e.q = qd;
    e.fb = bh;
    e.fr = bh | (qd != 0);
  }
  return e;
}

__global__ __launch_bounds__(kBlock) void k_lcp_reduce(LcpArgs A, LState* tiles) {
  __shared__ LState sh[kBlock];
  const u64 b0 = (u64)blockIdx.x * kLTile + (u64)threadIdx.x * kLItems;
  LState acc = lzero();
  if (b0 < A.n) {
    const LBytes lb = load_bytes(A, b0);
#pragma unroll
    for (u32 i = 0; i < kLItems; ++i)
      if (b0 + i < A.n) acc = lcombine(acc, lelem_b(A, lb, i));
  }
  LState tot;
  (void)block_excl(acc, sh, tot);
  if (threadIdx.x == 0) tiles[blockIdx.x] = tot;
}

// One workgroup: exclusive scan of the tile totals, in place.
__global__ __launch_bounds__(kBlock) void k_lcp_tiles(LState* tiles, u32 nt) {
  __shared__ LState sh[kBlock];
  LState carry = lzero();
  for (u32 base = 0; base < nt; base += kBlock) {
    const u32 i = base + threadIdx.x;
    const LState v = i < nt ? tiles[i] : lzero();
    LState tot;
    const LState ex = block_excl(v, sh, tot);
    if (i < nt) tiles[i] = lcombine(carry, ex);
    carry = lcombine(carry, tot);
  }
}

__device__ __forceinline__ u32 compact3(u64 c) {  // z_order_curve.hpp:32-40
  u64 x = c & 0x1249249249249249ull;
  x = (x | (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x | (x >> 4)) & 0x100f00f00f00f00full;
  x = (x | (x >> 8)) & 0x1f0000ff0000ffull;
  x = (x | (x >> 16)) & 0x1f00000000ffffull;
  x = (x | (x >> 32)) & 0x1fffff;
  return (u32)x;
}

// hilbertToMorton3(h, L) (hilbert.cpp:396-421): digits L..0 from the top
// through the machine; digits above L stay.  Digit 21 is bit 63 alone.
__device__ __forceinline__ u64 h2m(u64 h, u32 L, const u8* apply, const u8* next) {
  u32 st = 0;
  u64 r = h;
  for (int i = (int)L; i >= 0; --i) {
    const u32 sh = 3u * (u32)i;
    const u32 hd = (u32)((h >> sh) & 7);
    const u32 md = apply[st * 8 + hd];
    const u64 mask = sh < 63 ? (7ull << sh) : (1ull << 63);
    r = (r & ~mask) | (((u64)md << sh) & mask);
    st = next[st * 8 + hd];
  }
  return r;
}

// Unsigned 64-bit division by a divisor fixed for the launch, as a multiply
// and shifts (the branch-free round-up method: magic = 1 + floor(2^(65 + l) / d)
// with l = floor(log2 d), q = ((n - hi) / 2 + hi) >> l, hi = mulhi(magic, n));
// exact for every 64-bit n.  Powers of two take magic 0 and shift l - 1; d = 1
// is flagged.
struct DivU64 {
  u64 magic;
  u32 more, one;
};

DivU64 div_u64(u64 d) {
  DivU64 r{0, 0, d == 1 ? 1u : 0u};
  if (d <= 1) return r;
  const u32 l = 63 - (u32)__builtin_clzll(d);
  if ((d & (d - 1)) == 0) {
    r.more = l - 1;
    return r;
  }
  const unsigned __int128 num = (unsigned __int128)1 << (64 + l);
  u64 pm = (u64)(num / d);
  const u64 rem = (u64)(num % d);
  pm += pm;
  const u64 twice = rem + rem;
  if (twice >= d || twice < rem) pm += 1;
  r.magic = pm + 1;
  r.more = l;
  return r;
}

__device__ __forceinline__ u64 div_by(u64 n, const DivU64& m) {
  if (m.one) return n;
  const u64 t = __umul64hi(m.magic, n);
  return (((n - t) >> 1) + t) >> m.more;
}

// Launch constants of k_lcp_apply: the block-grid divisions, and the state
// of the inverse Hilbert machine after the 15 zero digits above a block-mode
// code's 6 (with the Morton digits they produce).
struct LcpConst {
  DivU64 dx, dy;
  u64 h_hi;
  u32 h_st;
};

// The tile's points are assembled in LDS and stored with coalesced writes
// (each thread's 16 consecutive elements would otherwise store 192-byte runs
// at a 192-byte lane stride, writing partial lines); RLE tiles stage their
// cells' output offsets and counts the same way.  LDS is reused phase by
// phase: the block scan, the offsets and counts, the points.
constexpr u32 kLOut = kLTile;  // points staged per tile (the rest stored directly)
constexpr u32 kLSlots = kLTile + kLTile / 16;  // staged offsets / counts, one pad per 16
constexpr u32 kLWords = 3 * kLOut > 2 * kLSlots ? 3 * kLOut : 2 * kLSlots;
static_assert(kLWords * 4 >= kBlock * sizeof(LState), "the scan fits the staging area");

__device__ __forceinline__ u32 lslot(u32 i) { return i + (i >> 4); }

__global__ __launch_bounds__(kBlock) void k_lcp_apply(LcpArgs A, const LState* tiles,
                                                      HilbertInv hi, LcpConst K) {
  __shared__ u32 lds[kLWords];
  __shared__ u8 ap[24 * 8], nx_[24 * 8];
  for (u32 i = threadIdx.x; i < 24 * 8; i += kBlock) {
    ap[i] = hi.apply[i];
    nx_[i] = hi.next[i];
  }
  const u64 tf = (u64)blockIdx.x * kLTile;  // the tile's first element
  const u32 tc = (u32)(A.n - tf < kLTile ? A.n - tf : kLTile);
  const u64 b0 = tf + (u64)threadIdx.x * kLItems;
  LBytes lb;
  LState acc = lzero();
  if (b0 < A.n) {
    lb = load_bytes(A, b0);
#pragma unroll
    for (u32 i = 0; i < kLItems; ++i)
      if (b0 + i < A.n) acc = lcombine(acc, lelem_b(A, lb, i));
  }
  LState tot;
  LState run = lcombine(tiles[blockIdx.x], block_excl(acc, reinterpret_cast<LState*>(lds), tot));
  // Output points of the tile: [at0, at0 + span); RLE cells' offsets relative
  // to at0 and their counts (clamped to the output) staged in LDS.
  u64 at0 = tf, span = tc;
  u32 rel[kLItems], cnt[kLItems];
  if (A.rle) {
    at0 = A.out_at[tf];
    const u64 at1 = tf + tc < A.n ? A.out_at[tf + tc] : A.total;
    span = (at1 < A.total ? at1 : A.total) - (at0 < A.total ? at0 : A.total);
    u32* srel = lds;
    u32* scnt = lds + kLSlots;
    for (u32 i = threadIdx.x; i < tc; i += kBlock) {
      const u64 a = A.out_at[tf + i], c = A.counts[tf + i];
      const u64 room = a < A.total ? A.total - a : 0;
      srel[lslot(i)] = (u32)(a - at0);
      scnt[lslot(i)] = (u32)(c < room ? c : room);
    }
    __syncthreads();
#pragma unroll
    for (u32 i = 0; i < kLItems; ++i) {
      const u32 e = threadIdx.x * kLItems + i;
      rel[i] = e < tc ? srel[lslot(e)] : 0u;
      cnt[i] = e < tc ? scnt[lslot(e)] : 0u;
    }
  } else {
#pragma unroll
    for (u32 i = 0; i < kLItems; ++i) {
      rel[i] = threadIdx.x * kLItems + i;
      cnt[i] = 1;
    }
  }
  __syncthreads();  // the staging area takes the points from here on
  float* sout = reinterpret_cast<float*>(lds);
  if (b0 < A.n) {
    // Four items at a time: their prefix states, then every gathered load
    // (block id, patch sums) in flight together, then the points.
    constexpr u32 kG = 4;  // one word of each byte array
    for (u32 g = 0; g < kLItems; g += kG) {
      const u32 sw = word_of(lb.sym, g / 4), hw = word_of(lb.head, g / 4),
                qw = word_of(lb.quad, g / 4);
      LState rs[kG];
#pragma unroll
      for (u32 k = 0; k < kG; ++k) {
        if (b0 + g + k < A.n)
          run = lcombine(run, lelem_v(A, (sw >> (8 * k)) & 0xFFu, (hw >> (8 * k)) & 0xFFu,
                                      (qw >> (8 * k)) & 0xFFu));
        rs[k] = run;
      }
      u64 id[kG], p1[kG], p2[kG];
#pragma unroll
      for (u32 k = 0; k < kG; ++k) {
        const u64 j = b0 + g + k;
        const bool in = j < A.n;
        p1[k] = A.ps[rs[k].nz < A.np ? rs[k].nz : A.np];
        if (!A.direct) {
          id[k] = in && rs[k].nb ? A.blkst[rs[k].nb - 1] : 0;
          p2[k] = A.ps[rs[k].zsr < A.np ? rs[k].zsr : A.np];
        }
      }
#pragma unroll
      for (u32 k = 0; k < kG; ++k) {
        const u64 j = b0 + g + k;
        if (j >= A.n) continue;
        const LState& r = rs[k];
        u64 cx, cy, cz;
        if (A.direct) {  // block_decompressor.hpp:82-101
          const u64 sfc = r.a + p1[k];
          const u64 m = A.hilbert ? h2m(sfc, A.bits, ap, nx_) : sfc;
          cx = compact3(m);
          cy = compact3(m >> 1);
          cz = compact3(m >> 2);
        } else {  // block_quantizer.hpp:166-203
          const u64 q1 = div_by(id[k], K.dx), q2 = div_by(q1, K.dy);
          const u64 pbx = ((id[k] - q1 * A.nx) * 2) * 64, pby = ((q1 - q2 * A.ny) * 2) * 64,
                    pbz = (q2 * 2) * 64;
          const u64 rj = r.a + (p1[k] - p2[k]);
          u64 m = rj;
          if (A.hilbert) {
            if (rj < (1ull << 18)) {  // six live digits from the precomputed state
              u32 st = K.h_st;
              m = K.h_hi;
#pragma unroll
              for (int i = 5; i >= 0; --i) {
                const u32 hd = (u32)(rj >> (3 * i)) & 7u;
                m |= (u64)ap[st * 8 + hd] << (3 * i);
                st = nx_[st * 8 + hd];
              }
            } else {
              m = h2m(rj, 20, ap, nx_);
            }
          }
          cx = pbx + (u64)(r.q & 1) * 64 + compact3(m);
          cy = pby + (u64)((r.q >> 1) & 1) * 64 + compact3(m >> 1);
          cz = pbz + (u64)((r.q >> 2) & 1) * 64 + compact3(m >> 2);
        }
        int n3[3] = {(int)cx, (int)cy, (int)cz};
        float v[3];
        if (A.to) {  // truncated_octahedron_quantizer.hpp:57-76
          if (n3[2] & 1) {
            n3[0] = (n3[0] << 1) - 1;
            n3[1] = (n3[1] << 1) - 1;
          } else {
            n3[0] <<= 1;
            n3[1] <<= 1;
          }
#pragma unroll
          for (int a = 0; a < 3; ++a)
            v[a] = div_const((float(n3[a]) * 2.0f) * A.b, A.rec_div, A.rec_inv);
        } else {  // cube_quantizer.hpp:29-38
#pragma unroll
          for (int a = 0; a < 3; ++a)
            v[a] = div_const(((float(n3[a]) + 0.5f) * 2.0f) * A.b, A.rec_div, A.rec_inv);
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = v[a] + A.off[a];  // shifting.hpp:39-50
        // block_decompressor_rle.hpp:173-187: count copies, in cell order
        // (one point without RLE)
        for (u32 t = 0; t < cnt[g + k]; ++t) {
          const u32 o = rel[g + k] + t;
          if (o < kLOut) {
            sout[3 * o] = v[0];
            sout[3 * o + 1] = v[1];
            sout[3 * o + 2] = v[2];
          } else {
            float* d = A.out + 3 * (at0 + o);
            d[0] = v[0];
            d[1] = v[1];
            d[2] = v[2];
          }
        }
      }
    }
  }
  __syncthreads();
  const u32 staged = 3 * (u32)(span < kLOut ? span : kLOut);
  float* dst = A.out + 3 * at0;
  for (u32 i = threadIdx.x; i < staged; i += kBlock) dst[i] = sout[i];
}

__global__ __launch_bounds__(kBlock) void k_h2m(const u64* in, u64 n, u32 L, HilbertInv hi,
                                                u64* out) {
  __shared__ u8 ap[24 * 8], nx_[24 * 8];
  for (u32 i = threadIdx.x; i < 24 * 8; i += kBlock) {
    ap[i] = hi.apply[i];
    nx_[i] = hi.next[i];
  }
  __syncthreads();
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (u64)gridDim.x * blockDim.x)
    out[i] = h2m(in[i], L, ap, nx_);
}

}  // namespace

template <typename T>
void decode_section(const uint8_t* src, const CodeDesc& c, u64 count, T* out, DecScratch& w,
                    Arena& ar, hipStream_t s) {
  if (!count) return;
  const size_t mark = ar.off;
  if (c.kind == kCodeConst) {
    const u32 g = blocks_for(count, kBlock * 16);
    k_fill_const<T><<<g < 16384 ? g : 16384, kBlock, 0, s>>>(out, count, (T)c.const_value);
    LAUNCH_CHECK("k_fill_const");
  } else if (c.kind == kCodeHuf) {
    if (c.lb < 1 || c.lb > kLutBits) throw Error(-7, "decoder: bad lookup width");
    decode_section_t<kCodeHuf>(src, c, count, out, w, ar, s);
  } else if (c.kind == kCodeEG) {
    decode_section_t<kCodeEG>(src, c, count, out, w, ar, s);
  } else {
    decode_section_t<kCodeGR>(src, c, count, out, w, ar, s);
  }
  ar.off = mark;
}

template void decode_section<uint8_t>(const uint8_t*, const CodeDesc&, u64, uint8_t*,
                                      DecScratch&, Arena&, hipStream_t);
template void decode_section<u64>(const uint8_t*, const CodeDesc&, u64, u64*, DecScratch&,
                                  Arena&, hipStream_t);

void launch_bitunpack(const uint8_t* stream, u64 base_bit, u32 w, u64 n, u64* out,
                      hipStream_t s) {
  if (!n) return;
  const u32 g = blocks_for(n, kBlock * 4);
  k_bitunpack<<<g < 16384 ? g : 16384, kBlock, 0, s>>>(stream, base_bit, w, n, out);
  LAUNCH_CHECK("k_bitunpack");
}

void launch_add_u64(u64* a, const u64* b, u64 n, hipStream_t s) {
  if (!n) return;
  const u32 g = blocks_for(n, kBlock * 4);
  k_add_u64<<<g < 16384 ? g : 16384, kBlock, 0, s>>>(a, b, n);
  LAUNCH_CHECK("k_add_u64");
}

void launch_mark_heads(const u64* start, u64 n, uint8_t* flags, hipStream_t s) {
  if (!n) return;
  const u32 g = blocks_for(n, kBlock * 4);
  k_mark_heads<<<g < 16384 ? g : 16384, kBlock, 0, s>>>(start, n, flags);
  LAUNCH_CHECK("k_mark_heads");
}

void launch_zero_u8(uint8_t* p, u64 n, hipStream_t s) {
  if (!n) return;
  const u32 g = blocks_for(n, kBlock * 16);
  k_fill_const<uint8_t><<<g < 16384 ? g : 16384, kBlock, 0, s>>>(p, n, (uint8_t)0);
  LAUNCH_CHECK("k_fill_const");
}

void launch_h2m(const u64* in, u64 n, u32 iteration, const HilbertInv& hi, u64* out,
                hipStream_t s) {
  if (!n) return;
  const u32 g = blocks_for(n, kBlock * 4);
  k_h2m<<<g < 16384 ? g : 16384, kBlock, 0, s>>>(in, n, iteration, hi, out);
  LAUNCH_CHECK("k_h2m");
}

size_t lcp_scratch_bytes(u64 n) { return (size_t)blocks_for(n, kLTile) * sizeof(LState) + 256; }

void launch_lcp_scan(const LcpArgs& a, void* scratch, hipStream_t s) {
  if (!a.n) return;
  const u32 nt = blocks_for(a.n, kLTile);
  LState* tiles = static_cast<LState*>(scratch);
  k_lcp_reduce<<<nt, kBlock, 0, s>>>(a, tiles);
  LAUNCH_CHECK("k_lcp_reduce");
  k_lcp_tiles<<<1, kBlock, 0, s>>>(tiles, nt);
  LAUNCH_CHECK("k_lcp_tiles");
}

void launch_lcp_apply(const LcpArgs& a, const HilbertInv& hi, void* scratch, hipStream_t s) {
  if (!a.n) return;
  const u32 nt = blocks_for(a.n, kLTile);
  LState* tiles = static_cast<LState*>(scratch);
  LcpConst K{div_u64(a.nx), div_u64(a.ny), 0, 0};
  for (int i = 20; i >= 6; --i) {  // hilbertToMorton3's loop over zero digits
    K.h_hi |= (u64)hi.apply[K.h_st * 8] << (3 * i);
    K.h_st = hi.next[K.h_st * 8];
  }
  k_lcp_apply<<<nt, kBlock, 0, s>>>(a, tiles, hi, K);
  LAUNCH_CHECK("k_lcp_apply");
}

}  // namespace xnyz
