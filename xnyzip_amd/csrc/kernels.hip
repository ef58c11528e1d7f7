// kernels.hip -- gfx950 kernels of the compression path.
//
// Stage map (reference file:line each stage reproduces):
//   k_bbox          shift_points minimum       preprocessor/shifting.hpp:11-36
//   k_quant_stats   quantize + range scan      quantizer/truncated_octahedron_quantizer.hpp:21-55,
//                   + RLE diagnostic pass      quantizer/cube_quantizer.hpp:15-27,
//                                              composer/block_compressor_rle.hpp:53-87,133-141
//   k_keygen        block id + Morton/Hilbert  quantizer/block_quantizer.hpp:91-115,
//                   (direct: global SFC)       composer/block_compressor.hpp:90-101
//   k_radix_*       LSD radix sort             quantizer/block_quantizer.hpp:33-76
//   runs_*          RLE cells / block runs     block_compressor_rle.hpp:107-120, block_quantizer.hpp:117-149
//   k_stats_*       PFD symbols, patches       encoder/p_for_delta_encoder.hpp:27-43
//   k_vl_*          Golomb-Rice / Huffman /    encoder/golomb_rice_encoder.hpp:92-148,
//                   Exp-Golomb bit packing     encoder/huffman_encoder.hpp:150-170,365-396
//   k_pack_fixed    BitPacker                  encoder/bitpacking_encoder.hpp:17-51
//
// Floating point: no contraction (this pragma and -ffp-contract=off), IEEE
// division, roundf = round half away from zero -- the reference's x86-64
// baseline arithmetic.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "common.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace xnyz {

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(-2, std::string(what) + ": " + hipGetErrorString(e));
}

thread_local Profiler* g_profiler = nullptr;

namespace {

inline u32 div_up(u64 a, u64 b) { return (u32)((a + b - 1) / b); }
inline u32 grid_for(u64 n, u32 per_block, u32 cap = 65535u) {
  u64 g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (u32)(g < cap ? g : cap);
}
#define LAUNCH_CHECK(name)                      \
  do {                                          \
    check_hip(hipGetLastError(), name);         \
    if (g_profiler) g_profiler->mark(name, s);  \
  } while (0)

struct F3 {
  float x, y, z;
};

// ---------------------------------------------------------------------------
// control block

__global__ void k_init_ctl(DevCtl* c) {
  u32* w = reinterpret_cast<u32*>(c);
  const u32 nw = sizeof(DevCtl) / 4;
  const u32 m0 = offsetof(DevCtl, minkey) / 4, m1 = m0 + 6;
  const u32 f0 = offsetof(DevCtl, first_sym) / 4, f1 = f0 + kSymBins;
  const u32 g0 = offsetof(DevCtl, first_quad) / 4, g1 = g0 + kSymBins;
  const u32 h0 = offsetof(DevCtl, first_cnt) / 4, h1 = h0 + kCntBins;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += gridDim.x * blockDim.x) {
    bool ones = (i >= m0 && i < m1) || (i >= f0 && i < f1) || (i >= g0 && i < g1) ||
                (i >= h0 && i < h1);
    w[i] = ones ? 0xFFFFFFFFu : 0u;
  }
}

// (point index << 32 | quad << 18 | sfc) -> separate arrays
__global__ void k_split_lo64(const u64* __restrict__ lo64, u32* __restrict__ lo,
                             u32* __restrict__ idx, u32 n) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const u64 v = lo64[i];
    lo[i] = (u32)v;
    idx[i] = (u32)(v >> 32);
  }
}

__global__ void k_fill_u32(u32* p, u64 count, u32 value) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (u64)gridDim.x * blockDim.x)
    p[i] = value;
}

// ---------------------------------------------------------------------------
// shift_points minimum (shifting.hpp:11-36): per axis the first point, in
// index order, holding the smallest value under '<'.  Values that are not
// '< FLT_MAX' (NaN, +inf, FLT_MAX) never replace the FLT_MAX start; +0 and
// -0 compare equal, so the earlier one wins.  Key = (order-preserving map of
// the value) << 32 | index, reduced with a 64-bit min.

__device__ __forceinline__ u64 min_key(float v, u32 i) {
  if (!(v < FLT_MAX)) return ~0ull;
  u32 u = __float_as_uint(v == 0.0f ? 0.0f : v);
  u32 o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((u64)o << 32) | i;
}

__global__ __launch_bounds__(kBlock) void k_bbox(const F3* __restrict__ p, u32 n, DevCtl* ctl) {
  u64 bx = ~0ull, by = ~0ull, bz = ~0ull;
  for (u32 i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    F3 q = p[i];
    u64 kx = min_key(q.x, i), ky = min_key(q.y, i), kz = min_key(q.z, i);
    bx = kx < bx ? kx : bx;
    by = ky < by ? ky : by;
    bz = kz < bz ? kz : bz;
  }
  bx = wave_min(bx);
  by = wave_min(by);
  bz = wave_min(bz);
  if (lane_id() == 0) {
    if (bx != ~0ull) atomicMin(&ctl->minkey[0], bx);
    if (by != ~0ull) atomicMin(&ctl->minkey[1], by);
    if (bz != ~0ull) atomicMin(&ctl->minkey[2], bz);
  }
}

// The offset is the winner's exact float (keeps the sign of a zero).
__global__ void k_offsets(const float* xyz, DevCtl* ctl) {
  const u32 a = threadIdx.x;
  if (a < 3) {
    const u64 k = ctl->minkey[a];
    ctl->offset[a] = (k == ~0ull) ? FLT_MAX : xyz[3ull * (u32)k + a];
  }
}

__device__ __forceinline__ float offset_of(const float*, const DevCtl* ctl, int a) {
  return ctl->offset[a];
}

__device__ __forceinline__ u32 dyn_n(u32 n, const u32* nd) { return nd ? *nd : n; }

// ---------------------------------------------------------------------------
// Quantizers.  truncated_octahedron_quantizer.hpp:29-50: q = s / div;
// E = round(q*0.5)*2, O = round((q-1)*0.5)*2+1; d = x^2 + (y^2 + z^2); E wins
// ties; then the x/y halving.  cube_quantizer.hpp:20-21: floor(s / div).

__device__ __forceinline__ void to_node(float qx, float qy, float qz, int& X, int& Y, int& Z) {
  const float ex = roundf(qx * 0.5f) * 2.0f, ey = roundf(qy * 0.5f) * 2.0f,
              ez = roundf(qz * 0.5f) * 2.0f;
  const float ox = roundf((qx - 1.0f) * 0.5f) * 2.0f + 1.0f;
  const float oy = roundf((qy - 1.0f) * 0.5f) * 2.0f + 1.0f;
  const float oz = roundf((qz - 1.0f) * 0.5f) * 2.0f + 1.0f;
  const float a = qx - ex, b = qy - ey, c = qz - ez;
  const float d = qx - ox, e = qy - oy, f = qz - oz;
  const float dE = __fadd_rn(__fmul_rn(a, a), __fadd_rn(__fmul_rn(b, b), __fmul_rn(c, c)));
  const float dO = __fadd_rn(__fmul_rn(d, d), __fadd_rn(__fmul_rn(e, e), __fmul_rn(f, f)));
  if (dE <= dO) {
    X = (int)ex;
    Y = (int)ey;
    Z = (int)ez;
  } else {
    X = (int)ox;
    Y = (int)oy;
    Z = (int)oz;
  }
  if (X & 1) {
    X = (X + 1) >> 1;
    Y = (Y + 1) >> 1;
  } else {
    X >>= 1;
    Y >>= 1;
  }
}

template <bool kTO>
__device__ __forceinline__ void quant_point(const F3& p, float ox, float oy, float oz,
                                            const QuantParams& qp, float& sx, float& sy,
                                            float& sz, int& X, int& Y, int& Z) {
  sx = p.x - ox;
  sy = p.y - oy;
  sz = p.z - oz;
  const float qx = sx / qp.div, qy = sy / qp.div, qz = sz / qp.div;
  if (kTO) {
    to_node(qx, qy, qz, X, Y, Z);
  } else {
    X = (int)floorf(qx);
    Y = (int)floorf(qy);
    Z = (int)floorf(qz);
  }
}

// quantizer recover (truncated_octahedron_quantizer.hpp:57-76,
// cube_quantizer.hpp:29-38) for the RLE diagnostic pass.
template <bool kTO>
__device__ __forceinline__ void recover_node(int X, int Y, int Z, const QuantParams& qp,
                                             float& rx, float& ry, float& rz) {
  if (kTO) {
    if (Z & 1) {
      X = (X << 1) - 1;
      Y = (Y << 1) - 1;
    } else {
      X <<= 1;
      Y <<= 1;
    }
    rx = ((float)X * 2.0f * qp.b) / qp.rec_div;
    ry = ((float)Y * 2.0f * qp.b) / qp.rec_div;
    rz = ((float)Z * 2.0f * qp.b) / qp.rec_div;
  } else {
    rx = (((float)X + 0.5f) * 2.0f * qp.b) / qp.rec_div;
    ry = (((float)Y + 0.5f) * 2.0f * qp.b) / qp.rec_div;
    rz = (((float)Z + 0.5f) * 2.0f * qp.b) / qp.rec_div;
  }
}

template <bool kTO, bool kDiag>
__global__ __launch_bounds__(kBlock) void k_quant_stats(const F3* __restrict__ p, u32 n,
                                                        QuantParams qp, DevCtl* ctl) {
  const float* xyz = reinterpret_cast<const float*>(p);
  const float ox = offset_of(xyz, ctl, 0), oy = offset_of(xyz, ctl, 1),
              oz = offset_of(xyz, ctl, 2);
  int mx = 0, my = 0, mz = 0;
  u32 wr = 0;
  double se = 0.0, me = 0.0;
  for (u32 i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    float sx, sy, sz;
    int X, Y, Z;
    quant_point<kTO>(p[i], ox, oy, oz, qp, sx, sy, sz, X, Y, Z);
    mx = X > mx ? X : mx;
    my = Y > my ? Y : my;
    mz = Z > mz ? Z : mz;
    wr += (Z == -1);
    if (kDiag) {
      float rx, ry, rz;
      recover_node<kTO>(X, Y, Z, qp, rx, ry, rz);
      const float a = sx - rx, b = sy - ry, c = sz - rz;
      // block_compressor_rle.hpp:60-62: float products, left-to-right sum
      const float m = __fadd_rn(__fadd_rn(__fmul_rn(a, a), __fmul_rn(b, b)), __fmul_rn(c, c));
      const double md = (double)m;
      se += md;
      me = md > me ? md : me;
    }
  }
  mx = wave_max(mx);
  my = wave_max(my);
  mz = wave_max(mz);
  wr = wave_sum(wr);
  if (kDiag) {
    se = wave_sum(se);
    me = wave_max(me);
  }
  if (lane_id() == 0) {
    atomicMax(&ctl->range[0], mx);
    atomicMax(&ctl->range[1], my);
    atomicMax(&ctl->range[2], mz);
    if (wr) atomicAdd(&ctl->wrapped, wr);
    if (kDiag) {
      atomicAdd(&ctl->sum_err2, se);
      atomicMax(&ctl->max_err2_bits, (u64)__double_as_longlong(me));
    }
  }
}

template <bool kTO>
__global__ __launch_bounds__(kBlock) void k_quantize_nodes(const F3* __restrict__ p, u32 n,
                                                           QuantParams qp, const DevCtl* ctl,
                                                           int32_t* nodes) {
  const float* xyz = reinterpret_cast<const float*>(p);
  const float ox = offset_of(xyz, ctl, 0), oy = offset_of(xyz, ctl, 1),
              oz = offset_of(xyz, ctl, 2);
  for (u32 i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    float sx, sy, sz;
    int X, Y, Z;
    quant_point<kTO>(p[i], ox, oy, oz, qp, sx, sy, sz, X, Y, Z);
    nodes[3ull * i] = X;
    nodes[3ull * i + 1] = Y;
    nodes[3ull * i + 2] = Z;
  }
}

// ---------------------------------------------------------------------------
// Space-filling curves.  Morton: z_order_curve.hpp:14-29 (21 bits per axis,
// x at 3i, y at 3i+1, z at 3i+2).  Hilbert: hilbert.cpp:423-448 run as a
// state machine (tables from the host, see HilbertSM).

__device__ __forceinline__ u64 split3_21(u32 v) {
  u64 x = v & 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__device__ __forceinline__ u32 split3_6(u32 x) {  // 6-bit coordinate
  x = (x | (x << 8)) & 0x0000F00Fu;
  x = (x | (x << 4)) & 0x000C30C3u;
  x = (x | (x << 2)) & 0x00249249u;
  return x;
}

struct SMTables {
  u8 apply[24 * 8];
  u8 next[24 * 8];
  u8 out[8];
};

__device__ __forceinline__ void load_sm(SMTables& t, const HilbertSM& hs) {
  for (u32 i = threadIdx.x; i < 24 * 8; i += blockDim.x) {
    t.apply[i] = hs.apply[i];
    t.next[i] = hs.next[i];
  }
  if (threadIdx.x < 8) t.out[threadIdx.x] = hs.out[threadIdx.x];
}

// mortonToHilbert3(m, 20) for m < 2^18: digits 20..6 are zero, stay zero and
// leave the state at `init` (composition of 15 copies of the digit-0
// permutation); digits 5..0 go through the machine.
__device__ __forceinline__ u32 hilbert18(u32 m, const SMTables& t, u32 init) {
  u32 s = init, r = 0;
#pragma unroll
  for (int i = 5; i >= 1; --i) {
    const u32 md = t.apply[s * 8 + ((m >> (3 * i)) & 7)];
    r |= (u32)t.out[md] << (3 * i);
    s = t.next[s * 8 + md];
  }
  return r | t.out[t.apply[s * 8 + (m & 7)]];
}

// mortonToHilbert3(m, L) for any 64-bit m and L <= 21: digits above L are
// left as they are; shifts of digit 21 keep only bit 63, as the reference's
// u64 arithmetic does.
__device__ __forceinline__ u64 hilbert_general(u64 m, u32 L, const SMTables& t, u32 identity) {
  u32 s = identity;
  u64 r = 0;
  for (u32 i = L; i >= 1; --i) {
    const u32 sh = 3 * i;
    const u32 md = t.apply[s * 8 + (u32)((m >> sh) & 7)];
    r |= (u64)t.out[md] << sh;
    s = t.next[s * 8 + md];
  }
  r |= t.out[t.apply[s * 8 + (u32)(m & 7)]];
  const u32 top = 3 * (L + 1);
  if (top < 64) r |= m & ~((1ull << top) - 1);
  return r;
}

// Sort keys.  Block mode (block_quantizer.hpp:97-115): the node's int
// coordinates are taken as size_t, c = v / 64, d = v % 64, id = cx/2 +
// (cy/2)*nx + (cz/2)*nx*ny, reid = SFC(d) | parity bits << 60.  The key packs
// (seg, id, quad, sfc): seg marks nodes with z == -1, whose id wraps to
// C + cx/2 + (cy/2)*nx, above every other id, so they sort last and are
// stored by their offset from C.  Direct mode (block_compressor.hpp:97-101):
// the 63-bit Morton code of the node, or its Hilbert index.
template <bool kTO, bool kDirect, bool kHilbert, bool kWide>
__global__ __launch_bounds__(kBlock) void k_keygen(const F3* __restrict__ p, u32 n,
                                                   QuantParams qp, KeyParams kp, HilbertSM hs,
                                                   const DevCtl* ctl, u64* __restrict__ key,
                                                   u32* __restrict__ lo, u32* __restrict__ idx,
                                                   u64* __restrict__ lo64) {
  __shared__ SMTables t;
  if (kHilbert) {
    load_sm(t, hs);
    __syncthreads();
  }
  const float* xyz = reinterpret_cast<const float*>(p);
  const float ox = offset_of(xyz, ctl, 0), oy = offset_of(xyz, ctl, 1),
              oz = offset_of(xyz, ctl, 2);
  for (u32 i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    float sx, sy, sz;
    int X, Y, Z;
    quant_point<kTO>(p[i], ox, oy, oz, qp, sx, sy, sz, X, Y, Z);
    if (idx) idx[i] = i;  // point index payload (permutation output)
    if (kDirect) {
      u64 m = split3_21((u32)X) | (split3_21((u32)Y) << 1) | (split3_21((u32)Z) << 2);
      key[i] = kHilbert ? hilbert_general(m, kp.L, t, hs.identity) : m;
    } else {
      const u64 ux = (u64)(i64)X, uy = (u64)(i64)Y, uz = (u64)(i64)Z;
      const u64 cx = ux >> 6, cy = uy >> 6, cz = uz >> 6;
      const u32 m = split3_6((u32)(ux & 63)) | (split3_6((u32)(uy & 63)) << 1) |
                    (split3_6((u32)(uz & 63)) << 2);
      const u32 sfc = kHilbert ? hilbert18(m, t, hs.block_init) : m;
      const u32 quad = (u32)((cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2));
      const u64 base = (cx >> 1) + (cy >> 1) * kp.nx;
      const bool seg = (Z == -1);
      const u64 v = seg ? base : base + (cz >> 1) * kp.nxny;
      const u64 hi = ((u64)seg << kp.idb) | v;
      const u32 k21 = (quad << 18) | sfc;
      if (kWide) {
        key[i] = hi;
        if (lo64) lo64[i] = ((u64)i << 32) | k21;  // wide keys + permutation
        else lo[i] = k21;
      } else {
        key[i] = (hi << 21) | k21;
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_m2h(const u64* in, u32 n, u32 L, HilbertSM hs,
                                                u64* out) {
  __shared__ SMTables t;
  load_sm(t, hs);
  __syncthreads();
  for (u32 i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    out[i] = hilbert_general(in[i], L, t, hs.identity);
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan (reduce, scan of partials, apply).

constexpr u32 kScanItems = 16;
constexpr u32 kScanTile = kBlock * kScanItems;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const T* in, u64 n, T* partial) {
  __shared__ T sh[kWavesPerBlock];
  const u64 base = (u64)blockIdx.x * kScanTile;
  T s = 0;
#pragma unroll
  for (u32 it = 0; it < kScanItems; ++it) {
    u64 i = base + it * kBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  T t = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// One workgroup scans the partials in place (exclusive) and writes the total.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_partials(T* part, u32 np, T* total) {
  __shared__ T sh[kWavesPerBlock];
  T carry = 0;
  for (u32 base = 0; base < np; base += kBlock * 4) {
    T v[4];
    T s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      u32 i = base + threadIdx.x * 4 + k;
      v[k] = i < np ? part[i] : 0;
      s += v[k];
    }
    T tot;
    T ex = block_excl_scan(s, sh, tot);
    T run = carry + ex;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      u32 i = base + threadIdx.x * 4 + k;
      if (i < np) part[i] = run;
      run += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_apply(const T* in, T* out, u64 n,
                                                       const T* part) {
  __shared__ T sh[kWavesPerBlock];
  const u64 base = (u64)blockIdx.x * kScanTile + (u64)threadIdx.x * kScanItems;
  T v[kScanItems];
  T s = 0;
#pragma unroll
  for (u32 k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  T tot;
  T run = block_excl_scan(s, sh, tot) + part[blockIdx.x];
#pragma unroll
  for (u32 k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

template <typename T>
void scan_excl(const T* in, T* out, u64 n, T* total, Arena& ar, hipStream_t s) {
  const u32 np = div_up(n ? n : 1, kScanTile);
  T* part = ar.take<T>(np);
  k_scan_reduce<T><<<np, kBlock, 0, s>>>(in, n, part);
  LAUNCH_CHECK("k_scan_reduce");
  k_scan_partials<T><<<1, kBlock, 0, s>>>(part, np, total);
  LAUNCH_CHECK("k_scan_partials");
  k_scan_apply<T><<<np, kBlock, 0, s>>>(in, out, n, part);
  LAUNCH_CHECK("k_scan_apply");
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits.  Per pass: per-tile digit counts
// (k_radix_hist), an exclusive scan of the digit-major count matrix, then a
// stable scatter (k_radix_scatter): each wavefront ranks its 64-key slots in
// order with ballot-based digit matching and running per-wave counters in
// LDS, so equal digits keep their input order.

constexpr u32 kSortItems = 16;
constexpr u32 kSortTile = kBlock * kSortItems;
constexpr u32 kWaveSub = kWave * kSortItems;  // keys per wavefront

template <typename K>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const K* __restrict__ keys, u32 n,
                                                       u32 shift, u32 ntiles,
                                                       u32* __restrict__ counts) {
  __shared__ u32 h[kWavesPerBlock][256];
  for (u32 i = threadIdx.x; i < kWavesPerBlock * 256; i += kBlock) (&h[0][0])[i] = 0;
  __syncthreads();
  const u32 w = threadIdx.x / kWave;
  const u64 base = (u64)blockIdx.x * kSortTile;
#pragma unroll 4
  for (u32 it = 0; it < kSortItems; ++it) {
    u64 i = base + it * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[w][(u32)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  const u32 d = threadIdx.x;
  u32 c = 0;
#pragma unroll
  for (int k = 0; k < kWavesPerBlock; ++k) c += h[k][d];
  counts[(u64)d * ntiles + blockIdx.x] = c;
}

template <typename K, typename V, bool kHasV>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const K* __restrict__ kin,
                                                          K* __restrict__ kout,
                                                          const V* __restrict__ vin,
                                                          V* __restrict__ vout, u32 n,
                                                          u32 shift, u32 ntiles,
                                                          const u32* __restrict__ offs) {
  __shared__ u32 wc[kWavesPerBlock][256];
  __shared__ u32 wb[kWavesPerBlock][256];
  for (u32 i = threadIdx.x; i < kWavesPerBlock * 256; i += kBlock) (&wc[0][0])[i] = 0;
  __syncthreads();
  const u32 w = threadIdx.x / kWave, l = lane_id();
  const u64 lt = (1ull << l) - 1;
  const u64 wstart = (u64)blockIdx.x * kSortTile + (u64)w * kWaveSub;
  K k[kSortItems];
  V v[kSortItems];
  u32 rk[kSortItems];
  u32 dg[kSortItems];
#pragma unroll
  for (u32 s = 0; s < kSortItems; ++s) {
    const u64 i = wstart + s * kWave + l;
    const bool valid = i < n;
    k[s] = valid ? kin[i] : (K)0;
    if (kHasV) v[s] = valid ? vin[i] : (V)0;
    const u32 d = (u32)(k[s] >> shift) & 255u;
    u64 peers = __ballot(valid);
#pragma unroll
    for (u32 b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const u64 bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const u32 r = (u32)__popcll(peers & lt);
    const u32 prev = wc[w][d];
    if (valid && r == 0) wc[w][d] = prev + (u32)__popcll(peers);
    rk[s] = prev + r;
    dg[s] = d;
  }
  __syncthreads();
  {
    const u32 d = threadIdx.x;
    u32 acc = offs[(u64)d * ntiles + blockIdx.x];
#pragma unroll
    for (int ww = 0; ww < kWavesPerBlock; ++ww) {
      wb[ww][d] = acc;
      acc += wc[ww][d];
    }
  }
  __syncthreads();
#pragma unroll
  for (u32 s = 0; s < kSortItems; ++s) {
    const u64 i = wstart + s * kWave + l;
    if (i < n) {
      const u32 dst = wb[w][dg[s]] + rk[s];
      kout[dst] = k[s];
      if (kHasV) vout[dst] = v[s];
    }
  }
}

template <typename K, typename V, bool kHasV>
void radix_pass(const K* kin, K* kout, const V* vin, V* vout, u32 n, u32 shift, Arena& ar,
                hipStream_t s) {
  const u32 ntiles = div_up(n, kSortTile);
  const size_t mark = ar.off;
  u32* counts = ar.take<u32>((u64)256 * ntiles);
  k_radix_hist<K><<<ntiles, kBlock, 0, s>>>(kin, n, shift, ntiles, counts);
  LAUNCH_CHECK("k_radix_hist");
  scan_excl<u32>(counts, counts, (u64)256 * ntiles, nullptr, ar, s);
  k_radix_scatter<K, V, kHasV><<<ntiles, kBlock, 0, s>>>(kin, kout, vin, vout, n, shift, ntiles,
                                                          counts);
  LAUNCH_CHECK("k_radix_scatter");
  ar.off = mark;
}

// ---------------------------------------------------------------------------
// Run detection on sorted keys: flag functors and stream compaction.

struct FlagUniqueNarrow {
  const u64* a;
  __device__ bool operator()(u32 j) const { return j == 0 || a[j] != a[j - 1]; }
};
struct FlagUniqueWide {
  const u64* a;
  const u32* b;
  __device__ bool operator()(u32 j) const {
    return j == 0 || a[j] != a[j - 1] || b[j] != b[j - 1];
  }
};
struct FlagBlockNarrow {
  const u64* a;
  __device__ bool operator()(u32 j) const { return j == 0 || (a[j] >> 21) != (a[j - 1] >> 21); }
};
struct FlagBlockWide {
  const u64* a;
  __device__ bool operator()(u32 j) const { return j == 0 || a[j] != a[j - 1]; }
};

template <class F>
__global__ __launch_bounds__(kBlock) void k_flag_count(F f, u32 n_bound, const u32* n_dev,
                                                       u32* partial) {
  __shared__ u32 sh[kWavesPerBlock];
  const u32 n = dyn_n(n_bound, n_dev);
  const u64 base = (u64)blockIdx.x * kScanTile;
  u32 c = 0;
#pragma unroll 4
  for (u32 it = 0; it < kScanItems; ++it) {
    u64 j = base + it * kBlock + threadIdx.x;
    if (j < n) c += f((u32)j) ? 1u : 0u;
  }
  u32 t = block_sum(c, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_flag_scatter(F f, u32 n_bound, const u32* n_dev,
                                                         const u32* part, u32* pos) {
  __shared__ u32 sh[kWavesPerBlock];
  const u32 n = dyn_n(n_bound, n_dev);
  const u64 base = (u64)blockIdx.x * kScanTile + (u64)threadIdx.x * kScanItems;
  u32 bits = 0, c = 0;
#pragma unroll
  for (u32 k = 0; k < kScanItems; ++k) {
    const u64 j = base + k;
    const bool fl = j < n && f((u32)j);
    bits |= (u32)fl << k;
    c += fl;
  }
  u32 tot;
  u32 run = block_excl_scan(c, sh, tot) + part[blockIdx.x];
#pragma unroll
  for (u32 k = 0; k < kScanItems; ++k)
    if ((bits >> k) & 1u) pos[run++] = (u32)(base + k);
}

template <class F>
void compact(F f, u32 n, const u32* n_dev, u32* pos, u32* count, Arena& ar, hipStream_t s) {
  const u32 np = div_up(n ? n : 1, kScanTile);
  const size_t mark = ar.off;
  u32* part = ar.take<u32>(np);
  k_flag_count<F><<<np, kBlock, 0, s>>>(f, n, n_dev, part);
  LAUNCH_CHECK("k_flag_count");
  k_scan_partials<u32><<<1, kBlock, 0, s>>>(part, np, count);
  LAUNCH_CHECK("k_scan_partials");
  k_flag_scatter<F><<<np, kBlock, 0, s>>>(f, n, n_dev, part, pos);
  LAUNCH_CHECK("k_flag_scatter");
  ar.off = mark;
}

// ---------------------------------------------------------------------------
// RLE cells (block_compressor_rle.hpp:107-120 deduplicates with a hash map;
// the stream only depends on the sorted cells and their counts).

template <bool kWide>
__global__ __launch_bounds__(kBlock) void k_rle_finish(const u64* __restrict__ ka,
                                                       const u32* __restrict__ kb,
                                                       const u32* __restrict__ upos,
                                                       const u32* U_dev, u32 N,
                                                       u64* __restrict__ ua,
                                                       u32* __restrict__ ub,
                                                       u32* __restrict__ cnt, DevCtl* ctl,
                                                       u32* large, u32 large_cap) {
  __shared__ u32 hs[256], fs[256];
  const u32 U = *U_dev;
  hs[threadIdx.x] = 0;
  fs[threadIdx.x] = 0xFFFFFFFFu;
  __syncthreads();
  for (u32 u = blockIdx.x * kBlock + threadIdx.x; u < U; u += gridDim.x * kBlock) {
    const u32 st = upos[u];
    const u32 en = (u + 1 < U) ? upos[u + 1] : N;
    const u32 c = en - st;
    ua[u] = ka[st];
    if (kWide) ub[u] = kb[st];
    cnt[u] = c;
    if (c < 256) {
      atomicAdd(&hs[c], 1u);
      if (u < fs[c]) atomicMin(&fs[c], u);
    } else if (c < kCntBins) {
      atomicAdd(&ctl->hist_cnt[c], 1u);
      atomicMin(&ctl->first_cnt[c], u);
    } else {
      const u32 slot = atomicAdd(&ctl->n_large, 1u);
      if (slot < large_cap) {
        large[2 * slot] = c;
        large[2 * slot + 1] = u;
      }
    }
  }
  __syncthreads();
  const u32 d = threadIdx.x;
  if (hs[d]) {
    atomicAdd(&ctl->hist_cnt[d], hs[d]);
    atomicMin(&ctl->first_cnt[d], fs[d]);
  }
}

// ---------------------------------------------------------------------------
// Block-mode streams (block_quantizer.hpp:117-149): with the keys sorted by
// (id, quad, sfc), element j's values depend only on keys j-1 and j.

template <bool kWide>
struct KeysT {
  KeyView v;
  __device__ __forceinline__ u64 hi(u32 j) const { return kWide ? v.a[j] : (v.a[j] >> 21); }
  __device__ __forceinline__ u32 lo(u32 j) const {
    return kWide ? v.b[j] : (u32)(v.a[j] & 0x1FFFFFu);
  }
  __device__ __forceinline__ u64 id_of(u64 h) const {
    const u64 seg = h >> v.idb, val = h & ((1ull << v.idb) - 1);
    return seg ? v.C + val : val;
  }
  // repos[j] (reid - prevRe with the resets) and quads[j] (quad - prevQuad)
  __device__ __forceinline__ void lcp(u32 j, u64& repos, u32& qd) const {
    const u32 l = lo(j), quad = l >> 18, sfc = l & 0x3FFFFu;
    if (j == 0 || hi(j - 1) != hi(j)) {
      repos = sfc;
      qd = quad;
      return;
    }
    const u32 pl = lo(j - 1), pq = pl >> 18;
    qd = (quad - pq) & 0xFFu;
    repos = (pq == quad) ? (u64)sfc - (u64)(pl & 0x3FFFFu) : (u64)sfc;
  }
};

template <bool kWide>
__global__ __launch_bounds__(kBlock) void k_stats_block(KeysT<kWide> K, u32 n_bound,
                                                        const u32* n_dev, DevCtl* ctl) {
  __shared__ u32 hs[256], fs[256], hq[256], fq[256];
  const u32 n = dyn_n(n_bound, n_dev);
  hs[threadIdx.x] = 0;
  hq[threadIdx.x] = 0;
  fs[threadIdx.x] = 0xFFFFFFFFu;
  fq[threadIdx.x] = 0xFFFFFFFFu;
  __syncthreads();
  u32 np = 0, nqp = 0;
  u64 ps = 0;
  for (u32 j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
    u64 repos;
    u32 qd;
    K.lcp(j, repos, qd);
    const u32 sym = repos > 254 ? 0u : (u32)repos + 1u;
    atomicAdd(&hs[sym], 1u);
    if (j < fs[sym]) atomicMin(&fs[sym], j);
    if (repos > 254) {
      ++np;
      ps += repos;
    }
    const u32 qs = qd > 254 ? 0u : qd + 1u;
    nqp += qd > 254;
    atomicAdd(&hq[qs], 1u);
    if (j < fq[qs]) atomicMin(&fq[qs], j);
  }
  np = wave_sum(np);
  nqp = wave_sum(nqp);
  ps = wave_sum(ps);
  if (lane_id() == 0) {
    if (np) atomicAdd(&ctl->n_patches, np);
    if (nqp) atomicAdd(&ctl->n_qpatches, nqp);
    if (ps) atomicAdd(&ctl->patch_sum, ps);
  }
  __syncthreads();
  const u32 d = threadIdx.x;
  if (hs[d]) {
    atomicAdd(&ctl->hist_sym[d], hs[d]);
    atomicMin(&ctl->first_sym[d], fs[d]);
  }
  if (hq[d]) {
    atomicAdd(&ctl->hist_quad[d], hq[d]);
    atomicMin(&ctl->first_quad[d], fq[d]);
  }
}

// Direct mode (block_compressor.hpp:110-117): deltas of the sorted SFC codes,
// the first one absolute.
__global__ __launch_bounds__(kBlock) void k_stats_direct(const u64* __restrict__ k,
                                                        u32 n_bound, const u32* n_dev,
                                                        DevCtl* ctl) {
  __shared__ u32 hs[256], fs[256];
  const u32 n = dyn_n(n_bound, n_dev);
  hs[threadIdx.x] = 0;
  fs[threadIdx.x] = 0xFFFFFFFFu;
  __syncthreads();
  u32 np = 0;
  u64 ps = 0;
  for (u32 j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
    const u64 d = j ? k[j] - k[j - 1] : k[0];
    const u32 sym = d > 254 ? 0u : (u32)d + 1u;
    atomicAdd(&hs[sym], 1u);
    if (j < fs[sym]) atomicMin(&fs[sym], j);
    if (d > 254) {
      ++np;
      ps += d;
    }
  }
  np = wave_sum(np);
  ps = wave_sum(ps);
  if (lane_id() == 0) {
    if (np) atomicAdd(&ctl->n_patches, np);
    if (ps) atomicAdd(&ctl->patch_sum, ps);
  }
  __syncthreads();
  const u32 d = threadIdx.x;
  if (hs[d]) {
    atomicAdd(&ctl->hist_sym[d], hs[d]);
    atomicMin(&ctl->first_sym[d], fs[d]);
  }
}

// Block-id and block-count deltas (block_compressor.hpp:146-147): u64
// wrapping differences, the first value absolute.
template <bool kWide>
struct BlockDelta {
  KeysT<kWide> K;
  const u32* bpos;
  u32 B, n;
  int which;  // 0: ids, 1: counts
  __device__ __forceinline__ u64 operator()(u32 b) const {
    if (which == 0) {
      const u64 id = K.id_of(K.hi(bpos[b]));
      return b ? id - K.id_of(K.hi(bpos[b - 1])) : id;
    }
    const u32 st = bpos[b];
    const u64 c = (u64)((b + 1 < B ? bpos[b + 1] : n) - st);
    return b ? c - (u64)(st - bpos[b - 1]) : c;
  }
};

template <bool kWide>
__global__ __launch_bounds__(kBlock) void k_block_stats(BlockDelta<kWide> f, const u32* B_dev,
                                                        const u32* n_dev, DevCtl* ctl) {
  u64 m0 = 0, m1 = 0;
  BlockDelta<kWide> g = f;
  g.B = *B_dev;
  g.n = dyn_n(f.n, n_dev);
  for (u32 b = blockIdx.x * kBlock + threadIdx.x; b < g.B; b += gridDim.x * kBlock) {
    g.which = 0;
    const u64 a = g(b);
    g.which = 1;
    const u64 c = g(b);
    m0 = a > m0 ? a : m0;
    m1 = c > m1 ? c : m1;
  }
  m0 = wave_max(m0);
  m1 = wave_max(m1);
  if (lane_id() == 0) {
    if (m0) atomicMax(&ctl->max_blkd, m0);
    if (m1) atomicMax(&ctl->max_cntd, m1);
  }
}

// BitPacker (bitpacking_encoder.hpp:17-51): value i occupies stream bits
// [i*w, i*w + w), least significant first.  One thread per output word,
// gathering the values that overlap it.
template <class Val>
__global__ __launch_bounds__(kBlock) void k_pack_fixed(Val f, u32 n, u32 w, u64 sec_byte0,
                                                       u32* out) {
  const u64 bit0 = sec_byte0 * 8;
  const u64 nbits = (u64)n * w;
  const u64 wfirst = bit0 >> 5, wlast = (bit0 + nbits - 1) >> 5;
  for (u64 g = wfirst + (u64)blockIdx.x * kBlock + threadIdx.x; g <= wlast;
       g += (u64)gridDim.x * kBlock) {
    const i64 p0 = (i64)(g * 32) - (i64)bit0;  // stream bit at the word's bit 0
    const u64 ilo = p0 < 0 ? 0 : (u64)p0 / w;
    u64 ihi = (u64)(p0 + 31) / w;
    if (ihi > n - 1) ihi = n - 1;
    u32 acc = 0;
    for (u64 i = ilo; i <= ihi; ++i) {
      const u64 v = f((u32)i);
      const i64 sh = (i64)(i * w) - p0;
      acc |= sh >= 0 ? (u32)(v << sh) : (u32)(v >> (-sh));
    }
    out[g] = acc;
  }
}

// ---------------------------------------------------------------------------
// Variable-length coders.  Each yields the Code of element j.

template <bool kWide>
struct CoderGRBlockT {  // golomb_rice_encoder.hpp:92-148 on repos patches
  KeysT<kWide> K;
  u32 k;
  static constexpr bool kMsb = false;
  __device__ __forceinline__ Code operator()(u32 j) const {
    u64 r;
    u32 qd;
    K.lcp(j, r, qd);
    if (r <= 254) return Code{0, 0, 0};
    return Code{(r & ((1ull << k) - 1)) << 1, k + 1, (u32)(r >> k)};
  }
};
struct CoderGRDirect {
  const u64* s;
  u32 k;
  static constexpr bool kMsb = false;
  __device__ __forceinline__ Code operator()(u32 j) const {
    const u64 d = j ? s[j] - s[j - 1] : s[0];
    if (d <= 254) return Code{0, 0, 0};
    return Code{(d & ((1ull << k) - 1)) << 1, k + 1, (u32)(d >> k)};
  }
};
template <bool kWide>
struct CoderHufBlockT {  // huffman_encoder.hpp:150-170 on repos PFD symbols
  KeysT<kWide> K;
  const HufEntry* lut;
  static constexpr bool kMsb = true;
  __device__ __forceinline__ Code operator()(u32 j) const {
    u64 r;
    u32 qd;
    K.lcp(j, r, qd);
    const HufEntry e = lut[r > 254 ? 0u : (u32)r + 1u];
    return Code{e.bits, e.len, 0};
  }
};
template <bool kWide>
struct CoderQuadHufT {  // quads PFD symbols (quad - prevQuad + 1)
  KeysT<kWide> K;
  const HufEntry* lut;
  static constexpr bool kMsb = true;
  __device__ __forceinline__ Code operator()(u32 j) const {
    u64 r;
    u32 qd;
    K.lcp(j, r, qd);
    const HufEntry e = lut[qd > 254 ? 0u : qd + 1u];
    return Code{e.bits, e.len, 0};
  }
};
struct CoderHufDirect {
  const u64* s;
  const HufEntry* lut;
  static constexpr bool kMsb = true;
  __device__ __forceinline__ Code operator()(u32 j) const {
    const u64 d = j ? s[j] - s[j - 1] : s[0];
    const HufEntry e = lut[d > 254 ? 0u : (u32)d + 1u];
    return Code{e.bits, e.len, 0};
  }
};
struct CoderCntHuf {  // Huffman<int> over the RLE counts
  const u32* cnt;
  const HufEntry* lut;
  const u32* big_vals;
  const HufEntry* big_codes;
  u32 n_big;
  static constexpr bool kMsb = true;
  __device__ __forceinline__ Code operator()(u32 j) const {
    const u32 c = cnt[j];
    HufEntry e;
    if (c < kCntBins) {
      e = lut[c];
    } else {
      u32 lo = 0, hi = n_big;  // first index with big_vals[idx] >= c
      while (lo < hi) {
        u32 mid = (lo + hi) >> 1;
        if (big_vals[mid] < c) lo = mid + 1; else hi = mid;
      }
      e = big_codes[lo < n_big ? lo : 0];
    }
    return Code{e.bits, e.len, 0};
  }
};
struct CoderCntEG {  // huffman_encoder.hpp:365-396 (Exp-Golomb, zigzag of int)
  const u32* cnt;
  u32 k;
  static constexpr bool kMsb = true;
  __device__ __forceinline__ Code operator()(u32 j) const {
    const u64 z = (u64)cnt[j] << 1;  // zigzag of a positive int
    const u64 n1 = (z >> k) + 1;
    const u32 nb = bitlen64(n1);
    const u64 low = k ? (__builtin_bitreverse64(z) >> (64 - k)) : 0;  // k low bits, LSB first
    return Code{(n1 << k) | low, nb + k, nb - 1};
  }
};

constexpr u32 kVlItems = 8;
constexpr u32 kVlTile = kBlock * kVlItems;
constexpr u32 kVlWin = 4096;  // LDS words per tile window (16 KiB)

template <class Coder>
__global__ __launch_bounds__(kBlock) void k_vl_bits(Coder c, u32 n, u64* tile_bits) {
  __shared__ u64 sh[kWavesPerBlock];
  const u64 base = (u64)blockIdx.x * kVlTile + (u64)threadIdx.x * kVlItems;
  u64 s = 0;
#pragma unroll
  for (u32 t = 0; t < kVlItems; ++t) {
    const u64 j = base + t;
    if (j < n) {
      const Code cd = c((u32)j);
      s += (u64)cd.lead + cd.len;
    }
  }
  const u64 tot = block_sum(s, sh);
  if (threadIdx.x == 0) tile_bits[blockIdx.x] = tot;
}

template <class Coder>
__global__ __launch_bounds__(kBlock) void k_vl_write(Coder c, u32 n, const u64* tile_off,
                                                     u64 sec_bit0, u32* out) {
  constexpr bool kMsb = Coder::kMsb;
  __shared__ u32 win[kVlWin];
  __shared__ u64 sh[kWavesPerBlock];
  const u64 base = (u64)blockIdx.x * kVlTile + (u64)threadIdx.x * kVlItems;
  Code cd[kVlItems];
  u64 s = 0;
#pragma unroll
  for (u32 t = 0; t < kVlItems; ++t) {
    const u64 j = base + t;
    cd[t] = j < n ? c((u32)j) : Code{0, 0, 0};
    s += (u64)cd[t].lead + cd[t].len;
  }
  u64 tot;
  const u64 ex = block_excl_scan(s, sh, tot);
  if (tot == 0) return;
  const u64 abs0 = sec_bit0 + tile_off[blockIdx.x];
  const u64 w0 = abs0 >> 5;
  const u32 o0 = (u32)(abs0 & 31);
  const u64 nwords = (o0 + tot + 31) >> 5;
  if (nwords <= kVlWin) {
    for (u32 i = threadIdx.x; i < nwords; i += kBlock) win[i] = 0;
    __syncthreads();
    u64 p = o0 + ex;
#pragma unroll
    for (u32 t = 0; t < kVlItems; ++t) {
      put_code<true, kMsb>(win, p, cd[t]);
      p += (u64)cd[t].lead + cd[t].len;
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < nwords; i += kBlock) {
      u32 v = win[i];
      if (kMsb) v = bswap32(v);
      if (i == 0 || i == nwords - 1) {
        if (v) atomicOr(&out[w0 + i], v);
      } else {
        out[w0 + i] = v;
      }
    }
  } else {
    u64 p = abs0 + ex;
#pragma unroll
    for (u32 t = 0; t < kVlItems; ++t) {
      put_code<false, kMsb>(out, p, cd[t]);
      p += (u64)cd[t].lead + cd[t].len;
    }
  }
}

template <class Coder>
void vl_count_t(const Coder& c, u32 n, const VlPlan& p, u64* sec_total, Arena& ar,
                hipStream_t s) {
  k_vl_bits<Coder><<<p.ntiles, kBlock, 0, s>>>(c, n, p.tile_bits);
  LAUNCH_CHECK("k_vl_bits");
  scan_excl<u64>(p.tile_bits, p.tile_off, p.ntiles, sec_total, ar, s);
}

template <class Coder>
void vl_write_t(const Coder& c, u32 n, const VlPlan& p, u64 sec_byte0, u32* out,
                hipStream_t s) {
  k_vl_write<Coder><<<p.ntiles, kBlock, 0, s>>>(c, n, p.tile_off, sec_byte0 * 8, out);
  LAUNCH_CHECK("k_vl_write");
}

template <typename F>
void with_coder(CoderKind kind, const CoderArgs& a, F&& f) {
  switch (kind) {
    case kCoderGRBlock:
      if (a.wide) f(CoderGRBlockT<true>{KeysT<true>{a.kv}, a.k});
      else f(CoderGRBlockT<false>{KeysT<false>{a.kv}, a.k});
      return;
    case kCoderGRDirect:
      f(CoderGRDirect{a.sfc, a.k});
      return;
    case kCoderHufBlock:
      if (a.wide) f(CoderHufBlockT<true>{KeysT<true>{a.kv}, a.lut});
      else f(CoderHufBlockT<false>{KeysT<false>{a.kv}, a.lut});
      return;
    case kCoderHufDirect:
      f(CoderHufDirect{a.sfc, a.lut});
      return;
    case kCoderQuadHuf:
      if (a.wide) f(CoderQuadHufT<true>{KeysT<true>{a.kv}, a.lut});
      else f(CoderQuadHufT<false>{KeysT<false>{a.kv}, a.lut});
      return;
    case kCoderCntHuf:
      f(CoderCntHuf{a.cnt, a.lut, a.big_vals, a.big_codes, a.n_big});
      return;
    case kCoderCntEG:
      f(CoderCntEG{a.cnt, a.k});
      return;
  }
  throw Error(-7, "unknown coder");
}

}  // namespace

// ---------------------------------------------------------------------------
// launch wrappers

void launch_init_ctl(DevCtl* ctl, hipStream_t s) {
  k_init_ctl<<<grid_for(sizeof(DevCtl) / 4, kBlock), kBlock, 0, s>>>(ctl);
  LAUNCH_CHECK("k_init_ctl");
}

void launch_split_lo64(const u64* lo64, u32* lo, u32* idx, u32 n, hipStream_t s) {
  k_split_lo64<<<grid_for(n, kBlock * 8, 16384), kBlock, 0, s>>>(lo64, lo, idx, n);
  LAUNCH_CHECK("k_split_lo64");
}

void launch_fill_u32(u32* p, u64 count, u32 value, hipStream_t s) {
  if (!count) return;
  k_fill_u32<<<grid_for(count, kBlock * 8, 16384), kBlock, 0, s>>>(p, count, value);
  LAUNCH_CHECK("k_fill_u32");
}

void launch_bbox(const float* xyz, u32 n, DevCtl* ctl, hipStream_t s) {
  k_bbox<<<grid_for(n, kBlock * 8, 8192), kBlock, 0, s>>>(reinterpret_cast<const F3*>(xyz), n,
                                                          ctl);
  LAUNCH_CHECK("k_bbox");
  k_offsets<<<1, 64, 0, s>>>(xyz, ctl);
  LAUNCH_CHECK("k_offsets");
}

void launch_quant_stats(const float* xyz, u32 n, const QuantParams& qp, bool diag,
                        DevCtl* ctl, hipStream_t s) {
  const F3* p = reinterpret_cast<const F3*>(xyz);
  const u32 g = grid_for(n, kBlock * 8, 8192);
  if (qp.to) {
    if (diag) k_quant_stats<true, true><<<g, kBlock, 0, s>>>(p, n, qp, ctl);
    else k_quant_stats<true, false><<<g, kBlock, 0, s>>>(p, n, qp, ctl);
  } else {
    if (diag) k_quant_stats<false, true><<<g, kBlock, 0, s>>>(p, n, qp, ctl);
    else k_quant_stats<false, false><<<g, kBlock, 0, s>>>(p, n, qp, ctl);
  }
  LAUNCH_CHECK("k_quant_stats");
}

void launch_quantize_nodes(const float* xyz, u32 n, const QuantParams& qp, const DevCtl* ctl,
                           int32_t* nodes, hipStream_t s) {
  const F3* p = reinterpret_cast<const F3*>(xyz);
  const u32 g = grid_for(n, kBlock * 8, 8192);
  if (qp.to) k_quantize_nodes<true><<<g, kBlock, 0, s>>>(p, n, qp, ctl, nodes);
  else k_quantize_nodes<false><<<g, kBlock, 0, s>>>(p, n, qp, ctl, nodes);
  LAUNCH_CHECK("k_quantize_nodes");
}

void launch_keygen(const float* xyz, u32 n, const QuantParams& qp, const KeyParams& kp,
                   const HilbertSM& hs, const DevCtl* ctl, u64* key, u32* lo, u32* idx,
                   u64* lo64, hipStream_t s) {
  const F3* p = reinterpret_cast<const F3*>(xyz);
  const u32 g = grid_for(n, kBlock * 4, 16384);
#define KG(TO, DI, HI, WI) \
  k_keygen<TO, DI, HI, WI><<<g, kBlock, 0, s>>>(p, n, qp, kp, hs, ctl, key, lo, idx, lo64)
  const bool to = qp.to != 0, hi = kp.hilbert != 0, wi = kp.wide != 0;
  if (kp.direct) {
    if (to) { if (hi) KG(true, true, true, false); else KG(true, true, false, false); }
    else { if (hi) KG(false, true, true, false); else KG(false, true, false, false); }
  } else if (wi) {
    if (to) { if (hi) KG(true, false, true, true); else KG(true, false, false, true); }
    else { if (hi) KG(false, false, true, true); else KG(false, false, false, true); }
  } else {
    if (to) { if (hi) KG(true, false, true, false); else KG(true, false, false, false); }
    else { if (hi) KG(false, false, true, false); else KG(false, false, false, false); }
  }
#undef KG
  LAUNCH_CHECK("k_keygen");
}

void launch_m2h(const u64* in, u32 n, u32 iteration, const HilbertSM& hs, u64* out,
                hipStream_t s) {
  k_m2h<<<grid_for(n, kBlock * 4, 16384), kBlock, 0, s>>>(in, n, iteration, hs, out);
  LAUNCH_CHECK("k_m2h");
}

void scan_excl_u32(const u32* in, u32* out, u64 n, u32* total, Arena& ar, hipStream_t s) {
  const size_t mark = ar.off;
  scan_excl<u32>(in, out, n, total, ar, s);
  ar.off = mark;
}
void scan_excl_u64(const u64* in, u64* out, u64 n, u64* total, Arena& ar, hipStream_t s) {
  const size_t mark = ar.off;
  scan_excl<u64>(in, out, n, total, ar, s);
  ar.off = mark;
}

int radix_sort_narrow(u64* a, u64* a2, u32 n, u32 bits, Arena& ar, hipStream_t s) {
  u64* src = a;
  u64* dst = a2;
  int cur = 0;
  for (u32 shift = 0; shift < bits; shift += 8) {
    radix_pass<u64, u32, false>(src, dst, nullptr, nullptr, n, shift, ar, s);
    u64* t = src;
    src = dst;
    dst = t;
    cur ^= 1;
  }
  return cur;
}

int radix_sort_narrow_v(u64* a, u64* a2, u32* v, u32* v2, u32 n, u32 bits, Arena& ar,
                        hipStream_t s) {
  int cur = 0;
  for (u32 shift = 0; shift < bits; shift += 8) {
    radix_pass<u64, u32, true>(a, a2, v, v2, n, shift, ar, s);
    std::swap(a, a2);
    std::swap(v, v2);
    cur ^= 1;
  }
  return cur;
}

int radix_sort_wide_v(u64* hi, u64* hi2, u64* lo64, u64* lo64b, u32 n, u32 hi_bits, Arena& ar,
                      hipStream_t s) {
  int cur = 0;
  for (u32 shift = 0; shift < 21; shift += 8) {
    radix_pass<u64, u64, true>(lo64, lo64b, hi, hi2, n, shift, ar, s);
    std::swap(hi, hi2);
    std::swap(lo64, lo64b);
    cur ^= 1;
  }
  for (u32 shift = 0; shift < hi_bits; shift += 8) {
    radix_pass<u64, u64, true>(hi, hi2, lo64, lo64b, n, shift, ar, s);
    std::swap(hi, hi2);
    std::swap(lo64, lo64b);
    cur ^= 1;
  }
  return cur;
}

int radix_sort_wide(u64* hi, u64* hi2, u32* lo, u32* lo2, u32 n, u32 lo_bits, u32 hi_bits,
                    Arena& ar, hipStream_t s) {
  u64* h0 = hi;
  u64* h1 = hi2;
  u32* l0 = lo;
  u32* l1 = lo2;
  int cur = 0;
  for (u32 shift = 0; shift < lo_bits; shift += 8) {
    radix_pass<u32, u64, true>(l0, l1, h0, h1, n, shift, ar, s);
    std::swap(h0, h1);
    std::swap(l0, l1);
    cur ^= 1;
  }
  for (u32 shift = 0; shift < hi_bits; shift += 8) {
    radix_pass<u64, u32, true>(h0, h1, l0, l1, n, shift, ar, s);
    std::swap(h0, h1);
    std::swap(l0, l1);
    cur ^= 1;
  }
  return cur;
}

void runs_unique(const KeyView& kv, bool wide, u32 n, u32* pos, u32* count, Arena& ar,
                 hipStream_t s) {
  if (wide) compact(FlagUniqueWide{kv.a, kv.b}, n, nullptr, pos, count, ar, s);
  else compact(FlagUniqueNarrow{kv.a}, n, nullptr, pos, count, ar, s);
}

void runs_blocks(const KeyView& kv, bool wide, u32 n_bound, const u32* n_dev, u32* pos,
                 u32* count, Arena& ar, hipStream_t s) {
  if (wide) compact(FlagBlockWide{kv.a}, n_bound, n_dev, pos, count, ar, s);
  else compact(FlagBlockNarrow{kv.a}, n_bound, n_dev, pos, count, ar, s);
}

void launch_rle_finish(const KeyView& in, bool wide, const u32* upos, const u32* U_dev, u32 N,
                       u64* ua, u32* ub, u32* cnt, DevCtl* ctl, u32* large, u32 large_cap,
                       hipStream_t s) {
  const u32 g = grid_for(N, kBlock * 8, 8192);
  if (wide) k_rle_finish<true><<<g, kBlock, 0, s>>>(in.a, in.b, upos, U_dev, N, ua, ub, cnt, ctl,
                                                    large, large_cap);
  else k_rle_finish<false><<<g, kBlock, 0, s>>>(in.a, in.b, upos, U_dev, N, ua, ub, cnt, ctl,
                                                large, large_cap);
  LAUNCH_CHECK("k_rle_finish");
}

void launch_stats_block(const KeyView& kv, bool wide, const u32* bpos, const u32* B_dev,
                        u32 n_bound, const u32* n_dev, DevCtl* ctl, hipStream_t s) {
  const u32 g = grid_for(n_bound, kBlock * 8, 8192);
  if (wide) {
    k_stats_block<true><<<g, kBlock, 0, s>>>(KeysT<true>{kv}, n_bound, n_dev, ctl);
    LAUNCH_CHECK("k_stats_block");
    k_block_stats<true><<<g, kBlock, 0, s>>>(
        BlockDelta<true>{KeysT<true>{kv}, bpos, 0, n_bound, 0}, B_dev, n_dev, ctl);
  } else {
    k_stats_block<false><<<g, kBlock, 0, s>>>(KeysT<false>{kv}, n_bound, n_dev, ctl);
    LAUNCH_CHECK("k_stats_block");
    k_block_stats<false><<<g, kBlock, 0, s>>>(
        BlockDelta<false>{KeysT<false>{kv}, bpos, 0, n_bound, 0}, B_dev, n_dev, ctl);
  }
  LAUNCH_CHECK("k_block_stats");
}

void launch_stats_direct(const u64* sfc, u32 n_bound, const u32* n_dev, DevCtl* ctl,
                         hipStream_t s) {
  k_stats_direct<<<grid_for(n_bound, kBlock * 8, 8192), kBlock, 0, s>>>(sfc, n_bound, n_dev,
                                                                        ctl);
  LAUNCH_CHECK("k_stats_direct");
}

VlPlan vl_plan(u32 n, Arena& ar) {
  VlPlan p;
  p.ntiles = div_up(n ? n : 1, kVlTile);
  p.tile_bits = ar.take<u64>(p.ntiles);
  p.tile_off = ar.take<u64>(p.ntiles);
  return p;
}

void vl_count(CoderKind kind, const CoderArgs& a, u32 n, const VlPlan& p, u64* sec_total,
              Arena& ar, hipStream_t s) {
  const size_t mark = ar.off;
  with_coder(kind, a, [&](const auto& c) { vl_count_t(c, n, p, sec_total, ar, s); });
  ar.off = mark;
}

void vl_write(CoderKind kind, const CoderArgs& a, u32 n, const VlPlan& p, u64 sec_byte0,
              u32* out_words, hipStream_t s) {
  with_coder(kind, a, [&](const auto& c) { vl_write_t(c, n, p, sec_byte0, out_words, s); });
}

void pack_block_deltas(int which, const KeyView& kv, bool wide, const u32* bpos, u32 B,
                       u32 n, u32 width, u64 sec_byte0, u32* out_words, hipStream_t s) {
  if (!B) return;
  const u64 nwords = ((u64)B * width + 63) / 32;
  const u32 g = grid_for(nwords, kBlock, 16384);
  if (wide) k_pack_fixed<<<g, kBlock, 0, s>>>(BlockDelta<true>{KeysT<true>{kv}, bpos, B, n, which},
                                              B, width, sec_byte0, out_words);
  else k_pack_fixed<<<g, kBlock, 0, s>>>(BlockDelta<false>{KeysT<false>{kv}, bpos, B, n, which},
                                         B, width, sec_byte0, out_words);
  LAUNCH_CHECK("k_pack_fixed");
}

}  // namespace xnyz
