// kernels.hip -- gfx950 kernels of the compression path, segmented: one
// launch covers every cloud of a batch (see kernels.hpp).
//
// Stage map (reference file:line each stage reproduces):
//   k_bbox, k_offsets  shift_points minimum       preprocessor/shifting.hpp:11-36
//   k_quant_stats      quantize + range scan      quantizer/truncated_octahedron_quantizer.hpp:21-55,
//                      + RLE diagnostic pass      quantizer/cube_quantizer.hpp:15-27,
//                                                 composer/block_compressor_rle.hpp:53-87,133-141
//   k_keygen           block id + Morton/Hilbert  quantizer/block_quantizer.hpp:91-115,
//                      (direct: global SFC)       composer/block_compressor.hpp:90-101
//   k_radix_*          LSD radix sort             quantizer/block_quantizer.hpp:33-76
//   k_flag_*, k_rle_*  RLE cells, block runs      block_compressor_rle.hpp:107-120,
//                                                 block_quantizer.hpp:117-149
//   k_stats            PFD symbols, patches       encoder/p_for_delta_encoder.hpp:27-43
//   k_sec_*            BitPacker, Golomb-Rice,    encoder/bitpacking_encoder.hpp:17-51,
//                      Huffman, Exp-Golomb        encoder/golomb_rice_encoder.hpp:92-148,
//                                                 encoder/huffman_encoder.hpp:150-170,365-396
//
// Floating point: no contraction (this pragma and -ffp-contract=off), IEEE
// quotients (exact_div.hpp), roundf = round half away from zero (fround.hpp)
// -- the reference's x86-64 baseline arithmetic.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <mutex>
#include <type_traits>
#include <unordered_map>
#include <utility>

#include "common.hpp"
#include "kernels.hpp"
#include "xnyz.h"

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

constexpr u32 kItems = kTile / kBlock;  // 16 per thread

// The range of `local`-th tile inside [first, first + count).
struct Range {
  u32 first, count;
};
__device__ __forceinline__ Range tile_range(u32 first, u32 count, u32 local) {
  const u64 f = (u64)local * kTile;
  if (f >= count) return Range{first + count, 0};
  const u64 c = count - f;
  return Range{first + (u32)f, (u32)(c < kTile ? c : kTile)};
}


// Reduction kernels run persistent workgroups: workgroup b walks the
// contiguous tile span [b*T/G, (b+1)*T/G), accumulates in registers / LDS
// while the segment stays the same, and flushes to the segment's counters
// once per segment it touches -- a few thousand device atomics per launch
// instead of one per wavefront and tile.
constexpr u32 kPersistBlocks = 2048;  // 8 workgroups per CU on 256 CUs (default)
inline u32 persist_grid(const Batch& bt) {
  const u32 cap = bt.grid_cap ? bt.grid_cap : kPersistBlocks;
  return bt.ntiles < cap ? bt.ntiles : cap;
}

// The grid of one persistent kernel: persist_grid, capped at the workgroups
// that fit on the device at once (its occupancy per CU times the CUs), so no
// partial second round of workgroups starts as the first one ends.  Used for
// the instruction-bound ones (quantizer statistics, the cell/block start pass
// with statistics: 0.284 -> 0.262 and 0.266 -> 0.260 ms on 50M points); the
// memory-bound reductions ran slower with the cap (their extra workgroups
// wait ready to run as others stall).
// The occupancy queries run once per kernel (prime_persist_grids, from
// launch_init, outside any graph capture); 0 = unknown, no cap.
template <typename... A>
u32 resident_blocks(void (*kernel)(A...)) {
  static std::mutex mu;
  static std::unordered_map<const void*, u32> fit;
  std::lock_guard<std::mutex> lk(mu);
  auto it = fit.find(reinterpret_cast<const void*>(kernel));
  if (it != fit.end()) return it->second;
  int per_cu = 0, dev = 0, cus = 0;
  u32 f = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) == hipSuccess &&
      hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
    f = (u32)(per_cu > 0 ? per_cu : 0) * (u32)(cus > 0 ? cus : 0);
  else
    (void)hipGetLastError();
  fit[reinterpret_cast<const void*>(kernel)] = f;
  return f;
}

template <typename... A>
u32 persist_grid_for(void (*kernel)(A...), const Batch& bt) {
  const u32 f = resident_blocks(kernel), g = persist_grid(bt);
  return f && g > f ? f : g;
}

__device__ __forceinline__ void block_tiles(u32 ntiles, u32& t0, u32& t1) {
  t0 = (u32)((u64)blockIdx.x * ntiles / gridDim.x);
  t1 = (u32)((u64)(blockIdx.x + 1) * ntiles / gridDim.x);
}

// Block-wide min / max; every thread gets the result.  `sh` holds
// kWavesPerBlock entries and may be reused right after the call.
template <typename T>
__device__ __forceinline__ T block_min(T v, T* sh) {
  v = wave_min(v);
  if (lane_id() == 0) sh[threadIdx.x / kWave] = v;
  __syncthreads();
  T r = sh[0];
#pragma unroll
  for (int i = 1; i < kWavesPerBlock; ++i) r = sh[i] < r ? sh[i] : r;
  __syncthreads();
  return r;
}
template <typename T>
__device__ __forceinline__ T block_max(T v, T* sh) {
  v = wave_max(v);
  if (lane_id() == 0) sh[threadIdx.x / kWave] = v;
  __syncthreads();
  T r = sh[0];
#pragma unroll
  for (int i = 1; i < kWavesPerBlock; ++i) r = sh[i] > r ? sh[i] : r;
  __syncthreads();
  return r;
}

// Monotone device counters: skip the atomic when the (possibly stale) current
// value already dominates -- a stale read only costs an extra atomic.
__device__ __forceinline__ void min_to(u64* a, u64 v) {
  if (v < __atomic_load_n(a, __ATOMIC_RELAXED)) atomicMin(a, v);
}
__device__ __forceinline__ void min_to(u32* a, u32 v) {
  if (v < __atomic_load_n(a, __ATOMIC_RELAXED)) atomicMin(a, v);
}
__device__ __forceinline__ void max_to(u64* a, u64 v) {
  if (v > __atomic_load_n(a, __ATOMIC_RELAXED)) atomicMax(a, v);
}
__device__ __forceinline__ void max_to(int* a, int v) {
  if (v > __atomic_load_n(a, __ATOMIC_RELAXED)) atomicMax(a, v);
}

// ---------------------------------------------------------------------------
// initialisation

__global__ void k_init(Batch bt) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u32 nc = sizeof(SegCtl) / 4, nh = sizeof(SegHist) / 4;
  for (u64 i = tid; i < (u64)bt.S * nc; i += stride) {
    const u32 s = (u32)(i / nc), w = (u32)(i % nc);
    u32 v = 0;
    if (w < 6) v = 0xFFFFFFFFu;  // minkey
    if (w == offsetof(SegCtl, e0) / 4) v = bt.sp[s].p0;
    if (w == offsetof(SegCtl, ne) / 4) v = bt.sp[s].n;
    reinterpret_cast<u32*>(bt.sc)[i] = v;
  }
  for (u64 i = tid; i < (u64)bt.S * nh; i += stride) {
    const u32 w = (u32)(i % nh);
    const bool first = (w >= offsetof(SegHist, first_sym) / 4 &&
                        w < offsetof(SegHist, quad) / 4) ||
                       (w >= offsetof(SegHist, first_quad) / 4 &&
                        w < offsetof(SegHist, cnt) / 4) ||
                       (w >= offsetof(SegHist, first_cnt) / 4 &&
                        w < offsetof(SegHist, bitpop) / 4);
    reinterpret_cast<u32*>(bt.sh)[i] = first ? 0xFFFFFFFFu : 0u;
  }
  if (tid == 0) *bt.n_large = 0;
}

// Word copy between device memory and pinned host memory (either side):
// the small tables a call exchanges with the host move through the shader,
// not the copy engine, so they never queue behind an earlier call's stream
// copy to the host.
__global__ void k_copy_words(const u32* __restrict__ src, u32* __restrict__ dst, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (u64)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_fill_u32(u32* p, u64 count, u32 value) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (u64)gridDim.x * blockDim.x)
    p[i] = value;
}

__global__ void k_patch(const Patch* __restrict__ recs, u32 n, const uint8_t* __restrict__ bytes,
                        uint8_t* __restrict__ out) {
  for (u32 r = blockIdx.x; r < n; r += gridDim.x) {
    const Patch q = recs[r];
    for (u32 i = threadIdx.x; i < q.len; i += blockDim.x) out[q.dst + i] = bytes[q.src + i];
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

// ---------------------------------------------------------------------------
// shift_points minimum (shifting.hpp:11-36): per axis the first point, in
// index order, holding the smallest value under '<'.  Values that are not
// '< FLT_MAX' (NaN, +inf, FLT_MAX) never replace the FLT_MAX start; +0 and
// -0 compare equal, so the earlier one wins.  Key = (order-preserving map of
// the value) << 32 | index, reduced with a 64-bit min.

__device__ __forceinline__ u64 min_key(float v, u32 i) {
  if (!(v < FLT_MAX)) return ~0ull;
  const u32 u = __float_as_uint(v == 0.0f ? 0.0f : v);
  const u32 o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((u64)o << 32) | i;
}

__device__ __forceinline__ void flush_bbox(SegCtl& c, u64 (&b)[3], u64* sh) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const u64 m = block_min(b[a], sh);
    if (threadIdx.x == 0 && m != ~0ull) min_to(&c.minkey[a], m);
    b[a] = ~0ull;
  }
}

__global__ __launch_bounds__(kBlock) void k_bbox(Batch bt, const F3* __restrict__ p) {
  __shared__ u64 sh[kWavesPerBlock];
  u32 t0, t1;
  block_tiles(bt.ntiles, t0, t1);
  if (t0 >= t1) return;
  u64 b[3] = {~0ull, ~0ull, ~0ull};
  u32 cur = bt.tiles[t0].seg;
  for (u32 ti = t0; ti < t1; ++ti) {
    const Tile t = bt.tiles[ti];
    if (t.seg != cur) {
      flush_bbox(bt.sc[cur], b, sh);
      cur = t.seg;
    }
    const SegParam& sp = bt.sp[t.seg];
    const Range r = tile_range(sp.p0, sp.n, t.local);
#pragma unroll 4
    for (u32 k = 0; k < kItems; ++k) {
      const u32 o = k * kBlock + threadIdx.x;
      if (o < r.count) {
        const u32 i = r.first + o;
        const F3 q = p[i];
        const u64 kx = min_key(q.x, i), ky = min_key(q.y, i), kz = min_key(q.z, i);
        b[0] = kx < b[0] ? kx : b[0];
        b[1] = ky < b[1] ? ky : b[1];
        b[2] = kz < b[2] ? kz : b[2];
      }
    }
  }
  flush_bbox(bt.sc[cur], b, sh);
}

// The offset is the winner's exact float (keeps the sign of a zero).
__global__ void k_offsets(Batch bt, const float* xyz) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < bt.S * 3; i += gridDim.x * blockDim.x) {
    const u32 s = i / 3, a = i % 3;
    const u64 k = bt.sc[s].minkey[a];
    bt.sc[s].offset[a] = (k == ~0ull) ? FLT_MAX : xyz[3ull * (u32)k + a];
  }
}

// ---------------------------------------------------------------------------
// Quantizers.  truncated_octahedron_quantizer.hpp:29-50: q = s / div;
// E = round(q*0.5)*2, O = round((q-1)*0.5)*2+1; d = x^2 + (y^2 + z^2); E wins
// ties; then the x/y halving.  cube_quantizer.hpp:20-21: floor(s / div).

__device__ __forceinline__ void to_node(float qx, float qy, float qz, int& X, int& Y, int& Z) {
  // round_half_away: roundf in three operations (fround.hpp)
  const float ex = round_half_away(qx * 0.5f) * 2.0f, ey = round_half_away(qy * 0.5f) * 2.0f,
              ez = round_half_away(qz * 0.5f) * 2.0f;
  const float ox = round_half_away((qx - 1.0f) * 0.5f) * 2.0f + 1.0f;
  const float oy = round_half_away((qy - 1.0f) * 0.5f) * 2.0f + 1.0f;
  const float oz = round_half_away((qz - 1.0f) * 0.5f) * 2.0f + 1.0f;
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

__device__ __forceinline__ void quant_point(const F3& p, const float* off,
                                            const QuantParams& qp, float& sx, float& sy,
                                            float& sz, int& X, int& Y, int& Z) {
  sx = p.x - off[0];
  sy = p.y - off[1];
  sz = p.z - off[2];
  float qx, qy, qz;
  if (qp.fast) {  // uniform
    qx = div_const(sx, qp.div, qp.inv);
    qy = div_const(sy, qp.div, qp.inv);
    qz = div_const(sz, qp.div, qp.inv);
  } else {
    qx = sx / qp.div;
    qy = sy / qp.div;
    qz = sz / qp.div;
  }
  if (qp.to) {
    to_node(qx, qy, qz, X, Y, Z);
  } else {
    X = (int)floorf(qx);
    Y = (int)floorf(qy);
    Z = (int)floorf(qz);
  }
}

// quantizer recover (truncated_octahedron_quantizer.hpp:57-76,
// cube_quantizer.hpp:29-38) for the RLE diagnostic pass.
__device__ __forceinline__ void recover_node(int X, int Y, int Z, const QuantParams& qp,
                                             float& rx, float& ry, float& rz) {
  if (qp.to) {
    if (Z & 1) {
      X = (X << 1) - 1;
      Y = (Y << 1) - 1;
    } else {
      X <<= 1;
      Y <<= 1;
    }
    rx = div_const((float)X * 2.0f * qp.b, qp.rec_div, qp.rec_inv);
    ry = div_const((float)Y * 2.0f * qp.b, qp.rec_div, qp.rec_inv);
    rz = div_const((float)Z * 2.0f * qp.b, qp.rec_div, qp.rec_inv);
  } else {
    rx = div_const(((float)X + 0.5f) * 2.0f * qp.b, qp.rec_div, qp.rec_inv);
    ry = div_const(((float)Y + 0.5f) * 2.0f * qp.b, qp.rec_div, qp.rec_inv);
    rz = div_const(((float)Z + 0.5f) * 2.0f * qp.b, qp.rec_div, qp.rec_inv);
  }
}

struct QStat {
  int m[3];
  u32 wr;
  double se, me;
};

__device__ __forceinline__ void flush_qstat(SegCtl& c, QStat& q, bool diag, void* shv) {
  int* shi = reinterpret_cast<int*>(shv);
  u32* shu = reinterpret_cast<u32*>(shv);
  double* shd = reinterpret_cast<double*>(shv);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int m = block_max(q.m[a], shi);
    if (threadIdx.x == 0) max_to(&c.range[a], m);
    q.m[a] = 0;
  }
  const u32 wr = block_sum(q.wr, shu);
  if (threadIdx.x == 0 && wr) atomicAdd(&c.wrapped, wr);
  q.wr = 0;
  if (diag) {
    const double se = block_sum(q.se, shd);
    const double me = block_max(q.me, shd);
    if (threadIdx.x == 0) {
      atomicAdd(&c.sum_err2, se);
      max_to(&c.max_err2_bits, (u64)__double_as_longlong(me));
    }
    q.se = q.me = 0.0;
  }
}

// The points of kGroup of a thread's elements (element k * kBlock + thread
// of the range, k = k0 .. k0 + kGroup - 1), loaded together before any is
// used, indices clamped into the range (r.count >= 1): the thread waits on
// memory once per group.
constexpr u32 kGroup = 8;
__device__ __forceinline__ void load_points(const F3* __restrict__ p, const Range& r, u32 k0,
                                            F3* v) {
#pragma unroll
  for (u32 k = 0; k < kGroup; ++k) {
    const u32 o = (k0 + k) * kBlock + threadIdx.x;
    v[k] = p[r.first + (o < r.count ? o : r.count - 1)];
  }
}

__global__ __launch_bounds__(kBlock) void k_quant_stats(Batch bt, const F3* __restrict__ p) {
  __shared__ double sh[kWavesPerBlock];
  u32 t0, t1;
  block_tiles(bt.ntiles, t0, t1);
  if (t0 >= t1) return;
  QStat q{{0, 0, 0}, 0, 0.0, 0.0};
  u32 cur = bt.tiles[t0].seg;
  for (u32 ti = t0; ti < t1; ++ti) {
    const Tile t = bt.tiles[ti];
    if (t.seg != cur) {
      flush_qstat(bt.sc[cur], q, bt.sp[cur].diag != 0, sh);
      cur = t.seg;
    }
    const SegParam& sp = bt.sp[t.seg];
    const SegCtl& sc = bt.sc[t.seg];
    const Range r = tile_range(sp.p0, sp.n, t.local);
    const QuantParams qp = sp.qp;
    const float off[3] = {sc.offset[0], sc.offset[1], sc.offset[2]};
    const bool diag = sp.diag != 0;
    if (!r.count) continue;
    for (u32 k0 = 0; k0 < kItems; k0 += kGroup) {
      F3 pt[kGroup];
      load_points(p, r, k0, pt);
#pragma unroll
      for (u32 kk = 0; kk < kGroup; ++kk) {
        const u32 o = (k0 + kk) * kBlock + threadIdx.x;
        if (o < r.count) {
          float sx, sy, sz;
          int X, Y, Z;
          quant_point(pt[kk], off, qp, sx, sy, sz, X, Y, Z);
          q.m[0] = X > q.m[0] ? X : q.m[0];
          q.m[1] = Y > q.m[1] ? Y : q.m[1];
          q.m[2] = Z > q.m[2] ? Z : q.m[2];
          q.wr += (Z == -1);
          if (diag) {
            float rx, ry, rz;
            recover_node(X, Y, Z, qp, rx, ry, rz);
            const float a = sx - rx, b = sy - ry, c = sz - rz;
            // block_compressor_rle.hpp:60-62: float products, left-to-right sum
            const float m = __fadd_rn(__fadd_rn(__fmul_rn(a, a), __fmul_rn(b, b)), __fmul_rn(c, c));
            const double md = (double)m;
            q.se += md;
            q.me = md > q.me ? md : q.me;
          }
        }
      }
    }
  }
  flush_qstat(bt.sc[cur], q, bt.sp[cur].diag != 0, sh);
}

__global__ __launch_bounds__(kBlock) void k_quantize_nodes(Batch bt, const F3* __restrict__ p,
                                                           int32_t* nodes) {
  const Tile t = bt.tiles[blockIdx.x];
  const SegParam& sp = bt.sp[t.seg];
  const SegCtl& sc = bt.sc[t.seg];
  const Range r = tile_range(sp.p0, sp.n, t.local);
  const float off[3] = {sc.offset[0], sc.offset[1], sc.offset[2]};
  const QuantParams qp = sp.qp;  // in registers: the stores may alias the table
#pragma unroll 4
  for (u32 k = 0; k < kItems; ++k) {
    const u32 o = k * kBlock + threadIdx.x;
    if (o < r.count) {
      const u32 i = r.first + o;
      float sx, sy, sz;
      int X, Y, Z;
      quant_point(p[i], off, qp, sx, sy, sz, X, Y, Z);
      nodes[3ull * i] = X;
      nodes[3ull * i + 1] = Y;
      nodes[3ull * i + 2] = Z;
    }
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
  // Two digits at once: two[s * 64 + (d1 << 3 | d2)] = next state after both
  // (low byte) | the two output digits (h1 << 3 | h2) << 8.
  uint16_t two[24 * 64];
};

// Every thread of the block calls this; synchronises before returning.
__device__ __forceinline__ void load_sm(SMTables& t, const HilbertSM& hs) {
  for (u32 i = threadIdx.x; i < 24 * 8; i += blockDim.x) {
    t.apply[i] = hs.apply[i];
    t.next[i] = hs.next[i];
  }
  if (threadIdx.x < 8) t.out[threadIdx.x] = hs.out[threadIdx.x];
  __syncthreads();
  for (u32 i = threadIdx.x; i < 24 * 64; i += blockDim.x) {
    const u32 s = i >> 6;
    const u32 m1 = t.apply[s * 8 + ((i >> 3) & 7)], s1 = t.next[s * 8 + m1];
    const u32 m2 = t.apply[s1 * 8 + (i & 7)], s2 = t.next[s1 * 8 + m2];
    t.two[i] = (uint16_t)(s2 | (((u32)t.out[m1] << 3 | t.out[m2]) << 8));
  }
  __syncthreads();
}

// mortonToHilbert3(m, 20) for m < 2^18: digits 20..6 are zero, stay zero and
// leave the state at `init` (composition of 15 copies of the digit-0
// permutation); digits 5..0 go through the machine, two at a time.
__device__ __forceinline__ u32 hilbert18(u32 m, const SMTables& t, u32 init) {
  u32 s = init, r = 0;
#pragma unroll
  for (int i = 2; i >= 0; --i) {
    const u32 e = t.two[s * 64 + ((m >> (6 * i)) & 63)];
    r |= (e >> 8) << (6 * i);
    s = e & 0xFFu;
  }
  return r;
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
// the 63-bit Morton code of the node, or its Hilbert index (wide layout:
// in the high word, low word 0).
__global__ __launch_bounds__(kBlock) void k_keygen(Batch bt, const F3* __restrict__ p,
                                                   HilbertSM hs, u64* __restrict__ key,
                                                   u32* __restrict__ lo, u32* __restrict__ idx,
                                                   u64* __restrict__ lo64) {
  __shared__ SMTables t;
  load_sm(t, hs);
  const Tile tl = bt.tiles[blockIdx.x];
  const SegParam& sp = bt.sp[tl.seg];
  const SegCtl& sc = bt.sc[tl.seg];
  const Range r = tile_range(sp.p0, sp.n, tl.local);
  const KeyParams kp = sp.kp;
  const QuantParams qp = sp.qp;  // in registers: the key stores may alias the table
  const float off[3] = {sc.offset[0], sc.offset[1], sc.offset[2]};
  if (!r.count) return;
  for (u32 k0 = 0; k0 < kItems; k0 += kGroup) {
    F3 pt[kGroup];
    load_points(p, r, k0, pt);
#pragma unroll
    for (u32 kk = 0; kk < kGroup; ++kk) {
      const u32 o = (k0 + kk) * kBlock + threadIdx.x;
      if (o < r.count) {
        const u32 i = r.first + o;
        float sx, sy, sz;
        int X, Y, Z;
        quant_point(pt[kk], off, qp, sx, sy, sz, X, Y, Z);
        if (idx) idx[i] = i;  // point index payload (permutation output)
        u64 hi;
        u32 k21;
        if (kp.direct) {
          const u64 m = split3_21((u32)X) | (split3_21((u32)Y) << 1) | (split3_21((u32)Z) << 2);
          const u64 sfc = kp.hilbert ? hilbert_general(m, kp.L, t, hs.identity) : m;
          if (!bt.wide) {
            key[i] = sfc;
            continue;
          }
          hi = sfc;
          k21 = 0;
        } else {
          const u64 ux = (u64)(i64)X, uy = (u64)(i64)Y, uz = (u64)(i64)Z;
          const u64 cx = ux >> 6, cy = uy >> 6, cz = uz >> 6;
          const u32 m = split3_6((u32)(ux & 63)) | (split3_6((u32)(uy & 63)) << 1) |
                        (split3_6((u32)(uz & 63)) << 2);
          const u32 sfc = kp.hilbert ? hilbert18(m, t, hs.block_init) : m;
          const u32 quad = (u32)((cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2));
          const u64 base = (cx >> 1) + (cy >> 1) * kp.nx;
          const bool seg = (Z == -1);
          const u64 v = seg ? base : base + (cz >> 1) * kp.nxny;
          hi = ((u64)seg << kp.idb) | v;
          k21 = (quad << 18) | sfc;
          if (!bt.wide) {
            key[i] = (hi << 21) | k21;
            continue;
          }
        }
        key[i] = hi;
        if (lo64) lo64[i] = ((u64)i << 32) | k21;  // wide keys + permutation
        else lo[i] = k21;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Selector features (see SegFeat).  HyperLogLog: a 64-bit mix of the key
// picks one of 256 registers by its top 8 bits; the register keeps the
// largest 1 + (leading zeros of the remaining 56 bits).

__device__ __forceinline__ u64 mix64(u64 z) {  // SplitMix64 finaliser
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ void hll_add(u32* reg, u64 key, u64 salt) {
  const u64 h = mix64(key * 0x9E3779B97F4A7C15ull + salt);
  const u32 r = (u32)__clzll((i64)((h << 8) | 0x80ull)) + 1u;
  const u32 j = (u32)(h >> 56);
  if (r > reg[j]) atomicMax(&reg[j], r);
}

struct FeatLds {
  u32 cell[kHll], sub[kHll], blk[kHll];
  u32 xh[16];
  u32 red[kWavesPerBlock];
};

__device__ __forceinline__ void flush_feat(FeatLds& L, SegFeat& f, u32 (&c)[4]) {
  u32 tot[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) tot[i] = block_sum(c[i], L.red);  // synchronises
  const u32 d = threadIdx.x;
  if (d < 4 && tot[d]) atomicAdd(d == 0 ? &f.same_cell : d == 1 ? &f.same_sub
                                        : d == 2 ? &f.same_blk : &f.pairs, tot[d]);
  if (L.cell[d]) atomicMax(&f.hll_cell[d], L.cell[d]);
  if (L.sub[d]) atomicMax(&f.hll_sub[d], L.sub[d]);
  if (L.blk[d]) atomicMax(&f.hll_blk[d], L.blk[d]);
  if (d < 16 && L.xh[d]) atomicAdd(&f.xor_hist[d], L.xh[d]);
  __syncthreads();
  L.cell[d] = L.sub[d] = L.blk[d] = 0;
  if (d < 16) L.xh[d] = 0;
  c[0] = c[1] = c[2] = c[3] = 0;
  __syncthreads();
}

__device__ __forceinline__ u64 node_morton(const F3& q, const float* off, const QuantParams& qp,
                                           int& X, int& Y, int& Z) {
  float sx, sy, sz;
  quant_point(q, off, qp, sx, sy, sz, X, Y, Z);
  return split3_21((u32)X) | (split3_21((u32)Y) << 1) | (split3_21((u32)Z) << 2);
}

__global__ __launch_bounds__(kBlock) void k_features(Batch bt, const F3* __restrict__ p,
                                                     SegFeat* __restrict__ feat) {
  static_assert(kHll == kBlock, "one register of each table per thread");
  __shared__ FeatLds L;
  u32 t0, t1;
  block_tiles(bt.ntiles, t0, t1);
  if (t0 >= t1) return;
  const u32 d = threadIdx.x;
  L.cell[d] = L.sub[d] = L.blk[d] = 0;
  if (d < 16) L.xh[d] = 0;
  __syncthreads();
  u32 c[4] = {0, 0, 0, 0};  // same cell, same sub-block, same block, pairs
  u32 cur = bt.tiles[t0].seg;
  for (u32 ti = t0; ti < t1; ++ti) {
    const Tile t = bt.tiles[ti];
    if (t.seg != cur) {
      flush_feat(L, feat[cur], c);
      cur = t.seg;
    }
    const SegParam& sp = bt.sp[t.seg];
    const SegCtl& sc = bt.sc[t.seg];
    const Range r = tile_range(sp.p0, sp.n, t.local);
    const float off[3] = {sc.offset[0], sc.offset[1], sc.offset[2]};
    const QuantParams qp = sp.qp;  // in registers: the atomics may alias the table
    const u32 p0 = sp.p0;
#pragma unroll 4
    for (u32 k = 0; k < kItems; ++k) {
      const u32 o = k * kBlock + threadIdx.x;
      if (o < r.count) {
        const u32 i = r.first + o;
        int X, Y, Z;
        const u64 m = node_morton(p[i], off, qp, X, Y, Z);
        const u64 cell = ((u64)(u32)X << 42) ^ ((u64)(u32)Y << 21) ^ (u64)(u32)Z;
        const u64 sub = ((u64)(u32)(X >> 6) << 42) ^ ((u64)(u32)(Y >> 6) << 21) ^ (u32)(Z >> 6);
        const u64 blk = ((u64)(u32)(X >> 7) << 42) ^ ((u64)(u32)(Y >> 7) << 21) ^ (u32)(Z >> 7);
        hll_add(L.cell, cell, 1);
        hll_add(L.sub, sub, 2);
        hll_add(L.blk, blk, 3);
        if (i > p0) {
          int PX, PY, PZ;
          const u64 pm = node_morton(p[i - 1], off, qp, PX, PY, PZ);
          const u64 x = m ^ pm;
          const u32 bl = bitlen64(x);
          atomicAdd(&L.xh[bl / 4 < 15 ? bl / 4 : 15], 1u);
          c[0] += x == 0;
          c[1] += (X >> 6) == (PX >> 6) && (Y >> 6) == (PY >> 6) && (Z >> 6) == (PZ >> 6);
          c[2] += (X >> 7) == (PX >> 7) && (Y >> 7) == (PY >> 7) && (Z >> 7) == (PZ >> 7);
          c[3] += 1;
        }
      }
    }
  }
  flush_feat(L, feat[cur], c);
}

__global__ __launch_bounds__(kBlock) void k_m2h(const u64* in, u32 n, u32 L, HilbertSM hs,
                                                u64* out) {
  __shared__ SMTables t;
  load_sm(t, hs);
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

// Up to this many partials, each apply workgroup sums the partials before
// its own (one load per thread) instead of a separate kernel scanning them.
constexpr u32 kScanDirect = 2048;

// Exclusive scan of one tile of kScanTile values plus the tile's carry.  The
// tile is loaded and stored wave-striped (coalesced) and transposed through
// LDS, so that each thread scans kScanItems consecutive values.  kDirect: the
// carry is the sum of partials [0, blockIdx.x) (np <= kScanDirect), and the
// last workgroup writes the total.
template <typename T, bool kDirect>
__global__ __launch_bounds__(kBlock) void k_scan_apply(const T* in, T* out, u64 n,
                                                       const T* part, u32 np, T* total) {
  __shared__ T sh[kWavesPerBlock];
  __shared__ T st[kScanTile + kScanTile / 16];  // one pad word per 16 values
  const u64 base = (u64)blockIdx.x * kScanTile;
  auto slot = [](u32 i) { return i + (i >> 4); };
  T carry;
  if (kDirect) {
    T c = 0;
    for (u32 i = threadIdx.x; i < blockIdx.x; i += kBlock) c += part[i];
    carry = block_sum(c, sh);
  } else {
    carry = part[blockIdx.x];
  }
  T v[kScanItems];
#pragma unroll
  for (u32 k = 0; k < kScanItems; ++k) {
    const u64 i = base + k * kBlock + threadIdx.x;
    v[k] = i < n ? in[i] : 0;
  }
#pragma unroll
  for (u32 k = 0; k < kScanItems; ++k) st[slot(k * kBlock + threadIdx.x)] = v[k];
  __syncthreads();
  T s = 0;
#pragma unroll
  for (u32 k = 0; k < kScanItems; ++k) {
    v[k] = st[slot(threadIdx.x * kScanItems + k)];
    s += v[k];
  }
  T tot;
  T run = block_excl_scan(s, sh, tot) + carry;  // synchronises: the reads are done
#pragma unroll
  for (u32 k = 0; k < kScanItems; ++k) {
    st[slot(threadIdx.x * kScanItems + k)] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (u32 k = 0; k < kScanItems; ++k) {
    const u64 i = base + k * kBlock + threadIdx.x;
    if (i < n) out[i] = st[slot(k * kBlock + threadIdx.x)];
  }
  if (kDirect && total && blockIdx.x + 1 == np && threadIdx.x == 0) *total = carry + tot;
}

template <typename T>
void scan_excl(const T* in, T* out, u64 n, T* total, Arena& ar, hipStream_t s) {
  const u32 np = div_up(n ? n : 1, kScanTile);
  T* part = ar.take<T>(np);
  k_scan_reduce<T><<<np, kBlock, 0, s>>>(in, n, part);
  LAUNCH_CHECK("k_scan_reduce");
  if (np <= kScanDirect) {
    k_scan_apply<T, true><<<np, kBlock, 0, s>>>(in, out, n, part, np, total);
    LAUNCH_CHECK("k_scan_apply");
    return;
  }
  k_scan_partials<T><<<1, kBlock, 0, s>>>(part, np, total);
  LAUNCH_CHECK("k_scan_partials");
  k_scan_apply<T, false><<<np, kBlock, 0, s>>>(in, out, n, part, np, nullptr);
  LAUNCH_CHECK("k_scan_apply");
}

// ---------------------------------------------------------------------------
// Segmented LSD radix sort, kBits-bit digits (6, 7 or 8), each tile of the
// tile table cut into kSub radix tiles of kTile / kSub keys.  The digit-count
// matrix is laid out per segment ([segment][digit][radix tile in segment]),
// so one exclusive scan over it yields, for every radix tile and digit, the
// output position inside the tile's own segment range.  The scatter is
// stable: each wavefront ranks its 64-key slots in order with ballot-based
// digit matching and running per-wave counters in LDS, so equal digits keep
// their input order.  Digit width and radix tile size are run-time choices
// (xnyz_tuning); every choice gives the same sorted order.

constexpr u32 kWaveSub = kWave * kItems;  // elements per wavefront in a tile

template <u32 kBits, u32 kSub>
struct RadixShapeT {
  static constexpr u32 D = 1u << kBits;           // digits
  static constexpr u32 kRT = kTile / kSub;        // keys per radix tile
  static constexpr u32 kIt = kItems / kSub;       // keys per thread
  static constexpr u32 kWSub = kWave * kIt;       // keys per wavefront
  static_assert(D <= kBlock, "one digit per thread in the digit scans");
};

// Index of radix tile (t, sub) of a segment's digit d in the count matrix.
__device__ __forceinline__ u64 count_index(const SegParam& sp, u32 local, u32 sub, u32 d, u32 D,
                                           u32 nsub) {
  return (u64)D * nsub * sp.tile0 + (u64)d * sp.ntiles * nsub + (u64)local * nsub + sub;
}

// Whether pass `pass` moves this segment's keys: a segment whose keys need
// fewer passes than the batch's keeps them in place from then on (its digits
// there are all zero); if the passes it skips are odd in number, the batch's
// last pass copies them across so that they end in the same buffer as the
// rest.
__device__ __forceinline__ bool pass_active(const SegParam& sp, u32 pass) {
  return sp.passes == 0 || pass < sp.passes;
}

// Digit counts of one radix tile.  The LDS histogram is replicated per lane
// group (copy lane & 15, copies 2^kBits + 1 words apart, odd modulo the 32
// banks): in the passes over the keys' top digits most keys of a tile share
// a few digits, and one copy per wavefront would serialise their atomics.
constexpr u32 kHistRep = 16;

template <typename K, u32 kBits, u32 kSub>
__global__ __launch_bounds__(kBlock) void k_radix_hist(Batch bt, const K* __restrict__ keys,
                                                       u32 shift, u32 pass,
                                                       u32* __restrict__ counts) {
  using RS = RadixShapeT<kBits, kSub>;
  constexpr u32 kStride = RS::D + 1;
  __shared__ u32 h[kHistRep * kStride];
  const u32 tix = blockIdx.x / kSub, sub = blockIdx.x % kSub;
  const Tile t = bt.tiles[tix];
  const SegParam& sp = bt.sp[t.seg];
  const Range r = tile_range(sp.p0, sp.n, t.local);
  const u32 base = sub * RS::kRT;
  if (!pass_active(sp, pass)) {  // uniform: every key counts as digit 0
    const u32 n = r.count > base ? (r.count - base < RS::kRT ? r.count - base : RS::kRT) : 0u;
    if (threadIdx.x < RS::D)
      counts[count_index(sp, t.local, sub, threadIdx.x, RS::D, kSub)] = threadIdx.x ? 0u : n;
    return;
  }
  for (u32 i = threadIdx.x; i < kHistRep * kStride; i += kBlock) h[i] = 0;
  __syncthreads();
  u32* mine = h + (lane_id() & (kHistRep - 1)) * kStride;
  if (r.count > base) {  // uniform; an empty radix tile still writes its zero counts
    // every key in flight before the first count: loads clamped to the
    // tile's last key, so none is conditional
    const u32 last = r.count - 1;
    K kk[RS::kIt];
#pragma unroll
    for (u32 it = 0; it < RS::kIt; ++it) {
      const u32 o = base + it * kBlock + threadIdx.x;
      kk[it] = keys[r.first + (o < r.count ? o : last)];
    }
#pragma unroll
    for (u32 it = 0; it < RS::kIt; ++it)
      if (base + it * kBlock + threadIdx.x < r.count)
        atomicAdd(&mine[(u32)(kk[it] >> shift) & (RS::D - 1)], 1u);
  }
  __syncthreads();
  const u32 d = threadIdx.x;
  if (d < RS::D) {
    u32 c = 0;
#pragma unroll
    for (u32 k = 0; k < kHistRep; ++k) c += h[k * kStride + d];
    counts[count_index(sp, t.local, sub, d, RS::D, kSub)] = c;
  }
}

template <typename K, typename V, bool kHasV, u32 kBits, u32 kSub>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(Batch bt, const K* __restrict__ kin,
                                                          K* __restrict__ kout,
                                                          const V* __restrict__ vin,
                                                          V* __restrict__ vout, u32 shift,
                                                          u32 pass, u32 npasses,
                                                          const u32* __restrict__ offs) {
  using RS = RadixShapeT<kBits, kSub>;
  constexpr u32 D = RS::D, kIt = RS::kIt;
  __shared__ u32 wc[kWavesPerBlock][D];
  __shared__ u32 wb[kWavesPerBlock][D];
  __shared__ u32 ds[D], gb[D], sc4[kWavesPerBlock];
  __shared__ u64 stage[RS::kRT];  // the radix tile's keys, then its values, in output order
  static_assert(sizeof(K) <= 8 && sizeof(V) <= 8, "staging holds 64-bit items");
  K* stage_k = reinterpret_cast<K*>(stage);
  for (u32 i = threadIdx.x; i < kWavesPerBlock * D; i += kBlock) (&wc[0][0])[i] = 0;
  __syncthreads();
  const u32 tix = blockIdx.x / kSub, sub = blockIdx.x % kSub;
  const Tile t = bt.tiles[tix];
  const SegParam& sp = bt.sp[t.seg];
  const Range r0 = tile_range(sp.p0, sp.n, t.local);
  const u32 base = sub * RS::kRT;
  // this radix tile: keys [first, first + count) of the sorted range
  const u32 count = r0.count > base ? (r0.count - base < RS::kRT ? r0.count - base : RS::kRT) : 0u;
  const u32 first = r0.first + base;
  const u32 w = threadIdx.x / kWave, l = lane_id();
  const u64 lt = (1ull << l) - 1;
  const u32 wstart = w * RS::kWSub;
  // The radix tile's global digit offsets are strided in memory; start that
  // load now so its latency overlaps the key loads and the ranking.
  if (!pass_active(sp, pass)) {  // uniform: keys stay, or cross over in the last pass
    if (pass + 1 == npasses && ((npasses - sp.passes) & 1u)) {
      for (u32 i = threadIdx.x; i < count; i += kBlock) {
        kout[first + i] = kin[first + i];
        if (kHasV) vout[first + i] = vin[first + i];
      }
    }
    return;
  }
  const u32 my_off = threadIdx.x < D ? offs[count_index(sp, t.local, sub, threadIdx.x, D, kSub)] : 0u;
  if (count == 0) return;  // uniform: an empty radix tile moves nothing
  K k[kIt];
  V v[kIt];
  u32 rk[kIt];
  u32 dg[kIt];
  // every key (and value) in flight before the ranking: loads clamped to
  // the tile's last item, so none is conditional
#pragma unroll
  for (u32 s = 0; s < kIt; ++s) {
    const u32 o = wstart + s * kWave + l;
    const u32 oc = o < count ? o : count - 1;
    k[s] = kin[first + oc];
    if (kHasV) v[s] = vin[first + oc];
  }
#pragma unroll
  for (u32 s = 0; s < kIt; ++s) {
    const u32 o = wstart + s * kWave + l;
    const bool valid = o < count;
    const u32 d = (u32)(k[s] >> shift) & (D - 1);
    // peers = the lanes whose digit equals this lane's: for every bit b, the
    // lanes that agree on it are ballot(bit) where the bit is set and its
    // complement elsewhere, i.e. ~(ballot ^ m) with m all ones where the bit
    // is set (one three-input bit operation per 32-bit half).
    u64 peers = __ballot(valid);
#pragma unroll
    for (u32 b = 0; b < kBits; ++b) {
      const u32 m = (u32)((int)(d << (31 - b)) >> 31);
      const u64 bb = __ballot(m != 0);
      peers &= ~(bb ^ (((u64)m << 32) | m));
    }
    const u32 rr = (u32)__popcll(peers & lt);
    const u32 prev = wc[w][d];
    if (valid && rr == 0) wc[w][d] = prev + (u32)__popcll(peers);
    rk[s] = prev + rr;
    dg[s] = d;
  }
  __syncthreads();
  // Tile-local order: digit-major, then (wave, slot, lane) = input order, so
  // the permutation is stable.  wb[w][d] = tile-local start of wave w's run
  // of digit d; gb[d] - ds[d] maps a tile-local slot of digit d to its
  // global position.
  {
    const u32 d = threadIdx.x;
    u32 tot = 0;
    if (d < D) {
#pragma unroll
      for (int ww = 0; ww < kWavesPerBlock; ++ww) tot += wc[ww][d];
    }
    u32 all;
    u32 acc = block_excl_scan(tot, sc4, all);
    if (d < D) {
      ds[d] = acc;
      gb[d] = my_off;
#pragma unroll
      for (int ww = 0; ww < kWavesPerBlock; ++ww) {
        wb[ww][d] = acc;
        acc += wc[ww][d];
      }
    }
  }
  __syncthreads();
  u32 lp[kIt];
#pragma unroll
  for (u32 s = 0; s < kIt; ++s) {
    lp[s] = wb[w][dg[s]] + rk[s];
    if (wstart + s * kWave + l < count) stage_k[lp[s]] = k[s];
  }
  __syncthreads();
  // Slot i of the staged tile goes to gb[d] + (i - ds[d]): runs of equal
  // digits are written by consecutive lanes.
  u32 dst[kIt];
#pragma unroll
  for (u32 s = 0; s < kIt; ++s) {
    const u32 i = s * kBlock + threadIdx.x;
    if (i < count) {
      const K key = stage_k[i];
      const u32 d = (u32)(key >> shift) & (D - 1);
      dst[s] = gb[d] + (i - ds[d]);
      kout[dst[s]] = key;
    }
  }
  if (kHasV) {
    __syncthreads();
    V* stage_v = reinterpret_cast<V*>(stage_k);
#pragma unroll
    for (u32 s = 0; s < kIt; ++s)
      if (wstart + s * kWave + l < count) stage_v[lp[s]] = v[s];
    __syncthreads();
#pragma unroll
    for (u32 s = 0; s < kIt; ++s) {
      const u32 i = s * kBlock + threadIdx.x;
      if (i < count) vout[dst[s]] = stage_v[i];
    }
  }
}

template <typename K, typename V, bool kHasV, u32 kBits, u32 kSub>
void radix_pass_t(const Batch& bt, const K* kin, K* kout, const V* vin, V* vout, u32 shift,
                  u32 pass, u32 npasses, Arena& ar, hipStream_t s) {
  const size_t mark = ar.off;
  const u64 ncounts = (u64)(1u << kBits) * bt.ntiles * kSub;
  u32* counts = ar.take<u32>(ncounts);
  k_radix_hist<K, kBits, kSub><<<bt.ntiles * kSub, kBlock, 0, s>>>(bt, kin, shift, pass, counts);
  LAUNCH_CHECK("k_radix_hist");
  scan_excl<u32>(counts, counts, ncounts, nullptr, ar, s);
  k_radix_scatter<K, V, kHasV, kBits, kSub><<<bt.ntiles * kSub, kBlock, 0, s>>>(
      bt, kin, kout, vin, vout, shift, pass, npasses, counts);
  LAUNCH_CHECK("k_radix_scatter");
  ar.off = mark;
}

// One pass of `shape.bits`-bit digits at `shift` (digits past the key's top
// bit are zero).
template <typename K, typename V, bool kHasV>
void radix_pass(const Batch& bt, const K* kin, K* kout, const V* vin, V* vout, u32 shift,
                RadixShape shape, u32 pass, u32 npasses, Arena& ar, hipStream_t s) {
  const bool half = shape.sub == 2;
#define XNYZ_RADIX_CASE(B)                                                                    \
  if (shape.bits == B) {                                                                      \
    if (half)                                                                                 \
      radix_pass_t<K, V, kHasV, B, 2>(bt, kin, kout, vin, vout, shift, pass, npasses, ar, s); \
    else                                                                                      \
      radix_pass_t<K, V, kHasV, B, 1>(bt, kin, kout, vin, vout, shift, pass, npasses, ar, s); \
    return;                                                                                   \
  }
  XNYZ_RADIX_CASE(6)
  XNYZ_RADIX_CASE(7)
  XNYZ_RADIX_CASE(8)
#undef XNYZ_RADIX_CASE
  throw Error(XNYZ_EINVAL, "radix digit width must be 6, 7 or 8 bits");
}

// ---------------------------------------------------------------------------
// Runs on sorted keys: flags, per-tile counts, compaction, per-segment
// starts.  A segment's first element always starts a run.

struct KV {
  const u64* a;
  const u32* b;
  bool wide;
  // Block-mode fields: hi = seg<<idb | id-or-base, lo = quad<<18 | sfc18.
  __device__ __forceinline__ u64 hi(u32 j) const { return wide ? a[j] : (a[j] >> 21); }
  __device__ __forceinline__ u32 lo(u32 j) const {
    return wide ? b[j] : (u32)(a[j] & 0x1FFFFFu);
  }
  __device__ __forceinline__ bool same(u32 i, u32 j) const {
    return a[i] == a[j] && (!wide || b[i] == b[j]);
  }
};

__device__ __forceinline__ u64 id_of(u64 h, const KeyParams& kp) {
  const u64 seg = h >> kp.idb, val = h & ((1ull << kp.idb) - 1);
  return seg ? kp.C + val : val;
}

// repos[j] (reid - prevRe with the resets) and quads[j] (quad - prevQuad),
// block_quantizer.hpp:128-149, for keys sorted by (id, quad, sfc).
__device__ __forceinline__ void lcp(const KV& k, u32 j, u32 e0, u64& repos, u32& qd) {
  const u32 l = k.lo(j), quad = l >> 18, sfc = l & 0x3FFFFu;
  if (j == e0 || k.hi(j - 1) != k.hi(j)) {
    repos = sfc;
    qd = quad;
    return;
  }
  const u32 pl = k.lo(j - 1), pq = pl >> 18;
  qd = (quad - pq) & 0xFFu;
  repos = (pq == quad) ? (u64)sfc - (u64)(pl & 0x3FFFFu) : (u64)sfc;
}

// lcp() (or the direct-mode delta) of kN elements first + o[q] of a range of
// `count` elements whose segment starts at e0: the keys at j and j - 1 are
// loaded for every item first, indices clamped into the range, so the
// thread waits on memory once; items with o[q] >= count give v = qd = 0.
template <u32 kN>
__device__ __forceinline__ void lcp_items(const KV& k, bool direct, u32 e0, u32 first,
                                          u32 count, const u32* o, u64* v, u32* qd) {
  const bool wide = k.wide;
  u64 a0[kN], a1[kN];
  u32 w0[kN], w1[kN];
#pragma unroll
  for (u32 q = 0; q < kN; ++q) {
    const u32 j = first + (o[q] < count ? o[q] : count - 1), jp = j > e0 ? j - 1 : j;
    a0[q] = k.a[j];
    a1[q] = k.a[jp];
    if (wide && !direct) {
      w0[q] = k.b[j];
      w1[q] = k.b[jp];
    }
  }
#pragma unroll
  for (u32 q = 0; q < kN; ++q) {
    const u32 j = first + (o[q] < count ? o[q] : count - 1);
    u64 r;
    u32 d = 0;
    if (direct) {
      r = j == e0 ? a0[q] : a0[q] - a1[q];
    } else {
      const u64 hc = wide ? a0[q] : a0[q] >> 21, hp = wide ? a1[q] : a1[q] >> 21;
      const u32 lc = wide ? w0[q] : (u32)(a0[q] & 0x1FFFFFu);
      const u32 lp = wide ? w1[q] : (u32)(a1[q] & 0x1FFFFFu);
      const u32 quad = lc >> 18, sfc = lc & 0x3FFFFu;
      if (j == e0 || hp != hc) {
        r = sfc;
        d = quad;
      } else {
        const u32 pq = lp >> 18;
        d = (quad - pq) & 0xFFu;
        r = pq == quad ? (u64)sfc - (u64)(lp & 0x3FFFFu) : (u64)sfc;
      }
    }
    v[q] = o[q] < count ? r : 0;
    qd[q] = o[q] < count ? d : 0;
  }
}

// lcp() (or the direct-mode delta) of the kN consecutive elements
// first + b .. first + b + kN - 1 of a range of `count` elements whose
// segment starts at e0: each key is loaded once, the one before the run
// as well; items at or past `count` give v = qd = 0.
template <u32 kN>
__device__ __forceinline__ void lcp_run(const KV& k, bool direct, u32 e0, u32 first, u32 count,
                                        u32 b, u64* v, u32* qd) {
  const bool wide = k.wide && !direct;
  const u32 last = first + count - 1;
  u64 a[kN + 1];
  u32 w[kN + 1];
  {
    const u32 j = first + b;  // b < count for a thread with elements
    const u32 jp = j > e0 ? j - 1 : j;
    a[0] = k.a[jp];
    if (wide) w[0] = k.b[jp];
  }
#pragma unroll
  for (u32 q = 0; q < kN; ++q) {
    const u32 j = first + b + q;
    const u32 jc = j < last ? j : last;
    a[q + 1] = k.a[jc];
    if (wide) w[q + 1] = k.b[jc];
  }
#pragma unroll
  for (u32 q = 0; q < kN; ++q) {
    const u32 j = first + b + q;
    const u64 c0 = a[q + 1], c1 = a[q];
    u64 r;
    u32 d = 0;
    if (direct) {
      r = j == e0 ? c0 : c0 - c1;
    } else {
      const u64 hc = wide ? c0 : c0 >> 21, hp = wide ? c1 : c1 >> 21;
      const u32 lc = wide ? w[q + 1] : (u32)(c0 & 0x1FFFFFu);
      const u32 lp = wide ? w[q] : (u32)(c1 & 0x1FFFFFu);
      const u32 quad = lc >> 18, sfc = lc & 0x3FFFFu;
      if (j == e0 || hp != hc) {
        r = sfc;
        d = quad;
      } else {
        const u32 pq = lp >> 18;
        d = (quad - pq) & 0xFFu;
        r = pq == quad ? (u64)sfc - (u64)(lp & 0x3FFFFu) : (u64)sfc;
      }
    }
    v[q] = b + q < count ? r : 0;
    qd[q] = b + q < count ? d : 0;
  }
}

enum FlagKind { kFlagCell = 0, kFlagBlock = 1 };

// In a batch mixing RLE and plain segments, the plain segments' points all
// pass through the cell stage (each its own cell, count 1).
template <int kKind>
__device__ __forceinline__ bool run_flag(const KV& k, u32 j, u32 first, bool all) {
  if (all || j == first) return true;
  if (kKind == kFlagCell) return !k.same(j, j - 1);
  return k.hi(j) != k.hi(j - 1);
}

// run_flag() of kN elements r.first + o[q] (false past the range), the keys
// at j and j - 1 loaded for every item before any is compared.
template <int kKind, u32 kN>
__device__ __forceinline__ void run_flags(const KV& k, const Range& r, u32 first, bool all,
                                          const u32* o, bool* f) {
  if (r.count == 0 || all) {
#pragma unroll
    for (u32 q = 0; q < kN; ++q) f[q] = o[q] < r.count;
    return;
  }
  const bool wide = k.wide;
  u64 a0[kN], a1[kN];
  u32 w0[kN], w1[kN];
#pragma unroll
  for (u32 q = 0; q < kN; ++q) {
    const u32 j = r.first + (o[q] < r.count ? o[q] : r.count - 1), jp = j > first ? j - 1 : j;
    a0[q] = k.a[j];
    a1[q] = k.a[jp];
    if (kKind == kFlagCell && wide) {
      w0[q] = k.b[j];
      w1[q] = k.b[jp];
    }
  }
#pragma unroll
  for (u32 q = 0; q < kN; ++q) {
    const u32 j = r.first + (o[q] < r.count ? o[q] : r.count - 1);
    bool fl;
    if (j == first) fl = true;
    else if (kKind == kFlagCell) fl = !(a0[q] == a1[q] && (!wide || w0[q] == w1[q]));
    else fl = (wide ? a0[q] : a0[q] >> 21) != (wide ? a1[q] : a1[q] >> 21);
    f[q] = o[q] < r.count && fl;
  }
}

template <int kKind>
__device__ __forceinline__ Range flag_range(const Batch& bt, const Tile& t) {
  const SegParam& sp = bt.sp[t.seg];
  if (kKind == kFlagCell) return tile_range(sp.p0, sp.n, t.local);  // over sorted points
  const SegCtl& sc = bt.sc[t.seg];
  if (sp.kp.direct) return Range{0, 0};  // no blocks in direct mode
  return tile_range(sc.e0, sc.ne, t.local);
}

template <int kKind>
__global__ __launch_bounds__(kBlock) void k_flag_count(Batch bt, KV k, u32* partial) {
  __shared__ u32 sh[kWavesPerBlock];
  const Tile t = bt.tiles[blockIdx.x];
  const Range r = flag_range<kKind>(bt, t);
  const u32 first = r.count ? r.first - (u32)((u64)t.local * kTile) : 0;
  const bool all = kKind == kFlagCell && !bt.sp[t.seg].rle;
  u32 o[kItems], c = 0;
  bool f[kItems];
#pragma unroll
  for (u32 it = 0; it < kItems; ++it) o[it] = it * kBlock + threadIdx.x;
  run_flags<kKind, kItems>(k, r, first, all, o, f);
#pragma unroll
  for (u32 it = 0; it < kItems; ++it) c += f[it] ? 1u : 0u;
  const u32 tot = block_sum(c, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

template <int kKind>
__global__ __launch_bounds__(kBlock) void k_flag_scatter(Batch bt, KV k, const u32* part,
                                                         u32* pos) {
  __shared__ u32 sh[kWavesPerBlock];
  const Tile t = bt.tiles[blockIdx.x];
  const Range r = flag_range<kKind>(bt, t);
  const u32 first = r.count ? r.first - (u32)((u64)t.local * kTile) : 0;
  const bool all = kKind == kFlagCell && !bt.sp[t.seg].rle;
  // Wavefront w owns elements [w*kWaveSub, (w+1)*kWaveSub) of the tile,
  // visited 64 at a time (coalesced); ballots give the in-order ranks.
  const u32 w = threadIdx.x / kWave, l = lane_id();
  const u64 lt = (1ull << l) - 1;
  const u32 wstart = w * kWaveSub;
  u64 m[kItems];
  u32 c = 0, o[kItems];
  bool f[kItems];
#pragma unroll
  for (u32 q = 0; q < kItems; ++q) o[q] = wstart + q * kWave + l;
  run_flags<kKind, kItems>(k, r, first, all, o, f);
#pragma unroll
  for (u32 q = 0; q < kItems; ++q) {
    m[q] = __ballot(f[q]);
    c += (u32)__popcll(m[q]);
  }
  if (l == 0) sh[w] = c;
  __syncthreads();
  u32 run = part[blockIdx.x];
#pragma unroll
  for (u32 i = 0; i < (u32)kWavesPerBlock; ++i)
    if (i < w) run += sh[i];
#pragma unroll
  for (u32 q = 0; q < kItems; ++q) {
    if ((m[q] >> l) & 1ull) pos[run + (u32)__popcll(m[q] & lt)] = r.first + wstart + q * kWave + l;
    run += (u32)__popcll(m[q]);
  }
}

// Per-segment start and count of the compacted array from the scanned
// per-tile counts: which 0 = RLE cells (e0, ne), 1 = blocks (b0, nb).
__global__ void k_seg_starts(Batch bt, const u32* part, const u32* total, int which) {
  for (u32 s = blockIdx.x * blockDim.x + threadIdx.x; s < bt.S; s += gridDim.x * blockDim.x) {
    const SegParam& sp = bt.sp[s];
    const u32 st = part[sp.tile0];
    const u32 en = (s + 1 < bt.S) ? part[bt.sp[s + 1].tile0] : *total;
    if (which == 0) {
      bt.sc[s].e0 = st;
      bt.sc[s].ne = en - st;
    } else {
      bt.sc[s].b0 = st;
      bt.sc[s].nb = en - st;
    }
  }
}

template <int kKind>
void compact(const Batch& bt, const KV& k, u32* pos, int which, Arena& ar, hipStream_t s) {
  const size_t mark = ar.off;
  u32* part = ar.take<u32>(bt.ntiles);
  u32* total = ar.take<u32>(1);
  k_flag_count<kKind><<<bt.ntiles, kBlock, 0, s>>>(bt, k, part);
  LAUNCH_CHECK("k_flag_count");
  scan_excl<u32>(part, part, bt.ntiles, total, ar, s);
  k_flag_scatter<kKind><<<bt.ntiles, kBlock, 0, s>>>(bt, k, part, pos);
  LAUNCH_CHECK("k_flag_scatter");
  k_seg_starts<<<grid_for(bt.S, kBlock), kBlock, 0, s>>>(bt, part, total, which);
  LAUNCH_CHECK("k_seg_starts");
  ar.off = mark;
}

// Cell and block starts of the sorted points in one pass (batches with RLE
// segments).  A point starts a cell where its key differs from the one before
// it (every point of a plain segment is a cell of its own), and a block where
// its block id does (block-mode segments); a block start is a cell start, so
// the block's first cell is the cell rank of that point.  Results:
// upos[cell] = its first point, bpos[block] = its first cell -- what
// compact<kFlagCell> over the points and then compact<kFlagBlock> over the
// cells give, without the second pass over the keys.  part[0, nt) counts the
// cells of each tile, part[nt, 2 nt) its blocks.
__device__ __forceinline__ void cell_block_flags(const KV& k, const Range& r, u32 first,
                                                 bool all, bool blocks, const u32* o, bool* fc,
                                                 bool* fb) {
  const bool wide = k.wide;
  u64 a0[kItems], a1[kItems];
  u32 w0[kItems], w1[kItems];
#pragma unroll
  for (u32 q = 0; q < kItems; ++q) {
    const u32 j = r.first + (o[q] < r.count ? o[q] : r.count - 1), jp = j > first ? j - 1 : j;
    a0[q] = k.a[j];
    a1[q] = k.a[jp];
    if (wide) {
      w0[q] = k.b[j];
      w1[q] = k.b[jp];
    }
  }
#pragma unroll
  for (u32 q = 0; q < kItems; ++q) {
    const u32 j = r.first + (o[q] < r.count ? o[q] : r.count - 1);
    const bool in = o[q] < r.count, head = j == first;
    const bool hi_diff = (wide ? a0[q] : a0[q] >> 21) != (wide ? a1[q] : a1[q] >> 21);
    const bool key_diff = a0[q] != a1[q] || (wide && w0[q] != w1[q]);
    fc[q] = in && (all || head || key_diff);
    fb[q] = in && blocks && (head || hi_diff);
  }
}

__global__ __launch_bounds__(kBlock) void k_flag_count2(Batch bt, KV k, u32* part) {
  __shared__ u32 sh[kWavesPerBlock];
  const Tile t = bt.tiles[blockIdx.x];
  const SegParam& sp = bt.sp[t.seg];
  const Range r = tile_range(sp.p0, sp.n, t.local);
  u32 cc = 0, cb = 0;
  if (r.count) {  // uniform
    const u32 first = r.first - (u32)((u64)t.local * kTile);
    u32 o[kItems];
    bool fc[kItems], fb[kItems];
#pragma unroll
    for (u32 q = 0; q < kItems; ++q) o[q] = q * kBlock + threadIdx.x;
    cell_block_flags(k, r, first, !sp.rle, !sp.kp.direct, o, fc, fb);
#pragma unroll
    for (u32 q = 0; q < kItems; ++q) {
      cc += fc[q] ? 1u : 0u;
      cb += fb[q] ? 1u : 0u;
    }
  }
  const u32 a = block_sum(cc, sh), b = block_sum(cb, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[bt.ntiles + blockIdx.x] = b;
  }
}

// Per-segment starts and counts of cells (e0, ne) and blocks (b0, nb) from
// the scan over [cells | blocks] (total: its sum).
__global__ void k_seg_starts2(Batch bt, const u32* part, const u32* total) {
  const u32 nt = bt.ntiles, cells = part[nt];
  for (u32 s = blockIdx.x * blockDim.x + threadIdx.x; s < bt.S; s += gridDim.x * blockDim.x) {
    const SegParam& sp = bt.sp[s];
    const bool last = s + 1 == bt.S;
    const u32 c0 = part[sp.tile0], c1 = last ? cells : part[bt.sp[s + 1].tile0];
    const u32 b0 = part[nt + sp.tile0] - cells;
    const u32 b1 = (last ? *total : part[nt + bt.sp[s + 1].tile0]) - cells;
    bt.sc[s].e0 = c0;
    bt.sc[s].ne = c1 - c0;
    bt.sc[s].b0 = b0;
    bt.sc[s].nb = b1 - b0;
  }
}

// RLE cells (block_compressor_rle.hpp:107-120 deduplicates with a hash map;
// the stream only depends on the sorted cells and their counts).  Cells of
// RLE segments only; counts below kCntBinsSeg go to the segment's histogram
// (with first occurrences), larger ones to a (segment, count, cell) list.
// LDS histogram with first occurrences -> the segment's device histogram.
// Every thread of the block calls this (bin d = threadIdx.x); clears the LDS
// copy for the next segment.
// LDS histograms replicated kRep times (copy c at c * stride words, the
// stride odd modulo the 32 banks), so that the lanes of a wavefront adding to
// one bin mostly hit distinct addresses and banks.
constexpr u32 kRep = 8;
constexpr u32 kRepStride256 = 257;

// Adds 1 to bin s for every valid lane: a single atomic when all valid lanes
// of the wavefront share the bin (skewed symbol distributions: RLE counts of
// 1, quad deltas of 0), else one per lane on the lane's copy.
template <u32 kStride>
__device__ __forceinline__ void hist_add(u32* h, u32 s, bool valid) {
  const u64 act = __ballot(valid);
  if (!act) return;
  const u32 first = (u32)__ffsll((unsigned long long)act) - 1;
  const u32 s0 = (u32)__builtin_amdgcn_readlane((int)s, (int)first);
  const u64 same = __ballot(valid && s == s0);
  if (same == act) {
    if (lane_id() == first) atomicAdd(&h[s0], (u32)__popcll(act));
  } else if (valid) {
    atomicAdd(&h[(lane_id() & (kRep - 1)) * kStride + s], 1u);
  }
}

// Folds the copies of a replicated histogram into copy 0 and clears the rest
// (synchronises first; the caller synchronises before reading copy 0).
template <u32 kBins, u32 kStride>
__device__ __forceinline__ void fold_hist(u32* h) {
  __syncthreads();
  for (u32 d = threadIdx.x; d < kBins; d += kBlock) {
    u32 sum = h[d];
#pragma unroll
    for (u32 c = 1; c < kRep; ++c) {
      sum += h[c * kStride + d];
      h[c * kStride + d] = 0;
    }
    h[d] = sum;
  }
}

__device__ __forceinline__ void flush_hist(u32* h, u32* f, u32* gh, u32* gf) {
  __syncthreads();
  const u32 d = threadIdx.x;
  if (h[d]) {
    atomicAdd(&gh[d], h[d]);
    min_to(&gf[d], f[d]);
  }
  h[d] = 0;
  f[d] = 0xFFFFFFFFu;
  __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_rle_finish(Batch bt, KV in, const u32* upos,
                                                       u64* __restrict__ ua,
                                                       u32* __restrict__ ub,
                                                       u32* __restrict__ cnt) {
  __shared__ u32 hs[kRep * kRepStride256], fs[kCntBinsSeg];
  static_assert(kCntBinsSeg == kBlock, "one count bin per thread");
  u32 t0, t1;
  block_tiles(bt.ntiles, t0, t1);
  if (t0 >= t1) return;
  for (u32 i = threadIdx.x; i < kRep * kRepStride256; i += kBlock) hs[i] = 0;
  fs[threadIdx.x] = 0xFFFFFFFFu;
  __syncthreads();
  u32 cur = bt.tiles[t0].seg;
  for (u32 ti = t0; ti < t1; ++ti) {
    const Tile t = bt.tiles[ti];
    if (t.seg != cur) {
      fold_hist<kCntBinsSeg, kRepStride256>(hs);
      flush_hist(hs, fs, bt.sh[cur].cnt, bt.sh[cur].first_cnt);
      cur = t.seg;
    }
    const SegParam& sp = bt.sp[t.seg];
    const SegCtl& sc = bt.sc[t.seg];
    const Range r = tile_range(sc.e0, sc.ne, t.local);
    const u32 end_cells = sc.e0 + sc.ne, end_pts = sp.p0 + sp.n;
    if (!r.count) continue;
    // run starts of every item, then their keys, loaded before any is used
    u32 st[kItems], en[kItems], kb[kItems];
    u64 ka[kItems];
#pragma unroll
    for (u32 it = 0; it < kItems; ++it) {
      const u32 o = it * kBlock + threadIdx.x;
      const u32 u = r.first + (o < r.count ? o : r.count - 1);
      st[it] = upos[u];
      const u32 nx = upos[u + 1 < end_cells ? u + 1 : u];
      en[it] = u + 1 < end_cells ? nx : end_pts;
    }
#pragma unroll
    for (u32 it = 0; it < kItems; ++it) {
      ka[it] = in.a[st[it]];
      if (in.wide) kb[it] = in.b[st[it]];
    }
#pragma unroll
    for (u32 it = 0; it < kItems; ++it) {
      const u32 o = it * kBlock + threadIdx.x;
      const u32 c = en[it] - st[it];
      hist_add<kRepStride256>(hs, c, o < r.count && c < kCntBinsSeg);
      if (o < r.count) {
        const u32 u = r.first + o;
        ua[u] = ka[it];
        if (in.wide) ub[u] = kb[it];
        cnt[u] = c;
        const u32 rel = u - sc.e0;
        if (c < kCntBinsSeg) {
          if (rel < fs[c]) atomicMin(&fs[c], rel);
        } else {
          const u32 slot = atomicAdd(bt.n_large, 1u);
          if (slot < bt.large_cap) {
            bt.large[3 * slot] = t.seg;
            bt.large[3 * slot + 1] = c;
            bt.large[3 * slot + 2] = rel;
          }
        }
      }
    }
  }
  fold_hist<kCntBinsSeg, kRepStride256>(hs);
  flush_hist(hs, fs, bt.sh[cur].cnt, bt.sh[cur].first_cnt);
}

// Symbol statistics over each segment's coded elements: PFD symbols of repos
// (block mode) or of the SFC deltas (direct mode, block_compressor.hpp:110-117),
// quads symbols, patch count and wrapping sum.  First occurrences are
// relative to the segment's first element.
//
// A thread codes elements t, t + 256, ... of each tile and walks its tiles
// in order, so the elements it sees come in increasing position: a thread
// records a symbol's first occurrence once (an LDS atomicMin) and remembers
// that in a register bitmap, instead of reading the LDS minimum for every
// element.  The repos / delta histogram takes one atomic per element on the
// lane's copy (kRep copies), except for the patch symbol 0, which is counted
// in the patch count (np) and added once per flush: where most values are
// patches (one point per block) every lane would hit the same bin.  The
// quads symbols -- 0, 1..8 or 250..255, quad deltas being differences of
// 3-bit values -- go to 16 private bins per thread, bin (s + 8) & 15, 17
// words apart so that the lanes of a wave never share a bank on one bin.
constexpr u32 kQBins = 16, kQStride = 17;

__device__ __forceinline__ u32 quad_bin(u32 qs) { return (qs + 8) & 15; }
// The quads symbol of bin b (b = 1 would be 249, which never occurs).
__device__ __forceinline__ u32 quad_sym(u32 b) { return b == 0 ? 8u : b < 8 ? 248u + b : b - 8; }

struct StatLds {
  u32 hs[kRep * kRepStride256];  // replicated
  u32 hq[256];                   // quads histogram, filled at the flush
  u32 hq16[kBlock * kQStride];   // per-thread quads bins
  u32 fs[256], fq[256];
  u64 bitpop[64];
  u64 red[kWavesPerBlock];
};

__device__ __forceinline__ void init_stats(StatLds& L) {
  const u32 d = threadIdx.x;
  for (u32 i = d; i < kRep * kRepStride256; i += kBlock) L.hs[i] = 0;
  for (u32 i = d; i < kBlock * kQStride; i += kBlock) L.hq16[i] = 0;
  L.hq[d] = 0;
  L.fs[d] = L.fq[d] = 0xFFFFFFFFu;
  if (d < 64) L.bitpop[d] = 0;
  __syncthreads();
}

__device__ __forceinline__ void flush_stats(StatLds& L, SegCtl& sc, SegHist& h, u32& np,
                                            u32& nqp, u64& ps) {
  const u32 a = block_sum(np, reinterpret_cast<u32*>(L.red));
  const u32 b = block_sum(nqp, reinterpret_cast<u32*>(L.red));
  const u64 c = block_sum(ps, L.red);
  const u32 d = threadIdx.x;
  if (d == 0) {
    if (a) atomicAdd(&sc.n_patches, a);
    if (b) atomicAdd(&sc.n_qpatches, b);
    if (c) atomicAdd(&sc.patch_sum, c);
    L.hs[0] += a;  // the patch symbol
  }
  np = nqp = 0;
  ps = 0;
  {  // thread d sums bin d & 15 over rows d / 16 + 16 i
    const u32 bin = d & (kQBins - 1);
    u32 sum = 0;
#pragma unroll
    for (u32 i = 0; i < kBlock / kQBins; ++i) {
      u32& x = L.hq16[(d / kQBins + kQBins * i) * kQStride + bin];
      sum += x;
      x = 0;
    }
    if (sum) atomicAdd(&L.hq[quad_sym(bin)], sum);
  }
  fold_hist<256, kRepStride256>(L.hs);
  __syncthreads();
  if (d < 64) {
    if (L.bitpop[d]) atomicAdd(&h.bitpop[d], L.bitpop[d]);
    L.bitpop[d] = 0;
  }
  flush_hist(L.hq, L.fq, h.quad, h.first_quad);  // synchronises first
  flush_hist(L.hs, L.fs, h.sym, h.first_sym);
}

// Symbols whose first occurrence a thread has recorded: 256 bits for the
// repos / delta symbols (w0..w3), and for the quads symbols -- which are 0,
// 1..8 or 250..255, quad deltas being differences of 3-bit values -- bit
// (s + 8) & 15 of q, distinct for each of them.
__device__ __forceinline__ bool seen_sym(u64 w0, u64 w1, u64 w2, u64 w3, u32 s) {
  const u64 x = (s & 128) ? ((s & 64) ? w3 : w2) : ((s & 64) ? w1 : w0);
  return (x >> (s & 63)) & 1ull;
}

constexpr u32 kStatGroup = 8;  // elements per thread loaded together

// SegHist::bitpop of one wavefront's values: lane b adds how many of the
// lanes with `hi` (a value > 254) have bit b of v set.  Bits at or above
// the wave's longest such value are all zero, so only those below it are
// counted (18 for block-mode repos values instead of 64).
__device__ __forceinline__ void bitpop_add(bool hi, u64 v, u32 l, u64& bp) {
  if (!__ballot(hi)) return;  // uniform
  const u32 bl = (u32)__builtin_amdgcn_readfirstlane(
      (int)wave_max(hi ? 64u - (u32)__clzll((long long)(v | 1ull)) : 0u));  // hi: v > 254
  for (u32 bb = 0; bb < bl; ++bb) {
    const u32 c = (u32)__popcll(__ballot(hi && ((v >> bb) & 1ull)));
    if (l == bb) bp += c;
  }
}

template <bool kFull>
__device__ __forceinline__ void stats_items(StatLds& L, u64& w0, u64& w1, u64& w2, u64& w3,
                                            u32& sq, u32 rep, u32 qrow, bool direct, u32 lim,
                                            u32 count,
                                            u32 rel0, const u32* o, const u64* v, const u32* qd,
                                            u32& np, u32& nqp, u64& ps) {
#pragma unroll
  for (u32 it = 0; it < kStatGroup; ++it) {
    if (!kFull && o[it] >= count) break;
    const u32 rel = rel0 + o[it];
    const u32 qs = qd[it] > 254 ? 0u : qd[it] + 1u;
    const bool big = v[it] > lim;
    const u32 sym = big ? 0u : (u32)v[it] + 1u;
    if (!direct) {
      const u32 bin = quad_bin(qs);
      atomicAdd(&L.hq16[qrow + bin], 1u);
      nqp += qd[it] > 254;
      if (!(sq >> bin & 1u)) {
        sq |= 1u << bin;
        atomicMin(&L.fq[qs], rel);
      }
    }
    if (!big) atomicAdd(&L.hs[rep + sym], 1u);
    if (!seen_sym(w0, w1, w2, w3, sym)) {
      const u64 m = 1ull << (sym & 63);
      const u32 h = sym >> 6;
      w0 |= h == 0 ? m : 0ull;
      w1 |= h == 1 ? m : 0ull;
      w2 |= h == 2 ? m : 0ull;
      w3 |= h == 3 ? m : 0ull;
      atomicMin(&L.fs[sym], rel);
    }
    np += big;
    ps += big ? v[it] : 0ull;
  }
}

// Four waves per SIMD: the compiler keeps it within 128 registers.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_stats(
    Batch bt, KV k) {
  __shared__ StatLds L;
  u32 t0, t1;
  block_tiles(bt.ntiles, t0, t1);
  if (t0 >= t1) return;
  const u32 d = threadIdx.x, l = lane_id();
  const u32 rep = (l & (kRep - 1)) * kRepStride256;  // this lane's histogram copy
  const u32 qrow = d * kQStride;                      // this thread's quads bins
  init_stats(L);
  u32 np = 0, nqp = 0;
  u64 ps = 0;
  u64 w0 = 0, w1 = 0, w2 = 0, w3 = 0;  // symbols recorded (seen_sym)
  u32 sq = 0;
  u32 cur = bt.tiles[t0].seg;
  for (u32 ti = t0; ti < t1; ++ti) {
    const Tile t = bt.tiles[ti];
    if (t.seg != cur) {
      flush_stats(L, bt.sc[cur], bt.sh[cur], np, nqp, ps);
      w0 = w1 = w2 = w3 = 0;
      sq = 0;
      cur = t.seg;
    }
    const SegParam& sp = bt.sp[t.seg];
    const SegCtl& sc = bt.sc[t.seg];
    const Range r = tile_range(sc.e0, sc.ne, t.local);
    if (!r.count) continue;
    const bool direct = sp.kp.direct != 0;
    const u32 lim = sp.lim;
    // Element it * kBlock + thread of the tile (coalesced loads; a thread's
    // elements still come in increasing position), a group's loads in
    // flight before its LDS updates.
    const u32 rel0 = r.first - sc.e0;
    const bool full = r.count == kTile;
    u64 bp = 0;  // mode search: this lane's share of bitpop (lane b counts bit b)
    for (u32 g0 = 0; g0 < kItems; g0 += kStatGroup) {
      u64 v[kStatGroup];
      u32 qd[kStatGroup], o[kStatGroup];
#pragma unroll
      for (u32 it = 0; it < kStatGroup; ++it) o[it] = (g0 + it) * kBlock + threadIdx.x;
      lcp_items<kStatGroup>(k, direct, sc.e0, r.first, r.count, o, v, qd);
      if (full)  // uniform
        stats_items<true>(L, w0, w1, w2, w3, sq, rep, qrow, direct, lim, r.count, rel0, o, v, qd, np,
                          nqp, ps);
      else
        stats_items<false>(L, w0, w1, w2, w3, sq, rep, qrow, direct, lim, r.count, rel0, o, v, qd, np,
                           nqp, ps);
      if (sp.search) {  // bit populations of the values > 254 (mode search only)
#pragma unroll
        for (u32 it = 0; it < kStatGroup; ++it)
          bitpop_add(o[it] < r.count && v[it] > 254, v[it], l, bp);
      }
    }
    if (bp) atomicAdd(&L.bitpop[l], bp);
  }
  flush_stats(L, bt.sc[cur], bt.sh[cur], np, nqp, ps);
}

// The scatter of the cell and block starts (k_flag_count2 counted them),
// with the cells' symbol statistics -- what k_stats takes over the cell
// array -- from the same keys: a cell start's previous point is the last
// point of the previous cell, so the keys at j and j - 1 give the cell's
// repos / quads (or direct-mode delta) as lcp() does over the cells, and the
// cells are not read back.  Persistent workgroups over contiguous tile
// spans; the statistics flush once per segment a workgroup touches.
// Three waves per SIMD: the compiler keeps it within 168 registers.
// kWide: (u64 hi, u32 lo) keys; the narrow form carries no low words.
template <bool kWide>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_flag_scatter2(
    Batch bt, KV k, const u32* part, u32* upos, u32* bpos) {
  __shared__ u32 shc[kWavesPerBlock], shb[kWavesPerBlock];
  __shared__ StatLds L;
  u32 t0, t1;
  block_tiles(bt.ntiles, t0, t1);
  if (t0 >= t1) return;
  const u32 d = threadIdx.x, w = d / kWave, l = lane_id();
  const u32 rep = (l & (kRep - 1)) * kRepStride256;  // this lane's histogram copy
  const u32 qrow = d * kQStride;                      // this thread's quads bins
  init_stats(L);
  const u64 lt = (1ull << l) - 1;
  const u32 wstart = w * kWaveSub, nt = bt.ntiles, cells_total = part[nt];
  constexpr bool wide = kWide;
  u32 np = 0, nqp = 0;
  u64 ps = 0;
  u64 w0 = 0, w1 = 0, w2 = 0, w3 = 0;  // symbols recorded (seen_sym)
  u32 sq = 0;
  u32 cur = bt.tiles[t0].seg;
  for (u32 ti = t0; ti < t1; ++ti) {
    const Tile t = bt.tiles[ti];
    if (t.seg != cur) {
      flush_stats(L, bt.sc[cur], bt.sh[cur], np, nqp, ps);
      w0 = w1 = w2 = w3 = 0;
      sq = 0;
      cur = t.seg;
    }
    const SegParam& sp = bt.sp[t.seg];
    const Range r = tile_range(sp.p0, sp.n, t.local);
    if (!r.count) continue;  // uniform
    const u32 first = r.first - (u32)((u64)t.local * kTile);
    // segment fields in registers: the stores below may alias them
    const bool all = !sp.rle, direct = sp.kp.direct != 0, search = sp.search != 0;
    const u32 lim = sp.lim, e0 = part[sp.tile0];  // the segment's first cell
    // Wavefront w owns elements [w*kWaveSub, (w+1)*kWaveSub) of the tile,
    // visited 64 at a time (coalesced); ballots give the in-order ranks.
    // Per group of kStatGroup elements: the keys at j and j - 1, the flags,
    // and the cell's lcp() against the previous cell (block_quantizer.hpp:
    // 128-149) or the direct-mode delta -- a cell start's previous point is
    // the last point of the previous cell.  Only v and dq cross the rank
    // barrier.
    u64 mc[kItems], mb[kItems], v[kItems];
    u32 o[kItems], dq[kItems], nc = 0, nb = 0;
#pragma unroll
    for (u32 g0 = 0; g0 < kItems; g0 += kStatGroup) {
      u64 a0[kStatGroup], a1[kStatGroup];
      u32 b0[kStatGroup], b1[kStatGroup];
#pragma unroll
      for (u32 i = 0; i < kStatGroup; ++i) {
        const u32 q = g0 + i;
        o[q] = wstart + q * kWave + l;
        const u32 j = r.first + (o[q] < r.count ? o[q] : r.count - 1),
                  jp = j > first ? j - 1 : j;
        a0[i] = k.a[j];
        a1[i] = k.a[jp];
        if (wide) {
          b0[i] = k.b[j];
          b1[i] = k.b[jp];
        }
      }
#pragma unroll
      for (u32 i = 0; i < kStatGroup; ++i) {
        const u32 q = g0 + i;
        const u32 j = r.first + o[q];
        const bool in = o[q] < r.count, head = j == first;
        const u64 hcur = wide ? a0[i] : a0[i] >> 21, hprev = wide ? a1[i] : a1[i] >> 21;
        const bool key_diff = a0[i] != a1[i] || (wide && b0[i] != b1[i]);
        mc[q] = __ballot(in && (all || head || key_diff));
        mb[q] = __ballot(in && !direct && (head || hcur != hprev));
        nc += (u32)__popcll(mc[q]);
        nb += (u32)__popcll(mb[q]);
        dq[q] = 0;
        if (direct) {
          v[q] = head ? a0[i] : a0[i] - a1[i];
        } else {
          const u32 lc = wide ? b0[i] : (u32)(a0[i] & 0x1FFFFFu);
          const u32 lp = wide ? b1[i] : (u32)(a1[i] & 0x1FFFFFu);
          const u32 quad = lc >> 18, sfc = lc & 0x3FFFFu, pq = lp >> 18;
          const bool restart = head || hprev != hcur;
          dq[q] = restart ? quad : (quad - pq) & 0xFFu;
          v[q] = !restart && pq == quad ? (u64)sfc - (u64)(lp & 0x3FFFFu) : (u64)sfc;
        }
      }
    }
    if (l == 0) {
      shc[w] = nc;
      shb[w] = nb;
    }
    __syncthreads();
    // block ranks: the scan over [cells | blocks] puts every block after all
    // cells; part[nt] is the cell total
    u32 rc = part[ti], rb = part[nt + ti] - cells_total;
#pragma unroll
    for (u32 i = 0; i < (u32)kWavesPerBlock; ++i)
      if (i < w) {
        rc += shc[i];
        rb += shb[i];
      }
    __syncthreads();  // shc / shb are free for the next tile
    u64 bp = 0;  // mode search: this lane's share of bitpop (lane b counts bit b)
#pragma unroll
    for (u32 q = 0; q < kItems; ++q) {
      const bool fc = (mc[q] >> l) & 1ull, fb = (mb[q] >> l) & 1ull;
      const u32 c = rc + (u32)__popcll(mc[q] & lt);
      if (fc) upos[c] = r.first + o[q];
      if (fb) bpos[rb + (u32)__popcll(mb[q] & lt)] = c;
      rc += (u32)__popcll(mc[q]);
      rb += (u32)__popcll(mb[q]);
      if (fc) {
        const u32 rel = c - e0;
        const u32 qs = dq[q] > 254 ? 0u : dq[q] + 1u;
        const bool big = v[q] > lim;
        const u32 sym = big ? 0u : (u32)v[q] + 1u;
        if (!direct) {
          const u32 bin = quad_bin(qs);
          atomicAdd(&L.hq16[qrow + bin], 1u);
          nqp += dq[q] > 254;
          if (!(sq >> bin & 1u)) {
            sq |= 1u << bin;
            atomicMin(&L.fq[qs], rel);
          }
        }
        if (!big) atomicAdd(&L.hs[rep + sym], 1u);
        if (!seen_sym(w0, w1, w2, w3, sym)) {
          const u64 m = 1ull << (sym & 63);
          const u32 h = sym >> 6;
          w0 |= h == 0 ? m : 0ull;
          w1 |= h == 1 ? m : 0ull;
          w2 |= h == 2 ? m : 0ull;
          w3 |= h == 3 ? m : 0ull;
          atomicMin(&L.fs[sym], rel);
        }
        np += big;
        ps += big ? v[q] : 0ull;
      }
      if (search)  // bit populations of the values > 254 (mode search only)
        bitpop_add(fc && v[q] > 254, v[q], l, bp);
    }
    if (bp) atomicAdd(&L.bitpop[l], bp);
  }
  flush_stats(L, bt.sc[cur], bt.sh[cur], np, nqp, ps);
}

__device__ __forceinline__ void flush_block_stats(SegCtl& sc, u64& m0, u64& m1, u64* sh) {
  const u64 a = block_max(m0, sh), b = block_max(m1, sh);
  if (threadIdx.x == 0) {
    if (a) max_to(&sc.max_blkd, a);
    if (b) max_to(&sc.max_cntd, b);
  }
  m0 = m1 = 0;
}

__global__ __launch_bounds__(kBlock) void k_block_stats(Batch bt, KV k, const u32* bpos) {
  __shared__ u64 sh[kWavesPerBlock];
  u32 t0, t1;
  block_tiles(bt.ntiles, t0, t1);
  if (t0 >= t1) return;
  u64 m0 = 0, m1 = 0;
  u32 cur = bt.tiles[t0].seg;
  for (u32 ti = t0; ti < t1; ++ti) {
    const Tile t = bt.tiles[ti];
    if (t.seg != cur) {
      flush_block_stats(bt.sc[cur], m0, m1, sh);
      cur = t.seg;
    }
    const SegParam& sp = bt.sp[t.seg];
    if (sp.kp.direct) continue;
    const SegCtl& sc = bt.sc[t.seg];
    const Range r = tile_range(sc.b0, sc.nb, t.local);
    if (!r.count) continue;
    // Block-id and block-count deltas (block_compressor.hpp:146-147: u64
    // wrapping differences, the first value absolute) of kGroup items at a
    // time: block positions first, then their keys, each step's loads issued
    // together
    const u32 b0 = sc.b0, bend = sc.b0 + sc.nb, cend = sc.e0 + sc.ne;
    for (u32 g0 = 0; g0 < kItems; g0 += kGroup) {
      u32 st[kGroup], pv[kGroup], nx[kGroup];
      u64 h0[kGroup], h1[kGroup];
#pragma unroll
      for (u32 it = 0; it < kGroup; ++it) {
        const u32 o = (g0 + it) * kBlock + threadIdx.x;
        const u32 b = r.first + (o < r.count ? o : r.count - 1);
        st[it] = bpos[b];
        pv[it] = bpos[b > b0 ? b - 1 : b];
        const u32 after = bpos[b + 1 < bend ? b + 1 : b];
        nx[it] = b + 1 < bend ? after : cend;
      }
#pragma unroll
      for (u32 it = 0; it < kGroup; ++it) {
        h0[it] = k.a[st[it]];
        h1[it] = k.a[pv[it]];
      }
#pragma unroll
      for (u32 it = 0; it < kGroup; ++it) {
        const u32 o = (g0 + it) * kBlock + threadIdx.x;
        if (o >= r.count) continue;
        const u32 b = r.first + o;
        const u64 id = id_of(k.wide ? h0[it] : h0[it] >> 21, sp.kp);
        const u64 a = b > b0 ? id - id_of(k.wide ? h1[it] : h1[it] >> 21, sp.kp) : id;
        const u64 cc = (u64)(nx[it] - st[it]);
        const u64 c = b > b0 ? cc - (u64)(st[it] - pv[it]) : cc;
        m0 = a > m0 ? a : m0;
        m1 = c > m1 ? c : m1;
      }
    }
  }
  flush_block_stats(bt.sc[cur], m0, m1, sh);
}

// ---------------------------------------------------------------------------
// Sections: sizes (k_sec_bits), then the writers (k_sec_write_fixed for the
// BitPacker sections, k_sec_write_cells for the variable-length ones).

struct SecCtx {
  KV k;
  const u32* cnt;
  const u32* bpos;
};

template <u32 kKind>
__device__ __forceinline__ bool sec_active(const SegParam& sp, const SegCtl& sc) {
  if (kKind == kSecBlk || kKind == kSecCnt || kKind == kSecQuad) return !sp.kp.direct;
  if (kKind == kSecGR) return sc.n_patches != 0;
  if (kKind == kSecRle) return sp.rle != 0;
  return true;
}

__device__ __forceinline__ HufEntry big_code(const Batch& bt, const SegParam& sp, u32 c) {
  u32 lo = sp.big0, hi = sp.big0 + sp.nbig;  // first index with value >= c
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if (bt.big_vals[mid] < c) lo = mid + 1; else hi = mid;
  }
  return bt.big_codes[lo < sp.big0 + sp.nbig ? lo : sp.big0];
}

constexpr u32 kVlWin = 8448;  // LDS words: 4096 codes of 64 bits plus the edges

// Phase A: the bits of every section in one pass over the tile.  The
// BitPacker sections are fixed-width (count x width); the cell-space sections
// share one repos / quads derivation per element.  tile_bits is laid out
// [section][tile], so one exclusive scan over it serves every section: a
// section's in-segment offsets are differences of scan values, and the
// constant base each section inherits from the ones before it cancels.
__global__ __launch_bounds__(kBlock) void k_sec_bits(Batch bt, SecCtx x, u64* tile_bits) {
  static_assert(kSymBins == kBlock && kCntBinsSeg == kBlock, "one table entry per thread");
  __shared__ u64 sh[kWavesPerBlock];
  __shared__ u32 len_sym[kBlock], len_quad[kBlock], len_cnt[kBlock];  // code lengths
  const Tile t = bt.tiles[blockIdx.x];
  const SegParam& sp = bt.sp[t.seg];
  const SegCtl& sc = bt.sc[t.seg];
  const SegLut& lut = bt.lut[t.seg];
  const KV& k = x.k;
  const bool direct = sp.kp.direct != 0, rle = sp.rle != 0, eg = sp.cnt_eg != 0;
  const u32 lim = sp.lim, grk = sp.grk, egk = sp.egk;
  len_sym[threadIdx.x] = lut.sym[threadIdx.x].len;
  len_quad[threadIdx.x] = lut.quad[threadIdx.x].len;
  len_cnt[threadIdx.x] = lut.cnt[threadIdx.x].len;
  __syncthreads();
  u64 sg = 0, sh_ = 0, sq = 0, sr = 0;
  const Range r = tile_range(sc.e0, sc.ne, t.local);
  if (r.count) {
    // every element's inputs loaded before any is used (see lcp_items)
    u64 v[kItems];
    u32 qd[kItems], oo[kItems], c[kItems];
#pragma unroll
    for (u32 it = 0; it < kItems; ++it) oo[it] = it * kBlock + threadIdx.x;
    if (rle) {
#pragma unroll
      for (u32 it = 0; it < kItems; ++it) c[it] = x.cnt[r.first + (oo[it] < r.count ? oo[it] : r.count - 1)];
    }
    lcp_items<kItems>(k, direct, sc.e0, r.first, r.count, oo, v, qd);
#pragma unroll
    for (u32 it = 0; it < kItems; ++it) {
      if (oo[it] >= r.count) continue;
      if (!direct) sq += len_quad[qd[it] > 254 ? 0u : qd[it] + 1u];
      sh_ += len_sym[v[it] > lim ? 0u : (u32)v[it] + 1u];
      if (v[it] > lim) sg += (v[it] >> grk) + grk + 1;
      if (rle) {  // coder<kSecRle>'s lead + len
        if (eg) {
          const u32 nb = bitlen64((((u64)c[it] << 1) >> egk) + 1);
          sr += (u64)(nb - 1) + nb + egk;
        } else {
          sr += c[it] < kCntBinsSeg ? len_cnt[c[it]] : big_code(bt, sp, c[it]).len;
        }
      }
    }
  }
  u64 tot[kNumSec];
  tot[kSecGR] = block_sum(sg, sh);
  tot[kSecHuf] = block_sum(sh_, sh);
  tot[kSecQuad] = block_sum(sq, sh);
  tot[kSecRle] = block_sum(sr, sh);
  if (threadIdx.x == 0) {
    const u64 nb = direct ? 0 : tile_range(sc.b0, sc.nb, t.local).count;
    tot[kSecBlk] = nb * sp.w1;
    tot[kSecCnt] = nb * sp.w2;
#pragma unroll
    for (u32 q = 0; q < kNumSec; ++q) tile_bits[(u64)q * bt.ntiles + blockIdx.x] = tot[q];
  }
}

// Per segment and section: the scan offset at its first tile and its size.
__global__ void k_seg_sections(Batch bt, const u64* tile_off, const u64* total) {
  const u64 nt = bt.ntiles;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < bt.S * kNumSec;
       i += gridDim.x * blockDim.x) {
    const u32 s = i / kNumSec, q = i % kNumSec;
    const u64* off = tile_off + q * nt;
    const u64 st = off[bt.sp[s].tile0];
    const u64 en = (s + 1 < bt.S) ? off[bt.sp[s + 1].tile0] : (q + 1 < kNumSec ? off[nt] : *total);
    bt.sc[s].sec_bit0[q] = st;
    bt.sc[s].sec_bits[q] = en - st;
  }
}

// A thread's run of bits inside a section: its codes are consecutive in the
// stream, so whole 32-bit words are assembled in a register and stored
// plainly; only the run's first and last words, which neighbouring runs
// share, are ORed.  Zero words are skipped: the target (the LDS window or the
// zero-filled output) already holds zeros.  `words` is the LDS window
// (kLds) or the output (big-endian byte order applied here for MSB-first
// streams).
template <bool kMsb, bool kLds>
struct BitRun {
  using Idx = typename std::conditional<kLds, u32, u64>::type;  // window words fit 32 bits
  u32* words;
  Idx cw, first, last;  // current word; the run's first and last words
  u64 acc;              // bits not yet stored; MSB-first: word cw in acc's top half
  u32 fill;             // bits of word cw already taken (by this run or before it)

  __device__ __forceinline__ BitRun(u32* w, u64 p, u64 nbits)
      : words(w), cw((Idx)(p >> 5)), first((Idx)(p >> 5)), last((Idx)((p + nbits - 1) >> 5)),
        acc(0), fill((u32)(p & 31)) {}

  __device__ __forceinline__ void store(u32 v) {
    if (v) {
      if (!kLds && kMsb) v = bswap32(v);
      if (cw == first || cw == last) atomicOr(&words[cw], v);
      else words[cw] = v;
    }
    ++cw;
  }
  // `n` <= 32 bits of `v` (its low bits) at the current position.
  __device__ __forceinline__ void put(u32 v, u32 n) {
    if (!n) return;
    if (kMsb) {
      acc |= ((u64)v << (64 - n)) >> fill;
    } else {
      acc |= (u64)(n == 32 ? v : v & ((1u << n) - 1u)) << fill;
    }
    fill += n;
    if (fill >= 32) {
      store(kMsb ? (u32)(acc >> 32) : (u32)acc);
      acc = kMsb ? acc << 32 : acc >> 32;
      fill -= 32;
    }
  }
  // `n` zero bits (MSB-first Exp-Golomb prefixes).
  __device__ __forceinline__ void zeros(u32 n) {
    fill += n;
    while (fill >= 32) {
      store(kMsb ? (u32)(acc >> 32) : (u32)acc);
      acc = kMsb ? acc << 32 : acc >> 32;
      fill -= 32;
    }
  }
  // `n` one bits (LSB-first Golomb-Rice quotients; any length).
  __device__ __forceinline__ void ones(u32 n) {
    while (n) {
      const u32 t = n < 32 ? n : 32;
      put(t == 32 ? 0xFFFFFFFFu : (1u << t) - 1u, t);
      n -= t;
    }
  }
  __device__ __forceinline__ void code(const Code& c) {
    if (kMsb) {
      zeros(c.lead);
      if (c.len > 32) {
        put((u32)(c.val >> 32), c.len - 32);
        put((u32)c.val, 32);
      } else {
        put((u32)c.val, c.len);
      }
    } else {
      ones(c.lead);
      if (c.len > 32) {
        put((u32)c.val, 32);
        put((u32)(c.val >> 32), c.len - 32);
      } else {
        put((u32)c.val, c.len);
      }
    }
  }
  __device__ __forceinline__ void finish() {
    if (fill) store(kMsb ? (u32)(acc >> 32) : (u32)acc);
  }
};

// A thread's run of bits in the LDS window: every piece of at most 32 bits
// is shifted to its bit position and ORed into the (at most two) window
// words it touches, with no state but the position and no branches.  The
// window starts zeroed; neighbouring runs share edge words through the OR.
template <bool kMsb>
struct OrRun {
  u32* words;
  u32 pos;  // bit position in the window

  __device__ __forceinline__ OrRun(u32* w, u32 p) : words(w), pos(p) {}

  // `n` <= 32 bits of `v` (its low bits).
  __device__ __forceinline__ void put(u32 v, u32 n) {
    const u32 w = pos >> 5, sh = pos & 31;
    if (kMsb) {
      const u64 x = n ? ((u64)v << (64 - n)) >> sh : 0ull;
      atomicOr(&words[w], (u32)(x >> 32));
      atomicOr(&words[w + 1], (u32)x);
    } else {
      const u64 x = (u64)(n >= 32 ? v : v & ((1u << n) - 1u)) << sh;
      atomicOr(&words[w], (u32)x);
      atomicOr(&words[w + 1], (u32)(x >> 32));
    }
    pos += n;
  }
  // `n` <= 64 bits of `v`.
  __device__ __forceinline__ void put64(u64 v, u32 n) {
    const u32 nh = n > 32 ? n - 32 : 0u, nl = n > 32 ? 32u : n;
    if (kMsb) {
      put((u32)(v >> 32), nh);
      put((u32)v, nl);
    } else {
      put((u32)v, nl);
      put((u32)(v >> 32), nh);
    }
  }
  // `n` one bits (LSB-first Golomb-Rice quotients; any length).
  __device__ __forceinline__ void ones(u32 n) {
    while (n) {
      const u32 t = n < 32 ? n : 32u;
      put(t == 32 ? 0xFFFFFFFFu : (1u << t) - 1u, t);
      n -= t;
    }
  }
  // A Huffman code: a 32-bit value; a code longer than 32 bits (a tree deeper
  // than the reference's uint32_t prefix holds) is len - 32 zero bits, then
  // the value.
  __device__ __forceinline__ void huff(const Code& c) {
    pos += c.len > 32 ? c.len - 32 : 0u;
    put((u32)c.val, c.len > 32 ? 32u : c.len);
  }
  // MSB-first: `lead` zero bits; LSB-first: `lead` one bits; then the value.
  // LSB-first codes of at most 64 bits in all (Golomb-Rice codes with short
  // quotients) go in as one piece: up to three words.
  __device__ __forceinline__ void code(const Code& c) {
    const u32 n = c.lead + c.len;
    if (!kMsb && n <= 64) {
      const u64 unary = c.lead >= 64 ? ~0ull : (1ull << c.lead) - 1ull;
      const u64 x = unary | (c.len ? c.val << c.lead : 0ull);  // len > 0: lead < 64
      const u32 w = pos >> 5, sh = pos & 31;
      const u64 y = x << sh;
      atomicOr(&words[w], (u32)y);
      atomicOr(&words[w + 1], (u32)(y >> 32));
      if (sh + n > 64) atomicOr(&words[w + 2], (u32)(x >> (64 - sh)));
      pos += n;
      return;
    }
    if (kMsb) pos += c.lead;
    else ones(c.lead);
    put64(c.val, c.len);
  }
  __device__ __forceinline__ void finish() {}
};

// MSB-first codes of a run of at most 64 bits, shifted together: each code
// is `lead` zero bits, then `len` bits of val (a Huffman code longer than 32
// bits keeps its leading zeros implicit, val < 2^32).
struct MsbAcc {
  u64 a = 0;
  __device__ __forceinline__ void add(const Code& c) {
    const u32 n = c.lead + c.len;  // the whole run is <= 64 bits
    a = (n >= 64 ? 0ull : a << n) | c.val;
  }
};

// The four cell-space sections (Golomb-Rice patches, Huffman symbols,
// quads, RLE counts) of one tile in one pass.  Each thread owns 16
// consecutive elements, so its codes form one run of each section.  The
// tile's keys (and counts) are loaded coalesced into LDS and read back per
// thread; the repos / quads are derived once; every thread then emits its
// runs of all four sections together, into the LDS window when the four
// spans fit in it (stored to the output afterwards), otherwise straight into
// the output.  kDirect: the direct-mode tiles (64-bit deltas, no quads);
// each launch skips the other mode's tiles.  kWide: block-mode keys split
// into (u64 hi, u32 lo), loaded straight from memory.
//
// LDS staging: element e of the tile (e = -1: the element before the tile)
// sits at slot e + 1, halves of a 64-bit key in two 32-bit arrays, with one
// pad word per 16 slots, so the 64 lanes reading 16 consecutive slots each
// hit distinct banks.
__device__ __forceinline__ u32 stage_slot(u32 s) { return s + (s >> 4); }
constexpr u32 kStageWords = kTile + 1 + (kTile + 1) / 16 + 1;  // one 32-bit array
constexpr u32 kCellLds = 2 * kStageWords > kVlWin ? 2 * kStageWords : kVlWin;

// The thread's Golomb-Rice run: only patches have codes.
template <typename R, typename F>
__device__ __forceinline__ void emit_gr(R& run, F f) {
#pragma unroll
  for (u32 q = 0; q < kItems; ++q) {
    const Code c = f(q);
    if (c.len) run.code(c);
  }
  run.finish();
}

template <bool kDirect, bool kWide>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_sec_write_cells(Batch bt, SecCtx x,
                                                            const u64* tile_off, u32* out) {
  static_assert(kSymBins == kBlock && kCntBinsSeg == kBlock, "one table entry per thread");
  using V = typename std::conditional<kDirect, u64, u32>::type;  // block-mode repos < 2^18
  __shared__ u32 win[kCellLds];
  __shared__ u64 sh[4 * kWavesPerBlock];
  // quads codes by quad_bin: the quads symbols are 0, 1..8 or 250..255
  __shared__ HufEntry tsym[kBlock], tquad[kDirect ? 1 : kQBins], tcnt[kBlock];
  const Tile t = bt.tiles[blockIdx.x];
  const SegParam& sp = bt.sp[t.seg];
  if ((sp.kp.direct != 0) != kDirect) return;
  const SegCtl& sc = bt.sc[t.seg];
  const Range r = tile_range(sc.e0, sc.ne, t.local);
  if (r.count == 0) return;
  const bool gr = sec_active<kSecGR>(sp, sc), quad = !kDirect;
  const bool rle = sec_active<kSecRle>(sp, sc);
  const SegLut& lut = bt.lut[t.seg];
  tsym[threadIdx.x] = lut.sym[threadIdx.x];
  if (!kDirect && threadIdx.x < kQBins) tquad[threadIdx.x] = lut.quad[quad_sym(threadIdx.x)];
  if (rle) tcnt[threadIdx.x] = lut.cnt[threadIdx.x];
  const u32 b = threadIdx.x * kItems;  // this thread's first element in the tile
  const bool has = b < r.count;
  const u32 lim = sp.lim, grk = sp.grk;
  const u64 nt = bt.ntiles, blk = blockIdx.x;
  const u32 last = r.first + r.count - 1;
  V v[kItems];
  u32 qd4[kItems / 4];  // quad deltas, four bytes per word
  u32 c[kItems];
#pragma unroll
  for (u32 q = 0; q < kItems; ++q) {
    v[q] = 0;
    c[q] = 0;
  }
#pragma unroll
  for (u32 q = 0; q < kItems / 4; ++q) qd4[q] = 0;
  if (kWide) {
    // (hi, lo) keys: strided loads straight from memory
    __syncthreads();
    if (has) {
      u64 vv[kItems];
      u32 dd[kItems];
      lcp_run<kItems>(x.k, false, sc.e0, r.first, r.count, b, vv, dd);
#pragma unroll
      for (u32 q = 0; q < kItems; ++q) {
        v[q] = (V)vv[q];
        qd4[q / 4] |= dd[q] << (8 * (q % 4));
      }
    }
  } else {
    u32* lo32 = win;
    u32* hi32 = win + kStageWords;
    {
      // slots 0 .. count: elements -1 .. count - 1, all loads in flight first
      constexpr u32 kLd = (kTile + 1 + kBlock - 1) / kBlock;
      u64 kk[kLd];
#pragma unroll
      for (u32 i = 0; i < kLd; ++i) {
        const u32 s = i * kBlock + threadIdx.x;
        const u32 j = s == 0 ? (r.first > sc.e0 ? r.first - 1 : r.first) : r.first + s - 1;
        kk[i] = x.k.a[j < last ? j : last];
      }
#pragma unroll
      for (u32 i = 0; i < kLd; ++i) {
        const u32 s = i * kBlock + threadIdx.x;
        if (s <= r.count) {
          lo32[stage_slot(s)] = (u32)kk[i];
          hi32[stage_slot(s)] = (u32)(kk[i] >> 32);
        }
      }
    }
    __syncthreads();
    if (has) {
      u64 a[kItems + 1];
#pragma unroll
      for (u32 q = 0; q <= kItems; ++q) {
        const u32 s = stage_slot(b + q);
        a[q] = ((u64)hi32[s] << 32) | lo32[s];
      }
#pragma unroll
      for (u32 q = 0; q < kItems; ++q) {
        const u32 j = r.first + b + q;
        const u64 c0 = a[q + 1], c1 = a[q];
        V rr;
        u32 d = 0;
        if (kDirect) {
          rr = (V)(j == sc.e0 ? c0 : c0 - c1);
        } else {
          const u32 lc = (u32)(c0 & 0x1FFFFFu), lp = (u32)(c1 & 0x1FFFFFu);
          const u32 qc = lc >> 18, sfc = lc & 0x3FFFFu;
          if (j == sc.e0 || (c0 >> 21) != (c1 >> 21)) {
            rr = sfc;
            d = qc;
          } else {
            const u32 pq = lp >> 18;
            d = (qc - pq) & 0xFFu;
            rr = pq == qc ? sfc - (lp & 0x3FFFFu) : sfc;
          }
        }
        const bool in = b + q < r.count;
        v[q] = in ? rr : 0;
        qd4[q / 4] |= (in ? d : 0u) << (8 * (q % 4));
      }
    }
  }
  if (rle) {
    // counts, staged the same way (element e at slot e of the low array)
    u32 cc[kItems];
    __syncthreads();  // the key slots have been read
#pragma unroll
    for (u32 i = 0; i < kItems; ++i) {
      const u32 j = r.first + i * kBlock + threadIdx.x;
      cc[i] = x.cnt[j < last ? j : last];
    }
#pragma unroll
    for (u32 i = 0; i < kItems; ++i) {
      const u32 e = i * kBlock + threadIdx.x;
      if (e < r.count) win[stage_slot(e)] = cc[i];
    }
    __syncthreads();
    if (has) {
#pragma unroll
      for (u32 q = 0; q < kItems; ++q)
        if (b + q < r.count) c[q] = win[stage_slot(b + q)];
    }
  }
  __syncthreads();  // staging done: the area becomes the window

  auto valid = [&](u32 q) { return b + q < r.count; };
  auto code_gr = [&](u32 q) {  // Golomb-Rice of the patches, LSB-first
    const u64 y = v[q];
    return valid(q) && y > lim ? Code{(y & ((1ull << grk) - 1)) << 1, grk + 1, (u32)(y >> grk)}
                               : Code{0, 0, 0};
  };
  auto code_huf = [&](u32 q) {
    const HufEntry e = tsym[v[q] > lim ? 0u : (u32)v[q] + 1u];
    return valid(q) ? Code{e.bits, e.len, 0} : Code{0, 0, 0};
  };
  auto code_quad = [&](u32 q) {
    const u32 d = (qd4[q / 4] >> (8 * (q % 4))) & 0xFFu;
    const HufEntry e = tquad[kDirect ? 0u : quad_bin(d > 254 ? 0u : d + 1u)];
    return valid(q) ? Code{e.bits, e.len, 0} : Code{0, 0, 0};
  };
  const bool eg = sp.cnt_eg != 0;
  const u32 egk = sp.egk;
  auto code_rle = [&](u32 q) {
    Code e;
    if (eg) {  // Exp-Golomb of zigzag(int)
      const u64 z = (u64)c[q] << 1;
      const u64 n1 = (z >> egk) + 1;
      const u32 nb = bitlen64(n1);
      const u64 low = egk ? (__builtin_bitreverse64(z) >> (64 - egk)) : 0;
      e = Code{(n1 << egk) | low, nb + egk, nb - 1};
    } else {
      const HufEntry h = c[q] < kCntBinsSeg ? tcnt[c[q]] : big_code(bt, sp, c[q]);
      e = Code{h.bits, h.len, 0};
    }
    return valid(q) ? e : Code{0, 0, 0};
  };

  // run lengths of the four sections and their block scans, together
  constexpr u32 kKinds[4] = {kSecGR, kSecHuf, kSecQuad, kSecRle};
  u64 mine[4] = {0, 0, 0, 0};
#pragma unroll
  for (u32 q = 0; q < kItems; ++q) {
    if (gr && valid(q) && v[q] > lim) mine[0] += ((u64)v[q] >> grk) + grk + 1;
    mine[1] += code_huf(q).len;
    if (quad) mine[2] += code_quad(q).len;
    if (rle) {
      const Code e = code_rle(q);
      mine[3] += (u64)e.lead + e.len;
    }
  }
  u64 ex[4], tot[4];
  {
    const u32 l = lane_id(), w = threadIdx.x / kWave;
    u64 inc[4];
#pragma unroll
    for (u32 s = 0; s < 4; ++s) inc[s] = mine[s];
#pragma unroll
    for (u32 o = 1; o < kWave; o <<= 1) {
      u64 up[4];
#pragma unroll
      for (u32 s = 0; s < 4; ++s) up[s] = __shfl_up(inc[s], o, kWave);
#pragma unroll
      for (u32 s = 0; s < 4; ++s)
        if (l >= o) inc[s] += up[s];
    }
    if (l == kWave - 1) {
#pragma unroll
      for (u32 s = 0; s < 4; ++s) sh[s * kWavesPerBlock + w] = inc[s];
    }
    __syncthreads();
#pragma unroll
    for (u32 s = 0; s < 4; ++s) {
      u64 base = 0, all = 0;
#pragma unroll
      for (u32 i = 0; i < (u32)kWavesPerBlock; ++i) {
        const u64 y = sh[s * kWavesPerBlock + i];
        if (i < w) base += y;
        all += y;
      }
      ex[s] = base + inc[s] - mine[s];
      tot[s] = all;
    }
  }
  // each section's span: bit position in the output, words, place in the window
  u64 abs0[4];
  u32 nw[4], woff[4];
  u32 wtot = 0;
#pragma unroll
  for (u32 s = 0; s < 4; ++s) {
    const u32 kind = kKinds[s];
    abs0[s] = tot[s] ? sp.sec_byte[kind] * 8 + (tile_off[kind * nt + blk] - sc.sec_bit0[kind]) : 0;
    const u64 n = tot[s] ? ((abs0[s] & 31) + tot[s] + 31) >> 5 : 0;
    nw[s] = n > kVlWin ? kVlWin + 1 : (u32)n;
    woff[s] = wtot;
    wtot = wtot + nw[s] > kVlWin + 1 ? kVlWin + 1 : wtot + nw[s];
  }
  const bool g = gr && mine[0], h = mine[1] != 0, qu = quad && mine[2], cn = rle && mine[3];
  if (wtot > kVlWin) {
    // spans too long for the window: runs straight into the output
    // (one section at a time)
    if (g) {
      BitRun<false, false> run(out, abs0[0] + ex[0], mine[0]);
      emit_gr(run, code_gr);
    }
    if (h) {
      BitRun<true, false> run(out, abs0[1] + ex[1], mine[1]);
#pragma unroll
      for (u32 q = 0; q < kItems; ++q) run.code(code_huf(q));
      run.finish();
    }
    if (qu) {
      BitRun<true, false> run(out, abs0[2] + ex[2], mine[2]);
#pragma unroll
      for (u32 q = 0; q < kItems; ++q) run.code(code_quad(q));
      run.finish();
    }
    if (cn) {
      BitRun<true, false> run(out, abs0[3] + ex[3], mine[3]);
#pragma unroll
      for (u32 q = 0; q < kItems; ++q) run.code(code_rle(q));
      run.finish();
    }
    return;
  }
  for (u32 i = threadIdx.x; i < wtot; i += kBlock) win[i] = 0;
  __syncthreads();
  {
    // Huffman-coded symbols and quads; RLE counts Huffman or Exp-Golomb coded
    // A thread's run of one section that fits in 64 bits (short codes, the
    // usual case) is shifted together in a register and ORed in once;
    // longer runs OR each code at its position.
    OrRun<true> rh(win + woff[1], (u32)((abs0[1] & 31) + ex[1]));
    OrRun<true> rq(win + woff[2], (u32)((abs0[2] & 31) + ex[2]));
    OrRun<true> rc(win + woff[3], (u32)((abs0[3] & 31) + ex[3]));
    const bool sh = mine[1] <= 64, sq = mine[2] <= 64, sc = mine[3] <= 64;
    MsbAcc ah, aq, ac;
#pragma unroll
    for (u32 q = 0; q < kItems; ++q) {
      if (h) {
        const Code e = code_huf(q);
        if (sh) ah.add(e);
        else rh.huff(e);
      }
      if (qu) {
        const Code e = code_quad(q);
        if (sq) aq.add(e);
        else rq.huff(e);
      }
      if (cn) {
        const Code e = code_rle(q);
        if (sc) ac.add(e);
        else if (eg) rc.code(e);
        else rc.huff(e);
      }
    }
    if (h && sh) rh.put64(ah.a, (u32)mine[1]);
    if (qu && sq) rq.put64(aq.a, (u32)mine[2]);
    if (cn && sc) rc.put64(ac.a, (u32)mine[3]);
    if (h) rh.finish();
    if (qu) rq.finish();
    if (cn) rc.finish();
  }
  if (g) {
    OrRun<false> rg(win + woff[0], (u32)((abs0[0] & 31) + ex[0]));
#pragma unroll
    for (u32 q = 0; q < kItems; ++q) rg.code(code_gr(q));
    rg.finish();
  }
  __syncthreads();
  for (u32 i = threadIdx.x; i < wtot; i += kBlock) {
    const u32 s = i >= woff[3] && nw[3] ? 3u : i >= woff[2] && nw[2] ? 2u : i >= woff[1] && nw[1] ? 1u : 0u;
    const u32 li = i - woff[s];
    u32 y = win[i];
    if (s != 0) y = bswap32(y);  // the MSB-first sections
    u32* dst = out + (abs0[s] >> 5) + li;
    if (li == 0 || li == nw[s] - 1) {
      if (y) atomicOr(dst, y);
    } else {
      *dst = y;
    }
  }
}

// The two BitPacker sections of a tile of blocks (block-id deltas and
// block-count deltas, block_compressor.hpp:146-147, bitpacking_encoder.hpp:
// 17-51) in one pass.  Every delta of a section has the section's width w,
// so element e of the tile lies at bits [o0 + e w, o0 + (e + 1) w) of the
// tile's span, LSB-first: the deltas are staged in LDS and each output word
// gathers the (at most 32 / w + 2) deltas that overlap it -- no scan and no
// atomics but on the span's two edge words, which the neighbouring tiles or
// the section's meta share.
__global__ __launch_bounds__(kBlock) void k_sec_write_fixed(Batch bt, SecCtx x, u32* out) {
  __shared__ u64 vals[kTile];
  const Tile t = bt.tiles[blockIdx.x];
  const SegParam& sp = bt.sp[t.seg];
  const SegCtl& sc = bt.sc[t.seg];
  if (sp.kp.direct) return;  // no blocks in direct mode
  const Range r = tile_range(sc.b0, sc.nb, t.local);
  if (r.count == 0) return;
  const KV& k = x.k;
  const u32* bpos = x.bpos;
  // the deltas of every element (o = it * kBlock + thread), block
  // positions first, then their keys, each group's loads issued together
  const u32 b0 = sc.b0, bend = sc.b0 + sc.nb, cend = sc.e0 + sc.ne;
  u64 db[kItems], dc[kItems];
  for (u32 g0 = 0; g0 < kItems; g0 += kGroup) {
    u32 st[kGroup], pv[kGroup], nx[kGroup];
    u64 h0[kGroup], h1[kGroup];
#pragma unroll
    for (u32 it = 0; it < kGroup; ++it) {
      const u32 o = (g0 + it) * kBlock + threadIdx.x;
      const u32 b = r.first + (o < r.count ? o : r.count - 1);
      st[it] = bpos[b];
      pv[it] = bpos[b > b0 ? b - 1 : b];
      const u32 after = bpos[b + 1 < bend ? b + 1 : b];
      nx[it] = b + 1 < bend ? after : cend;
    }
#pragma unroll
    for (u32 it = 0; it < kGroup; ++it) {
      h0[it] = k.a[st[it]];
      h1[it] = k.a[pv[it]];
    }
#pragma unroll
    for (u32 it = 0; it < kGroup; ++it) {
      const u32 b = r.first + (g0 + it) * kBlock + threadIdx.x;
      const u64 id = id_of(k.wide ? h0[it] : h0[it] >> 21, sp.kp);
      db[g0 + it] = b > b0 ? id - id_of(k.wide ? h1[it] : h1[it] >> 21, sp.kp) : id;
      const u64 cc = (u64)(nx[it] - st[it]);
      dc[g0 + it] = b > b0 ? cc - (u64)(st[it] - pv[it]) : cc;
    }
  }
#pragma unroll
  for (u32 sec = 0; sec < 2; ++sec) {
    if (sec) __syncthreads();  // the first section's words are gathered
#pragma unroll
    for (u32 it = 0; it < kItems; ++it) {
      const u32 o = it * kBlock + threadIdx.x;
      if (o < r.count) vals[o] = sec ? dc[it] : db[it];
    }
    __syncthreads();
    const u32 kind = sec ? kSecCnt : kSecBlk;
    const u32 w = sec ? sp.w2 : sp.w1;
    const u64 mask = w >= 64 ? ~0ull : (1ull << w) - 1;
    const u64 abs0 = sp.sec_byte[kind] * 8 + (u64)(r.first - b0) * w;
    const u32 o0 = (u32)(abs0 & 31);
    u32* dst = out + (abs0 >> 5);
    const u32 nwords = (u32)((o0 + (u64)r.count * w + 31) >> 5);
    for (u32 kw = threadIdx.x; kw < nwords; kw += kBlock) {
      // tile bits [32 kw, 32 kw + 32); element e starts at o0 + e w
      const u32 lo = 32 * kw;
      const u32 e0 = lo > o0 ? (lo - o0) / w : 0u;
      u32 e1 = (lo + 31 - o0) / w;
      e1 = e1 < r.count - 1 ? e1 : r.count - 1;
      u32 word = 0;
      for (u32 e = e0; e <= e1; ++e) {
        const u64 v = vals[e] & mask;
        const u32 p = o0 + e * w;
        word |= p >= lo ? (u32)(v << (p - lo)) : (u32)(v >> (lo - p));
      }
      if (kw == 0 || kw + 1 == nwords) {
        if (word) atomicOr(&dst[kw], word);
      } else {
        dst[kw] = word;
      }
    }
  }
}

}  // namespace

// Packs the segments' outputs back to back: segment s's `len` bytes at
// src + src_off (256-byte aligned) go to dst + dst_off (any alignment).  A
// workgroup copies one of `slices` parts of a segment's whole destination
// words, reading each as a funnel shift of two aligned source words; slice 0
// also writes the partial words at the segment's two ends byte by byte (the
// neighbouring segments own their other bytes).
__global__ __launch_bounds__(kBlock) void k_gather(const uint8_t* __restrict__ src,
                                                   const GatherSeg* __restrict__ segs,
                                                   u32 slices, uint8_t* __restrict__ dst) {
  const GatherSeg g = segs[blockIdx.x / slices];
  const u32 k = blockIdx.x % slices;
  if (g.len == 0) return;
  const u64 d0 = g.dst_off, d1 = g.dst_off + g.len;
  const u64 w0 = (d0 + 3) >> 2, w1 = d1 >> 2;  // whole destination words [w0, w1)
  if (k == 0) {
    const u64 head_end = w0 < w1 ? (w0 << 2) : d1;
    for (u64 b = d0 + threadIdx.x; b < head_end; b += kBlock) dst[b] = src[g.src_off + (b - d0)];
    if (w0 < w1)
      for (u64 b = (w1 << 2) + threadIdx.x; b < d1; b += kBlock) dst[b] = src[g.src_off + (b - d0)];
  }
  if (w0 >= w1) return;
  const u64 nw = w1 - w0;
  const u64 a = w0 + nw * k / slices, e = w0 + nw * (k + 1) / slices;
  // destination word w holds source bytes from s = src_off + 4w - d0; with
  // sh = s & 3 it is bytes [sh, sh + 4) of the aligned pair at s - sh
  const u64 s0 = g.src_off + (w0 << 2) - d0;
  const u32 sh = (u32)(s0 & 3);
  const u32* sw = reinterpret_cast<const u32*>(src + (s0 - sh));
  u32* dw = reinterpret_cast<u32*>(dst) + w0;
  for (u64 w = a + threadIdx.x; w < e; w += kBlock) {
    const u64 i = w - w0;
    const u32 lo = sw[i];
    dw[i] = sh ? __builtin_amdgcn_alignbyte(sw[i + 1], lo, sh) : lo;
  }
}

// ---------------------------------------------------------------------------
// launch wrappers

// Fills resident_blocks' cache for every persistent kernel.
void prime_persist_grids() {
  static std::once_flag once;
  std::call_once(once, [] {
    (void)resident_blocks(k_quant_stats);
    (void)resident_blocks(k_flag_scatter2<true>);
    (void)resident_blocks(k_flag_scatter2<false>);
  });
}

void launch_init(const Batch& bt, hipStream_t s) {
  prime_persist_grids();
  const u64 words = (u64)bt.S * (sizeof(SegHist) / 4);
  k_init<<<grid_for(words, kBlock * 4, 16384), kBlock, 0, s>>>(bt);
  LAUNCH_CHECK("k_init");
}

void launch_gather(const uint8_t* src, const GatherSeg* segs, u32 count, u64 bytes,
                   uint8_t* dst, hipStream_t s) {
  if (!count || !bytes) return;
  // about 64 KiB per workgroup, at least one slice per segment
  u64 slices = (bytes / count) >> 16;
  slices = slices < 1 ? 1 : slices > 1024 ? 1024 : slices;
  k_gather<<<(u32)(count * slices), kBlock, 0, s>>>(src, segs, (u32)slices, dst);
  LAUNCH_CHECK("k_gather");
}

// Up to four ranges in one launch (blockIdx.y picks the range).
__global__ void k_copy_ranges(CopyRanges r) {
  const u32 y = blockIdx.y;
  const u32* __restrict__ src = static_cast<const u32*>(r.src[y]);
  u32* __restrict__ dst = static_cast<u32*>(r.dst[y]);
  const u64 n = (r.bytes[y] + 3) / 4;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (u64)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void launch_copy_ranges(const CopyRanges& r, hipStream_t s) {
  if (r.n <= 0) return;
  u64 most = 0;
  for (int i = 0; i < r.n; ++i) most = r.bytes[i] > most ? r.bytes[i] : most;
  const u64 n = (most + 3) / 4;
  if (!n) return;
  k_copy_ranges<<<dim3(grid_for(n, kBlock * 4, 2048), (u32)r.n), kBlock, 0, s>>>(r);
  LAUNCH_CHECK("k_copy_words");
}

void launch_copy_words(const void* src, void* dst, u64 bytes, hipStream_t s) {
  const u64 n = (bytes + 3) / 4;
  if (!n) return;
  k_copy_words<<<grid_for(n, kBlock * 4, 2048), kBlock, 0, s>>>(static_cast<const u32*>(src),
                                                                 static_cast<u32*>(dst), n);
  LAUNCH_CHECK("k_copy_words");
}

void launch_fill_u32(u32* p, u64 count, u32 value, hipStream_t s) {
  if (!count) return;
  k_fill_u32<<<grid_for(count, kBlock * 8, 16384), kBlock, 0, s>>>(p, count, value);
  LAUNCH_CHECK("k_fill_u32");
}

void launch_patch(const Patch* recs, u32 n, const uint8_t* bytes, uint8_t* out, hipStream_t s) {
  if (!n) return;
  k_patch<<<n < 16384 ? n : 16384, 64, 0, s>>>(recs, n, bytes, out);
  LAUNCH_CHECK("k_patch");
}

void launch_split_lo64(const u64* lo64, u32* lo, u32* idx, u32 n, hipStream_t s) {
  k_split_lo64<<<grid_for(n, kBlock * 8, 16384), kBlock, 0, s>>>(lo64, lo, idx, n);
  LAUNCH_CHECK("k_split_lo64");
}

void launch_bbox(const Batch& bt, const float* xyz, hipStream_t s) {
  k_bbox<<<persist_grid(bt), kBlock, 0, s>>>(bt, reinterpret_cast<const F3*>(xyz));
  LAUNCH_CHECK("k_bbox");
  k_offsets<<<grid_for((u64)bt.S * 3, kBlock), kBlock, 0, s>>>(bt, xyz);
  LAUNCH_CHECK("k_offsets");
}

void launch_quant_stats(const Batch& bt, const float* xyz, hipStream_t s) {
  k_quant_stats<<<persist_grid_for(k_quant_stats, bt), kBlock, 0, s>>>(bt, reinterpret_cast<const F3*>(xyz));
  LAUNCH_CHECK("k_quant_stats");
}

void launch_features(const Batch& bt, const float* xyz, SegFeat* feat, hipStream_t s) {
  launch_fill_u32(reinterpret_cast<u32*>(feat), (u64)bt.S * (sizeof(SegFeat) / 4), 0, s);
  k_features<<<persist_grid(bt), kBlock, 0, s>>>(bt, reinterpret_cast<const F3*>(xyz),
                                                        feat);
  LAUNCH_CHECK("k_features");
}

void launch_quantize_nodes(const Batch& bt, const float* xyz, int32_t* nodes, hipStream_t s) {
  k_quantize_nodes<<<bt.ntiles, kBlock, 0, s>>>(bt, reinterpret_cast<const F3*>(xyz), nodes);
  LAUNCH_CHECK("k_quantize_nodes");
}

void launch_keygen(const Batch& bt, const float* xyz, const HilbertSM& hs, u64* key, u32* lo,
                   u32* idx, u64* lo64, hipStream_t s) {
  k_keygen<<<bt.ntiles, kBlock, 0, s>>>(bt, reinterpret_cast<const F3*>(xyz), hs, key, lo, idx,
                                        lo64);
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

u32 radix_passes(u32 bits, u32 hi_bits, bool wide, RadixShape shape) {
  const u32 w = shape.bits;
  return wide ? (21 + w - 1) / w + (hi_bits + w - 1) / w : (bits + w - 1) / w;
}

int radix_sort(const Batch& bt, u64* a, u64* a2, u32* v, u32* v2, u32 bits, RadixShape shape,
               Arena& ar, hipStream_t s) {
  const u32 P = radix_passes(bits, 0, false, shape);
  int cur = 0;
  u32 j = 0;
  for (u32 shift = 0; shift < bits; shift += shape.bits, ++j) {
    if (v) radix_pass<u64, u32, true>(bt, a, a2, v, v2, shift, shape, j, P, ar, s);
    else radix_pass<u64, u32, false>(bt, a, a2, nullptr, nullptr, shift, shape, j, P, ar, s);
    std::swap(a, a2);
    std::swap(v, v2);
    cur ^= 1;
  }
  return cur;
}
int radix_sort_wide(const Batch& bt, u64* hi, u64* hi2, u32* lo, u32* lo2, u32 hi_bits,
                    RadixShape shape, Arena& ar, hipStream_t s) {
  const u32 P = radix_passes(0, hi_bits, true, shape);
  int cur = 0;
  u32 j = 0;
  for (u32 shift = 0; shift < 21; shift += shape.bits, ++j) {
    radix_pass<u32, u64, true>(bt, lo, lo2, hi, hi2, shift, shape, j, P, ar, s);
    std::swap(hi, hi2);
    std::swap(lo, lo2);
    cur ^= 1;
  }
  for (u32 shift = 0; shift < hi_bits; shift += shape.bits, ++j) {
    radix_pass<u64, u32, true>(bt, hi, hi2, lo, lo2, shift, shape, j, P, ar, s);
    std::swap(hi, hi2);
    std::swap(lo, lo2);
    cur ^= 1;
  }
  return cur;
}
int radix_sort_wide_v(const Batch& bt, u64* hi, u64* hi2, u64* lo64, u64* lo64b, u32 hi_bits,
                      RadixShape shape, Arena& ar, hipStream_t s) {
  const u32 P = radix_passes(0, hi_bits, true, shape);
  int cur = 0;
  u32 j = 0;
  for (u32 shift = 0; shift < 21; shift += shape.bits, ++j) {
    radix_pass<u64, u64, true>(bt, lo64, lo64b, hi, hi2, shift, shape, j, P, ar, s);
    std::swap(hi, hi2);
    std::swap(lo64, lo64b);
    cur ^= 1;
  }
  for (u32 shift = 0; shift < hi_bits; shift += shape.bits, ++j) {
    radix_pass<u64, u64, true>(bt, hi, hi2, lo64, lo64b, shift, shape, j, P, ar, s);
    std::swap(hi, hi2);
    std::swap(lo64, lo64b);
    cur ^= 1;
  }
  return cur;
}

void rle_cells(const Batch& bt, const KeyView& in, u32* upos, u32* bpos, u64* ua, u32* ub,
               u32* cnt, Arena& ar, hipStream_t s) {
  const KV k{in.a, in.b, bt.wide};
  const size_t mark = ar.off;
  u32* part = ar.take<u32>(2ull * bt.ntiles);
  u32* total = ar.take<u32>(1);
  k_flag_count2<<<bt.ntiles, kBlock, 0, s>>>(bt, k, part);
  LAUNCH_CHECK("k_flag_count2");
  scan_excl<u32>(part, part, 2ull * bt.ntiles, total, ar, s);
  if (bt.wide)
    k_flag_scatter2<true><<<persist_grid_for(k_flag_scatter2<true>, bt), kBlock, 0, s>>>(bt, k, part, upos, bpos);
  else
    k_flag_scatter2<false><<<persist_grid_for(k_flag_scatter2<false>, bt), kBlock, 0, s>>>(bt, k, part, upos, bpos);
  LAUNCH_CHECK("k_flag_scatter2");
  k_seg_starts2<<<grid_for(bt.S, kBlock), kBlock, 0, s>>>(bt, part, total);
  LAUNCH_CHECK("k_seg_starts2");
  ar.off = mark;
  k_rle_finish<<<persist_grid(bt), kBlock, 0, s>>>(bt, k, upos, ua, ub, cnt);
  LAUNCH_CHECK("k_rle_finish");
}

void block_runs_and_stats(const Batch& bt, const KeyView& kv, bool direct_any, bool block_any,
                          bool rle_done, u32* bpos, Arena& ar, hipStream_t s) {
  (void)direct_any;
  const KV k{kv.a, kv.b, bt.wide};
  if (!rle_done) {
    if (block_any) compact<kFlagBlock>(bt, k, bpos, 1, ar, s);
    k_stats<<<persist_grid(bt), kBlock, 0, s>>>(bt, k);
    LAUNCH_CHECK("k_stats");
  }
  if (block_any) {
    k_block_stats<<<persist_grid(bt), kBlock, 0, s>>>(bt, k, bpos);
    LAUNCH_CHECK("k_block_stats");
  }
}

SecPlan plan_sections(const Batch& bt, Arena& ar) {
  SecPlan p{};
  p.tile_bits = ar.take<u64>((u64)kNumSec * bt.ntiles);
  p.tile_off = ar.take<u64>((u64)kNumSec * bt.ntiles);
  return p;
}

void size_sections(const Batch& bt, const KeyView& kv, const u32* cnt, const u32* bpos,
                   const SecPlan& p, Arena& ar, hipStream_t s) {
  const SecCtx x{KV{kv.a, kv.b, bt.wide}, cnt, bpos};
  const size_t mark = ar.off;
  u64* total = ar.take<u64>(1);
  k_sec_bits<<<bt.ntiles, kBlock, 0, s>>>(bt, x, p.tile_bits);
  LAUNCH_CHECK("k_sec_bits");
  scan_excl<u64>(p.tile_bits, p.tile_off, (u64)kNumSec * bt.ntiles, total, ar, s);
  k_seg_sections<<<grid_for((u64)bt.S * kNumSec, kBlock), kBlock, 0, s>>>(bt, p.tile_off, total);
  LAUNCH_CHECK("k_seg_sections");
  ar.off = mark;
}

void write_sections(const Batch& bt, const KeyView& kv, const u32* cnt, const u32* bpos,
                    const SecPlan& p, u32* out_words, bool any_block, bool any_direct,
                    hipStream_t s, hipEvent_t fixed_done) {
  const SecCtx x{KV{kv.a, kv.b, bt.wide}, cnt, bpos};
  if (any_block) {
    k_sec_write_fixed<<<bt.ntiles, kBlock, 0, s>>>(bt, x, out_words);
    LAUNCH_CHECK("k_sec_write<blk,cnt>");
    if (fixed_done) check_hip(hipEventRecord(fixed_done, s), "hipEventRecord");
    if (bt.wide) k_sec_write_cells<false, true><<<bt.ntiles, kBlock, 0, s>>>(bt, x, p.tile_off, out_words);
    else k_sec_write_cells<false, false><<<bt.ntiles, kBlock, 0, s>>>(bt, x, p.tile_off, out_words);
    LAUNCH_CHECK("k_sec_write<cells>");
  }
  if (any_direct) {
    k_sec_write_cells<true, false><<<bt.ntiles, kBlock, 0, s>>>(bt, x, p.tile_off, out_words);
    LAUNCH_CHECK("k_sec_write<cells_direct>");
  }
}

}  // namespace xnyz
