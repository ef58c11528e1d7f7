// kernels.hpp -- launch wrappers of the gfx950 kernels (kernels.hip) and the
// parameter blocks they take.  Wrappers enqueue on the given stream and throw
// xnyz::Error on a launch failure.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "common.hpp"

namespace xnyz {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check_hip(hipError_t e, const char* what);

// Optional per-kernel timing: when set (per host thread), every launch
// wrapper calls mark(kernel name, stream) right after enqueueing, so events
// recorded on the launch stream bracket each kernel.
struct Profiler {
  virtual void mark(const char* name, hipStream_t s) = 0;
  virtual ~Profiler() = default;
};
extern thread_local Profiler* g_profiler;

// Bump allocator over one device buffer owned by the context.
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t off = 0;
  template <typename T>
  T* take(size_t count) {
    size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
    if (off + bytes > cap) throw Error(-3, "device scratch arena exhausted");
    T* p = reinterpret_cast<T*>(base + off);
    off += bytes;
    return p;
  }
};

// Morton -> Hilbert digit transform as a finite-state machine over the
// rotation group: state s is a composition of the per-digit permutations of
// hilbert.cpp; apply[s][d] is s applied to digit d, next[s][md] the state
// after emitting digit md, out[md] = mortonToHilbertPerm[md].
struct HilbertSM {
  u8 apply[24 * 8];
  u8 next[24 * 8];
  u8 out[8];
  u8 identity;    // state of the identity permutation
  u8 block_init;  // state after the 15 zero digits 20..6 of an 18-bit code
  u8 nstates;
  u8 pad;
};

struct QuantParams {
  float div;       // quantizer divisor float(double(2b)/sqrt(5 or 3))
  float b;         // L2 bound (recover, diagnostics)
  float rec_div;   // float(sqrt(5)) or float(sqrt(3)) (recover)
  int to;          // 1: truncated octahedron, 0: cube
};

struct KeyParams {
  u64 nx;          // blocks along x: range_x / 128 + 1
  u64 nxny;        // nx * ny
  u64 C;           // id offset of the z = -1 segment: (2^57-1)*nx*ny mod 2^64
  u32 idb;         // bits of the block id field
  u32 L;           // direct mode: Hilbert iteration (bit count of max range)
  int direct;
  int hilbert;
  int wide;        // key split into (u64 hi, u32 lo)
};

// Key view used after sorting (block mode): hi = seg<<idb | id-or-base,
// lo = quad<<18 | sfc18.
struct KeyView {
  const u64* a;
  const u32* b;  // wide keys only
  u32 idb;
  u64 C;
};

void launch_init_ctl(DevCtl* ctl, hipStream_t s);
void launch_bbox(const float* xyz, u32 n, DevCtl* ctl, hipStream_t s);
void launch_quant_stats(const float* xyz, u32 n, const QuantParams& qp, bool diag,
                        DevCtl* ctl, hipStream_t s);
void launch_quantize_nodes(const float* xyz, u32 n, const QuantParams& qp, const DevCtl* ctl,
                           int32_t* nodes, hipStream_t s);
// idx (nullable): point index payload; lo64 (nullable, wide keys only):
// index << 32 | low key word, written instead of lo.
void launch_keygen(const float* xyz, u32 n, const QuantParams& qp, const KeyParams& kp,
                   const HilbertSM& hs, const DevCtl* ctl, u64* key, u32* lo, u32* idx,
                   u64* lo64, hipStream_t s);
void launch_split_lo64(const u64* lo64, u32* lo, u32* idx, u32 n, hipStream_t s);
void launch_m2h(const u64* in, u32 n, u32 iteration, const HilbertSM& hs, u64* out,
                hipStream_t s);

// Exclusive scan of n values; *total (device, nullable) gets the sum.
void scan_excl_u32(const u32* in, u32* out, u64 n, u32* total, Arena& ar, hipStream_t s);
void scan_excl_u64(const u64* in, u64* out, u64 n, u64* total, Arena& ar, hipStream_t s);

// LSD radix sort of n keys over bits [0, bits).  Narrow: u64 keys only.
// Wide: sorts (hi, lo) pairs by lo (lo_bits) then hi (hi_bits).  Results
// end in the returned buffer index (0: a/b, 1: a2/b2).
int radix_sort_narrow(u64* a, u64* a2, u32 n, u32 bits, Arena& ar, hipStream_t s);
int radix_sort_wide(u64* hi, u64* hi2, u32* lo, u32* lo2, u32 n, u32 lo_bits, u32 hi_bits,
                    Arena& ar, hipStream_t s);
// Same with a point-index payload: narrow keys carry v; wide keys carry the
// index in the upper half of lo64 (index << 32 | 21-bit low key).
int radix_sort_narrow_v(u64* a, u64* a2, u32* v, u32* v2, u32 n, u32 bits, Arena& ar,
                        hipStream_t s);
int radix_sort_wide_v(u64* hi, u64* hi2, u64* lo64, u64* lo64b, u32 n, u32 hi_bits, Arena& ar,
                      hipStream_t s);

// Element counts produced on the device (RLE cells, blocks) are passed as
// device pointers (`*_dev`), with a host-side upper bound for grid sizing, so
// no host round trip is needed between the sort and the statistics.

// Positions of the first key of every run of equal keys (RLE cells) or of
// equal block ids; the count goes to *count (device).
void runs_unique(const KeyView& kv, bool wide, u32 n, u32* pos, u32* count, Arena& ar,
                 hipStream_t s);
void runs_blocks(const KeyView& kv, bool wide, u32 n_bound, const u32* n_dev, u32* pos,
                 u32* count, Arena& ar, hipStream_t s);

// RLE: gather the unique keys and per-cell counts; histogram the counts.
void launch_rle_finish(const KeyView& in, bool wide, const u32* upos, const u32* U_dev, u32 N,
                       u64* ua, u32* ub, u32* cnt, DevCtl* ctl, u32* large, u32 large_cap,
                       hipStream_t s);

void launch_stats_block(const KeyView& kv, bool wide, const u32* bpos, const u32* B_dev,
                        u32 n_bound, const u32* n_dev, DevCtl* ctl, hipStream_t s);
void launch_stats_direct(const u64* sfc, u32 n_bound, const u32* n_dev, DevCtl* ctl,
                         hipStream_t s);

// Variable-length sections: phase A (per-tile bit totals, scanned into tile
// offsets; section total into ctl->sec_bits[sec]) and phase B (write).
struct VlPlan {
  u64* tile_bits;
  u64* tile_off;
  u32 ntiles;
};
enum CoderKind { kCoderGRBlock, kCoderGRDirect, kCoderHufBlock, kCoderHufDirect,
                 kCoderQuadHuf, kCoderCntHuf, kCoderCntEG };
struct CoderArgs {
  KeyView kv;
  bool wide;
  const u64* sfc;         // direct mode keys
  const u32* cnt;         // RLE counts
  const HufEntry* lut;    // 256 entries (symbols) or kCntBins entries (counts)
  const u32* big_vals;    // RLE counts >= kCntBins: sorted values ...
  const HufEntry* big_codes;  // ... and their codes
  u32 n_big;
  u32 k;                  // Golomb-Rice or Exp-Golomb parameter
};
VlPlan vl_plan(u32 n, Arena& ar);
void vl_count(CoderKind kind, const CoderArgs& a, u32 n, const VlPlan& p, u64* sec_total,
              Arena& ar, hipStream_t s);
void vl_write(CoderKind kind, const CoderArgs& a, u32 n, const VlPlan& p, u64 sec_byte0,
              u32* out_words, hipStream_t s);

// BitPacker sections: block-id deltas (which = 0) or block-count deltas (1).
void pack_block_deltas(int which, const KeyView& kv, bool wide, const u32* bpos, u32 B,
                       u32 n, u32 width, u64 sec_byte0, u32* out_words, hipStream_t s);

void launch_fill_u32(u32* p, u64 count, u32 value, hipStream_t s);

}  // namespace xnyz
