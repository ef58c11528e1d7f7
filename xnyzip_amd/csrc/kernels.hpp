// kernels.hpp -- the gfx950 kernels (kernels.hip) as one segmented pipeline:
// S clouds ("segments") of uneven size are processed together, every launch
// covering all of them.  Work is split into tiles of kTile points; a tile
// belongs to one segment, and the tile table maps a workgroup to its
// segment and its index inside the segment.  Each stage then derives the
// tile's range in the array it works on (points, RLE cells, or blocks).
// Launch wrappers enqueue on the given stream and throw xnyz::Error.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "common.hpp"
#include "exact_div.hpp"
#include "fround.hpp"

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

constexpr u32 kTile = 4096;  // points (cells, blocks) per tile, every stage

// The no-allocation entry point's device scratch (noalloc.hip): one array of
// this many bytes per GPU, held from the library's load until the process
// ends.  noalloc_scratch(device) is its device address on `device`, looked up
// once per device.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
void* noalloc_scratch(int device);

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
  float div;      // quantizer divisor float(double(2b)/sqrt(5 or 3))
  float b;        // L2 bound (recover, diagnostics)
  float rec_div;  // float(sqrt(5)) or float(sqrt(3)) (recover)
  int to;         // 1: truncated octahedron, 0: cube
  float inv;      // RN(1/div)      (div_const)
  float rec_inv;  // RN(1/rec_div)  (div_const)
  int fast;       // div_const_ok(div): divide by multiplication (exact_div.hpp)
};

struct KeyParams {
  u64 nx;    // blocks along x: range_x / 128 + 1
  u64 nxny;  // nx * ny
  u64 C;     // id offset of the z = -1 segment: (2^57-1)*nx*ny mod 2^64
  u32 idb;   // bits of the block id field
  u32 L;     // direct mode: Hilbert iteration (bit count of max range)
  int direct;
  int hilbert;
};

// Variable-length / fixed-width sections of one stream, in stream order.
enum SecKind : u32 {
  kSecBlk = 0,   // S1 BitPacker: block-id deltas            (block mode)
  kSecCnt = 1,   // S2 BitPacker: block-count deltas         (block mode)
  kSecGR = 2,    // S3 Golomb-Rice: repos / sfc-delta patches
  kSecHuf = 3,   // S4 Huffman: repos / sfc-delta PFD symbols
  kSecQuad = 4,  // S6 Huffman: quads PFD symbols            (block mode)
  kSecRle = 5,   // S7 Huffman / Exp-Golomb: RLE counts      (RLE)
  kNumSec = 6
};

// Per-segment parameters (host-written, uploaded before each stage).
struct SegParam {
  u32 p0, n;          // points [p0, p0 + n) of the concatenated input
  u32 tile0, ntiles;  // the segment's tiles
  QuantParams qp;
  KeyParams kp;       // after the range scan
  u32 rle, diag;
  u32 grk;            // Golomb-Rice k                 (after statistics)
  u32 cnt_eg, egk;    // RLE counts: Exp-Golomb and its k
  u32 w1, w2;         // BitPacker widths
  u32 big0, nbig;     // RLE counts >= kCntBinsSeg: range in the big tables
  u32 lim;            // PFD patch threshold of repos / sfc deltas: v > lim is a
                      // patch (254 = p_for_delta_encoder.hpp:31; lower values
                      // only come from a mode override)
  u32 search;         // also collect SegHist::bitpop (mode search)
  u32 passes;         // key sort: radix passes this segment's keys need (0: all);
                      // later passes leave its keys in place
  u64 sec_byte[kNumSec];  // byte offset of each data section (after sizing)
};

constexpr u32 kCntBinsSeg = 256;  // direct-indexed RLE count bins per segment

// Per-segment device reductions and counters.
struct SegCtl {
  u64 minkey[3];     // per-axis minimum: (ordered float << 32) | index
  float offset[3];   // the shift (exact float of the minimum, FLT_MAX if none)
  u32 wrapped;       // nodes with z == -1
  int range[3];      // max quantized x, y, z, from 0
  u32 pad0;
  u64 max_err2_bits; // RLE diagnostic: max squared error (double bits)
  double sum_err2;   // RLE diagnostic: sum of squared errors
  u32 e0, ne;        // coded elements: RLE cells [e0, e0+ne) (or the points)
  u32 b0, nb;        // blocks [b0, b0+nb)
  u32 n_patches, n_qpatches;
  u64 patch_sum;     // wrapping u64 sum of the patches
  u64 max_blkd, max_cntd;
  u64 sec_bits[kNumSec];   // bits of each section
  u64 sec_bit0[kNumSec];   // scan offset of the segment's first tile
};


struct SegHist {  // per-segment symbol counts and first occurrences
  u32 sym[kSymBins], first_sym[kSymBins];
  u32 quad[kSymBins], first_quad[kSymBins];
  u32 cnt[kCntBinsSeg], first_cnt[kCntBinsSeg];
  // Over the repos (or sfc-delta) values v > 254: bitpop[b] = how many have
  // bit b set, so sum(v >> k) = sum_{b >= k} bitpop[b] << (b - k) exactly,
  // for every Golomb-Rice k.
  u64 bitpop[64];
};

// Per-segment statistics for the learned mode selector, from one pass over
// the points in input order after the range scan: HyperLogLog registers
// (256 each) of the lattice cells, the 64^3 sub-blocks and the 128^3 blocks,
// and, over consecutive input points, a histogram of bitlen(m_i xor m_i-1)
// of their 63-bit node Morton codes plus how often both share a cell,
// sub-block or block.
constexpr u32 kHll = 256;
struct SegFeat {
  u32 hll_cell[kHll], hll_sub[kHll], hll_blk[kHll];
  u32 xor_hist[16];  // bin = min(15, bitlen / 4)
  u32 same_cell, same_sub, same_blk, pairs;
};

struct SegLut {   // per-segment codebooks (host-built)
  HufEntry sym[kSymBins];
  HufEntry quad[kSymBins];
  HufEntry cnt[kCntBinsSeg];
};

struct Tile {
  u32 seg;
  u32 local;
};

// Device pointers shared by the stages of one call.
struct Batch {
  u32 S;           // segments
  u32 ntiles;      // tiles in the table
  u32 N;           // points in all segments
  const Tile* tiles;
  const SegParam* sp;
  SegCtl* sc;
  SegHist* sh;
  const SegLut* lut;
  const u32* big_vals;        // sorted per segment
  const HufEntry* big_codes;
  u32* large;                 // (segment, count, cell) triples of large RLE counts
  u32* n_large;
  u32 large_cap;
  u32 grid_cap;               // host side: workgroups of the persistent kernels (0: 2048)
  bool wide;                  // keys split into (u64 hi, u32 lo)
};

// Sorted keys (block mode: hi = seg<<idb | id-or-base, lo = quad<<18 | sfc).
struct KeyView {
  const u64* a;
  const u32* b;  // wide keys only
};

void launch_init(const Batch& bt, hipStream_t s);
void launch_bbox(const Batch& bt, const float* xyz, hipStream_t s);
void launch_quant_stats(const Batch& bt, const float* xyz, hipStream_t s);
// Selector statistics of every segment (after launch_quant_stats).
void launch_features(const Batch& bt, const float* xyz, SegFeat* feat, hipStream_t s);
void launch_quantize_nodes(const Batch& bt, const float* xyz, int32_t* nodes, hipStream_t s);
// idx (nullable): point index payload; lo64 (nullable, wide keys only):
// index << 32 | low key word, written instead of lo.
void launch_keygen(const Batch& bt, const float* xyz, const HilbertSM& hs, u64* key, u32* lo,
                   u32* idx, u64* lo64, hipStream_t s);
void launch_split_lo64(const u64* lo64, u32* lo, u32* idx, u32 n, hipStream_t s);
void launch_m2h(const u64* in, u32 n, u32 iteration, const HilbertSM& hs, u64* out,
                hipStream_t s);

void scan_excl_u32(const u32* in, u32* out, u64 n, u32* total, Arena& ar, hipStream_t s);
void scan_excl_u64(const u64* in, u64* out, u64 n, u64* total, Arena& ar, hipStream_t s);

// Segmented LSD radix sort of each segment's keys over bits [0, bits);
// narrow: u64 keys (+ optional u32 payload); wide: (hi u64, lo u32) by lo then
// hi (+ optional index payload folded into lo64).  Returns the buffer index
// holding the result (0: first arrays, 1: second).
// Radix sort shape: digit width (6, 7 or 8 bits) and radix tiles per tile of
// the tile table (1 or 2).
struct RadixShape {
  u32 bits = 8;
  u32 sub = 1;
};

// Digit passes of a key sort: narrow keys of `bits` bits, or wide keys (a
// 21-bit low word, then `hi_bits` of the high word).
u32 radix_passes(u32 bits, u32 hi_bits, bool wide, RadixShape shape);
int radix_sort(const Batch& bt, u64* a, u64* a2, u32* v, u32* v2, u32 bits, RadixShape shape,
               Arena& ar, hipStream_t s);
int radix_sort_wide(const Batch& bt, u64* hi, u64* hi2, u32* lo, u32* lo2, u32 hi_bits,
                    RadixShape shape, Arena& ar, hipStream_t s);
int radix_sort_wide_v(const Batch& bt, u64* hi, u64* hi2, u64* lo64, u64* lo64b, u32 hi_bits,
                      RadixShape shape, Arena& ar, hipStream_t s);

// RLE: first position of every run of equal keys per segment, then the cells
// (keys) and their counts, segment by segment, with count histograms.
// Also the block starts (bpos: each block's first cell) and the cells' symbol
// statistics, found in the same pass over the sorted points.
void rle_cells(const Batch& bt, const KeyView& in, u32* upos, u32* bpos, u64* ua, u32* ub,
               u32* cnt, Arena& ar, hipStream_t s);
// Block runs over the coded elements and their symbol statistics (unless
// rle_cells found both: rle_done), then the BitPacker widths.
void block_runs_and_stats(const Batch& bt, const KeyView& kv, bool direct_any, bool block_any,
                          bool rle_done, u32* bpos, Arena& ar, hipStream_t s);

// Phase A of every section: per-tile bit counts -> tile offsets; per-segment
// section bits and first-tile offsets into SegCtl.
struct SecPlan {
  u64* tile_bits;  // [section][tile]
  u64* tile_off;   // exclusive scan of tile_bits
};
SecPlan plan_sections(const Batch& bt, Arena& ar);
void size_sections(const Batch& bt, const KeyView& kv, const u32* cnt, const u32* bpos,
                   const SecPlan& p, Arena& ar, hipStream_t s);
// Phase B: write every section at SegParam::sec_byte.  fixed_done (nullable)
// is recorded on s once the BitPacker sections are written, before the
// cell-space sections.
void write_sections(const Batch& bt, const KeyView& kv, const u32* cnt, const u32* bpos,
                    const SecPlan& p, u32* out_words, bool any_block, bool any_direct,
                    hipStream_t s, hipEvent_t fixed_done = nullptr);

// Copies ceil(bytes / 4) words between device memory and pinned host memory
// with a kernel (both pointers 4-byte aligned; the last word may read and
// write up to 3 bytes past `bytes`, inside the buffers' slack).
void launch_copy_words(const void* src, void* dst, u64 bytes, hipStream_t s);
// The same for up to four (src, dst, bytes) ranges in one launch.
struct CopyRanges {
  const void* src[4];
  void* dst[4];
  u64 bytes[4];
  int n;
};
void launch_copy_ranges(const CopyRanges& r, hipStream_t s);
void launch_fill_u32(u32* p, u64 count, u32 value, hipStream_t s);

// Segment s of a gather: `len` bytes from src + src_off to dst + dst_off.
struct GatherSeg {
  u64 src_off, dst_off, len;
};
// Packs `count` segments (their source offsets 256-byte aligned, the source
// readable 8 bytes past each segment's end); `bytes` is their total.
void launch_gather(const uint8_t* src, const GatherSeg* segs, u32 count, u64 bytes,
                   uint8_t* dst, hipStream_t s);

// Byte runs copied into the stream buffer: out[dst, dst + len) =
// bytes[src, src + len) per record.
struct Patch {
  u64 dst;
  u32 src, len;
};
void launch_patch(const Patch* recs, u32 n, const uint8_t* bytes, uint8_t* out, hipStream_t s);

}  // namespace xnyz
