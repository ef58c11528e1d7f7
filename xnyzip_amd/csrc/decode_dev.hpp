// decode_dev.hpp -- device decoder (decode_dev.hip): parallel decode of the
// entropy-coded sections and the LCP / direct reconstruction of the points.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace xnyz {

enum CodeKind : u32 { kCodeHuf = 0, kCodeEG = 1, kCodeGR = 2, kCodeConst = 3 };

constexpr u32 kSub = 512;      // bits per sub-sequence of the parallel decode (the shortest)
// Golomb-Rice codes (unary quotient, then k remainder bits) fall into step
// slowly after a wrong start -- their lengths hardly differ -- so their
// sub-sequences are longer: fewer boundaries to settle, in fewer rounds.
constexpr u32 kSubGR = 2048;
constexpr u32 kLutBits = 11;   // Huffman first-level lookup: at most 2^11 entries
constexpr u32 kTreeNull = 0xFFFFFFFFu;
constexpr u32 kTreeLeaf = 0x80000000u;

// One entropy-coded section: where its bits are and how to read a code word.
struct CodeDesc {
  u64 base_bit;  // first bit of the section's data in the stream buffer
  u64 bits;      // data bits (data bytes * 8)
  u32 kind;      // CodeKind
  u32 k;         // Exp-Golomb / Golomb-Rice k
  u32 lb;        // Huffman lookup bits
  u32 zigzag;    // signed symbols (int RLE counts): zigzag-decode the value
  const u32* lut;     // Huffman: 2^lb entries, (length << 24 | value index), or
                      // (0 << 24 | tree node) for codes longer than lb bits
  const u32* tree;    // Huffman: tree[2 * node + bit] = child node, or
                      // kTreeLeaf | value index, or kTreeNull
  const u64* values;  // Huffman: symbol of each value index
  u32 direct_values;  // Huffman: the lookup table and tree carry the symbols
                      // themselves (every symbol below 2^24), not indices
  u64 const_value;    // kCodeConst: a one-entry code table (every symbol equal)
};

struct DecScratch {
  u32* host_flag;  // pinned word for the round flag and totals
};

// Decodes `count` values of one section into out (device memory).
template <typename T>
void decode_section(const uint8_t* stream, const CodeDesc& c, u64 count, T* out, DecScratch& w,
                    Arena& ar, hipStream_t s);

// Morton <- Hilbert digit machine (hilbertToMorton3, hilbert.cpp:396-421,
// run from the top digit down): apply[s][h] is the Morton digit of Hilbert
// digit h in state s, next[s][h] the state for the digits below.
struct HilbertInv {
  u8 apply[24 * 8];
  u8 next[24 * 8];
};
HilbertInv build_hilbert_inv();

// BitPacker sections (bitpacking_encoder.hpp:17-51): n values of w bits,
// LSB-first, from bit `base_bit`.
void launch_bitunpack(const uint8_t* stream, u64 base_bit, u32 w, u64 n, u64* out, hipStream_t s);
// a[i] += b[i]
void launch_add_u64(u64* a, const u64* b, u64 n, hipStream_t s);
// flags[start[i]] = 1 for i < n (flags zeroed by the caller)
void launch_mark_heads(const u64* start, u64 n, uint8_t* flags, hipStream_t s);
void launch_zero_u8(uint8_t* p, u64 n, hipStream_t s);

// Reconstruction of the points from the decoded streams.
struct LcpArgs {
  u64 n;                    // coded elements (points, or RLE cells)
  int direct, hilbert, to, rle;
  u32 bits;                 // direct mode: Hilbert iteration
  u64 nx, ny;               // block mode: blocks per axis
  float b, rec_div;         // L2 bound, float(sqrt(5)) or float(sqrt(3))
  float rec_inv;            // RN(1/rec_div) (div_const, exact_div.hpp)
  float off[3];             // shift offset
  const uint8_t* sym;       // PFD symbols of repos / sfc deltas
  const uint8_t* quad;      // PFD symbols of quads (block mode)
  const uint8_t* head;      // block-start flags (block mode)
  const u64* ps;            // exclusive prefix sums of the patch values (P + 1)
  u64 np;                   // P: patch values
  u64 total;                // points written (RLE: sum of the counts)
  const u64* blkst;         // block ids (block mode)
  const u64* out_at;        // RLE: first output point of each cell
  const u64* counts;        // RLE: copies of each cell
  float* out;               // points, x y z interleaved
};
// Scratch: lcp_scratch_bytes(n) bytes.  The scan (the per-tile prefix
// states) needs the symbols, quads and block-start flags; the apply pass
// also the block ids, patch sums and RLE offsets, and the scan's scratch.
size_t lcp_scratch_bytes(u64 n);
void launch_lcp_scan(const LcpArgs& a, void* scratch, hipStream_t s);
void launch_lcp_apply(const LcpArgs& a, const HilbertInv& hi, void* scratch, hipStream_t s);
// out[i] = hilbertToMorton3(in[i], iteration) for n codes (device arrays).
void launch_h2m(const u64* in, u64 n, u32 iteration, const HilbertInv& hi, u64* out,
                hipStream_t s);

}  // namespace xnyz
