// modes.hpp -- exact stream-size model of the coding modes of one chunk.
//
// The assembled stream of a chunk is fixed by its curve and RLE flag up to
// two encoder choices that the decoder does not need to be told about:
//   * the Golomb-Rice k of the repos / sfc-delta patches (stored in the GR
//     meta, golomb_rice_encoder.hpp:135-139), and
//   * the PFD patch limit of the same values (p_for_delta_encoder.hpp:27-43
//     uses 254; its decoder, :45-57, only looks for the patch symbol 0).
// For width w (1..8) the limit is 2^w - 2: values v <= limit are Huffman
// symbols v + 1, larger ones are patches (symbol 0 + a GR-coded value).
//
// From the statistics of one pass at the reference limit 254 -- the symbol
// histogram (symbol v + 1 for v <= 254, symbol 0 for the patches), the patch
// count and wrapping sum, and bitpop[b] = number of patches with bit b set --
// every (k, w) costs out exactly:
//   patches(w)     = hist[0] + sum_{lim(w) < v <= 254} hist[v + 1]
//   sum(v >> k)    = sum_b>=k bitpop[b] << (b - k)  +  sum_{lim(w) < v <= 254} hist[v + 1] (v >> k)
//   GR bits        = patches * (k + 1) + sum(v >> k)       (golomb_rice_encoder.hpp:92-148)
//   Huffman bits   = optimal prefix-code cost of the folded histogram
//                    (huffman_encoder.hpp:91-123 builds a Huffman tree; any
//                    tie-breaking gives the same total), meta 4 + 13 K bytes
//                    for K symbols (u64 symbols, huffman_encoder.hpp:126-146).
#pragma once

#include <cstdint>
#include <functional>
#include <queue>
#include <vector>

#include "codebook.hpp"  // golomb_rice_k: the reference's k rule

namespace xnyz {

// Total bits of a Huffman code over `freqs` (zero entries ignored); a single
// symbol costs one bit per occurrence (huffman_encoder.hpp:269-281).
inline uint64_t huffman_bits(const std::vector<uint64_t>& freqs) {
  std::priority_queue<uint64_t, std::vector<uint64_t>, std::greater<uint64_t>> q;
  for (uint64_t f : freqs)
    if (f) q.push(f);
  if (q.empty()) return 0;
  if (q.size() == 1) return q.top();
  uint64_t total = 0;
  while (q.size() > 1) {
    const uint64_t a = q.top();
    q.pop();
    const uint64_t b = q.top();
    q.pop();
    total += a + b;
    q.push(a + b);
  }
  return total;
}

inline uint32_t pfd_limit(int width) { return (1u << width) - 2u; }

struct ModeStats {
  uint64_t hist[256];    // symbol counts at limit 254
  uint64_t bitpop[64];   // over the values > 254
  uint64_t patch_sum;    // wrapping sum of the values > 254
  uint64_t fixed_bytes;  // every other byte of the stream (header, trailer,
                         // block / quads / RLE sections with their headers)
};

struct ModeCost {
  uint64_t bytes;  // whole stream
  uint64_t patches;
  uint64_t patch_sum;  // wrapping, for the reference k rule
};

// The Huffman section (with its two size fields) and the patch set at width w.
struct WidthCost {
  uint64_t huf_bytes, patches, patch_sum;
};

inline WidthCost width_cost(const ModeStats& m, int w) {
  const uint32_t lim = pfd_limit(w);
  uint64_t np = m.hist[0], psum = m.patch_sum;
  std::vector<uint64_t> folded;
  folded.reserve(256);
  for (uint32_t v = 0; v <= 254; ++v) {
    const uint64_t c = m.hist[v + 1];
    if (!c) continue;
    if (v > lim) {
      np += c;
      psum += c * v;
    } else {
      folded.push_back(c);
    }
  }
  if (np) folded.push_back(np);
  const uint64_t k_syms = folded.size();
  return WidthCost{16 + (k_syms ? 4 + 13 * k_syms : 0) + (huffman_bits(folded) + 7) / 8, np,
                   psum};
}

// Sizes at or above this are out of reach (no stream of that size can be
// built); the model saturates there instead of wrapping.
constexpr uint64_t kHugeBytes = 1ull << 62;

// The Golomb-Rice section (with its two size fields) at width w and k.  The
// unary part is summed in 128 bits and saturates at kHugeBytes: a patch near
// 2^63 (the direct-mode delta of a wrapped z = -1 node) costs ~2^63 bits at
// k = 0, which a 64-bit sum would wrap into a small, wrong size.
inline uint64_t gr_bytes(const ModeStats& m, const WidthCost& wc, int w, int k) {
  if (!wc.patches) return 16;
  const uint32_t lim = pfd_limit(w);
  unsigned __int128 shifted = 0;
  for (int b = k; b < 64; ++b) shifted += (unsigned __int128)m.bitpop[b] << (b - k);
  for (uint32_t v = lim + 1; v <= 254; ++v) shifted += (unsigned __int128)m.hist[v + 1] * (v >> k);
  const unsigned __int128 bits = (unsigned __int128)wc.patches * (uint64_t)(k + 1) + shifted;
  const unsigned __int128 bytes = 16 + 5 + (bits + 7) / 8;
  return bytes >= kHugeBytes ? kHugeBytes : (uint64_t)bytes;
}

// Stream bytes with Golomb-Rice k and PFD width w.
inline ModeCost mode_cost(const ModeStats& m, int k, int w) {
  const WidthCost wc = width_cost(m, w);
  return ModeCost{m.fixed_bytes + wc.huf_bytes + gr_bytes(m, wc, w, k), wc.patches,
                  wc.patch_sum};
}

struct ModeChoice {
  int k, w;
  uint64_t bytes;
};

// The reference's choice (k by its rule, w = 8) and the exhaustive optimum
// over w in 1..8 and k in 0..62 (every k the Golomb-Rice meta can carry,
// golomb_rice_encoder.hpp:121-123 clamps to 62); ties keep the smaller w,
// then the smaller k.
inline void search_modes(const ModeStats& m, ModeChoice& ref, ModeChoice& best) {
  best = ModeChoice{0, 0, ~0ull};
  for (int w = 1; w <= 8; ++w) {
    const WidthCost wc = width_cost(m, w);
    const int rule_k = golomb_rice_k(wc.patches, wc.patch_sum, false);
    if (w == 8) ref = ModeChoice{rule_k, 8, m.fixed_bytes + wc.huf_bytes + gr_bytes(m, wc, 8, rule_k)};
    auto consider = [&](int k) {
      const uint64_t b = m.fixed_bytes + wc.huf_bytes + gr_bytes(m, wc, w, k);
      if (b < best.bytes) best = ModeChoice{k, w, b};
    };
    if (!wc.patches) {  // no GR data: k only reaches the meta, keep the rule's
      consider(rule_k);
      continue;
    }
    for (int k = 0; k <= 62; ++k) consider(k);
  }
}

}  // namespace xnyz
