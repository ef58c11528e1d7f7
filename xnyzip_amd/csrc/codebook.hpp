// codebook.hpp -- Huffman / Exp-Golomb codebooks from device histograms.
//
// The reference builds its codebooks with standard containers
// (source/encoder/huffman_encoder.hpp:33-146): a frequency std::unordered_map
// filled in data order, entropy summed in that map's iteration order, a
// std::priority_queue min-heap on frequency fed in the same order, codes
// assigned left = 0 / right = 1 into a member std::unordered_map whose
// iteration order is the meta entry order -- and whose bucket array survives
// clear(), which matters for the RLE counts stream that is built twice on one
// object (source/composer/block_compressor_rle.hpp:290-303).  Those orders
// decide bytes of the stream, so this builder drives the same containers:
// the kernels deliver, per symbol, its count and the index of its first
// occurrence; replaying the distinct symbols in first-occurrence order
// reproduces the map the reference's `freq[v]++` loop builds.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <queue>
#include <type_traits>
#include <unordered_map>
#include <vector>

namespace xnyz {

struct SymCount {
  uint64_t sym;    // symbol value (as the stream's integer type)
  uint64_t freq;   // occurrences
  uint64_t first;  // index of the first occurrence
};

template <typename T>
class Codebook {
 public:
  struct Code {
    uint32_t bits;
    uint8_t length;
  };
  std::unordered_map<T, Code> code_table;  // kept across build() calls
  bool use_eg = false;
  uint8_t eg_k = 0;
  double entropy = 0.0;

  // huffman_encoder.hpp:33-61 (build) and :91-123 (build_from_freq).
  void build(std::vector<SymCount> syms, uint64_t total) {
    code_table.clear();
    counts_ = syms;
    if (total == 0 || syms.empty()) return;
    std::sort(syms.begin(), syms.end(),
              [](const SymCount& a, const SymCount& b) { return a.first < b.first; });
    std::unordered_map<T, uint64_t> freq;
    for (const auto& s : syms) freq[static_cast<T>(s.sym)] += s.freq;
    double h = 0.0;
    for (auto& kv : freq) {
      const double p = double(kv.second) / total;
      h -= p * std::log2(p);
    }
    entropy = h;
    if (h > 10.0 || use_eg) {
      use_eg = true;
      eg_k = select_eg_k();
      return;
    }
    Node* root = build_tree(freq);
    assign(root, 0u, 0u);
  }

  // huffman_encoder.hpp:126-146
  std::vector<uint8_t> meta() const {
    std::vector<uint8_t> out;
    if (use_eg) return {0xE1, 0xC1, 0x01, 0x01, eg_k};
    put_be32(out, static_cast<uint32_t>(code_table.size()));
    for (auto& kv : code_table) {
      put_symbol(out, kv.first);
      out.push_back(kv.second.length);
      put_be32(out, kv.second.bits);
    }
    return out;
  }

  // Bits the data section takes (huffman_encoder.hpp:150-170, :365-396).
  uint64_t data_bits() const {
    uint64_t bits = 0;
    for (const auto& s : counts_) {
      if (use_eg) {
        bits += s.freq * eg_len(zigzag(static_cast<T>(s.sym)), eg_k);
      } else {
        bits += s.freq * code_table.at(static_cast<T>(s.sym)).length;
      }
    }
    return bits;
  }

  uint8_t max_length() const {
    uint8_t m = 0;
    for (auto& kv : code_table) m = kv.second.length > m ? kv.second.length : m;
    return m;
  }

  static uint64_t zigzag(T v) {
    if constexpr (std::is_signed<T>::value) {
      const int64_t s = int64_t(v);
      return (uint64_t(s) << 1) ^ uint64_t(s >> 63);
    } else {
      return uint64_t(v);
    }
  }
  static uint32_t bit_length(uint64_t v) {  // huffman_encoder.hpp:335-342
    uint32_t n = 0;
    while (v) {
      ++n;
      v >>= 1;
    }
    return n ? n : 1;
  }
  static uint64_t eg_len(uint64_t n, uint8_t k) {
    return 2 * bit_length((n >> k) + 1) - 1 + k;
  }

 private:
  struct Node {
    T symbol{};
    uint64_t freq = 0;
    Node* left = nullptr;
    Node* right = nullptr;
  };
  std::vector<std::unique_ptr<Node>> pool_;
  std::vector<SymCount> counts_;

  Node* build_tree(const std::unordered_map<T, uint64_t>& freq) {
    pool_.clear();
    auto cmp = [](Node* const& a, Node* const& b) { return a->freq > b->freq; };
    std::priority_queue<Node*, std::vector<Node*>, decltype(cmp)> pq(cmp);
    for (auto& kv : freq) {
      pool_.push_back(std::make_unique<Node>());
      pool_.back()->symbol = kv.first;
      pool_.back()->freq = kv.second;
      pq.push(pool_.back().get());
    }
    while (pq.size() > 1) {
      Node* a = pq.top();
      pq.pop();
      Node* b = pq.top();
      pq.pop();
      pool_.push_back(std::make_unique<Node>());
      Node* p = pool_.back().get();
      p->freq = a->freq + b->freq;
      p->left = a;
      p->right = b;
      pq.push(p);
    }
    return pq.top();
  }

  // huffman_encoder.hpp:269-281: a lone leaf gets {0, 1}; the prefix is a
  // uint32_t (deeper codes keep its low 32 bits) and the length a uint8_t.
  void assign(const Node* node, uint32_t prefix, uint8_t len) {
    if (!node->left && !node->right) {
      code_table[node->symbol] = len == 0 ? Code{0, 1} : Code{prefix, len};
      return;
    }
    assign(node->left, prefix << 1, static_cast<uint8_t>(len + 1));
    assign(node->right, (prefix << 1) | 1u, static_cast<uint8_t>(len + 1));
  }

  // huffman_encoder.hpp:344-363: k in 0..16 minimising the total length,
  // the first minimum wins.
  uint8_t select_eg_k() const {
    uint64_t best = std::numeric_limits<uint64_t>::max();
    uint8_t best_k = 0;
    for (uint8_t k = 0; k <= 16; ++k) {
      uint64_t cost = 0;
      for (const auto& s : counts_) cost += s.freq * eg_len(zigzag(static_cast<T>(s.sym)), k);
      if (cost < best) {
        best = cost;
        best_k = k;
      }
    }
    return best_k;
  }

  static void put_be32(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back(uint8_t(v >> 24));
    o.push_back(uint8_t(v >> 16));
    o.push_back(uint8_t(v >> 8));
    o.push_back(uint8_t(v));
  }
  // huffman_encoder.hpp:301-305: 8-byte symbols big-endian, others as their
  // in-memory (little-endian) bytes.
  static void put_symbol(std::vector<uint8_t>& o, T s) {
    if constexpr (sizeof(T) == 8) {
      for (int sh = 56; sh >= 0; sh -= 8) o.push_back(uint8_t(uint64_t(s) >> sh));
    } else {
      const uint8_t* p = reinterpret_cast<const uint8_t*>(&s);
      o.insert(o.end(), p, p + sizeof(T));
    }
  }
};

// golomb_rice_encoder.hpp:92-123: k = floor(log2(mean)) (0 when every value
// is zero or the mean is below 1), capped at 62; the sum is a wrapping u64.
inline uint8_t golomb_rice_k(uint64_t count, uint64_t sum_wrapped, bool all_zero) {
  if (count == 0 || all_zero) return 0;
  const double avg = static_cast<double>(sum_wrapped) / count;
  uint8_t k = avg < 1.0 ? 0 : static_cast<uint8_t>(std::floor(std::log2(avg)));
  return k >= 63 ? 62 : k;
}

}  // namespace xnyz
