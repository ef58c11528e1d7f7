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
#include <cstring>
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

  // huffman_encoder.hpp:69-84 (build_complete): the data's symbols, then
  // every symbol 0..alphabet_max it lacks with frequency 1, inserted in
  // ascending order after them; never Exp-Golomb.  The data bits still
  // count the data's symbols only.
  void build_complete(std::vector<SymCount> syms, uint64_t total, uint64_t alphabet_max) {
    code_table.clear();
    use_eg = false;
    eg_k = 0;
    counts_ = syms;
    std::sort(syms.begin(), syms.end(),
              [](const SymCount& a, const SymCount& b) { return a.first < b.first; });
    std::unordered_map<T, uint64_t> freq;
    for (const auto& s : syms) freq[static_cast<T>(s.sym)] += s.freq;
    for (uint64_t s = 0; s <= alphabet_max; ++s)
      if (freq.find(static_cast<T>(s)) == freq.end()) freq[static_cast<T>(s)] = 1;
    (void)total;
    Node* root = build_tree(freq);
    assign(root, 0u, 0u);
  }

  // huffman_encoder.hpp:172-203 (load_meta) on a fresh encoder: the entries
  // inserted into code_table in meta order, so meta() writes them in that
  // map's iteration order, as the reference's get_meta() after load_meta().
  // Returns false for an Exp-Golomb meta or a malformed one.
  bool load_meta(const std::vector<uint8_t>& m, std::vector<SymCount> syms) {
    code_table.clear();
    use_eg = false;
    eg_k = 0;
    counts_ = std::move(syms);
    if (m.size() >= 4 && m[0] == 0xE1 && m[1] == 0xC1 && m[2] == 0x01) return false;
    if (m.size() < 4) return false;
    const uint32_t n = uint32_t(m[0]) << 24 | uint32_t(m[1]) << 16 | uint32_t(m[2]) << 8 | m[3];
    const size_t esz = (sizeof(T) == 8 ? 8 : sizeof(T)) + 5;
    if (4 + (size_t)n * esz > m.size()) return false;
    for (uint32_t i = 0; i < n; ++i) {
      const uint8_t* e = m.data() + 4 + i * esz;
      T sym;
      if constexpr (sizeof(T) == 8) {
        uint64_t v = 0;
        for (int k = 0; k < 8; ++k) v = (v << 8) | e[k];
        sym = static_cast<T>(v);
      } else {
        std::memcpy(&sym, e, sizeof(T));
      }
      const uint8_t* t = e + (sizeof(T) == 8 ? 8 : sizeof(T));
      const uint32_t bits = uint32_t(t[1]) << 24 | uint32_t(t[2]) << 16 | uint32_t(t[3]) << 8 | t[4];
      code_table[sym] = Code{bits, t[0]};
    }
    return true;
  }

  // Whether every symbol the data holds has a code.
  bool covers_data() const {
    for (const auto& s : counts_)
      if (!code_table.count(static_cast<T>(s.sym))) return false;
    return true;
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
