// decode.cpp -- xnyz_decompress: frame -> points, on the host.
//
// Replaces BlockDecompressor<float>::decompress and
// BlockDecompressorRLE<float>::decompress
// (source/composer/block_decompressor.hpp:31-192,
// block_decompressor_rle.hpp:32-207).  Decoding is not part of the measured
// compression path; it runs here on the host (the Huffman and Golomb-Rice
// streams carry no restart points, so a stream is one sequential decode).
// The CLI uses it to write the decompressed file and check the error bound.
#include <zstd.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/xnyz.h"
#include "kernels.hpp"

namespace xnyz {

// hilbert_sm.cpp: Hilbert -> Morton with the same tables as the encoder.
uint64_t hilbert_to_morton3(uint64_t h, unsigned iteration);

namespace {

struct Reader {
  const uint8_t* p;
  const uint8_t* end;
  template <typename T>
  T get() {
    if (size_t(end - p) < sizeof(T)) throw Error(XNYZ_ECORRUPT, "truncated stream");
    T v;
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    return v;
  }
  const uint8_t* take(uint64_t n) {
    if (uint64_t(end - p) < n) throw Error(XNYZ_ECORRUPT, "truncated section");
    const uint8_t* q = p;
    p += n;
    return q;
  }
};

struct Sec {
  const uint8_t* meta;
  uint64_t ml;
  const uint8_t* data;
  uint64_t dl;
};

Sec section(Reader& r) {  // [u64 data size][u64 meta size][meta][data]
  Sec s;
  s.dl = r.get<uint64_t>();
  s.ml = r.get<uint64_t>();
  s.meta = r.take(s.ml);
  s.data = r.take(s.dl);
  return s;
}

uint32_t le32(const uint8_t* m) {
  return uint32_t(m[0]) | uint32_t(m[1]) << 8 | uint32_t(m[2]) << 16 | uint32_t(m[3]) << 24;
}
uint32_t be32(const uint8_t* m) {
  return uint32_t(m[0]) << 24 | uint32_t(m[1]) << 16 | uint32_t(m[2]) << 8 | uint32_t(m[3]);
}

// bitpacking_encoder.hpp:53-80
std::vector<uint64_t> bit_unpack(const Sec& s) {
  if (s.ml != 5) throw Error(XNYZ_ECORRUPT, "BitPacker meta must be 5 bytes");
  const uint32_t n = le32(s.meta);
  const uint32_t w = s.meta[4];
  if (w > 64 || (uint64_t)n * w > s.dl * 8) throw Error(XNYZ_ECORRUPT, "BitPacker data too short");
  std::vector<uint64_t> v(n);
  uint64_t bit = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t x = 0;
    for (uint32_t b = 0; b < w; ++b, ++bit)
      if (s.data[bit >> 3] & (1u << (bit & 7))) x |= 1ull << b;
    v[i] = x;
  }
  return v;
}

// golomb_rice_encoder.hpp:150-193
std::vector<uint64_t> gr_unpack(const Sec& s) {
  if (s.ml == 0 && s.dl == 0) return {};
  if (s.ml != 5) throw Error(XNYZ_ECORRUPT, "Golomb-Rice meta must be 5 bytes");
  const uint32_t n = le32(s.meta);
  const uint32_t k = s.meta[4];
  if (k >= 63) throw Error(XNYZ_ECORRUPT, "Golomb-Rice k too large");
  std::vector<uint64_t> v;
  v.reserve(n);
  uint64_t bit = 0;
  const uint64_t total = s.dl * 8;
  auto rd = [&]() -> bool {
    if (bit >= total) throw Error(XNYZ_ECORRUPT, "Golomb-Rice data too short");
    const bool b = (s.data[bit >> 3] >> (bit & 7)) & 1;
    ++bit;
    return b;
  };
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t q = 0;
    while (rd()) ++q;
    uint64_t r = 0;
    for (uint32_t b = 0; b < k; ++b)
      if (rd()) r |= 1ull << b;
    v.push_back((q << k) + r);
  }
  return v;
}

// huffman_encoder.hpp:172-243 (Huffman) and :398-442 (Exp-Golomb).  Symbols
// are decoded MSB-first; a code table of one entry repeats that symbol.
template <typename T>
std::vector<T> huffman_decode(const Sec& s, uint64_t count) {
  std::vector<T> out;
  out.reserve(count);
  const uint8_t* m = s.meta;
  if (s.ml >= 4 && m[0] == 0xE1 && m[1] == 0xC1 && m[2] == 0x01 && m[3] == 0x01) {
    const uint32_t k = s.ml > 4 ? m[4] : 0;
    uint64_t bit = 0;
    const uint64_t total = s.dl * 8;
    auto rd = [&](uint32_t& b) -> bool {
      if (bit >= total) return false;
      b = (s.data[bit >> 3] >> (7 - (bit & 7))) & 1;
      ++bit;
      return true;
    };
    while (out.size() < count) {
      uint32_t b = 0, lz = 0;
      for (;;) {
        if (!rd(b)) return out;
        if (b) break;
        ++lz;
      }
      uint64_t n1 = 1;
      for (uint32_t i = 0; i < lz; ++i) {
        if (!rd(b)) return out;
        n1 = (n1 << 1) | b;
      }
      uint64_t low = 0;
      for (uint32_t i = 0; i < k; ++i) {
        if (!rd(b)) return out;
        low |= uint64_t(b) << i;
      }
      const uint64_t n = ((n1 - 1) << k) | low;
      if constexpr (std::is_signed<T>::value)
        out.push_back(T(int64_t((n >> 1) ^ (~(n & 1) + 1))));
      else
        out.push_back(T(n));
    }
    return out;
  }
  if (s.ml < 4) throw Error(XNYZ_ECORRUPT, "Huffman meta too short");
  const uint32_t entries = be32(m);
  const uint64_t esz = (sizeof(T) == 8 ? 8 : sizeof(T)) + 5;
  if (4 + entries * esz > s.ml) throw Error(XNYZ_ECORRUPT, "Huffman meta truncated");
  std::unordered_map<uint64_t, T> codes;  // (len << 32 | bits) -> symbol
  T only{};
  for (uint32_t i = 0; i < entries; ++i) {
    const uint8_t* e = m + 4 + i * esz;
    T sym;
    if constexpr (sizeof(T) == 8) {
      uint64_t v = 0;
      for (int k = 0; k < 8; ++k) v = (v << 8) | e[k];
      sym = T(v);
    } else {
      std::memcpy(&sym, e, sizeof(T));
    }
    const uint8_t* t = e + (sizeof(T) == 8 ? 8 : sizeof(T));
    codes[(uint64_t(t[0]) << 32) | be32(t + 1)] = sym;
    only = sym;
  }
  if (entries == 1) {
    out.assign(count, only);
    return out;
  }
  uint32_t buf = 0;
  uint8_t len = 0;
  for (uint64_t i = 0; i < s.dl * 8 && out.size() < count; ++i) {
    buf = (buf << 1) | ((s.data[i >> 3] >> (7 - (i & 7))) & 1);
    ++len;
    auto it = codes.find((uint64_t(len) << 32) | buf);
    if (it != codes.end()) {
      out.push_back(it->second);
      buf = 0;
      len = 0;
    }
  }
  return out;
}

template <typename T>
std::vector<T> pfd_decode(const std::vector<T>& deltas, const std::vector<uint64_t>& patches) {
  std::vector<T> v;
  v.reserve(deltas.size());
  size_t pi = 0;
  for (T d : deltas) {
    if (d == 0) {
      if (pi >= patches.size()) throw Error(XNYZ_ECORRUPT, "missing PFD patch");
      v.push_back(T(patches[pi++]));
    } else {
      v.push_back(T(d - 1));
    }
  }
  return v;
}

uint32_t compact3(uint64_t c) {  // z_order_curve.hpp:32-40
  uint64_t x = c & 0x1249249249249249ull;
  x = (x | (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x | (x >> 4)) & 0x100f00f00f00f00full;
  x = (x | (x >> 8)) & 0x1f0000ff0000ffull;
  x = (x | (x >> 16)) & 0x1f00000000ffffull;
  x = (x | (x >> 32)) & 0x1fffff;
  return uint32_t(x);
}

struct Cell {
  int x, y, z;
};

std::vector<float> decode_buffer(const uint8_t* buf, uint64_t len, int quantizer, float b,
                                 bool rle) {
  Reader r{buf, buf + len};
  const uint64_t num = r.get<uint64_t>();
  const uint64_t params = r.get<uint64_t>();
  for (uint64_t i = 0; i < params; ++i) (void)r.get<float>();
  const int32_t rx = r.get<int32_t>(), ry = r.get<int32_t>(), rz = r.get<int32_t>();
  const bool hil = r.get<uint8_t>() != 0;
  const bool direct = r.get<uint8_t>() != 0;
  std::vector<Cell> cells;
  if (direct) {  // block_decompressor.hpp:62-102
    const auto patches = gr_unpack(section(r));
    const auto d = pfd_decode(huffman_decode<uint64_t>(section(r), num), patches);
    if (d.size() != num) throw Error(XNYZ_ECORRUPT, "short SFC stream");
    const int mr = std::max({rx, ry, rz});
    unsigned bits = 0;
    for (int t = mr; t > 0; t >>= 1) ++bits;
    if (!bits) bits = 1;
    uint64_t acc = 0;
    cells.resize(num);
    for (uint64_t i = 0; i < num; ++i) {
      acc = i ? acc + d[i] : d[0];
      const uint64_t m = hil ? hilbert_to_morton3(acc, bits) : acc;
      cells[i] = Cell{int(compact3(m)), int(compact3(m >> 1)), int(compact3(m >> 2))};
    }
  } else {  // block_decompressor.hpp:103-170, block_quantizer.hpp:153-207
    auto blk = bit_unpack(section(r));
    auto cnt = bit_unpack(section(r));
    const auto rp = gr_unpack(section(r));
    const auto repos = pfd_decode(huffman_decode<uint64_t>(section(r), num), rp);
    const auto qp = gr_unpack(section(r));
    const auto quads = pfd_decode(huffman_decode<uint8_t>(section(r), num), qp);
    for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
    for (size_t i = 1; i < blk.size(); ++i) blk[i] += blk[i - 1];
    const uint64_t nx = uint64_t(rx) / 128 + 1, ny = uint64_t(ry) / 128 + 1;
    cells.reserve(repos.size());
    size_t j = 0;
    for (size_t i = 0; i < blk.size(); ++i) {
      const uint64_t pbx = (blk[i] % nx * 2) * 64, pby = (blk[i] / nx % ny * 2) * 64,
                     pbz = (blk[i] / nx / ny * 2) * 64;
      uint64_t prequad = 0, prerepos = 0;
      for (uint64_t k = 0; k < cnt[i]; ++k, ++j) {
        if (j >= repos.size() || j >= quads.size()) throw Error(XNYZ_ECORRUPT, "short streams");
        if (quads[j] != 0) prerepos = 0;
        const uint64_t rj = repos[j] + prerepos, qj = quads[j] + prequad;
        prerepos = rj;
        prequad = qj;
        const uint64_t m = hil ? hilbert_to_morton3(rj, 20) : rj;
        const uint64_t ix = pbx + (qj & 1) * 64 + compact3(m);
        const uint64_t iy = pby + ((qj >> 1) & 1) * 64 + compact3(m >> 1);
        const uint64_t iz = pbz + ((qj >> 2) & 1) * 64 + compact3(m >> 2);
        cells.push_back(Cell{int(ix), int(iy), int(iz)});
      }
    }
  }
  std::vector<int> counts;
  if (rle) counts = huffman_decode<int>(section(r), num);
  const float ox = r.get<float>(), oy = r.get<float>(), oz = r.get<float>();
  std::vector<float> out;
  const size_t total_pts = rle ? 0 : cells.size();
  out.reserve(total_pts * 3);
  const float s5 = float(std::sqrt(5.0)), s3 = float(std::sqrt(3.0));
  auto emit = [&](Cell c) {  // quantizer recover, then unshift
    float v[3];
    if (quantizer == XNYZ_TRUNCATED_OCTAHEDRON) {
      if (c.z & 1) {
        c.x = (c.x << 1) - 1;
        c.y = (c.y << 1) - 1;
      } else {
        c.x <<= 1;
        c.y <<= 1;
      }
      const int n3[3] = {c.x, c.y, c.z};
      for (int a = 0; a < 3; ++a) v[a] = ((float(n3[a]) * 2.0f) * b) / s5;
    } else {
      const int n3[3] = {c.x, c.y, c.z};
      for (int a = 0; a < 3; ++a) v[a] = (((float(n3[a]) + 0.5f) * 2.0f) * b) / s3;
    }
    out.push_back(v[0] + ox);
    out.push_back(v[1] + oy);
    out.push_back(v[2] + oz);
  };
  if (rle) {  // block_decompressor_rle.hpp:173-187
    for (size_t i = 0; i < cells.size() && i < counts.size(); ++i)
      for (int k = 0; k < counts[i]; ++k) emit(cells[i]);
  } else {
    for (const Cell& c : cells) emit(c);
  }
  return out;
}

}  // namespace
}  // namespace xnyz

namespace xnyz {
int decompress_impl(const uint8_t* frame, uint64_t len, const xnyz_opts* o, xnyz_buf* out) {
  if (!frame || !o || !out) throw Error(XNYZ_EINVAL, "null argument");
  std::vector<uint8_t> raw;
  const uint8_t* buf = frame;
  uint64_t blen = len;
  if (o->frame_mode != XNYZ_FRAME_NONE) {  // source/composer/utils.hpp:13-29
    const unsigned long long fs = ZSTD_getFrameContentSize(frame, len);
    if (fs == ZSTD_CONTENTSIZE_ERROR) throw Error(XNYZ_ECORRUPT, "Not a zstd frame");
    if (fs == ZSTD_CONTENTSIZE_UNKNOWN) throw Error(XNYZ_ECORRUPT, "Original size unknown");
    raw.resize(fs);
    const size_t d = ZSTD_decompress(raw.data(), fs, frame, len);
    if (ZSTD_isError(d)) throw Error(XNYZ_ECORRUPT, ZSTD_getErrorName(d));
    if (d != fs) throw Error(XNYZ_ECORRUPT, "Decompressed size mismatch");
    buf = raw.data();
    blen = raw.size();
  }
  std::vector<float> pts = decode_buffer(buf, blen, o->quantizer, o->l2_bound, o->rle != 0);
  const size_t bytes = pts.size() * sizeof(float);
  uint8_t* p = static_cast<uint8_t*>(std::malloc(bytes + 16));
  if (!p) throw Error(XNYZ_ENOMEM, "malloc failed");
  if (bytes) std::memcpy(p, pts.data(), bytes);
  out->data = p;
  out->len = bytes;
  return XNYZ_OK;
}
}  // namespace xnyz
