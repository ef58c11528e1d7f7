// xnyz_oracle.cpp -- CPU restatement of XnYZip's compression path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline for
// the HIP path.  The product library never links or calls this file.
//
// Each function restates the reference function cited beside it (paths are
// relative to the reference repository root).  The code is single-threaded
// and keeps the reference's data structures where they decide output bytes:
// std::unordered_map insertion/iteration order and std::priority_queue
// tie-breaking for Huffman, std::sort for the reid runs, the same hash for the
// RLE cell map.  Floating point follows the reference's Eigen expressions op
// by op; build with -ffp-contract=off and no -march (x86-64 baseline has no
// FMA, matching the reference's flags in CMakePresets.json:97-98).
//
// Parity pins (tests/test_oracle_*.py): reference encoder and Hilbert sources
// compiled as-is into oracle/_ref (encoders, Hilbert tables and transforms);
// golden vectors under tests/golden; brute-force nearest-lattice-node search
// for the quantizers; round trips through the restated decoders; closed-form
// known answers for the bit packers.

#include "xnyz_oracle.h"

#include <zstd.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <queue>
#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

struct V3f {
  float v[3];
  float& operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
};
struct V3i {
  int v[3];
  int& operator[](int i) { return v[i]; }
  int operator[](int i) const { return v[i]; }
  bool operator==(const V3i& o) const {
    return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2];
  }
};

// ---------------------------------------------------------------------------
// source/preprocessor/shifting.hpp:11-36 (shift_points), :39-50 (unshift)
// Per-axis minimum with '<' (first occurrence wins; NaN never replaces),
// starting from FLT_MAX; then p - min per component in float.
std::vector<V3f> shift_points(const std::vector<V3f>& pts, V3f& off) {
  if (pts.empty()) return {};
  float mn[3] = {std::numeric_limits<float>::max(),
                 std::numeric_limits<float>::max(),
                 std::numeric_limits<float>::max()};
  for (const auto& p : pts)
    for (int a = 0; a < 3; ++a)
      if (p[a] < mn[a]) mn[a] = p[a];
  off = V3f{{mn[0], mn[1], mn[2]}};
  std::vector<V3f> out;
  out.reserve(pts.size());
  for (const auto& p : pts)
    out.push_back(V3f{{p[0] - off[0], p[1] - off[1], p[2] - off[2]}});
  return out;
}

std::vector<V3f> unshift_points(const std::vector<V3f>& pts, const V3f& off) {
  std::vector<V3f> out;
  out.reserve(pts.size());
  for (const auto& p : pts)
    out.push_back(V3f{{p[0] + off[0], p[1] + off[1], p[2] + off[2]}});
  return out;
}

// ---------------------------------------------------------------------------
// Quantizers.
// truncated_octahedron_quantizer.hpp:29: `pt / (2 * scale / sqrt(5))`: the
// int*float product is float, sqrt(5) is double, the quotient is double and
// Eigen converts that scalar to float before the per-component division.
float to_divisor(float b) {
  float two_b = 2 * b;
  return static_cast<float>(static_cast<double>(two_b) / std::sqrt(5.0));
}
// cube_quantizer.hpp:20 -- same shape with sqrt(3).
float cube_divisor(float b) {
  float two_b = 2 * b;
  return static_cast<float>(static_cast<double>(two_b) / std::sqrt(3.0));
}

// truncated_octahedron_quantizer.hpp:31-50 for one already-scaled point.
// E = round(q*0.5)*2, O = round((q-1)*0.5)*2+1 (round half away from zero);
// squared norms summed as x^2 + (y^2 + z^2) (Eigen's unrolled redux); E
// wins ties (dE <= dO); then the x/y halving of the BCC node.
V3i to_node(const V3f& q) {
  float E[3], O[3];
  for (int a = 0; a < 3; ++a) {
    E[a] = std::round(q[a] * 0.5f) * 2.0f;
    O[a] = std::round((q[a] - 1.0f) * 0.5f) * 2.0f + 1.0f;
  }
  float e0 = q[0] - E[0], e1 = q[1] - E[1], e2 = q[2] - E[2];
  float o0 = q[0] - O[0], o1 = q[1] - O[1], o2 = q[2] - O[2];
  float dE = e0 * e0 + (e1 * e1 + e2 * e2);
  float dO = o0 * o0 + (o1 * o1 + o2 * o2);
  V3i node = (dE <= dO)
                 ? V3i{{static_cast<int>(E[0]), static_cast<int>(E[1]),
                        static_cast<int>(E[2])}}
                 : V3i{{static_cast<int>(O[0]), static_cast<int>(O[1]),
                        static_cast<int>(O[2])}};
  if (node[0] & 1) {
    node[0] += 1;
    node[0] >>= 1;
    node[1] += 1;
    node[1] >>= 1;
  } else {
    node[0] >>= 1;
    node[1] >>= 1;
  }
  return node;
}

std::vector<V3i> quantize(const std::vector<V3f>& pts, int qtype, float b) {
  std::vector<V3i> out;
  out.reserve(pts.size());
  if (qtype == ORACLE_TO) {
    const float d = to_divisor(b);
    for (const auto& p : pts) {
      V3f q{{p[0] / d, p[1] / d, p[2] / d}};
      out.push_back(to_node(q));
    }
  } else {  // cube_quantizer.hpp:15-27: floor(p / d) cast to int
    const float d = cube_divisor(b);
    for (const auto& p : pts) {
      out.push_back(V3i{{static_cast<int>(std::floor(p[0] / d)),
                         static_cast<int>(std::floor(p[1] / d)),
                         static_cast<int>(std::floor(p[2] / d))}});
    }
  }
  return out;
}

// truncated_octahedron_quantizer.hpp:57-76 and cube_quantizer.hpp:29-38.
std::vector<V3f> recover(const std::vector<V3i>& nodes, int qtype, float b) {
  std::vector<V3f> out;
  out.reserve(nodes.size());
  if (qtype == ORACLE_TO) {
    const float s5 = static_cast<float>(std::sqrt(5.0));
    for (auto node : nodes) {
      if (node[2] & 1) {
        node[0] <<= 1;
        node[0] -= 1;
        node[1] <<= 1;
        node[1] -= 1;
      } else {
        node[0] <<= 1;
        node[1] <<= 1;
      }
      V3f r;
      for (int a = 0; a < 3; ++a)
        r[a] = ((static_cast<float>(node[a]) * 2.0f) * b) / s5;
      out.push_back(r);
    }
  } else {
    const float s3 = static_cast<float>(std::sqrt(3.0));
    for (const auto& v : nodes) {
      V3f r;
      for (int a = 0; a < 3; ++a)
        r[a] = (((static_cast<float>(v[a]) + 0.5f) * 2.0f) * b) / s3;
      out.push_back(r);
    }
  }
  return out;
}

// ---------------------------------------------------------------------------
// source/preprocessor/z_order_curve.hpp:14-47: 21 bits per axis, x at bit 3i,
// y at 3i+1, z at 3i+2; the int is taken as uint32 and masked (-1 -> 2^21-1).
uint64_t split_by_3(uint32_t n) {
  uint64_t x = n & 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}
uint64_t morton_code(const V3i& v) {
  return split_by_3(static_cast<uint32_t>(v[0])) |
         (split_by_3(static_cast<uint32_t>(v[1])) << 1) |
         (split_by_3(static_cast<uint32_t>(v[2])) << 2);
}
uint32_t compact_by_3(uint64_t code) {
  uint64_t x = code & 0x1249249249249249ULL;
  x = (x | (x >> 2)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x >> 4)) & 0x100f00f00f00f00fULL;
  x = (x | (x >> 8)) & 0x1f0000ff0000ffULL;
  x = (x | (x >> 16)) & 0x1f00000000ffffULL;
  x = (x | (x >> 32)) & 0x1fffff;
  return static_cast<uint32_t>(x);
}
V3i decode_morton_code(uint64_t m) {
  return V3i{{static_cast<int>(compact_by_3(m)),
              static_cast<int>(compact_by_3(m >> 1)),
              static_cast<int>(compact_by_3(m >> 2))}};
}

// ---------------------------------------------------------------------------
// source/quantizer/hilbert-curves-cpp/hilbert.cpp:17-341: the permutation
// tables are constructed, not hard-coded (DONT_CONSTRUCT_PERMUTATIONS is not
// defined).  Pattern A {0,1,3,2,6,7,5,4}, rotation 0; the 24 cube rotations
// are enumerated as 8-permutations and ordered as a std::set (lexicographic),
// then a depth-first search picks, for each of the 8 octants in curve order,
// the first rotation whose entry cell touches the previous exit cell.
using perm8 = std::array<uint8_t, 8>;
using vec3i = std::array<int, 3>;
using vec3u = std::array<unsigned, 3>;

struct HilbertTables {
  perm8 h2m_perm{}, m2h_perm{};
  std::array<perm8, 24> rots{};
  std::array<perm8, 8> h2m_perms{}, m2h_perms{};

  static perm8 invert(const perm8& p) {  // hilbert.cpp:89-101
    perm8 inv{};
    for (uint8_t val = 0; val < 8; ++val)
      for (uint8_t i = 0; i < 8; ++i)
        if (p[i] == val) {
          inv[val] = i;
          break;
        }
    return inv;
  }
  static perm8 matrix_to_perm(const vec3i rows[3]) {  // hilbert.cpp:150-162
    perm8 r{};
    for (int i = 0; i < 8; ++i) {
      vec3i v{((i >> 2) & 1) * 2 - 1, ((i >> 1) & 1) * 2 - 1, (i & 1) * 2 - 1};
      int rot[3];
      for (int k = 0; k < 3; ++k) {
        int dotv = rows[k][0] * v[0] + rows[k][1] * v[1] + rows[k][2] * v[2];
        rot[k] = (dotv + 1) / 2;
      }
      r[i] = static_cast<uint8_t>((rot[0] << 2) | (rot[1] << 1) | rot[2]);
    }
    return r;
  }
  static perm8 rotation(int rx, int ry, int rz) {  // hilbert.cpp:181-200
    const int sn[4] = {0, 1, 0, -1};
    const int cs[4] = {1, 0, -1, 0};
    int sx = sn[rx], sy = sn[ry], sz = sn[rz];
    int cx = cs[rx], cy = cs[ry], cz = cs[rz];
    vec3i rows[3] = {{cx * cy, sx * sz - cx * cz * sy, cz * sx + cx * sy * sz},
                     {sy, cy * cz, -cy * sz},
                     {-cy * sx, cx * sz + cz * sx * sy, cx * cz - sx * sy * sz}};
    return matrix_to_perm(rows);
  }
  static vec3u position(uint8_t v) {  // hilbert.cpp:234-246
    return vec3u{static_cast<unsigned>((v >> 2) & 1),
                 static_cast<unsigned>((v >> 1) & 1),
                 static_cast<unsigned>(v & 1)};
  }
  static bool neighbors(vec3u a, vec3u b) {  // hilbert.cpp:226-232
    int d0 = int(a[0]) - int(b[0]), d1 = int(a[1]) - int(b[1]),
        d2 = int(a[2]) - int(b[2]);
    return d0 * d0 + d1 * d1 + d2 * d2 == 1;
  }
  // hilbert.cpp:253-321 (makeHilbertToMortonPermutations_impl<false>)
  bool search(size_t i, vec3u prev_end) {
    const vec3u p0 = position(h2m_perm[0]);
    const vec3u p7 = position(h2m_perm[7]);
    const vec3u begin_g{p0[0] * 3u, p0[1] * 3u, p0[2] * 3u};
    const vec3u end_g{p7[0] * 3u, p7[1] * 3u, p7[2] * 3u};
    const vec3u pi = position(h2m_perm[i]);
    const vec3u curr{pi[0] * 2u, pi[1] * 2u, pi[2] * 2u};
    for (size_t k = 0; k < 24; ++k) {
      vec3u bl = position(rots[k][h2m_perm[0]]);
      vec3u bg{bl[0] + curr[0], bl[1] + curr[1], bl[2] + curr[2]};
      if (i == 0) {
        if (bg != begin_g) continue;
      } else if (!neighbors(bg, prev_end)) {
        continue;
      }
      vec3u el = position(rots[k][h2m_perm[7]]);
      vec3u eg{el[0] + curr[0], el[1] + curr[1], el[2] + curr[2]};
      if (i != 7) {
        if (search(i + 1, eg)) {
          h2m_perms[h2m_perm[i]] = rots[k];
          return true;
        }
        continue;
      }
      if (eg != end_g) continue;
      h2m_perms[h2m_perm[i]] = rots[k];
      return true;
    }
    return false;
  }

  HilbertTables() {
    h2m_perm = perm8{0, 1, 3, 2, 6, 7, 5, 4};  // pattern A, rotation 0
    m2h_perm = invert(h2m_perm);
    std::set<perm8> all;
    for (int rx = 0; rx < 4; ++rx)
      for (int ry = 0; ry < 4; ++ry)
        for (int rz = 0; rz < 4; ++rz) all.emplace(rotation(rx, ry, rz));
    size_t idx = 0;
    for (const auto& p : all) rots[idx++] = p;
    search(0, vec3u{0, 0, 0});
    for (int i = 0; i < 8; ++i) m2h_perms[i] = invert(h2m_perms[i]);
  }
};

const HilbertTables& htab() {
  static const HilbertTables t;
  return t;
}

// hilbert.cpp:423-448
uint64_t morton_to_hilbert3(uint64_t morton, unsigned iteration) {
  const auto& T = htab();
  uint64_t result = morton;
  for (unsigned i = iteration; i != 0; --i) {
    const unsigned shift = i * 3;
    const uint8_t md = (result >> shift) & 7;
    const uint8_t hd = T.m2h_perm[md];
    result ^= uint64_t{md} << shift;
    result |= uint64_t{hd} << shift;
    const perm8 perm = T.m2h_perms[md];
    for (unsigned j = i - 1; j < i; --j) {
      const unsigned cs = j * 3;
      const uint8_t c = (result >> cs) & 7;
      result ^= uint64_t{c} << cs;
      result |= uint64_t{perm[c]} << cs;
    }
  }
  const uint8_t lm = result & 7;
  return (result ^ lm) | T.m2h_perm[lm];
}

// hilbert.cpp:396-421
uint64_t hilbert_to_morton3(uint64_t hilbert, unsigned iteration) {
  const auto& T = htab();
  const uint8_t lh = hilbert & 7;
  uint64_t result = (hilbert ^ uint64_t{lh}) | uint64_t{T.h2m_perm[lh]};
  for (unsigned i = 1; i <= iteration; ++i) {
    const unsigned shift = i * 3;
    const uint8_t hd = (result >> shift) & 7;
    const uint8_t md = T.h2m_perm[hd];
    result ^= uint64_t{hd} << shift;
    result |= uint64_t{md} << shift;
    const perm8 perm = T.h2m_perms[md];
    for (unsigned j = i - 1; j < i; --j) {
      const unsigned cs = j * 3;
      const uint8_t c = (result >> cs) & 7;
      result ^= uint64_t{c} << cs;
      result |= uint64_t{perm[c]} << cs;
    }
  }
  return result;
}

// ---------------------------------------------------------------------------
// source/quantizer/block_quantizer.hpp:13-151
struct NodeWithOrder {
  uint64_t id;
  uint64_t reid;
  size_t ord;
};
struct LCPMeta {
  std::vector<uint64_t> blkst, blkcnt;
  std::vector<uint8_t> quads;
  std::vector<uint64_t> repos;
  std::vector<size_t> ords;
  bool is_hilbert = false;
};

// block_quantizer.hpp:33-76: 8 LSD passes of 8 bits over the whole 64-bit
// id, then std::sort by reid inside every run of equal ids.
void radix_sort(NodeWithOrder* start, NodeWithOrder* end) {
  size_t n = end - start;
  NodeWithOrder* buffer = new NodeWithOrder[n];
  NodeWithOrder* orig_buffer = buffer;
  NodeWithOrder* orig_start = start;
  for (int shift = 0; shift < 64; shift += 8) {
    size_t bucket[256] = {0};
    size_t offset[256] = {0};
    for (size_t i = 0; i < n; i++) bucket[(start[i].id >> shift) & 255]++;
    for (size_t i = 1; i < 256; i++) offset[i] = offset[i - 1] + bucket[i - 1];
    for (size_t i = 0; i < n; i++) {
      size_t d = (start[i].id >> shift) & 255;
      buffer[offset[d]++] = start[i];
    }
    std::swap(start, buffer);
  }
  (void)orig_start;
  delete[] orig_buffer;  // eight passes: the data is back in the caller's array
  NodeWithOrder *l = start, *r = l;
  while (l < end) {
    r = l;
    while (r + 1 < end && l->id == (r + 1)->id) ++r;
    if (l < r)
      std::sort(l, r + 1, [&](NodeWithOrder u, NodeWithOrder v) {
        return u.reid < v.reid;
      });
    l = r + 1;
  }
}

LCPMeta block_quantize(const std::vector<V3i>& pts, size_t bx, size_t by,
                       size_t bz, size_t nx, size_t ny, size_t nz,
                       bool is_hilbert) {
  size_t n = pts.size();
  std::vector<NodeWithOrder> vec(n);
  nx = nx / (2 * bx) + 1;
  ny = ny / (2 * by) + 1;
  nz = nz / (2 * bz) + 1;
  for (size_t i = 0; i < n; ++i) {
    const auto& p = pts[i];
    // int / size_t converts the int to size_t (z == -1 becomes 2^64-1)
    size_t cx = static_cast<size_t>(p[0]) / bx, dx = static_cast<size_t>(p[0]) % bx;
    size_t cy = static_cast<size_t>(p[1]) / by, dy = static_cast<size_t>(p[1]) % by;
    size_t cz = static_cast<size_t>(p[2]) / bz, dz = static_cast<size_t>(p[2]) % bz;
    uint64_t id = (cx / 2) + (cy / 2) * nx + (cz / 2) * nx * ny;
    uint64_t m = morton_code(V3i{{static_cast<int>(dx), static_cast<int>(dy),
                                  static_cast<int>(dz)}});
    uint64_t sfc = is_hilbert ? morton_to_hilbert3(m, 20) : m;
    vec[i] = {id, sfc | ((cx & 1) << 60) | ((cy & 1) << 61) | ((cz & 1) << 62), i};
  }
  radix_sort(vec.data(), vec.data() + n);
  size_t blknum = 1;
  for (size_t i = 1; i < n; ++i)
    if (vec[i].id != vec[i - 1].id) ++blknum;
  LCPMeta M;
  M.is_hilbert = is_hilbert;
  M.blkst.resize(blknum);
  M.blkcnt.resize(blknum);
  M.quads.resize(n);
  M.repos.resize(n);
  M.ords.resize(n);
  size_t i = static_cast<size_t>(-1);
  uint64_t prevId = uint64_t(-1);
  uint64_t prevQuad = 0, prevRe = 0;
  for (size_t j = 0; j < n; ++j) {
    auto& node = vec[j];
    M.ords[j] = node.ord;
    uint64_t id = node.id;
    size_t quad = node.reid >> 60;
    size_t reid = node.reid & 0x0fffffffffffffffULL;
    if (id != prevId) {
      M.blkst[++i] = prevId = id;
      prevQuad = 0;
      prevRe = 0;
    } else if (quad != prevQuad) {
      prevRe = 0;
    }
    M.blkcnt[i] += 1;
    M.quads[j] = uint8_t(quad - prevQuad);
    M.repos[j] = reid - prevRe;
    prevQuad = quad;
    prevRe = reid;
  }
  return M;
}

// block_quantizer.hpp:153-207
std::vector<V3i> recover_from_lcp_meta(const LCPMeta& meta, size_t bx,
                                       size_t by, size_t bz, size_t nx,
                                       size_t ny, size_t nz) {
  std::vector<V3i> coords;
  coords.reserve(meta.repos.size());
  nx = nx / (2 * bx) + 1;
  ny = ny / (2 * by) + 1;
  nz = nz / (2 * bz) + 1;
  (void)nz;
  size_t j = 0;
  for (size_t i = 0; i < meta.blkst.size(); i++) {
    size_t pbx = (meta.blkst[i] % nx * 2) * bx;
    size_t pby = (meta.blkst[i] / nx % ny * 2) * by;
    size_t pbz = (meta.blkst[i] / nx / ny * 2) * bz;
    size_t prequad = 0, prerepos = 0;
    for (size_t k = 0; k < meta.blkcnt[i]; k++) {
      if (j >= meta.repos.size()) throw std::runtime_error("lcp: short streams");
      if (meta.quads[j] != 0) prerepos = 0;
      size_t reposj = meta.repos[j] + prerepos;
      size_t quadj = meta.quads[j] + prequad;
      prerepos = reposj;
      prequad = quadj;
      uint64_t morton = meta.is_hilbert ? hilbert_to_morton3(reposj, 20) : reposj;
      V3i d = decode_morton_code(morton);
      size_t idx = pbx + ((quadj & 0x01) >> 0) * bx + static_cast<size_t>(d[0]);
      size_t idy = pby + ((quadj & 0x02) >> 1) * by + static_cast<size_t>(d[1]);
      size_t idz = pbz + ((quadj & 0x04) >> 2) * bz + static_cast<size_t>(d[2]);
      coords.push_back(V3i{{static_cast<int>(idx), static_cast<int>(idy),
                            static_cast<int>(idz)}});
      ++j;
    }
  }
  return coords;
}

// ---------------------------------------------------------------------------
// source/encoder/bitpacking_encoder.hpp:17-80
struct Packed {
  std::vector<uint8_t> meta, data;
};

Packed bitpack(const std::vector<uint64_t>& values) {
  if (values.empty()) throw std::invalid_argument("Input vector must not be empty");
  uint64_t maxv = 0;
  for (auto v : values)
    if (v > maxv) maxv = v;
  uint8_t w = (maxv == 0 ? 1 : static_cast<uint8_t>(64 - __builtin_clzll(maxv)));
  Packed out;
  uint32_t n = static_cast<uint32_t>(values.size());
  for (int i = 0; i < 4; ++i) out.meta.push_back(static_cast<uint8_t>((n >> (8 * i)) & 0xFF));
  out.meta.push_back(w);
  size_t total_bits = static_cast<size_t>(n) * w;
  out.data.assign((total_bits + 7) / 8, 0);
  size_t bit = 0;
  for (auto v : values) {
    for (uint8_t b = 0; b < w; ++b) {
      if (v & (uint64_t(1) << b)) out.data[bit >> 3] |= static_cast<uint8_t>(1u << (bit & 7));
      ++bit;
    }
  }
  return out;
}

std::vector<uint64_t> bitunpack(const Packed& in) {
  if (in.meta.size() != 5) throw std::runtime_error("Invalid meta size in bitpacked data");
  uint32_t n = 0;
  for (int i = 0; i < 4; ++i) n |= (uint32_t(in.meta[i]) << (8 * i));
  uint8_t w = in.meta[4];
  std::vector<uint64_t> values(n, 0);
  size_t bit = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t val = 0;
    for (uint8_t b = 0; b < w; ++b) {
      if ((bit >> 3) >= in.data.size()) throw std::runtime_error("bitpack: short data");
      if (in.data[bit >> 3] & (1u << (bit & 7))) val |= (uint64_t(1) << b);
      ++bit;
    }
    values[i] = val;
  }
  return values;
}

// source/encoder/p_for_delta_encoder.hpp:27-57 (lim = 254).  `lim` below 254
// and a forced Golomb-Rice k are the mode overrides of the repos / sfc-delta
// stream (oracle_compress_mode); the reference always uses 254 and its k rule.
struct ModeOverride {
  uint64_t lim = 254;
  int gr_k = -1;
};
thread_local ModeOverride g_mode;

template <typename T>
struct PFD {
  std::vector<T> patches, deltas;
};
template <typename T>
PFD<T> pfd_encode(const std::vector<T>& data, uint64_t lim_in = 254) {
  PFD<T> r;
  r.deltas.reserve(data.size());
  const T lim = static_cast<T>(lim_in);
  for (const auto& v : data) {
    if (v > lim) {
      r.patches.push_back(v);
      r.deltas.push_back(0);
    } else {
      r.deltas.push_back(static_cast<T>(v + 1));
    }
  }
  return r;
}
template <typename T>
std::vector<T> pfd_decode(const PFD<T>& d) {
  std::vector<T> out;
  out.reserve(d.deltas.size());
  size_t pi = 0;
  for (const auto& v : d.deltas) {
    if (v == 0) {
      if (pi >= d.patches.size()) throw std::runtime_error("pfd: missing patch");
      out.push_back(d.patches[pi++]);
    } else {
      out.push_back(static_cast<T>(v - 1));
    }
  }
  return out;
}

// source/encoder/golomb_rice_encoder.hpp:12-47 (BitWriter, LSB-first),
// :92-148 (pack), :150-193 (unpack)
struct BitWriterLSB {
  std::vector<uint8_t>& buf;
  size_t pos = 0;
  explicit BitWriterLSB(std::vector<uint8_t>& b) : buf(b) {}
  void bit(bool v) {
    size_t byte = pos / 8;
    if (byte >= buf.size()) buf.push_back(0);
    if (v) buf[byte] |= (1u << (pos % 8));
    ++pos;
  }
};

template <typename T>
Packed gr_pack(const std::vector<T>& values, int forced_k = -1) {
  if (values.empty()) return Packed{};
  uint64_t sum = 0;
  bool all_zero = true;
  for (const T& v : values) {
    if (v != 0) all_zero = false;
    sum += static_cast<uint64_t>(v);
  }
  uint8_t k;
  if (all_zero) {
    k = 0;
  } else {
    double avg = static_cast<double>(sum) / values.size();
    k = (avg < 1.0) ? 0 : static_cast<uint8_t>(std::floor(std::log2(avg)));
  }
  if (k >= 63) k = 62;
  if (forced_k >= 0) k = static_cast<uint8_t>(forced_k);
  Packed out;
  uint32_t n = static_cast<uint32_t>(values.size());
  out.meta.resize(5);
  for (int i = 0; i < 4; ++i) out.meta[i] = static_cast<uint8_t>((n >> (8 * i)) & 0xFF);
  out.meta[4] = k;
  BitWriterLSB w(out.data);
  uint64_t M = 1ULL << k;
  for (const T& vt : values) {
    uint64_t v = static_cast<uint64_t>(vt);
    uint64_t q = v / M, r = v % M;
    for (uint64_t i = 0; i < q; ++i) w.bit(true);
    w.bit(false);
    for (uint8_t b = 0; b < k; ++b) w.bit((r >> b) & 1);
  }
  return out;
}

template <typename T>
std::vector<T> gr_unpack(const Packed& in) {
  if (in.meta.empty() && in.data.empty()) return {};
  if (in.meta.size() != 5) throw std::runtime_error("Invalid meta size in Golomb-Rice packed data");
  uint32_t n = 0;
  for (int i = 0; i < 4; ++i) n |= (static_cast<uint32_t>(in.meta[i]) << (8 * i));
  if (n == 0) return {};
  uint8_t k = in.meta[4];
  if (k >= 63) throw std::runtime_error("Invalid k_param in Golomb-Rice meta");
  std::vector<T> out;
  out.reserve(n);
  size_t pos = 0, total = in.data.size() * 8;
  auto rd = [&]() -> bool {
    if (pos >= total) throw std::out_of_range("Attempting to read beyond bitstream boundary");
    bool b = ((in.data[pos / 8] >> (pos % 8)) & 1) != 0;
    ++pos;
    return b;
  };
  uint64_t M = 1ULL << k;
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t q = 0;
    while (rd()) q++;
    uint64_t r = 0;
    for (uint8_t b = 0; b < k; ++b)
      if (rd()) r |= (1ULL << b);
    uint64_t v = q * M + r;
    if (v > static_cast<uint64_t>(std::numeric_limits<T>::max()))
      throw std::runtime_error("Unpacked value exceeds capacity of target type T.");
    out.push_back(static_cast<T>(v));
  }
  return out;
}

// ---------------------------------------------------------------------------
// source/encoder/huffman_encoder.hpp:20-443.  The codebook is built with the
// same standard containers the reference uses, so the tree tie-breaks
// (std::priority_queue over unordered_map iteration order) and the meta
// entry order (unordered_map iteration order of code_table, which keeps its
// bucket array across build() calls) come out identical.
template <typename T>
class Huffman {
  struct Node {
    T symbol{};
    uint64_t freq = 0;
    Node* left = nullptr;
    Node* right = nullptr;
    bool leaf() const { return !left && !right; }
  };

 public:
  struct Code {
    uint32_t bits;
    uint8_t length;
  };
  std::unordered_map<T, Code> code_table;
  bool use_eg = false;
  uint8_t eg_k = 0;

  // :33-61
  void build(const std::vector<T>& data) {
    code_table.clear();
    if (data.empty()) return;
    std::unordered_map<T, uint64_t> freq;
    for (auto v : data) freq[v]++;
    double entropy = 0.0;
    for (auto& kv : freq) {
      double p = double(kv.second) / data.size();
      entropy -= p * std::log2(p);
    }
    if (entropy > 10.0 || use_eg) {
      use_eg = true;
      eg_k = select_k(data);
      return;
    }
    build_from_freq(freq);
  }

  // :69-84: the data's frequencies, then every symbol 0..alphabet_max not
  // seen with frequency 1 (inserted in ascending order after the data's
  // symbols); never Exp-Golomb.
  void build_complete(const std::vector<T>& data, T alphabet_max) {
    code_table.clear();
    use_eg = false;
    eg_k = 0;
    std::unordered_map<T, uint64_t> freq;
    for (auto v : data) freq[v]++;
    for (T s = 0; s <= alphabet_max; ++s) {
      if (freq.find(s) == freq.end()) freq[s] = 1;
      if (s == std::numeric_limits<T>::max()) break;
    }
    build_from_freq(freq);
  }

  // :126-146
  std::vector<uint8_t> meta() const {
    std::vector<uint8_t> out;
    if (use_eg) {
      out = {0xE1, 0xC1, 0x01, 0x01, eg_k};
      return out;
    }
    put32(out, uint32_t(code_table.size()));
    for (auto& kv : code_table) {
      put_symbol(out, kv.first);
      out.push_back(kv.second.length);
      put32(out, kv.second.bits);
    }
    return out;
  }

  // :150-170 (Huffman, MSB-first, last byte left-aligned); :365-396 (EG)
  std::vector<uint8_t> encode(const std::vector<T>& data) const {
    if (use_eg) return encode_eg(data);
    std::vector<uint8_t> out;
    uint64_t bitbuf = 0;
    int bits = 0;
    for (auto v : data) {
      const auto& c = code_table.at(v);
      bitbuf = (bitbuf << c.length) | c.bits;
      bits += c.length;
      while (bits >= 8) {
        out.push_back(uint8_t(bitbuf >> (bits - 8)));
        bits -= 8;
        bitbuf &= ((1ULL << bits) - 1);
      }
    }
    if (bits) out.push_back(uint8_t(bitbuf << (8 - bits)));
    return out;
  }

  // :172-203
  void load_meta(const std::vector<uint8_t>& m) {
    code_table.clear();
    use_eg = false;
    eg_k = 0;
    size_t pos = 0;
    if (m.size() >= 4 && m[0] == 0xE1 && m[1] == 0xC1 && m[2] == 0x01) {
      uint8_t mode = m[3];
      pos = 4;
      if (mode == 1) {
        if (m.size() > pos) eg_k = m[pos];
        use_eg = true;
        return;
      }
    }
    uint32_t count = get32(m, pos);
    for (uint32_t i = 0; i < count; ++i) {
      T sym = get_symbol(m, pos);
      if (pos + 5 > m.size()) throw std::runtime_error("huffman: short meta");
      uint8_t len = m[pos++];
      uint32_t bits = get32(m, pos);
      code_table[sym] = Code{bits, len};
    }
  }

  // :205-243 and :398-442
  std::vector<T> decode(const std::vector<uint8_t>& data, size_t n) const {
    if (use_eg) return decode_eg(data, n);
    std::vector<T> out;
    out.reserve(n);
    if (code_table.size() == 1) {
      out.assign(n, code_table.begin()->first);
      return out;
    }
    std::unordered_map<uint8_t, std::unordered_map<uint32_t, T>> dm;
    for (auto& kv : code_table) dm[kv.second.length][kv.second.bits] = kv.first;
    uint32_t buf = 0;
    uint8_t buflen = 0;
    for (size_t i = 0; i < data.size() * 8 && out.size() < n; ++i) {
      buf = (buf << 1) | ((data[i >> 3] >> (7 - (i & 7))) & 1);
      ++buflen;
      auto itl = dm.find(buflen);
      if (itl != dm.end()) {
        auto its = itl->second.find(buf);
        if (its != itl->second.end()) {
          out.push_back(its->second);
          buf = 0;
          buflen = 0;
        }
      }
    }
    return out;
  }

  static uint64_t to_unsigned(T v) {  // :311-325 zigzag for signed types
    if constexpr (std::is_signed<T>::value) {
      int64_t s = int64_t(v);
      return (uint64_t(s) << 1) ^ uint64_t(s >> 63);
    } else {
      return uint64_t(v);
    }
  }
  static T from_unsigned(uint64_t v) {
    if constexpr (std::is_signed<T>::value)
      return T(int64_t((v >> 1) ^ (~(v & 1) + 1)));
    else
      return T(v);
  }
  static int bit_length(uint64_t v) {  // :335-342
    int len = 0;
    while (v) {
      ++len;
      v >>= 1;
    }
    return len ? len : 1;
  }

 private:
  // :91-123
  void build_from_freq(const std::unordered_map<T, uint64_t>& freq) {
    if (freq.empty()) return;
    auto cmp = [](Node* const& a, Node* const& b) { return a->freq > b->freq; };
    std::priority_queue<Node*, std::vector<Node*>, decltype(cmp)> pq(cmp);
    std::vector<std::unique_ptr<Node>> pool;
    for (auto& kv : freq) {
      pool.push_back(std::make_unique<Node>());
      pool.back()->symbol = kv.first;
      pool.back()->freq = kv.second;
      pq.push(pool.back().get());
    }
    while (pq.size() > 1) {
      Node* a = pq.top();
      pq.pop();
      Node* b = pq.top();
      pq.pop();
      pool.push_back(std::make_unique<Node>());
      Node* p = pool.back().get();
      p->freq = a->freq + b->freq;
      p->left = a;
      p->right = b;
      pq.push(p);
    }
    build_code(pq.top(), 0u, 0u);
  }
  // :269-281
  void build_code(const Node* node, uint32_t prefix, uint8_t len) {
    if (node->leaf()) {
      if (len == 0)
        code_table[node->symbol] = Code{0, 1};
      else
        code_table[node->symbol] = Code{prefix, len};
      return;
    }
    build_code(node->left, prefix << 1, len + 1);
    build_code(node->right, (prefix << 1) | 1u, len + 1);
  }
  // :344-363
  static uint8_t select_k(const std::vector<T>& data) {
    uint64_t best = std::numeric_limits<uint64_t>::max();
    uint8_t best_k = 0;
    for (uint8_t k = 0; k <= 16; ++k) {
      uint64_t cost = 0;
      for (auto v : data) {
        uint64_t n1 = (to_unsigned(v) >> k) + 1;
        cost += uint64_t(2 * bit_length(n1) - 1 + k);
      }
      if (cost < best) {
        best = cost;
        best_k = k;
      }
    }
    return best_k;
  }
  std::vector<uint8_t> encode_eg(const std::vector<T>& data) const {
    std::vector<uint8_t> out;
    uint8_t bitbuf = 0, bits = 0;
    auto push = [&](uint8_t b) {
      bitbuf = (bitbuf << 1) | (b & 1u);
      if (++bits == 8) {
        out.push_back(bitbuf);
        bitbuf = 0;
        bits = 0;
      }
    };
    for (auto v : data) {
      uint64_t n = to_unsigned(v);
      uint64_t n1 = (n >> eg_k) + 1;
      int nb = bit_length(n1);
      for (int i = 0; i < nb - 1; ++i) push(0);
      for (int i = nb - 1; i >= 0; --i) push((n1 >> i) & 1u);
      for (uint8_t i = 0; i < eg_k; ++i) push((n >> i) & 1u);
    }
    if (bits) out.push_back(uint8_t(bitbuf << (8 - bits)));
    return out;
  }
  std::vector<T> decode_eg(const std::vector<uint8_t>& data, size_t n) const {
    std::vector<T> out;
    out.reserve(n);
    size_t pos = 0;
    auto rd = [&](int& b) -> bool {
      if ((pos >> 3) >= data.size()) return false;
      b = (data[pos >> 3] >> (7 - int(pos & 7))) & 1;
      ++pos;
      return true;
    };
    while (out.size() < n) {
      int b = 0, lz = 0;
      while (true) {
        if (!rd(b)) return out;
        if (b == 0) ++lz; else break;
      }
      uint64_t n1 = 1;
      for (int i = 0; i < lz; ++i) {
        if (!rd(b)) return out;
        n1 = (n1 << 1) | uint64_t(b);
      }
      uint64_t high = (n1 - 1) << eg_k, low = 0;
      for (uint8_t i = 0; i < eg_k; ++i) {
        if (!rd(b)) return out;
        low |= (uint64_t(b) << i);
      }
      out.push_back(from_unsigned(high | low));
    }
    return out;
  }
  static void put32(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back(v >> 24);
    o.push_back(v >> 16);
    o.push_back(v >> 8);
    o.push_back(v);
  }
  static uint32_t get32(const std::vector<uint8_t>& m, size_t& p) {
    if (p + 4 > m.size()) throw std::runtime_error("huffman: short meta");
    uint32_t v = (uint32_t(m[p]) << 24) | (uint32_t(m[p + 1]) << 16) |
                 (uint32_t(m[p + 2]) << 8) | uint32_t(m[p + 3]);
    p += 4;
    return v;
  }
  static void put_symbol(std::vector<uint8_t>& o, T s) {
    if constexpr (sizeof(T) == 8) {
      for (int sh = 56; sh >= 0; sh -= 8) o.push_back(uint8_t(uint64_t(s) >> sh));
    } else {
      const uint8_t* p = reinterpret_cast<const uint8_t*>(&s);
      o.insert(o.end(), p, p + sizeof(T));
    }
  }
  static T get_symbol(const std::vector<uint8_t>& m, size_t& p) {
    if (p + sizeof(T) > m.size()) throw std::runtime_error("huffman: short meta");
    if constexpr (sizeof(T) == 8) {
      uint64_t v = 0;
      for (int k = 0; k < 8; ++k) v = (v << 8) | m[p++];
      return T(v);
    } else {
      T v;
      std::memcpy(&v, m.data() + p, sizeof(T));
      p += sizeof(T);
      return v;
    }
  }
};

// ---------------------------------------------------------------------------
// source/composer/utils.hpp:35-53 and :13-29
std::vector<uint8_t> zstd_compress(const std::vector<uint8_t>& data) {
  size_t bound = ZSTD_compressBound(data.size());
  std::vector<uint8_t> out(bound);
  size_t r = ZSTD_compress(out.data(), bound, data.data(), data.size(), 3);
  if (ZSTD_isError(r)) throw std::runtime_error(ZSTD_getErrorName(r));
  out.resize(r);
  return out;
}
std::vector<uint8_t> zstd_decompress(const uint8_t* p, size_t len) {
  unsigned long long fs = ZSTD_getFrameContentSize(p, len);
  if (fs == ZSTD_CONTENTSIZE_ERROR) throw std::runtime_error("Not a zstd frame");
  if (fs == ZSTD_CONTENTSIZE_UNKNOWN) throw std::runtime_error("Original size unknown");
  std::vector<uint8_t> out(fs);
  size_t d = ZSTD_decompress(out.data(), fs, p, len);
  if (ZSTD_isError(d)) throw std::runtime_error(ZSTD_getErrorName(d));
  if (d != fs) throw std::runtime_error("Decompressed size mismatch");
  return out;
}

// ---------------------------------------------------------------------------
// The compressors: block_compressor.hpp:39-222 and
// block_compressor_rle.hpp:38-321.
struct Buf {
  std::vector<uint8_t> b;
  template <typename V>
  void put(V v) {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    b.insert(b.end(), p, p + sizeof(V));
  }
  void section(const std::vector<uint8_t>& meta, const std::vector<uint8_t>& data) {
    put(static_cast<size_t>(data.size()));
    put(static_cast<size_t>(meta.size()));
    b.insert(b.end(), meta.begin(), meta.end());
    b.insert(b.end(), data.begin(), data.end());
  }
};

struct Result {
  std::vector<uint8_t> pre;
  std::vector<size_t> perm;
  oracle_info info{};
  double t_diag = 0;
};

int bit_count_of(int max_range) {  // block_compressor.hpp:93-95
  int bits = 0;
  int tmp = max_range;
  while (tmp > 0) {
    bits++;
    tmp >>= 1;
  }
  return bits == 0 ? 1 : bits;
}

// Direct-mode section pair (block_compressor.hpp:90-136; RLE :154-208).
// Returns the order of the input cells (sfc_entries[i].second).
std::vector<size_t> encode_direct(const std::vector<V3i>& q, int max_range,
                                  bool curve, Buf& buf, oracle_info& info) {
  size_t n = q.size();
  int hb = bit_count_of(max_range);
  std::vector<std::pair<uint64_t, size_t>> e(n);
  for (size_t i = 0; i < n; i++) {
    uint64_t m = morton_code(q[i]);
    e[i] = {curve ? morton_to_hilbert3(m, hb) : m, i};
  }
  std::sort(e.begin(), e.end());
  std::vector<uint64_t> d(n);
  d[0] = e[0].first;
  for (size_t i = 1; i < n; i++) d[i] = e[i].first - e[i - 1].first;
  auto pf = pfd_encode(d, g_mode.lim);
  info.patches = pf.patches.size();
  auto gp = gr_pack(pf.patches, g_mode.gr_k);
  buf.section(gp.meta, gp.data);
  Huffman<uint64_t> h;
  h.build(pf.deltas);
  buf.section(h.meta(), h.encode(pf.deltas));
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; i++) order[i] = e[i].second;
  return order;
}

// Block-mode sections (block_compressor.hpp:137-211; RLE :208-288).
LCPMeta encode_block(const std::vector<V3i>& q, int rx, int ry, int rz,
                     bool curve, Buf& buf, oracle_info& info) {
  LCPMeta M = block_quantize(q, 64, 64, 64, rx, ry, rz, curve);
  auto blk = M.blkst;
  auto cnt = M.blkcnt;
  info.blocks = blk.size();
  for (size_t i = cnt.size() - 1; i > 0; i--) cnt[i] -= cnt[i - 1];
  for (size_t i = blk.size() - 1; i > 0; i--) blk[i] -= blk[i - 1];
  auto p1 = bitpack(blk);
  buf.section(p1.meta, p1.data);
  auto p2 = bitpack(cnt);
  buf.section(p2.meta, p2.data);
  auto pr = pfd_encode(M.repos, g_mode.lim);
  info.patches = pr.patches.size();
  auto g1 = gr_pack(pr.patches, g_mode.gr_k);
  buf.section(g1.meta, g1.data);
  Huffman<uint64_t> h1;
  h1.build(pr.deltas);
  buf.section(h1.meta(), h1.encode(pr.deltas));
  auto pq = pfd_encode(M.quads);
  info.quad_patches = pq.patches.size();
  auto g2 = gr_pack(pq.patches);
  buf.section(g2.meta, g2.data);
  Huffman<uint8_t> h2;
  h2.build(pq.deltas);
  buf.section(h2.meta(), h2.encode(pq.deltas));
  return M;
}

// Shared-codebook compression (block_compressor.hpp:241-431,
// compress_internal): block mode always, the repos and quads Huffman
// codebooks either built to cover every symbol 0..255 (fit_profile) or
// loaded from a profile (compress_with_profile); the embedded metas are the
// encoders' get_meta() -- after load_meta, that is the loaded map's order.
// Profile blob: serialize_profile (block_encoding_profile.hpp:34-44).
std::vector<uint8_t> compress_shared(const std::vector<V3f>& points, int qtype, float b,
                                     bool curve, const std::vector<uint8_t>* profile_in,
                                     std::vector<uint8_t>* profile_out) {
  V3f offset{};
  auto shifted = shift_points(points, offset);
  auto q = quantize(shifted, qtype, b);
  Buf buf;
  buf.put(static_cast<uint64_t>(points.size()));
  buf.put(static_cast<size_t>(0));
  int rx = 0, ry = 0, rz = 0;
  for (const auto& p : q) {
    rx = std::max(rx, p[0]);
    ry = std::max(ry, p[1]);
    rz = std::max(rz, p[2]);
  }
  buf.put(rx);
  buf.put(ry);
  buf.put(rz);
  buf.put(curve);
  buf.put(false);  // is_direct: the shared path always codes blocks
  LCPMeta M = block_quantize(q, 64, 64, 64, rx, ry, rz, curve);
  auto blk = M.blkst;
  auto cnt = M.blkcnt;
  for (size_t i = cnt.size() - 1; i > 0; i--) cnt[i] -= cnt[i - 1];
  for (size_t i = blk.size() - 1; i > 0; i--) blk[i] -= blk[i - 1];
  auto p1 = bitpack(blk);
  buf.section(p1.meta, p1.data);
  auto p2 = bitpack(cnt);
  buf.section(p2.meta, p2.data);
  auto pr = pfd_encode(M.repos);
  auto g1 = gr_pack(pr.patches);
  buf.section(g1.meta, g1.data);
  std::vector<uint8_t> pm[2];
  if (profile_in) {  // serialize_profile: u32 LE size + bytes, twice
    const auto& blob = *profile_in;
    size_t pos = 0;
    for (int i = 0; i < 2; ++i) {
      if (pos + 4 > blob.size()) throw std::runtime_error("short profile");
      uint32_t n = uint32_t(blob[pos]) | uint32_t(blob[pos + 1]) << 8 |
                   uint32_t(blob[pos + 2]) << 16 | uint32_t(blob[pos + 3]) << 24;
      pos += 4;
      if (pos + n > blob.size()) throw std::runtime_error("short profile");
      pm[i].assign(blob.begin() + pos, blob.begin() + pos + n);
      pos += n;
    }
  }
  Huffman<uint64_t> h1;
  if (profile_in) h1.load_meta(pm[0]);
  else h1.build_complete(pr.deltas, uint64_t(255));
  auto m1 = h1.meta();
  buf.section(m1, h1.encode(pr.deltas));
  auto pq = pfd_encode(M.quads);
  auto g2 = gr_pack(pq.patches);
  buf.section(g2.meta, g2.data);
  Huffman<uint8_t> h2;
  if (profile_in) h2.load_meta(pm[1]);
  else h2.build_complete(pq.deltas, uint8_t(255));
  auto m2 = h2.meta();
  buf.section(m2, h2.encode(pq.deltas));
  buf.put(offset[0]);
  buf.put(offset[1]);
  buf.put(offset[2]);
  if (profile_out) {
    profile_out->clear();
    for (const auto* m : {&m1, &m2}) {
      const uint32_t n = uint32_t(m->size());
      for (int k = 0; k < 4; ++k) profile_out->push_back(uint8_t(n >> (8 * k)));
      profile_out->insert(profile_out->end(), m->begin(), m->end());
    }
  }
  return std::move(buf.b);
}

struct V3iHasher {  // block_compressor_rle.hpp:98-105
  std::size_t operator()(const V3i& v) const {
    uint64_t h = (uint64_t)v[0] * 73856093u ^ (uint64_t)v[1] * 19349663u ^
                 (uint64_t)v[2] * 83492791u;
    return h;
  }
};

Result compress(const std::vector<V3f>& points, int qtype, float b, bool curve,
                bool rle, int direct_threshold) {
  using clk = std::chrono::steady_clock;
  Result R;
  oracle_info& info = R.info;
  info.input_points = points.size();
  V3f offset{};
  auto shifted = shift_points(points, offset);
  auto q = quantize(shifted, qtype, b);
  for (const auto& c : q)
    if (c[2] == -1) info.wrapped_points++;

  if (rle) {
    // block_compressor_rle.hpp:53-87: diagnostic recover pass (inside the
    // reference's timer).
    auto t0 = clk::now();
    auto rec = recover(q, qtype, b);
    double acc = 0, mx = 0;
    for (size_t i = 0; i < shifted.size(); i++) {
      double mse = ((shifted[i][0] - rec[i][0]) * (shifted[i][0] - rec[i][0])) +
                   ((shifted[i][1] - rec[i][1]) * (shifted[i][1] - rec[i][1])) +
                   ((shifted[i][2] - rec[i][2]) * (shifted[i][2] - rec[i][2]));
      acc += mse;
      mx = std::max(mx, std::sqrt(mse));
    }
    info.max_l2 = mx;
    info.mse = acc / shifted.size();
    R.t_diag = std::chrono::duration<double>(clk::now() - t0).count();
  }

  Buf buf;
  std::vector<int> cnts;
  if (rle) {  // block_compressor_rle.hpp:96-123
    std::unordered_map<V3i, int, V3iHasher> um;
    for (const auto& p : q) um[p]++;
    q.clear();
    for (const auto& [p, c] : um) {
      cnts.push_back(c);
      q.push_back(p);
    }
  }
  uint64_t num_points = q.size();
  info.num_points = num_points;
  buf.put(num_points);
  buf.put(static_cast<size_t>(0));  // params.size(): TO/cube carry none
  int rx = 0, ry = 0, rz = 0;
  for (const auto& p : q) {
    rx = std::max(rx, p[0]);
    ry = std::max(ry, p[1]);
    rz = std::max(rz, p[2]);
  }
  buf.put(rx);
  buf.put(ry);
  buf.put(rz);
  info.range[0] = rx;
  info.range[1] = ry;
  info.range[2] = rz;
  int max_range = std::max({rx, ry, rz});
  bool is_direct = (max_range < direct_threshold);
  buf.put(curve);
  buf.put(is_direct);
  info.is_direct = is_direct;

  std::vector<size_t> order;
  if (is_direct) {
    order = encode_direct(q, max_range, curve, buf, info);
  } else {
    order = encode_block(q, rx, ry, rz, curve, buf, info).ords;
  }
  if (rle) {
    // cnts follow the cells into coding order (:174-177, :217-220)
    std::vector<int> c2(cnts.size());
    for (size_t i = 0; i < cnts.size(); i++) c2[i] = cnts[order[i]];
    // :290-303: build, check meta against 0.25 x data, build again on the
    // same encoder object
    Huffman<int> h3;
    h3.build(c2);
    auto m5 = h3.meta();
    auto d5 = h3.encode(c2);
    if (static_cast<double>(m5.size()) > static_cast<double>(d5.size()) * 0.25)
      h3.use_eg = true;
    h3.build(c2);
    m5 = h3.meta();
    d5 = h3.encode(c2);
    info.counts_exp_golomb = h3.use_eg;
    info.counts_eg_k = h3.eg_k;
    buf.section(m5, d5);
  } else {
    R.perm = order;
  }
  buf.put(offset[0]);
  buf.put(offset[1]);
  buf.put(offset[2]);
  R.pre = std::move(buf.b);
  return R;
}

// The MPI driver's mapreduce mode on one rank (source/mpi_driver.cpp:365-648)
// followed by BlockCompressorRLE::encode_cells_with_counts
// (block_compressor_rle.hpp:341-470): the global per-axis minimum (strict
// '<' from FLT_MAX, :377-401), every point quantized against it and counted
// per cell (:346-360, an int64 sum clamped to INT32_MAX, :599-603), the
// largest quantized coordinate as the range of all three axes (:469-480,
// :608-611), block mode always, counts in the sorted cells' order, the RLE
// counts stream built twice as in compress(), the minimum as the trailer.
std::vector<uint8_t> compress_global_rle(const std::vector<V3f>& points, int qtype, float b,
                                         bool curve) {
  V3f mn{{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
          std::numeric_limits<float>::max()}};
  for (const auto& p : points)
    for (int a = 0; a < 3; ++a)
      if (p[a] < mn[a]) mn[a] = p[a];
  std::vector<V3f> shifted(points.size());
  for (size_t i = 0; i < points.size(); ++i)
    for (int a = 0; a < 3; ++a) shifted[i][a] = points[i][a] - mn[a];
  auto q = quantize(shifted, qtype, b);
  std::unordered_map<V3i, int64_t, V3iHasher> table;
  for (const auto& c : q) table[c] += 1;
  int gmax = 0;
  std::vector<V3i> cells;
  std::vector<int> counts;
  for (const auto& [c, n] : table) {
    gmax = std::max({gmax, c[0], c[1], c[2]});
    cells.push_back(c);
    counts.push_back(static_cast<int>(std::min<int64_t>(n, INT32_MAX)));
  }
  Buf buf;
  buf.put(static_cast<uint64_t>(cells.size()));
  buf.put(static_cast<uint64_t>(0));
  buf.put(gmax);
  buf.put(gmax);
  buf.put(gmax);
  buf.put(curve);
  buf.put(false);
  oracle_info info{};
  LCPMeta M = encode_block(cells, gmax, gmax, gmax, curve, buf, info);
  std::vector<int> c2(counts.size());
  for (size_t i = 0; i < counts.size(); i++) c2[i] = counts[M.ords[i]];
  Huffman<int> h3;
  h3.build(c2);
  auto m5 = h3.meta();
  auto d5 = h3.encode(c2);
  if (static_cast<double>(m5.size()) > static_cast<double>(d5.size()) * 0.25) h3.use_eg = true;
  h3.build(c2);
  buf.section(h3.meta(), h3.encode(c2));
  buf.put(mn[0]);
  buf.put(mn[1]);
  buf.put(mn[2]);
  return std::move(buf.b);
}

// ---------------------------------------------------------------------------
// Decompressors: block_decompressor.hpp:31-192, block_decompressor_rle.hpp:32-207
std::vector<V3f> decompress_buffer(const std::vector<uint8_t>& buffer, int qtype,
                                   float b, bool rle) {
  const uint8_t* p = buffer.data();
  const uint8_t* end = buffer.data() + buffer.size();
  auto rd = [&](auto& v) {
    if (p + sizeof(v) > end) throw std::runtime_error("decompress: truncated buffer");
    std::memcpy(&v, p, sizeof(v));
    p += sizeof(v);
  };
  auto take = [&](size_t len) {
    if (len > size_t(end - p)) throw std::runtime_error("decompress: truncated section");
    std::vector<uint8_t> v(p, p + len);
    p += len;
    return v;
  };
  auto section = [&]() {
    size_t ds, ms;
    rd(ds);
    rd(ms);
    Packed s;
    s.meta = take(ms);
    s.data = take(ds);
    return s;
  };
  uint64_t num_points, param_size;
  rd(num_points);
  rd(param_size);
  for (uint64_t i = 0; i < param_size; ++i) {
    float f;
    rd(f);
  }
  int32_t rx, ry, rz;
  rd(rx);
  rd(ry);
  rd(rz);
  bool is_hilbert, is_direct;
  rd(is_hilbert);
  rd(is_direct);
  std::vector<V3i> coords;
  if (is_direct) {
    auto s1 = section();
    auto patches = gr_unpack<uint64_t>(s1);
    auto s2 = section();
    Huffman<uint64_t> h;
    h.load_meta(s2.meta);
    PFD<uint64_t> pf{patches, h.decode(s2.data, num_points)};
    auto d = pfd_decode(pf);
    if (d.size() != num_points) throw std::runtime_error("decompress: short sfc stream");
    int hb = bit_count_of(std::max({rx, ry, rz}));
    uint64_t acc = 0;
    coords.resize(num_points);
    for (size_t i = 0; i < num_points; i++) {
      acc = (i == 0) ? d[0] : acc + d[i];
      uint64_t m = is_hilbert ? hilbert_to_morton3(acc, hb) : acc;
      coords[i] = decode_morton_code(m);
    }
  } else {
    auto blk = bitunpack(section());
    auto blkcnt = bitunpack(section());
    auto patches = gr_unpack<uint64_t>(section());
    auto s4 = section();
    Huffman<uint64_t> h1;
    h1.load_meta(s4.meta);
    PFD<uint64_t> pf{patches, h1.decode(s4.data, num_points)};
    auto repos = pfd_decode(pf);
    auto patches2 = gr_unpack<uint8_t>(section());
    auto s6 = section();
    Huffman<uint8_t> h2;
    h2.load_meta(s6.meta);
    PFD<uint8_t> pq{patches2, h2.decode(s6.data, num_points)};
    auto quads = pfd_decode(pq);
    for (size_t i = 1; i < blkcnt.size(); i++) blkcnt[i] += blkcnt[i - 1];
    for (size_t i = 1; i < blk.size(); i++) blk[i] += blk[i - 1];
    LCPMeta M;
    M.blkst = blk;
    M.blkcnt = blkcnt;
    M.quads = quads;
    M.repos = repos;
    M.is_hilbert = is_hilbert;
    coords = recover_from_lcp_meta(M, 64, 64, 64, rx, ry, rz);
  }
  std::vector<V3i> full;
  if (rle) {
    auto s7 = section();
    Huffman<int> h3;
    h3.load_meta(s7.meta);
    auto cnts = h3.decode(s7.data, num_points);
    for (size_t i = 0; i < coords.size() && i < cnts.size(); i++)
      for (int j = 0; j < cnts[i]; j++) full.push_back(coords[i]);
  } else {
    full = std::move(coords);
  }
  float ox, oy, oz;
  rd(ox);
  rd(oy);
  rd(oz);
  auto rec = recover(full, qtype, b);
  return unshift_points(rec, V3f{{ox, oy, oz}});
}

std::vector<V3f> to_points(const float* xyz, uint64_t n) {
  std::vector<V3f> pts(n);
  for (uint64_t i = 0; i < n; i++) pts[i] = V3f{{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}};
  return pts;
}

uint8_t* dup_bytes(const std::vector<uint8_t>& v) {
  uint8_t* p = static_cast<uint8_t*>(std::malloc(v.size() ? v.size() : 1));
  if (!v.empty()) std::memcpy(p, v.data(), v.size());
  return p;
}

int fail(const char* what) {
  g_err = what;
  return -1;
}

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

}  // namespace

extern "C" {

int oracle_compress(const float* xyz, uint64_t n, int quantizer, float l2_bound,
                    int is_hilbert, int rle, int direct_threshold, uint8_t** pre,
                    uint64_t* pre_len, uint8_t** frame, uint64_t* frame_len,
                    uint64_t* perm, oracle_info* info) {
  return guarded([&] {
    if (n == 0) return fail("empty input");
    auto R = compress(to_points(xyz, n), quantizer, l2_bound, is_hilbert != 0,
                      rle != 0, direct_threshold);
    if (pre) {
      *pre = dup_bytes(R.pre);
      *pre_len = R.pre.size();
    }
    if (frame) {
      auto f = zstd_compress(R.pre);
      *frame = dup_bytes(f);
      *frame_len = f.size();
    }
    if (perm && !rle)
      for (uint64_t i = 0; i < n; i++) perm[i] = R.perm[i];
    if (info) *info = R.info;
    return 0;
  });
}

int oracle_compress_mode(const float* xyz, uint64_t n, int quantizer, float l2_bound,
                         int is_hilbert, int rle, int direct_threshold, int gr_k, int pfd_width,
                         uint8_t** pre, uint64_t* pre_len, oracle_info* info) {
  return guarded([&] {
    if (n == 0) return fail("empty input");
    if (pfd_width < 0 || pfd_width > 8 || gr_k > 62) return fail("bad mode override");
    g_mode.lim = pfd_width ? (1ull << pfd_width) - 2 : 254;
    g_mode.gr_k = gr_k;
    struct Reset {
      ~Reset() { g_mode = ModeOverride{}; }
    } reset;
    auto R = compress(to_points(xyz, n), quantizer, l2_bound, is_hilbert != 0, rle != 0,
                      direct_threshold);
    *pre = dup_bytes(R.pre);
    *pre_len = R.pre.size();
    if (info) *info = R.info;
    return 0;
  });
}

int oracle_compress_timed(const float* xyz, uint64_t n, int quantizer,
                          float l2_bound, int is_hilbert, int rle,
                          int direct_threshold, int with_zstd,
                          uint64_t* out_bytes, oracle_info* info) {
  return guarded([&] {
    if (n == 0) return fail("empty input");
    auto pts = to_points(xyz, n);  // the caller's points, as main.cpp holds them
    auto t0 = std::chrono::steady_clock::now();
    auto R = compress(pts, quantizer, l2_bound, is_hilbert != 0, rle != 0,
                      direct_threshold);
    uint64_t bytes = R.pre.size();
    if (with_zstd) bytes = zstd_compress(R.pre).size();
    double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (out_bytes) *out_bytes = bytes;
    if (info) {
      *info = R.info;
      info->seconds_total = t;
      info->seconds_stream = t - R.t_diag;
    }
    return 0;
  });
}

int oracle_compress_shared(const float* xyz, uint64_t n, int quantizer, float l2_bound,
                           int is_hilbert, const uint8_t* profile_in, uint64_t profile_len,
                           uint8_t** pre, uint64_t* pre_len, uint8_t** profile_out,
                           uint64_t* profile_out_len) {
  return guarded([&] {
    if (n == 0) return fail("empty input");
    std::vector<uint8_t> pin;
    if (profile_in) pin.assign(profile_in, profile_in + profile_len);
    std::vector<uint8_t> pout;
    auto b = compress_shared(to_points(xyz, n), quantizer, l2_bound, is_hilbert != 0,
                             profile_in ? &pin : nullptr, profile_out ? &pout : nullptr);
    *pre = dup_bytes(b);
    *pre_len = b.size();
    if (profile_out) {
      *profile_out = dup_bytes(pout);
      *profile_out_len = pout.size();
    }
    return 0;
  });
}

int oracle_compress_global_rle(const float* xyz, uint64_t n, int quantizer, float l2_bound,
                               int is_hilbert, uint8_t** pre, uint64_t* pre_len) {
  return guarded([&] {
    if (n == 0) return fail("empty input");
    auto b = compress_global_rle(to_points(xyz, n), quantizer, l2_bound, is_hilbert != 0);
    *pre = dup_bytes(b);
    *pre_len = b.size();
    return 0;
  });
}

int oracle_decompress_buffer(const uint8_t* buf, uint64_t len, int quantizer,
                             float l2_bound, int rle, float** xyz, uint64_t* n) {
  return guarded([&] {
    std::vector<uint8_t> b(buf, buf + len);
    auto pts = decompress_buffer(b, quantizer, l2_bound, rle != 0);
    float* out = static_cast<float*>(std::malloc(pts.size() * 12 + 12));
    for (size_t i = 0; i < pts.size(); i++)
      for (int a = 0; a < 3; a++) out[3 * i + a] = pts[i][a];
    *xyz = out;
    *n = pts.size();
    return 0;
  });
}

int oracle_decompress_frame(const uint8_t* frame, uint64_t len, int quantizer,
                            float l2_bound, int rle, float** xyz, uint64_t* n) {
  return guarded([&] {
    auto b = zstd_decompress(frame, len);
    return oracle_decompress_buffer(b.data(), b.size(), quantizer, l2_bound, rle, xyz, n);
  });
}

void oracle_shift(const float* xyz, uint64_t n, float* shifted, float* offset) {
  V3f off{};
  auto s = shift_points(to_points(xyz, n), off);
  for (uint64_t i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) shifted[3 * i + a] = s[i][a];
  for (int a = 0; a < 3; a++) offset[a] = off[a];
}

void oracle_quantize(const float* xyz, uint64_t n, int quantizer, float l2_bound,
                     int32_t* nodes, float* offset) {
  V3f off{};
  auto s = shift_points(to_points(xyz, n), off);
  auto q = quantize(s, quantizer, l2_bound);
  for (uint64_t i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) nodes[3 * i + a] = q[i][a];
  if (offset)
    for (int a = 0; a < 3; a++) offset[a] = off[a];
}

void oracle_recover(const int32_t* nodes, uint64_t n, int quantizer, float l2_bound,
                    float* out) {
  std::vector<V3i> q(n);
  for (uint64_t i = 0; i < n; i++) q[i] = V3i{{nodes[3 * i], nodes[3 * i + 1], nodes[3 * i + 2]}};
  auto r = recover(q, quantizer, l2_bound);
  for (uint64_t i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) out[3 * i + a] = r[i][a];
}

uint64_t oracle_morton_code(int32_t x, int32_t y, int32_t z) {
  return morton_code(V3i{{x, y, z}});
}
void oracle_decode_morton(uint64_t m, int32_t* xyz) {
  V3i v = decode_morton_code(m);
  for (int a = 0; a < 3; a++) xyz[a] = v[a];
}
uint64_t oracle_morton_to_hilbert3(uint64_t m, unsigned it) { return morton_to_hilbert3(m, it); }
uint64_t oracle_hilbert_to_morton3(uint64_t h, unsigned it) { return hilbert_to_morton3(h, it); }

void oracle_hilbert_tables(uint8_t* m2h_perm, uint8_t* h2m_perm, uint8_t* m2h_perms,
                           uint8_t* h2m_perms) {
  const auto& T = htab();
  for (int i = 0; i < 8; i++) {
    m2h_perm[i] = T.m2h_perm[i];
    h2m_perm[i] = T.h2m_perm[i];
    for (int j = 0; j < 8; j++) {
      m2h_perms[8 * i + j] = T.m2h_perms[i][j];
      h2m_perms[8 * i + j] = T.h2m_perms[i][j];
    }
  }
}

static void out_packed(const Packed& p, uint8_t** meta, uint64_t* ml, uint8_t** data,
                       uint64_t* dl) {
  *meta = dup_bytes(p.meta);
  *ml = p.meta.size();
  *data = dup_bytes(p.data);
  *dl = p.data.size();
}

int oracle_bitpack_u64(const uint64_t* v, uint64_t n, uint8_t** meta, uint64_t* ml,
                       uint8_t** data, uint64_t* dl) {
  return guarded([&] {
    out_packed(bitpack(std::vector<uint64_t>(v, v + n)), meta, ml, data, dl);
    return 0;
  });
}
int oracle_gr_pack_u64(const uint64_t* v, uint64_t n, uint8_t** meta, uint64_t* ml,
                       uint8_t** data, uint64_t* dl) {
  return guarded([&] {
    out_packed(gr_pack(std::vector<uint64_t>(v, v + n)), meta, ml, data, dl);
    return 0;
  });
}
int oracle_gr_pack_u8(const uint8_t* v, uint64_t n, uint8_t** meta, uint64_t* ml,
                      uint8_t** data, uint64_t* dl) {
  return guarded([&] {
    out_packed(gr_pack(std::vector<uint8_t>(v, v + n)), meta, ml, data, dl);
    return 0;
  });
}

}  // extern "C"

template <typename T>
static int huff_impl(const T* v, uint64_t n, int mode, uint8_t** meta, uint64_t* ml,
                     uint8_t** data, uint64_t* dl) {
  std::vector<T> d(v, v + n);
  Huffman<T> h;
  if (mode == 3 || mode == 4) {  // build_complete(d, 255); 4: reloaded by load_meta
    h.build_complete(d, T(255));
    if (mode == 4) {
      Huffman<T> h2;
      h2.load_meta(h.meta());
      out_packed(Packed{h2.meta(), h2.encode(d)}, meta, ml, data, dl);
      return 0;
    }
    out_packed(Packed{h.meta(), h.encode(d)}, meta, ml, data, dl);
    return 0;
  }
  if (mode == 2) h.use_eg = true;
  h.build(d);
  auto m = h.meta();
  auto e = h.encode(d);
  if (mode == 1) {
    if (static_cast<double>(m.size()) > static_cast<double>(e.size()) * 0.25) h.use_eg = true;
    h.build(d);
    m = h.meta();
    e = h.encode(d);
  }
  out_packed(Packed{m, e}, meta, ml, data, dl);
  return 0;
}

extern "C" {

int oracle_huffman(const void* v, uint64_t n, int sym_type, int mode, uint8_t** meta,
                   uint64_t* ml, uint8_t** data, uint64_t* dl) {
  return guarded([&] {
    if (sym_type == 0) return huff_impl(static_cast<const uint64_t*>(v), n, mode, meta, ml, data, dl);
    if (sym_type == 1) return huff_impl(static_cast<const uint8_t*>(v), n, mode, meta, ml, data, dl);
    return huff_impl(static_cast<const int*>(v), n, mode, meta, ml, data, dl);
  });
}

}  // extern "C"

template <typename T>
static int huff_dec_impl(const uint8_t* meta, uint64_t ml, const uint8_t* data,
                         uint64_t dl, uint64_t count, T* out) {
  Huffman<T> h;
  h.load_meta(std::vector<uint8_t>(meta, meta + ml));
  auto r = h.decode(std::vector<uint8_t>(data, data + dl), count);
  if (r.size() != count) return fail("huffman: short data");
  std::copy(r.begin(), r.end(), out);
  return 0;
}

extern "C" {

int oracle_huffman_decode(const uint8_t* meta, uint64_t ml, const uint8_t* data,
                          uint64_t dl, uint64_t count, int sym_type, void* out) {
  return guarded([&] {
    if (sym_type == 0) return huff_dec_impl(meta, ml, data, dl, count, static_cast<uint64_t*>(out));
    if (sym_type == 1) return huff_dec_impl(meta, ml, data, dl, count, static_cast<uint8_t*>(out));
    return huff_dec_impl(meta, ml, data, dl, count, static_cast<int*>(out));
  });
}

int oracle_gr_unpack_u64(const uint8_t* meta, uint64_t ml, const uint8_t* data,
                         uint64_t dl, uint64_t* out, uint64_t* n) {
  return guarded([&] {
    Packed p{std::vector<uint8_t>(meta, meta + ml), std::vector<uint8_t>(data, data + dl)};
    auto r = gr_unpack<uint64_t>(p);
    std::copy(r.begin(), r.end(), out);
    *n = r.size();
    return 0;
  });
}

int oracle_bitunpack_u64(const uint8_t* meta, uint64_t ml, const uint8_t* data,
                         uint64_t dl, uint64_t* out, uint64_t* n) {
  return guarded([&] {
    Packed p{std::vector<uint8_t>(meta, meta + ml), std::vector<uint8_t>(data, data + dl)};
    auto r = bitunpack(p);
    std::copy(r.begin(), r.end(), out);
    *n = r.size();
    return 0;
  });
}

uint64_t oracle_zstd_version(void) { return ZSTD_versionNumber(); }
const char* oracle_last_error(void) { return g_err.c_str(); }
void oracle_free(void* p) { std::free(p); }

}  // extern "C"
