// hilbert_sm.cpp -- Morton -> Hilbert digit tables as a state machine.
//
// The reference builds its tables at static-init time
// (source/quantizer/hilbert-curves-cpp/hilbert.cpp:138-341): base pattern A
// {0,1,3,2,6,7,5,4}, rotation 0; the 24 cube rotations expressed as
// permutations of the 8 octant indices and ordered as a std::set; a depth-first
// search that gives each octant (in curve order) the first rotation whose
// entry cell touches the previous octant's exit cell.  Those tables must be
// the same here, so the construction follows the same steps.
//
// mortonToHilbert3 (hilbert.cpp:423-448) walks the digits from the top,
// emits mortonToHilbertPerm[d] for the current (already permuted) digit d
// and then permutes every lower digit by mortonToHilbertPermutations[d].
// The composition of those permutations is an element of the rotation group,
// so the walk is a finite-state machine: state = composed permutation,
// apply[s][digit], next[s][emitted digit].  The kernels run that machine, one
// table lookup per digit instead of a loop over all lower digits.
#include <array>
#include <map>
#include <set>
#include <vector>

#include "decode_dev.hpp"
#include "kernels.hpp"

namespace xnyz {
namespace {

using P8 = std::array<u8, 8>;
using V3 = std::array<int, 3>;

P8 inverse(const P8& p) {
  P8 r{};
  for (int i = 0; i < 8; ++i) r[p[i]] = static_cast<u8>(i);
  return r;
}

V3 octant(int v) { return V3{(v >> 2) & 1, (v >> 1) & 1, v & 1}; }

// Rotation about x, y, z by quarter turns (rx, ry, rz), as the permutation it
// induces on the octant indices x<<2 | y<<1 | z (hilbert.cpp:150-200).
P8 rotation_perm(int rx, int ry, int rz) {
  static const int sn[4] = {0, 1, 0, -1}, cs[4] = {1, 0, -1, 0};
  const int sx = sn[rx], sy = sn[ry], sz = sn[rz], cx = cs[rx], cy = cs[ry], cz = cs[rz];
  const int m[3][3] = {{cx * cy, sx * sz - cx * cz * sy, cz * sx + cx * sy * sz},
                       {sy, cy * cz, -cy * sz},
                       {-cy * sx, cx * sz + cz * sx * sy, cx * cz - sx * sy * sz}};
  P8 r{};
  for (int i = 0; i < 8; ++i) {
    const V3 o = octant(i);
    const int v[3] = {o[0] * 2 - 1, o[1] * 2 - 1, o[2] * 2 - 1};
    int out = 0;
    for (int k = 0; k < 3; ++k) {
      const int d = m[k][0] * v[0] + m[k][1] * v[1] + m[k][2] * v[2];
      out = (out << 1) | ((d + 1) / 2);
    }
    r[i] = static_cast<u8>(out);
  }
  return r;
}

struct Tables {
  P8 h2m{0, 1, 3, 2, 6, 7, 5, 4};  // pattern A
  P8 m2h{};
  std::vector<P8> rots;
  std::array<P8, 8> h2m_perms{}, m2h_perms{};

  bool touch(V3 a, V3 b) {
    int d = 0;
    for (int k = 0; k < 3; ++k) d += (a[k] - b[k]) * (a[k] - b[k]);
    return d == 1;
  }
  // Piece i of the 2x2x2 curve (hilbert.cpp:253-321).
  bool place(int i, V3 prev_exit) {
    const V3 first = octant(h2m[0]), last = octant(h2m[7]), cell = octant(h2m[i]);
    for (const P8& r : rots) {
      V3 in = octant(r[h2m[0]]), out = octant(r[h2m[7]]);
      for (int k = 0; k < 3; ++k) {
        in[k] += 2 * cell[k];
        out[k] += 2 * cell[k];
      }
      if (i == 0) {
        if (in != V3{3 * first[0], 3 * first[1], 3 * first[2]}) continue;
      } else if (!touch(in, prev_exit)) {
        continue;
      }
      if (i != 7) {
        if (!place(i + 1, out)) continue;
      } else if (out != V3{3 * last[0], 3 * last[1], 3 * last[2]}) {
        continue;
      }
      h2m_perms[h2m[i]] = r;
      return true;
    }
    return false;
  }
  Tables() {
    m2h = inverse(h2m);
    std::set<P8> all;
    for (int rx = 0; rx < 4; ++rx)
      for (int ry = 0; ry < 4; ++ry)
        for (int rz = 0; rz < 4; ++rz) all.insert(rotation_perm(rx, ry, rz));
    rots.assign(all.begin(), all.end());
    if (rots.size() != 24 || !place(0, V3{0, 0, 0}))
      throw Error(-7, "hilbert: table construction failed");
    for (int i = 0; i < 8; ++i) m2h_perms[i] = inverse(h2m_perms[i]);
  }
};

}  // namespace

// hilbertToMorton3 (hilbert.cpp:396-421), for the host decoder: undo the
// lowest digit, then walk up, mapping each digit back and re-permuting the
// digits below it.
uint64_t hilbert_to_morton3(uint64_t h, unsigned iteration) {
  static const Tables t;
  const uint8_t lh = h & 7;
  uint64_t r = (h ^ lh) | t.h2m[lh];
  for (unsigned i = 1; i <= iteration; ++i) {
    const unsigned sh = i * 3;
    const uint8_t hd = (r >> sh) & 7, md = t.h2m[hd];
    r ^= uint64_t{hd} << sh;
    r |= uint64_t{md} << sh;
    const P8& perm = t.h2m_perms[md];
    for (unsigned j = i - 1; j < i; --j) {
      const unsigned cs = j * 3;
      const uint8_t c = (r >> cs) & 7;
      r ^= uint64_t{c} << cs;
      r |= uint64_t{perm[c]} << cs;
    }
  }
  return r;
}

// hilbertToMorton3 as a machine over the same rotation group, walked from the
// top digit down.  The bottom-up loop of the reference maps Hilbert digit i
// to Morton digit md_i = h2m[h_i] (digit i is not yet touched when its turn
// comes) and then permutes every digit below i by h2m_perms[md_i]; so Morton
// digit j ends as P_L(...P_{j+1}(h2m[h_j])), P_i = h2m_perms[h2m[h_i]].  Top
// down: state Q (identity first), digit j -> Q(h2m[h_j]), then
// Q <- Q o P_j.
HilbertInv build_hilbert_inv() {
  const Tables t;
  std::vector<P8> states;
  std::map<P8, int> index;
  P8 id{};
  for (int i = 0; i < 8; ++i) id[i] = static_cast<u8>(i);
  states.push_back(id);
  index[id] = 0;
  HilbertInv inv{};
  for (size_t s = 0; s < states.size(); ++s) {
    for (int hd = 0; hd < 8; ++hd) {
      const u8 md = t.h2m[hd];
      inv.apply[s * 8 + hd] = states[s][md];
      P8 nxt{};
      for (int x = 0; x < 8; ++x) nxt[x] = states[s][t.h2m_perms[md][x]];
      auto it = index.find(nxt);
      int ni;
      if (it == index.end()) {
        ni = static_cast<int>(states.size());
        if (ni >= 24) throw Error(-7, "hilbert: more than 24 inverse states");
        index[nxt] = ni;
        states.push_back(nxt);
      } else {
        ni = it->second;
      }
      inv.next[s * 8 + hd] = static_cast<u8>(ni);
    }
  }
  return inv;
}

HilbertSM build_hilbert_sm() {
  const Tables t;
  std::vector<P8> states;
  std::map<P8, int> index;
  P8 id{};
  for (int i = 0; i < 8; ++i) id[i] = static_cast<u8>(i);
  states.push_back(id);
  index[id] = 0;
  HilbertSM sm{};
  for (size_t s = 0; s < states.size(); ++s) {
    for (int md = 0; md < 8; ++md) {
      // lower digits see the current permutation first, then m2h_perms[md]
      P8 nxt{};
      for (int x = 0; x < 8; ++x) nxt[x] = t.m2h_perms[md][states[s][x]];
      auto it = index.find(nxt);
      int ni;
      if (it == index.end()) {
        ni = static_cast<int>(states.size());
        if (ni >= 24) throw Error(-7, "hilbert: more than 24 states");
        index[nxt] = ni;
        states.push_back(nxt);
      } else {
        ni = it->second;
      }
      sm.next[s * 8 + md] = static_cast<u8>(ni);
    }
  }
  for (size_t s = 0; s < states.size(); ++s)
    for (int d = 0; d < 8; ++d) sm.apply[s * 8 + d] = states[s][d];
  for (int d = 0; d < 8; ++d) sm.out[d] = t.m2h[d];
  sm.identity = 0;
  sm.nstates = static_cast<u8>(states.size());
  // Block mode codes are below 2^18: digits 20..6 are zero.  They must stay
  // zero (emit 0) for the shortcut to hold; the state they leave behind is the
  // start state of the 6 live digits.
  int s = 0;
  for (int i = 20; i >= 6; --i) {
    const int md = sm.apply[s * 8 + 0];
    if (sm.out[md] != 0) throw Error(-7, "hilbert: zero digit does not map to zero");
    s = sm.next[s * 8 + md];
  }
  sm.block_init = static_cast<u8>(s);
  return sm;
}

}  // namespace xnyz
