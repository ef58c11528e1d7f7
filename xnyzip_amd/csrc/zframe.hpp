// zframe.hpp -- a zstd-format (RFC 8878) block encoder, written once for the
// host and the device (XZ_HD).  It emits valid zstd frames for any zstd
// decoder; it is not libzstd's compressor and its frames are not
// byte-identical to ZSTD_compress's (the assembled stream inside is).
//
// Block format used (RFC 8878 section 3.1.1):
//   * LZ: lazy parse of fixed 512-byte segments of a 16 KiB block.  Every
//     position's best fixed candidate -- offsets 1, 8, 16 and the hits of two
//     hash tables (4- and 8-byte keys) holding the EARLIEST position of each
//     hash in the block (order-independent, so a parallel build gives the
//     same tables) -- is chosen with lengths capped at kCap and kept as one
//     byte; the parse extends a capped winner and adds the segment's last
//     three offsets, scoring length minus an offset cost.  Matches stay
//     inside their segment and reference earlier bytes of the same block.
//   * Literals: Huffman-coded (1 stream up to 1023 literals, else 4 streams),
//     weights FSE-compressed (or direct 4-bit when that is smaller), code
//     lengths limited to 11 bits; raw or RLE literals when smaller.
//   * Sequences: per block and code kind (literal lengths, offsets, match
//     lengths) the cheaper of the predefined distribution and the block's own
//     FSE table by estimated bits (RLE mode for a single code); repeat
//     offsets (RFC 8878 3.1.1.5) only where the repeated offset was coded
//     earlier in the same block, so no block relies on the history left by
//     the blocks before it.
//   * A block whose encoding is not smaller than its bytes is stored raw.
#pragma once

#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define XZ_HD __host__ __device__
#else
#define XZ_HD
#endif

namespace xz {

using u8 = uint8_t;
using u16 = uint16_t;
using u32 = uint32_t;
using i16 = int16_t;
using i32 = int32_t;
using u64 = unsigned long long;

constexpr u32 kBlock = 1u << 14;   // bytes per zstd block
constexpr u32 kSeg = 512;          // bytes per parse segment
constexpr u32 kSegs = kBlock / kSeg;
constexpr u32 kHashLog = 12;
constexpr u32 kCap = 35;           // longest length a fixed candidate is measured to
constexpr u32 kMinMatch = 4;
constexpr u32 kMaxHufBits = 11;

XZ_HD inline u32 highbit(u32 v) { return 31u - (u32)__builtin_clz(v); }  // v > 0

XZ_HD inline u32 read32(const u8* p) {
  return (u32)p[0] | ((u32)p[1] << 8) | ((u32)p[2] << 16) | ((u32)p[3] << 24);
}
XZ_HD inline u32 hash4(u32 v) { return (v * 2654435761u) >> (32 - kHashLog); }
XZ_HD inline u64 read64(const u8* p) {
  return (u64)read32(p) | ((u64)read32(p + 4) << 32);
}
XZ_HD inline u32 hash8(u64 v) { return (u32)((v * 0x9E3779B97F4A7C15ull) >> (64 - kHashLog)); }

// ---------------------------------------------------------------------------
// Bit writer, LSB-first into little-endian bytes (the layout every zstd
// bitstream uses; decoders read it backwards from the final end-mark bit).
struct BitW {
  u8* out;
  u64 acc = 0;
  u32 nb = 0;
  u32 pos = 0;
  XZ_HD explicit BitW(u8* o) : out(o) {}
  XZ_HD void add(u64 v, u32 n) {  // n <= 56 - 7
    if (!n) return;
    acc |= (v & ((1ull << n) - 1)) << nb;
    nb += n;
    while (nb >= 8) {
      out[pos++] = (u8)acc;
      acc >>= 8;
      nb -= 8;
    }
  }
  XZ_HD u32 close() {  // end mark; returns the byte count
    add(1, 1);
    if (nb) out[pos++] = (u8)acc;
    return pos;
  }
};

// ---------------------------------------------------------------------------
// FSE (tANS) encoding tables from normalised counts (RFC 8878 4.1.1: symbols
// spread with step (size >> 1) + (size >> 3) + 3, "-1" symbols at the top).

constexpr u32 kFseMaxLog = 9;
constexpr u32 kFseMaxSym = 64;

struct FseTable {
  u32 log;
  u16 state[1u << kFseMaxLog];
  i32 dnb[kFseMaxSym];  // deltaNbBits
  i32 dfs[kFseMaxSym];  // deltaFindState
};

// spread: 2^log bytes, cumul: nsym + 1 words of working storage.
XZ_HD inline void fse_build(FseTable& t, const i16* norm, u32 nsym, u32 log, u8* spread,
                            u32* cumul) {
  const u32 size = 1u << log, mask = size - 1;
  u32 high = size - 1;
  cumul[0] = 0;
  for (u32 s = 0; s < nsym; ++s) {
    if (norm[s] == -1) {
      cumul[s + 1] = cumul[s] + 1;
      spread[high--] = (u8)s;
    } else {
      cumul[s + 1] = cumul[s] + (u32)norm[s];
    }
  }
  u32 p = 0;
  const u32 step = (size >> 1) + (size >> 3) + 3;
  for (u32 s = 0; s < nsym; ++s)
    for (i32 i = 0; i < norm[s]; ++i) {
      spread[p] = (u8)s;
      do p = (p + step) & mask;
      while (p > high);
    }
  for (u32 u = 0; u < size; ++u) {
    const u32 s = spread[u];
    t.state[cumul[s]++] = (u16)(size + u);
  }
  i32 total = 0;
  for (u32 s = 0; s < nsym; ++s) {
    const i32 n = norm[s];
    if (n == 0) {
      t.dnb[s] = (i32)(((log + 1) << 16) - size);
      t.dfs[s] = 0;
    } else if (n == -1 || n == 1) {
      t.dnb[s] = (i32)((log << 16) - size);
      t.dfs[s] = total - 1;
      total += 1;
    } else {
      const u32 maxb = log - highbit((u32)n - 1);
      const u32 minplus = (u32)n << maxb;
      t.dnb[s] = (i32)((maxb << 16) - minplus);
      t.dfs[s] = total - n;
      total += n;
    }
  }
  t.log = log;
}

struct FseState {
  u32 v;
};

// Working storage of the serial steps below (host: any memory; device: LDS,
// with one wavefront executing the steps uniformly).  The Huffman build's
// arrays are dead once it returns (its result is the HufTable), so the
// weight description and the FSE table builds that follow share their
// space.
struct Work {
  union {
    struct {  // huf_build
      u32 cnt[256];
      u16 sym[256], tmp[256];
      u32 pos[257];
      u32 w[512];  // tree weights: sums of one block's counts (< 2^15)
      i16 parent[512];
      u16 depth[512];
    };
    struct {  // huf_write_desc, the sequence tables
      u8 spread[1u << kFseMaxLog];
      u32 cumul[kFseMaxSym + 1];
      u8 wt[256];
      u8 dtmp[160];
      u32 wcount[16];
      i16 norm[kFseMaxSym];
      u8 ntmp[128];
      FseTable wtab;
    };
  };
};

XZ_HD inline void fse_init(FseState& st, const FseTable& t, u32 s) {
  const u32 nbo = (u32)((t.dnb[s] + (1 << 15)) >> 16);
  const u32 v = (nbo << 16) - (u32)t.dnb[s];
  st.v = t.state[(v >> nbo) + t.dfs[s]];
}
XZ_HD inline void fse_encode(BitW& bw, FseState& st, const FseTable& t, u32 s) {
  const u32 nbo = (st.v + (u32)t.dnb[s]) >> 16;
  bw.add(st.v, nbo);
  st.v = t.state[(st.v >> nbo) + t.dfs[s]];
}
XZ_HD inline void fse_flush(BitW& bw, const FseState& st, const FseTable& t) {
  bw.add(st.v, t.log);
}

// FSE table description (RFC 8878 4.1.1).  Returns bytes written.
XZ_HD inline u32 fse_write_ncount(u8* out, const i16* norm, u32 nsym, u32 log) {
  const u32 size = 1u << log;
  u64 bits = 0;
  u32 nbits = 0, pos = 0;
  auto flush = [&] {
    while (nbits >= 8) {
      out[pos++] = (u8)bits;
      bits >>= 8;
      nbits -= 8;
    }
  };
  bits |= (u64)(log - 5) << nbits;
  nbits += 4;
  i32 remaining = (i32)size + 1;
  i32 threshold = (i32)size;
  u32 nb = log + 1;
  u32 s = 0;
  bool prev0 = false;
  while (s < nsym && remaining > 1) {
    if (prev0) {
      u32 start = s;
      while (s < nsym && norm[s] == 0) ++s;
      while (s >= start + 3) {  // runs of three more zeros: repeat flag 3
        start += 3;
        bits |= 3ull << nbits;
        nbits += 2;
        flush();
      }
      bits |= (u64)(s - start) << nbits;
      nbits += 2;
      flush();
    }
    i32 count = norm[s++];
    const i32 maxv = (2 * threshold - 1) - remaining;
    remaining -= count < 0 ? -count : count;
    count++;  // 0 codes the "-1" probability
    if (count >= threshold) count += maxv;
    bits |= (u64)(u32)count << nbits;
    nbits += nb;
    if (count < maxv) nbits -= 1;
    flush();
    prev0 = count == 1;
    while (remaining < threshold) {
      nb--;
      threshold >>= 1;
    }
  }
  if (nbits) out[pos++] = (u8)bits;
  return pos;
}

// Normalised counts summing to 2^log, every present symbol >= 1.
XZ_HD inline void fse_normalize(const u32* count, u32 nsym, u32 total, u32 log, i16* norm) {
  const u32 size = 1u << log;
  i32 sum = 0;
  u32 big = 0;
  for (u32 s = 0; s < nsym; ++s) {
    if (!count[s]) {
      norm[s] = 0;
      continue;
    }
    u32 v = (u32)(((u64)count[s] * size + total / 2) / total);
    if (v < 1) v = 1;
    norm[s] = (i16)v;
    sum += (i32)v;
    if (count[s] > count[big] || !count[big]) big = s;
  }
  // fix the sum on the largest symbols, keeping every present symbol >= 1
  while (sum != (i32)size) {
    if (sum > (i32)size) {
      u32 best = nsym;
      for (u32 s = 0; s < nsym; ++s)
        if (norm[s] > 1 && (best == nsym || norm[s] > norm[best])) best = s;
      if (best == nsym) return;  // more symbols than slots: callers size log to avoid this
      norm[best]--;
      sum--;
    } else {
      norm[big]++;
      sum++;
    }
  }
}

// ---------------------------------------------------------------------------
// Huffman code lengths <= kMaxHufBits (Huffman on counts halved until the
// longest code fits), canonical codes and the tree description.

struct HufTable {
  u8 nbits[256];
  u16 code[256];
  u32 maxbits;
  u32 last;  // highest symbol present
};

// Returns false when fewer than two symbols are present.
XZ_HD inline bool huf_build(const u32* count_in, HufTable& h, Work& W) {
  u32* cnt = W.cnt;
  u32 nsym = 0;
  for (u32 s = 0; s < 256; ++s) {
    cnt[s] = count_in[s];
    if (cnt[s]) {
      ++nsym;
      h.last = s;
    }
  }
  if (nsym < 2) return false;
  for (;;) {
    // Two-queue Huffman over the present symbols sorted by count.
    // present symbols sorted by (count, symbol): stable LSD radix sort over
    // the count's bytes, symbols entered in increasing order
    u16* sym = W.sym;
    u16* tmp = W.tmp;
    u32 n = 0;
    u32 maxc = 0;
    for (u32 s = 0; s < 256; ++s)
      if (cnt[s]) {
        sym[n++] = (u16)s;
        maxc = cnt[s] > maxc ? cnt[s] : maxc;
      }
    for (u32 sh = 0; sh < 32 && (maxc >> sh); sh += 8) {
      u32* pos = W.pos;
      for (u32 i = 0; i < 257; ++i) pos[i] = 0;
      for (u32 i = 0; i < n; ++i) pos[((cnt[sym[i]] >> sh) & 255u) + 1]++;
      for (u32 i = 1; i < 257; ++i) pos[i] += pos[i - 1];
      for (u32 i = 0; i < n; ++i) tmp[pos[(cnt[sym[i]] >> sh) & 255u]++] = sym[i];
      for (u32 i = 0; i < n; ++i) sym[i] = tmp[i];
    }
    u32* w = W.w;
    i16* parent = W.parent;
    for (u32 i = 0; i < n; ++i) w[i] = cnt[sym[i]];
    u32 leaf = 0, node = n, inner = n;  // inner nodes created at [n, 2n-1)
    for (u32 k = 0; k + 1 < n; ++k) {
      u32 pick[2];
      for (int c = 0; c < 2; ++c) {
        if (leaf < n && (inner >= node || w[leaf] <= w[inner])) pick[c] = leaf++;
        else pick[c] = inner++;
      }
      w[node] = w[pick[0]] + w[pick[1]];
      parent[pick[0]] = parent[pick[1]] = (i16)node;
      ++node;
    }
    u16* depth = W.depth;
    depth[node - 1] = 0;
    u32 maxd = 0;
    for (i32 i = (i32)node - 2; i >= 0; --i) {
      depth[i] = (u16)(depth[parent[i]] + 1);
      if ((u32)i < n && depth[i] > maxd) maxd = depth[i];
    }
    if (maxd <= kMaxHufBits) {
      for (u32 s = 0; s < 256; ++s) h.nbits[s] = 0;
      for (u32 i = 0; i < n; ++i) h.nbits[sym[i]] = (u8)depth[i];
      h.maxbits = maxd;
      break;
    }
    for (u32 s = 0; s < 256; ++s)
      if (cnt[s]) cnt[s] = (cnt[s] >> 1) | 1u;
  }
  // canonical codes (RFC 8878 4.2.1.3): longest codes first, from 0, by symbol
  u32 next = 0;
  for (u32 len = h.maxbits; len >= 1; --len) {
    for (u32 s = 0; s < 256; ++s)
      if (h.nbits[s] == len) h.code[s] = (u16)next++;
    next >>= 1;
  }
  return true;
}

// Huffman tree description (RFC 8878 4.2.1.1): weights of symbols
// 0..last-1 (the last one is implied), FSE-compressed or direct.  Returns
// bytes written, 0 if not representable.
XZ_HD inline u32 huf_write_desc(u8* out, const HufTable& h, Work& W) {
  u8* wt = W.wt;
  for (u32 s = 0; s < h.last; ++s) wt[s] = h.nbits[s] ? (u8)(h.maxbits + 1 - h.nbits[s]) : 0;
  const u32 nw = h.last;
  // direct: 4 bits per weight, at most 128 weights
  const u32 direct = nw <= 128 ? 1 + (nw + 1) / 2 : 0xFFFFFFFFu;
  // FSE: accuracy log 6, two interleaved states
  u32 fse = 0xFFFFFFFFu;
  u8* tmp = W.dtmp;
  if (nw >= 2) {
    u32* count = W.wcount;
    for (u32 i = 0; i < 16; ++i) count[i] = 0;
    u32 maxw = 0;
    for (u32 i = 0; i < nw; ++i) {
      count[wt[i]]++;
      maxw = wt[i] > maxw ? wt[i] : maxw;
    }
    const u32 nsym = maxw + 1;
    u32 present = 0;
    for (u32 i = 0; i < nsym; ++i) present += count[i] != 0;
    if (present >= 2) {
      i16* norm = W.norm;
      const u32 log = 6;
      fse_normalize(count, nsym, nw, log, norm);
      FseTable& t = W.wtab;
      fse_build(t, norm, nsym, log, W.spread, W.cumul);
      u32 hp = fse_write_ncount(tmp, norm, nsym, log);
      BitW bw(tmp + hp);
      FseState s1, s2;
      i32 i = (i32)nw;
      if (nw & 1) {
        fse_init(s1, t, wt[--i]);
        fse_init(s2, t, wt[--i]);
        fse_encode(bw, s1, t, wt[--i]);
      } else {
        fse_init(s2, t, wt[--i]);
        fse_init(s1, t, wt[--i]);
      }
      while (i > 0) {
        fse_encode(bw, s2, t, wt[--i]);
        fse_encode(bw, s1, t, wt[--i]);
      }
      fse_flush(bw, s2, t);
      fse_flush(bw, s1, t);
      const u32 n = hp + bw.close();
      if (n < 128) fse = 1 + n;
    }
  }
  if (fse <= direct && fse != 0xFFFFFFFFu) {
    out[0] = (u8)(fse - 1);
    for (u32 i = 0; i + 1 < fse; ++i) out[1 + i] = tmp[i];
    return fse;
  }
  if (direct == 0xFFFFFFFFu) return 0;
  out[0] = (u8)(127 + nw);
  for (u32 i = 0; i < nw; i += 2)
    out[1 + i / 2] = (u8)((wt[i] << 4) | (i + 1 < nw ? wt[i + 1] : 0));
  return direct;
}

// One Huffman stream of n literals (encoded last to first).
XZ_HD inline u32 huf_stream(u8* out, const u8* lit, u32 n, const HufTable& h) {
  BitW bw(out);
  for (u32 i = n; i-- > 0;) bw.add(h.code[lit[i]], h.nbits[lit[i]]);
  return bw.close();
}

// ---------------------------------------------------------------------------
// Sequences with the predefined distributions (RFC 8878 3.1.1.3.2.2).

struct Seq {
  u32 ll, ml, off;  // literal length, match length (>= 3), offset (>= 1)
};

XZ_HD inline u32 ll_code(u32 ll) {
  if (ll < 16) return ll;
  if (ll < 24) return 16 + ((ll - 16) >> 1);
  if (ll < 32) return 20 + ((ll - 24) >> 2);
  if (ll < 48) return 22 + ((ll - 32) >> 3);
  if (ll < 64) return 24;
  return 19 + highbit(ll);  // 64..127 -> 25, 128..255 -> 26, ...
}
XZ_HD inline u32 ll_base(u32 c) {
  if (c < 16) return c;
  if (c < 20) return 16 + 2 * (c - 16);
  if (c < 22) return 24 + 4 * (c - 20);
  if (c < 24) return 32 + 8 * (c - 22);
  if (c == 24) return 48;
  return 1u << (c - 19);
}
XZ_HD inline u32 ll_bits(u32 c) {
  if (c < 16) return 0;
  if (c < 20) return 1;
  if (c < 22) return 2;
  if (c < 24) return 3;
  if (c == 24) return 4;
  return c - 19;
}
XZ_HD inline u32 ml_code(u32 ml) {  // ml >= 3
  const u32 m = ml - 3;
  if (m < 32) return m;
  if (m < 40) return 32 + ((m - 32) >> 1);
  if (m < 48) return 36 + ((m - 40) >> 2);
  if (m < 64) return 38 + ((m - 48) >> 3);
  if (m < 96) return 40 + ((m - 64) >> 4);
  if (m < 128) return 42;
  return 36 + highbit(m);  // 128..255 -> 43, ...
}
XZ_HD inline u32 ml_base(u32 c) {
  if (c < 32) return c + 3;
  if (c < 36) return 35 + 2 * (c - 32);
  if (c < 38) return 43 + 4 * (c - 36);
  if (c < 40) return 51 + 8 * (c - 38);
  if (c < 42) return 67 + 16 * (c - 40);
  if (c == 42) return 99;
  return (1u << (c - 36)) + 3;
}
XZ_HD inline u32 ml_bits(u32 c) {
  if (c < 32) return 0;
  if (c < 36) return 1;
  if (c < 38) return 2;
  if (c < 40) return 3;
  if (c < 42) return 4;
  if (c == 42) return 5;
  return c - 36;
}

// Predefined distributions (RFC 8878 3.1.1.3.2.2).
XZ_HD inline const i16* pre_norm(int which) {  // 0 LL (36, log 6), 1 OF (29, log 5), 2 ML (53, log 6)
  static const i16 ll[36] = {4, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 2, 2,
                             2, 2, 2, 2, 2, 2, 2, 3, 2, 1, 1, 1, 1, 1, -1, -1, -1, -1};
  static const i16 of[29] = {1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1,
                             1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1};
  static const i16 ml[53] = {1, 4, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                             1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                             1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, -1, -1};
  return which == 0 ? ll : which == 1 ? of : ml;
}
constexpr u32 kPreSyms[3] = {36, 29, 53};
constexpr u32 kPreLog[3] = {6, 5, 6};
constexpr u32 kMaxLog[3] = {9, 8, 9};  // LL, OF, ML accuracy limits
constexpr u32 kMaxSym[3] = {35, 31, 52};

// -log2 of n / 2^log in 1/16 bits, for cost estimates (n >= 1).
XZ_HD inline u32 cost16(u32 n, u32 log) {
  // 16 * (log - log2(n)), log2 by the top bits of n
  const u32 hb = highbit(n);
  const u32 frac = hb >= 4 ? (n >> (hb - 4)) & 15 : (n << (4 - hb)) & 15;
  return 16 * (log - hb) - frac;  // linear interpolation of log2 between powers of two
}

// Code streams of the sequences: offsets become Offset_Values with repeat
// codes (RFC 8878 3.1.1.5) where the repeated offset was itself coded as a
// new offset earlier in the same block -- the history left by earlier blocks
// is never relied on, so blocks encode independently.
struct SeqCodes {
  u8 ll, ml, of;
  u32 ov;
};

XZ_HD inline void seq_codes(const Seq* sq, u32 n, SeqCodes* c) {
  u32 rep[3] = {0, 0, 0};
  u32 known = 0;
  for (u32 i = 0; i < n; ++i) {
    const u32 o = sq[i].off, ll = sq[i].ll;
    u32 ov = 0;
    if (ll) {
      if (known >= 1 && o == rep[0]) {
        ov = 1;
      } else if (known >= 2 && o == rep[1]) {
        ov = 2;
        rep[1] = rep[0];
        rep[0] = o;
      } else if (known >= 3 && o == rep[2]) {
        ov = 3;
        rep[2] = rep[1];
        rep[1] = rep[0];
        rep[0] = o;
      }
    } else {
      if (known >= 2 && o == rep[1]) {
        ov = 1;
        rep[1] = rep[0];
        rep[0] = o;
      } else if (known >= 3 && o == rep[2]) {
        ov = 2;
        rep[2] = rep[1];
        rep[1] = rep[0];
        rep[0] = o;
      } else if (known >= 1 && rep[0] > 1 && o == rep[0] - 1) {
        ov = 3;
        rep[2] = rep[1];
        rep[1] = rep[0];
        rep[0] = o;
        known = known < 3 ? known + 1 : 3;
      }
    }
    if (!ov) {
      ov = o + 3;
      rep[2] = rep[1];
      rep[1] = rep[0];
      rep[0] = o;
      known = known < 3 ? known + 1 : 3;
    }
    c[i].ll = (u8)ll_code(ll);
    c[i].ml = (u8)ml_code(sq[i].ml);
    c[i].ov = ov;
    c[i].of = (u8)highbit(ov);
  }
}

// Per code stream: mode 0 predefined, 1 RLE, 2 FSE table written in the
// block; fills the encoding table and writes its description (if any).
XZ_HD inline u32 choose_table(int which, const u32* count, u32 n, FseTable& t, u8* desc,
                              u32& mode, Work& W) {
  const u32 ms = kMaxSym[which];
  u32 present = 0, last = 0, only = 0;
  for (u32 s = 0; s <= ms; ++s)
    if (count[s]) {
      ++present;
      last = s;
      only = s;
    }
  if (present == 1 && n > 1) {  // RLE: one symbol, no bits
    i16* norm = W.norm;
    for (u32 s = 0; s <= only; ++s) norm[s] = 0;
    norm[only] = 1;
    fse_build(t, norm, only + 1, 0, W.spread, W.cumul);
    desc[0] = (u8)only;
    mode = 1;
    return 1;
  }
  // predefined cost (every code is representable there except OF > 28)
  u64 pre = 0;
  bool pre_ok = true;
  const i16* pn = pre_norm(which);
  for (u32 s = 0; s <= last; ++s) {
    if (!count[s]) continue;
    if (s >= kPreSyms[which]) {
      pre_ok = false;
      break;
    }
    pre += (u64)count[s] * cost16(pn[s] < 0 ? 1u : (u32)pn[s], kPreLog[which]);
  }
  // FSE table for this block
  u32 log = n > 1 ? highbit(n - 1) + 1 : 5;
  log = log < 5 ? 5 : log > kMaxLog[which] ? kMaxLog[which] : log;
  i16* norm = W.norm;
  fse_normalize(count, last + 1, n, log, norm);
  u64 own = 0;
  for (u32 s = 0; s <= last; ++s)
    if (count[s]) own += (u64)count[s] * cost16((u32)norm[s], log);
  u8* tmp = W.ntmp;
  const u32 hl = fse_write_ncount(tmp, norm, last + 1, log);
  own += 16ull * 8 * hl;
  if (pre_ok && pre <= own) {
    fse_build(t, pn, kPreSyms[which], kPreLog[which], W.spread, W.cumul);
    mode = 0;
    return 0;
  }
  fse_build(t, norm, last + 1, log, W.spread, W.cumul);
  for (u32 i = 0; i < hl; ++i) desc[i] = tmp[i];
  mode = 2;
  return hl;
}

// The sequences section in steps the device can spread over a wavefront:
// count header, code histograms, the three tables (mode byte + table
// descriptions), the FSE bitstream.
struct SeqWork {
  FseTable t[3];  // literal lengths, offsets, match lengths
  u32 cnt[3][64];
};

XZ_HD inline u32 seq_count_header(u8* out, u32 n) {
  if (n < 128) {
    out[0] = (u8)n;
    return 1;
  }
  if (n < 0x7F00) {
    out[0] = (u8)((n >> 8) + 128);
    out[1] = (u8)n;
    return 2;
  }
  out[0] = 0xFF;
  out[1] = (u8)(n - 0x7F00);
  out[2] = (u8)((n - 0x7F00) >> 8);
  return 3;
}

// Mode byte and table descriptions from sw.cnt; fills sw.t.
XZ_HD inline u32 seq_write_tables(u8* out, u32 n, SeqWork& sw, Work& W) {
  u32 m[3];
  u32 p = 1;
  for (int k = 0; k < 3; ++k) p += choose_table(k, sw.cnt[k], n, sw.t[k], out + p, m[k], W);
  out[0] = (u8)((m[0] << 6) | (m[1] << 4) | (m[2] << 2));
  return p;
}

// One sequence's FSE symbols and extra bits (all but the last sequence).
XZ_HD inline void seq_encode_one(BitW& bw, FseState* st, const SeqWork& sw, const Seq& q,
                                 const SeqCodes& c) {
  fse_encode(bw, st[1], sw.t[1], c.of);
  fse_encode(bw, st[2], sw.t[2], c.ml);
  fse_encode(bw, st[0], sw.t[0], c.ll);
  bw.add(q.ll - ll_base(c.ll), ll_bits(c.ll));
  bw.add(q.ml - ml_base(c.ml), ml_bits(c.ml));
  bw.add(c.ov - (1u << c.of), c.of);
}

// The last sequence starts the states and writes its extra bits.
XZ_HD inline void seq_encode_last(BitW& bw, FseState* st, const SeqWork& sw, const Seq& q,
                                  const SeqCodes& c) {
  fse_init(st[2], sw.t[2], c.ml);
  fse_init(st[1], sw.t[1], c.of);
  fse_init(st[0], sw.t[0], c.ll);
  bw.add(q.ll - ll_base(c.ll), ll_bits(c.ll));
  bw.add(q.ml - ml_base(c.ml), ml_bits(c.ml));
  bw.add(c.ov - (1u << c.of), c.of);
}

XZ_HD inline u32 seq_encode_close(BitW& bw, const FseState* st, const SeqWork& sw) {
  fse_flush(bw, st[2], sw.t[2]);
  fse_flush(bw, st[1], sw.t[1]);
  fse_flush(bw, st[0], sw.t[0]);
  return bw.close();
}

// Sequences section (without the literals).  `codes`: scratch for n
// entries.  Returns bytes written.
XZ_HD inline u32 write_sequences(u8* out, const Seq* sq, u32 n, SeqCodes* codes, SeqWork& sw,
                                 Work& W) {
  u32 p = seq_count_header(out, n);
  if (!n) return p;
  seq_codes(sq, n, codes);
  for (u32 k = 0; k < 3; ++k)
    for (u32 s = 0; s < 64; ++s) sw.cnt[k][s] = 0;
  for (u32 i = 0; i < n; ++i) {
    sw.cnt[0][codes[i].ll]++;
    sw.cnt[1][codes[i].of]++;
    sw.cnt[2][codes[i].ml]++;
  }
  p += seq_write_tables(out + p, n, sw, W);
  BitW bw(out + p);
  FseState st[3];
  seq_encode_last(bw, st, sw, sq[n - 1], codes[n - 1]);
  for (u32 i = n - 1; i-- > 0;) seq_encode_one(bw, st, sw, sq[i], codes[i]);
  return p + seq_encode_close(bw, st, sw);
}

// ---------------------------------------------------------------------------
// Literals section.  Returns bytes written (at most 5 + n + 16 Huffman
// overhead, or 3 + n raw).

XZ_HD inline u32 write_raw_lits(u8* out, const u8* lit, u32 n, u32 type /*0 raw, 1 rle*/) {
  u32 p;
  if (n < 32) {
    out[0] = (u8)((n << 3) | type);
    p = 1;
  } else if (n < 4096) {
    const u32 h = (n << 4) | (1u << 2) | type;
    out[0] = (u8)h;
    out[1] = (u8)(h >> 8);
    p = 2;
  } else {
    const u32 h = (n << 4) | (3u << 2) | type;
    out[0] = (u8)h;
    out[1] = (u8)(h >> 8);
    out[2] = (u8)(h >> 16);
    p = 3;
  }
  if (type == 1) {
    out[p] = lit[0];
    return p + 1;
  }
  for (u32 i = 0; i < n; ++i) out[p + i] = lit[i];
  return p + n;
}

// Huffman literals with table h built over `lit`.  The body is written at
// out + 5 and moved down once the header size is known.  Returns 0 if the
// encoding is not representable or not smaller than raw literals.
XZ_HD inline u32 write_huf_lits(u8* out, const u8* lit, u32 n, const HufTable& h, Work& W) {
  const bool four = n > 1023;
  u8* q = out + 5;
  const u32 dl = huf_write_desc(q, h, W);
  if (!dl) return 0;
  u32 cs = dl;
  if (!four) {
    cs += huf_stream(q + cs, lit, n, h);
  } else {
    const u32 seg = (n + 3) / 4;
    u8* jt = q + cs;
    cs += 6;
    u32 sizes[4];
    for (u32 k = 0; k < 4; ++k) {
      const u32 a = k * seg, b = k < 3 ? a + seg : n;
      sizes[k] = huf_stream(q + cs, lit + a, b - a, h);
      cs += sizes[k];
    }
    for (u32 k = 0; k < 3; ++k) {
      if (sizes[k] > 0xFFFF) return 0;
      jt[2 * k] = (u8)sizes[k];
      jt[2 * k + 1] = (u8)(sizes[k] >> 8);
    }
  }
  const u32 big = n > cs ? n : cs;
  u32 hdr;
  u64 v;
  if (!four) {
    if (big > 1023) return 0;
    hdr = 3;
    v = 2u | (0u << 2) | ((u64)n << 4) | ((u64)cs << 14);
  } else if (big < 16384) {
    hdr = 4;
    v = 2u | (2u << 2) | ((u64)n << 4) | ((u64)cs << 18);
  } else {
    if (big >= (1u << 18)) return 0;
    hdr = 5;
    v = 2u | (3u << 2) | ((u64)n << 4) | ((u64)cs << 22);
  }
  const u32 total = hdr + cs;
  if (total >= n + (n < 32 ? 1u : n < 4096 ? 2u : 3u)) return 0;
  for (u32 i = 0; i < hdr; ++i) out[i] = (u8)(v >> (8 * i));
  if (hdr < 5)
    for (u32 i = 0; i < cs; ++i) out[hdr + i] = q[i];
  return total;
}

// ---------------------------------------------------------------------------
// LZ parse of one segment [a, b) of a block d[0, n) (see the file comment).
// ht4 / ht8: earliest position of every hash4 (positions [0, n - 4]) and
// hash8 (positions [0, n - 8]).  emit(literal start, match position,
// length, offset) per match; parse_segment returns the position after the
// segment's last match (its trailing literals start there).

XZ_HD inline u32 match_len(const u8* d, u32 p, u32 q, u32 end) {
  u32 l = 0;
  while (p + l < end && d[p + l] == d[q + l]) ++l;
  return l;
}

// How the match finder reads the block: 4-byte little-endian loads and
// common-prefix lengths.  This byte-wise reader is the definition; a faster
// reader (the device's word-wise one) must return the same values.
struct ByteReader {
  XZ_HD static u32 r32(const u8* d, u32 p) { return read32(d + p); }
  XZ_HD static u32 len(const u8* d, u32 p, u32 q, u32 end) { return match_len(d, p, q, end); }
  // out[c] = len(d, p, p - off[c], end) for every c (< 4) with bit c of `want`.
  XZ_HD static void lens(const u8* d, u32 p, const u32* off, u32 want, u32 end, u32* out) {
    for (u32 c = 0; c < 4; ++c)
      if (want >> c & 1u) out[c] = match_len(d, p, p - off[c], end);
  }
};

// Score of a match in 1/4-byte-ish units: length minus the offset's cost
// (small for a repeat of one of the last three offsets).
constexpr int kLenWeight = 4, kRepCost = 2, kOffCost = 4, kLitCost = 3;

XZ_HD inline u32 seg_end(u32 p, u32 n) {
  const u32 e = (p / kSeg + 1) * kSeg;
  return e < n ? e : n;
}

// The best fixed candidate at p: offsets 1, 8, 16, then the 4-byte and the
// 8-byte hash hits (candidate index 1..5), lengths measured up to kCap and
// bounded by p's segment, scored by length minus offset cost (no repeat
// discount; ties keep that order).  Returned as one byte, (index << 5) |
// (length - kMinMatch), 0 if none -- a function of p alone, so every
// position can be evaluated in parallel.
template <typename R = ByteReader>
XZ_HD inline u64 r64(const u8* d, u32 p) {
  return (u64)R::r32(d, p) | ((u64)R::r32(d, p + 4) << 32);
}

template <typename R = ByteReader>
XZ_HD inline u32 cand_offset(const u8* d, u32 n, const u32* ht4, const u32* ht8, u32 p, u32 idx) {
  switch (idx) {
    case 1: return 1;
    case 2: return 8;
    case 3: return 16;
    case 4: return p + 4 <= n ? p - ht4[hash4(R::r32(d, p))] : 0;
    default: return p + 8 <= n ? p - ht8[hash8(r64<R>(d, p))] : 0;
  }
}

template <typename R = ByteReader>
XZ_HD inline u8 best_fixed(const u8* d, u32 n, const u32* ht4, const u32* ht8, u32 p) {
  const u32 b = seg_end(p, n);
  if (p + kMinMatch > b) return 0;
  const u32 cap = p + kCap < b ? p + kCap : b;
  u32 best = 0, bidx = 0;
  i32 bs = -(1 << 30);
  for (u32 idx = 1; idx <= 5; ++idx) {
    if (idx == 4 && !(p + 4 <= n && ht4[hash4(R::r32(d, p))] < p)) continue;
    if (idx == 5 && !(p + 8 <= n && ht8[hash8(r64<R>(d, p))] < p)) continue;
    const u32 off = cand_offset<R>(d, n, ht4, ht8, p, idx);
    if (!off || off > p) continue;
    const u32 l = R::len(d, p, p - off, cap);
    if (l < kMinMatch) continue;
    const i32 sc = (i32)l * kLenWeight - ((i32)highbit(off + 3) + kOffCost);
    if (sc > bs) {
      bs = sc;
      best = l;
      bidx = idx;
    }
  }
  return best ? (u8)((bidx << 5) | (best - kMinMatch)) : (u8)0;
}

// The match a candidate byte stands for: (length << 16) | offset, the
// length capped at kCap (parse_segment measures a capped one to the segment
// end).
template <typename R = ByteReader>
XZ_HD inline u32 cand_match(const u8* d, u32 n, const u32* ht4, const u32* ht8, u32 p, u8 info) {
  if (!info) return 0;
  const u32 off = cand_offset<R>(d, n, ht4, ht8, p, info >> 5);
  return (((info & 31u) + kMinMatch) << 16) | off;
}

// Lazy parse of segment [a, b): at each position the precomputed best_at(p)
// -- (length << 16) | offset, the length capped at kCap, 0 if none; a kCap
// length is measured to the segment end here -- competes with the segment's
// last three offsets (scored with the repeat discount); a literal is emitted
// first when the next position scores better by more than a literal's cost.
template <typename R = ByteReader, typename Best, typename Emit>
XZ_HD inline u32 parse_segment(const u8* d, u32 a, u32 b, Best&& best_at, Emit&& emit) {
  u32 rep[3] = {0, 0, 0};
  // Lengths of the repeat offsets at position `lp` (kNone: not measured): a
  // match of length l >= 1 at p has length l - 1 at p + 1 under the same
  // offset and bound, so the lazy look-ahead need not measure it again.
  constexpr u32 kNone = 0xFFFFFFFFu;
  u32 lp = kNone, rl[3] = {kNone, kNone, kNone};
  auto find = [&](u32 p, u32& best, u32& boff) -> i32 {
    const u32 v = best_at(p);
    best = v >> 16;
    boff = v & 0xFFFF;
    auto sc_of = [&](u32 l, u32 off) {
      const bool r = off == rep[0] || off == rep[1] || off == rep[2];
      return (i32)l * kLenWeight - (r ? kRepCost : (i32)highbit(off + 3) + kOffCost);
    };
    const bool next = lp != kNone && lp + 1 == p;
    // The offsets to try are fixed by the candidate's offset (an offset
    // equal to it or to an earlier winner scores no better), so all lengths
    // still unknown -- the repeat offsets' and a capped candidate's -- are
    // measured together, then scored in order.
    const u32 offs[4] = {rep[0], rep[1], rep[2], boff};
    u32 l[4], use = 0, measure = best == kCap ? 8u : 0u;
    for (u32 c = 0; c < 3; ++c) {
      const u32 off = rep[c];
      l[c] = next && rl[c] != kNone && rl[c] >= 1 ? rl[c] - 1 : kNone;
      rl[c] = kNone;
      if (!off || off > p || off == boff) continue;
      use |= 1u << c;
      if (l[c] == kNone) measure |= 1u << c;
    }
    if (measure) R::lens(d, p, offs, measure, b, l);
    if (measure & 8u) best = l[3];
    i32 bs = best ? sc_of(best, boff) : -(1 << 30);
    for (u32 c = 0; c < 3; ++c) {
      if (!(use >> c & 1u)) continue;
      rl[c] = l[c];
      if (l[c] < kMinMatch) continue;
      const i32 sc = sc_of(l[c], rep[c]);
      if (sc > bs) {
        bs = sc;
        best = l[c];
        boff = rep[c];
      }
    }
    lp = p;
    return bs;
  };
  // One loop, one find() per step: at p, or -- while holding the match
  // found at p -- at p + 1 (lazy: a literal now if the next position matches
  // better).  Lanes of a wavefront parsing different segments then run the
  // same step code instead of diverging between two nested loops.
  u32 p = a, lit0 = a, best = 0, boff = 0;
  i32 sc = 0;
  bool held = false;
  for (;;) {
    const u32 q = held ? p + 1 : p;
    bool take = false;
    if (q + kMinMatch > b) {
      if (!held) break;
      take = true;
    } else {
      u32 b1, o1;
      const i32 sc1 = find(q, b1, o1);
      if (!held) {
        if (b1 >= kMinMatch && sc1 > 0) {
          best = b1;
          boff = o1;
          sc = sc1;
          held = true;
        } else {
          ++p;
        }
      } else if (b1 < kMinMatch || sc1 - kLitCost <= sc) {
        take = true;
      } else {
        ++p;
        best = b1;
        boff = o1;
        sc = sc1;
      }
    }
    if (take) {
      emit(lit0, p, best, boff);
      p += best;
      lit0 = p;
      lp = kNone;  // the history may change
      held = false;
      if (boff != rep[0]) {
        rep[2] = rep[1];
        rep[1] = rep[0];
        rep[0] = boff;
      }
    }
  }
  return lit0;
}

// ---------------------------------------------------------------------------
// A whole block from its parse: literals section (RLE, Huffman or raw,
// whichever is smallest), sequences section, 3-byte block header; stored
// raw when that is not smaller.  `codes` is scratch for nseq entries.
// Returns the bytes written (3 + size).  Not the last block (the caller sets
// bit 0 of out[0]).
XZ_HD inline u32 finish_block(u8* out, const u8* raw, u32 n, const u8* lit, u32 nl,
                              const u32* hist, const Seq* seqs, u32 nseq, SeqCodes* codes,
                              HufTable& h, Work& W, SeqWork& sw) {
  u8* body = out + 3;
  u32 ls = 0;
  bool rle = nl > 1;
  for (u32 i = 0; rle && i < 256; ++i) rle = hist[i] == 0 || hist[i] == nl;
  if (rle) {
    ls = write_raw_lits(body, lit, nl, 1);
  } else {
    if (nl >= 64 && huf_build(hist, h, W)) {
      u64 bits = 0;  // skip the encoding when the code cannot beat raw bytes
      for (u32 c = 0; c < 256; ++c) bits += (u64)hist[c] * h.nbits[c];
      if (bits / 8 + 16 < nl) ls = write_huf_lits(body, lit, nl, h, W);
    }
    if (!ls) ls = write_raw_lits(body, lit, nl, 0);
  }
  const u32 size = ls + write_sequences(body + ls, seqs, nseq, codes, sw, W);
  if (size >= n) {
    out[0] = (u8)(n << 3);
    out[1] = (u8)(n >> 5);
    out[2] = (u8)(n >> 13);
    for (u32 i = 0; i < n; ++i) out[3 + i] = raw[i];
    return 3 + n;
  }
  const u32 hdr = (2u << 1) | (size << 3);
  out[0] = (u8)hdr;
  out[1] = (u8)(hdr >> 8);
  out[2] = (u8)(hdr >> 16);
  return 3 + size;
}

// Bytes a block's encoding may take before finish_block falls back to raw
// (the body is written before the size test): literals at most n + 5, and
// at most n / 4 sequences of at most 12 bytes each plus the tables.
constexpr u32 kSlot = 4 * kBlock + 1024;

}  // namespace xz
