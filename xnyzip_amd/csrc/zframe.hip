// zframe.hip -- zstd-format frames built on the device from the assembled
// streams in HBM (format and parse rules in zframe.hpp).
//
//   k_zcand    one 512-thread workgroup per 16 KiB block: the block and its
//              two hash tables in LDS (40 KiB, four workgroups per CU: the tables are
//              built one after the other in one 32-bit area and narrowed to
//              16 bits, and byte-equality masks for the offsets 1, 8 and 16
//              take the rest); all threads evaluate every position's best
//              fixed candidate and write it, resolved to (capped length,
//              offset), to global scratch.
//   k_zparse   one 32-lane workgroup per block, one lane per 512-byte
//              segment: the block's bytes and a candidate window per lane in
//              LDS (20 KiB, so two such wavefronts share a SIMD); each lane
//              runs the lazy parse of its segment, writing its sequences and
//              literals into per-segment slots and the block's literal
//              histogram.
//   Both read the block with aligned 32-bit LDS loads and byte funnel
//   shifts (match lengths 16 bytes per round trip); results equal the
//   byte-wise definitions in zframe.hpp.
//   k_zlits    one wavefront per block: literals compacted in order, then
//              the literals section as zframe.hpp's finish_block chooses it
//              (Huffman table built across the lanes, streams bit-packed by
//              all lanes at offsets from a wavefront scan).
//   k_zseqs    one wavefront per block: sequences compacted in order (the
//              first sequence of a segment takes the literals left pending
//              by the segments before it), repeat-offset codes (a scan plus
//              a short serial walk), the three FSE tables built across the
//              lanes, the FSE state chains on three lanes and the bitstream
//              packed by all lanes; the block size (or raw) decided.
//   k_zgather  each block's header and sections (or its raw bytes) copied
//              behind its frame header, positions from a device-wide scan of
//              the block sizes.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "zframe.hpp"
#include "zframe_dev.hpp"

namespace xnyz {
namespace {

using xz::kBlock;
using xz::kHashLog;
using xz::kMinMatch;
using xz::kSeg;
using xz::kSegs;
using xz::kSlot;
using xz::Seq;
using xz::SeqCodes;
using xz::u16;
using xz::i16;
using xz::i32;

constexpr u32 kSegSeqs = kSeg / kMinMatch;  // most sequences one segment can hold

struct alignas(16) BlkScratch {  // per block, in global memory
  u32 cand[kBlock];            // (capped length << 16) | offset, 0 = none (k_zcand)
  Seq seqs[kSegs * kSegSeqs];  // per-segment slots (k_zparse)
  Seq cseq[kSegs * kSegSeqs];  // compacted (k_zseqs)
  u8 sym[3][kBlock / kMinMatch];   // LL, OF, ML codes per sequence (k_zseqs)
  u16 fse[3][kBlock / kMinMatch];  // FSE bits per sequence and table: (count << 10) | value
  u32 ls, ss, raw;  // literals / sequences section bytes, block stored raw (k_zlits, k_zseqs)
  u8 lits[kBlock];
  u8 clit[kBlock];  // compacted (k_zlits)
  u32 hist[256];
  u32 nseq[kSegs], nlit[kSegs];
  u32 ntail[kSegs];  // literals after a segment's last sequence (k_zparse)
  u16 ov[kBlock / kMinMatch];  // offset values (offset + 3 or a repeat code), < 2^15 (k_zseqs)
};

constexpr u32 kMaxSeqs = kBlock / kMinMatch;

// A block's compressed form in its slot: header (3 bytes), the literals
// section from byte 3 (k_zlits), the sequences section from kSeqOff
// (k_zseqs); k_zgather joins them.  Literals take at most 3 + 3 + 16 KiB;
// sequences at most 4 header bytes, 3 table descriptions of < 128 bytes and
// < 80 bits per sequence (three FSE codes of <= 9 bits, extra bits <= 16 +
// 16 + 14) plus the final states.
constexpr u32 kSeqOff = kBlock + 512;
static_assert(kSeqOff + 4 + 3 * 128 + kMaxSeqs * 10 + 8 <= kSlot, "slot too small");

struct LitLds {  // k_zlits
  xz::Work W;  // Huffman build and the weights' FSE table
  xz::HufTable h;
  u32 hist[256];
  u32 segl[kSegs], loff[kSegs + 1];
  u8 desc[160];
  u32 sbits[4];
};

struct FseLds {  // FSE table builds
  u8 spread[1u << xz::kFseMaxLog];
  i16 norm[xz::kFseMaxSym];
  u8 ntmp[128];
  u32 ncum[xz::kFseMaxSym + 1];  // cells of the symbols before
};

struct SeqLds {  // k_zseqs
  xz::SeqWork sw;
  FseLds F;
  u32 segn[kSegs], segl[kSegs], segt[kSegs], carry[kSegs], soff[kSegs + 1];
  u32 fin[3];  // FSE states after the first sequence (flushed at the end)
};

constexpr u32 kSeqBatch = 4;  // sequences each lane loads at once

__device__ __forceinline__ Seq shfl_seq(const Seq& q, u32 j) {
  return Seq{(u32)__shfl((int)q.ll, (int)j, 64), (u32)__shfl((int)q.ml, (int)j, 64),
             (u32)__shfl((int)q.off, (int)j, 64)};
}

// Offset_Values of a block's sequences in order (repeat codes against the
// running three-offset history; RFC 8878 3.1.1.5), as xz::seq_codes
// computes them, into ov (block scratch; kept out of LDS so more blocks
// share a CU): entries (offset | 0x8000 when the literal length is 0)
// rewritten in place.
__device__ void dev_offset_values(const Seq* cs, u32 ns, u16* ov) {
  const u32 lane = threadIdx.x;
  for (u32 i0 = 0; i0 < ns; i0 += 64 * kSeqBatch) {
    Seq q[kSeqBatch];
#pragma unroll
    for (u32 b = 0; b < kSeqBatch; ++b) {
      const u32 i = i0 + b * 64 + lane;
      q[b] = cs[i < ns ? i : ns - 1];
    }
#pragma unroll
    for (u32 b = 0; b < kSeqBatch; ++b) {
      const u32 i = i0 + b * 64 + lane;
      if (i < ns) ov[i] = (u16)(q[b].off | (q[b].ll ? 0u : 0x8000u));
    }
  }
  __threadfence_block();
  __syncthreads();
  // Every step puts its offset in slot 0 of the history; a step that repeats
  // slot 0 after literals ("stay") leaves the history as it is, and every
  // other step moves the old slot 0 to slot 1.  So slot 0 before step i is
  // o[i-1], and slot 1 is a copy-forward scan over the batch of 64 steps
  // (one per lane).  Only slot 2 and the count of known slots need a serial
  // walk: the wavefront runs it on uniform values, reading each step's
  // packed inputs with readlane; lane q keeps step q's match flags, from
  // which it forms its offset value.
  constexpr u32 kNone = 0xFFFFFFFFu;
  u32 prev_o = 0, slot1 = 0;  // o and slot 1 after the previous batch
  u32 r2 = 0, known = 0;      // the serial walk's state
  for (u32 i0 = 0; i0 < ns; i0 += 64) {
    const u32 g = i0 + lane;
    const u32 x = g < ns ? (u32)ov[g] : 0u;
    const u32 o = x & 0x7FFFu;
    const bool z = (x & 0x8000u) != 0;  // literal length 0
    u32 po = (u32)__shfl_up((int)o, 1, 64);
    if (lane == 0) po = prev_o;
    const bool stay = g >= 1 && !z && o == po;
    const bool em = g >= 1 && po > 1 && o == po - 1;
    u32 s1 = stay ? kNone : po;  // slot 1 after step g
    for (u32 d = 1; d < 64; d <<= 1) {
      const u32 u = (u32)__shfl_up((int)s1, d, 64);
      if (s1 == kNone && lane >= d) s1 = u;
    }
    if (s1 == kNone) s1 = slot1;
    u32 r1 = (u32)__shfl_up((int)s1, 1, 64);  // slot 1 before step g
    if (lane == 0) r1 = slot1;
    prev_o = (u32)__builtin_amdgcn_readlane((int)o, 63);
    slot1 = (u32)__builtin_amdgcn_readlane((int)s1, 63);
    const u32 pack = o | (r1 << 15) | (stay ? 1u << 30 : 0u);
    const u32 m = ns - i0 < 64 ? ns - i0 : 64;
    u32 flags = 0;
    for (u32 q = 0; q < m; ++q) {
      const u32 y = (u32)__builtin_amdgcn_readlane((int)pack, (int)q);
      const u32 oq = y & 0x7FFFu, r1q = (y >> 15) & 0x7FFFu;
      const bool e1 = known >= 2 && oq == r1q, e2 = known >= 3 && oq == r2;
      flags = lane == q ? (e1 ? 1u : 0u) | (e2 ? 2u : 0u) : flags;
      if (!(y >> 30)) {
        r2 = e1 ? r2 : r1q;
        known = !e1 && !e2 && known < 3 ? known + 1 : known;
      }
    }
    const bool e1 = flags & 1u, e2 = flags & 2u;
    u32 v;
    if (!z) v = stay ? 1u : e1 ? 2u : e2 ? 3u : o + 3;
    else v = e1 ? 1u : e2 ? 2u : em ? 3u : o + 3;
    if (g < ns) ov[g] = (u16)v;
  }
  __threadfence_block();
  __syncthreads();
}

__device__ __forceinline__ SeqCodes codes_of(const Seq& q, u32 ov) {
  SeqCodes c;
  c.ll = (u8)xz::ll_code(q.ll);
  c.ml = (u8)xz::ml_code(q.ml);
  c.ov = ov;
  c.of = (u8)xz::highbit(ov);
  return c;
}

// xz::huf_build on one wavefront, with the same tables: counts copied and
// tested in parallel; present symbols ranked by (count, symbol) through
// all-pairs comparison (the order the stable radix sort gives); the
// two-queue merge by one lane; leaf depths by pointer jumping up the tree;
// canonical codes from per-length wavefront scans.  Returns the same value
// on every lane.
__device__ bool dev_huf_build(const u32* count_in, xz::HufTable& h, xz::Work& W) {
  const u32 lane = threadIdx.x;
  u32* cnt = W.cnt;
  u32 np = 0, last = 0;
  for (u32 s = lane; s < 256; s += 64) {
    const u32 c = count_in[s];
    cnt[s] = c;
    if (c) {
      ++np;
      last = s;
    }
  }
  const u32 n = wave_sum(np);
  const u32 top = wave_max(np ? last : 0u);
  if (lane == 0 && n) h.last = top;
  __syncthreads();
  if (n < 2) return false;
  u32* key = W.pos;
  u16* sym = W.sym;
  u32* w = W.w;
  i16* anc = W.parent;
  u16* dep = W.depth;
  const u32 nodes = 2 * n - 1;
  u32 maxd;
  for (;;) {
    for (u32 s = lane; s < 256; s += 64) key[s] = cnt[s] ? (cnt[s] << 8) | s : 0xFFFFFFFFu;
    __syncthreads();
    u32 mk[4], rk[4] = {0, 0, 0, 0};
    for (u32 j = 0; j < 4; ++j) mk[j] = key[lane + 64 * j];
    for (u32 t = 0; t < 256; ++t) {
      const u32 kt = key[t];
#pragma unroll
      for (u32 j = 0; j < 4; ++j) rk[j] += kt < mk[j];
    }
    for (u32 j = 0; j < 4; ++j)
      if (mk[j] != 0xFFFFFFFFu) sym[rk[j]] = (u16)(lane + 64 * j);
    __syncthreads();
    for (u32 i = lane; i < n; i += 64) w[i] = cnt[sym[i]];
    __syncthreads();
    if (lane == 0) {  // two queues: leaves by weight, inner nodes in creation order
      u32 leaf = 0, node = n, inner = n;
      for (u32 k = 0; k + 1 < n; ++k) {
        u32 pick[2];
        for (int c = 0; c < 2; ++c) {
          if (leaf < n && (inner >= node || w[leaf] <= w[inner])) pick[c] = leaf++;
          else pick[c] = inner++;
        }
        w[node] = w[pick[0]] + w[pick[1]];
        anc[pick[0]] = anc[pick[1]] = (i16)node;
        ++node;
      }
      anc[nodes - 1] = (i16)(nodes - 1);  // the root is its own ancestor
    }
    __syncthreads();
    // distance to the root: d += d(anc), anc = anc(anc), 8 rounds (depth < 256)
    constexpr u32 kPer = 8;  // nodes per lane (2 * 256 - 1 < 64 * 8)
    u32 d[kPer], a[kPer];
    for (u32 k = 0; k < kPer; ++k) {
      const u32 i = lane + 64 * k;
      a[k] = i < nodes ? (u32)anc[i] : 0u;
      d[k] = i + 1 < nodes ? 1u : 0u;
      if (i < nodes) dep[i] = (u16)d[k];
    }
    __syncthreads();
    for (u32 r = 0; r < 8; ++r) {
      for (u32 k = 0; k < kPer; ++k)
        if (lane + 64 * k < nodes) {
          d[k] += dep[a[k]];
          a[k] = (u32)anc[a[k]];
        }
      __syncthreads();
      for (u32 k = 0; k < kPer; ++k)
        if (lane + 64 * k < nodes) {
          dep[lane + 64 * k] = (u16)d[k];
          anc[lane + 64 * k] = (i16)a[k];
        }
      __syncthreads();
    }
    u32 md = 0;
    for (u32 k = 0; k < kPer; ++k)
      if (lane + 64 * k < n) md = d[k] > md ? d[k] : md;
    maxd = wave_max(md);
    if (maxd <= xz::kMaxHufBits) {
      for (u32 s = lane; s < 256; s += 64) h.nbits[s] = 0;
      __syncthreads();
      for (u32 k = 0; k < kPer; ++k)
        if (lane + 64 * k < n) h.nbits[sym[lane + 64 * k]] = (u8)d[k];
      if (lane == 0) h.maxbits = maxd;
      __syncthreads();
      break;
    }
    for (u32 s = lane; s < 256; s += 64)
      if (cnt[s]) cnt[s] = (cnt[s] >> 1) | 1u;
    __syncthreads();
  }
  // canonical codes: longest first from 0, by symbol; lane holds 4l .. 4l+3
  u32 nb[4];
  for (u32 j = 0; j < 4; ++j) nb[j] = h.nbits[4 * lane + j];
  u32 base = 0;
  for (u32 len = maxd; len >= 1; --len) {
    u32 c = 0;
    for (u32 j = 0; j < 4; ++j) c += nb[j] == len;
    const u32 inc = wave_incl_scan(c);
    const u32 tot = (u32)__shfl((int)inc, 63, 64);
    u32 r = base + inc - c;
    for (u32 j = 0; j < 4; ++j)
      if (nb[j] == len) h.code[4 * lane + j] = (u16)r++;
    base = (base + tot) >> 1;
  }
  __syncthreads();
  return true;
}

// ---- sequence tables on one wavefront (lane s handles symbol s; <= 64) ----

// xz::fse_normalize: same normalised counts.
__device__ void dev_fse_normalize(const u32* count, u32 nsym, u32 total, u32 log, i16* norm) {
  const u32 lane = threadIdx.x, size = 1u << log;
  const u32 c = lane < nsym ? count[lane] : 0u;
  i32 v = 0;
  if (c) {
    v = (i32)(((u64)c * size + total / 2) / total);
    if (v < 1) v = 1;
  }
  // the first symbol with the largest count; sum fixed there or, when too
  // large, on the first symbol with the largest value above 1, one at a time
  const u32 bk = wave_max(c ? (c << 6) | (63u - lane) : 0u);
  const u32 big = 63u - (bk & 63u);
  i32 sum = wave_sum(v);
  while (sum > (i32)size) {
    const u32 m = wave_max(v > 1 ? ((u32)v << 6) | (63u - lane) : 0u);
    if (!m) break;
    if (lane == 63u - (m & 63u)) --v;
    --sum;
  }
  if (sum < (i32)size && lane == big) v += (i32)size - sum;
  if (lane < nsym) norm[lane] = (i16)v;
  __syncthreads();
}

// xz::fse_build: same table.  The k-th step of the spread walk visits
// position (k * step) mod size, so each lane takes a run of steps, counts
// the positions at or below the "-1" region, and places the cells whose
// ranks a wavefront scan gives it.
__device__ void dev_fse_build(xz::FseTable& t, const i16* norm, u32 nsym, u32 log, FseLds& L) {
  const u32 lane = threadIdx.x, size = 1u << log, mask = size - 1;
  const i32 n = lane < nsym ? (i32)norm[lane] : 0;
  const u32 m1 = n == -1 ? 1u : 0u, cn = n > 0 ? (u32)n : 0u;
  const u32 cells = m1 + cn;
  const u32 cum = wave_incl_scan(cells) - cells;  // xz::fse_build's cumul[s]
  const u32 m1x = wave_incl_scan(m1) - m1;
  const u32 nm1 = (u32)__shfl((int)(m1x + m1), 63, 64);
  const u32 high = size - 1 - nm1;
  const u32 ncx = wave_incl_scan(cn) - cn;
  u8* spread = L.spread;
  if (lane < nsym) L.ncum[lane] = ncx;
  if (lane == 0) L.ncum[nsym] = high + 1;
  if (m1) spread[size - 1 - m1x] = (u8)lane;
  __syncthreads();
  const u32 step = (size >> 1) + (size >> 3) + 3, per = (size + 63) / 64;
  const u32 k0 = lane * per, k1 = k0 + per < size ? k0 + per : size;
  u32 valid = 0;
  for (u32 k = k0; k < k1; ++k) valid += ((k * step) & mask) <= high;
  u32 rank = wave_incl_scan(valid) - valid;
  if (valid) {
    u32 lo = 0, hi = nsym - 1;  // the symbol whose cells cover `rank`
    while (lo < hi) {
      const u32 mid = (lo + hi + 1) / 2;
      if (L.ncum[mid] <= rank) lo = mid;
      else hi = mid - 1;
    }
    u32 sym = lo;
    for (u32 k = k0; k < k1; ++k) {
      const u32 pos = (k * step) & mask;
      if (pos > high) continue;
      while (L.ncum[sym + 1] <= rank) ++sym;
      spread[pos] = (u8)sym;
      ++rank;
    }
  }
  __syncthreads();
  // states: the cells of symbol s in position order take slots cum(s) ...
  if (lane < nsym && cells) {
    u32 slot = cum;
    for (u32 u = 0; u < size; u += 4) {
      const u32 q = *reinterpret_cast<const u32*>(spread + u);
#pragma unroll
      for (u32 b = 0; b < 4; ++b)
        if (u + b < size && ((q >> (8 * b)) & 255u) == lane) t.state[slot++] = (u16)(size + u + b);
    }
  }
  const u32 c1 = n == 0 ? 0u : (n == -1 || n == 1) ? 1u : (u32)n;
  const u32 tot = wave_incl_scan(c1) - c1;
  if (lane < nsym) {
    if (n == 0) {
      t.dnb[lane] = (i32)(((log + 1) << 16) - size);
      t.dfs[lane] = 0;
    } else if (n == -1 || n == 1) {
      t.dnb[lane] = (i32)((log << 16) - size);
      t.dfs[lane] = (i32)tot - 1;
    } else {
      const u32 maxb = log - xz::highbit((u32)n - 1);
      t.dnb[lane] = (i32)((maxb << 16) - ((u32)n << maxb));
      t.dfs[lane] = (i32)tot - n;
    }
  }
  if (lane == 0) t.log = log;
  __syncthreads();
}

// xz::choose_table: same choice, table and description.
__device__ u32 dev_choose_table(int which, const u32* count, u32 n, xz::FseTable& t, u8* desc,
                                u32& mode, FseLds& L) {
  const u32 lane = threadIdx.x;
  const u32 ms = xz::kMaxSym[which];
  const u32 c = lane <= ms ? count[lane] : 0u;
  const u32 present = wave_sum(c ? 1u : 0u);
  const u32 last = wave_max(c ? lane : 0u);
  i16* norm = L.norm;
  if (present == 1 && n > 1) {  // RLE
    if (lane <= last) norm[lane] = lane == last ? 1 : 0;
    __syncthreads();
    dev_fse_build(t, norm, last + 1, 0, L);
    if (lane == 0) desc[0] = (u8)last;
    mode = 1;
    return 1;
  }
  const i16* pn = xz::pre_norm(which);
  const bool bad = c && lane >= xz::kPreSyms[which];
  const bool pre_ok = wave_max(bad ? 1u : 0u) == 0;
  u64 pre = 0;
  if (pre_ok && c) pre = (u64)c * xz::cost16(pn[lane] < 0 ? 1u : (u32)pn[lane], xz::kPreLog[which]);
  pre = wave_sum(pre);
  u32 log = n > 1 ? xz::highbit(n - 1) + 1 : 5;
  log = log < 5 ? 5 : log > xz::kMaxLog[which] ? xz::kMaxLog[which] : log;
  dev_fse_normalize(count, last + 1, n, log, norm);
  u64 own = c ? (u64)c * xz::cost16((u32)norm[lane], log) : 0ull;
  own = wave_sum(own);
  u8* tmp = L.ntmp;
  const u32 hl = xz::fse_write_ncount(tmp, norm, last + 1, log);  // uniform
  own += 16ull * 8 * hl;
  __syncthreads();
  if (pre_ok && pre <= own) {
    dev_fse_build(t, pn, xz::kPreSyms[which], xz::kPreLog[which], L);
    mode = 0;
    return 0;
  }
  dev_fse_build(t, norm, last + 1, log, L);
  for (u32 i = lane; i < hl; i += 64) desc[i] = tmp[i];
  mode = 2;
  return hl;
}

// xz::seq_write_tables: mode byte and table descriptions from sw.cnt.
__device__ u32 dev_seq_write_tables(u8* out, u32 n, xz::SeqWork& sw, FseLds& L) {
  u32 m[3];
  u32 p = 1;
  for (int k = 0; k < 3; ++k) p += dev_choose_table(k, sw.cnt[k], n, sw.t[k], out + p, m[k], L);
  if (threadIdx.x == 0) out[0] = (u8)((m[0] << 6) | (m[1] << 4) | (m[2] << 2));
  return p;
}

// The sequences' FSE bitstream as xz::write_sequences writes it, at slot +
// base (slot 256-byte aligned): the three state chains (literal lengths,
// offsets, match lengths) are independent, so three lanes run them, from
// the last sequence down, recording each step's emitted bits; then every
// lane packs a run of sequences in stream order (last sequence first) at
// offsets from a wavefront scan, and lane 0 appends the final states and
// the end mark.  sym / ov / tables must be ready.  Returns bytes written.
__device__ u32 dev_seq_bits(u8* slot, u32 base, BlkScratch& S, u32 ns, SeqLds& L) {
  const u32 lane = threadIdx.x;
  if (lane < 3) {
    const xz::FseTable& t = L.sw.t[lane];
    const u8* sym = S.sym[lane];
    u16* fo = S.fse[lane];
    xz::FseState st;
    xz::fse_init(st, t, sym[ns - 1]);
    // Step k encodes sequence ns - 2 - k.  The next batch's codes are
    // loaded before this batch's records are stored, so waiting for those
    // loads never waits for the stores.
    constexpr u32 kB = 16;
    auto code = [&](u32 k) -> u32 { return k + 1 < ns ? sym[ns - 2 - k] : 0u; };
    u32 nxt[kB];
#pragma unroll
    for (u32 q = 0; q < kB; ++q) nxt[q] = code(q);
    for (u32 k0 = 0; k0 + 1 < ns; k0 += kB) {
      u32 dnb[kB], dfs[kB], rec[kB];
#pragma unroll
      for (u32 q = 0; q < kB; ++q) {
        dnb[q] = (u32)t.dnb[nxt[q]];
        dfs[q] = (u32)t.dfs[nxt[q]];
      }
#pragma unroll
      for (u32 q = 0; q < kB; ++q) nxt[q] = code(k0 + kB + q);
#pragma unroll
      for (u32 q = 0; q < kB; ++q) {
        const u32 nbo = (st.v + dnb[q]) >> 16;
        rec[q] = (nbo << 10) | (st.v & ((1u << nbo) - 1u));
        if (k0 + q + 1 < ns) st.v = t.state[(st.v >> nbo) + dfs[q]];
      }
#pragma unroll
      for (u32 q = 0; q < kB; ++q)
        if (k0 + q + 1 < ns) fo[ns - 2 - k0 - q] = (u16)rec[q];
    }
    L.fin[lane] = st.v;
  }
  __threadfence_block();
  __syncthreads();
  // stream order r = 0 .. ns - 1 is sequence ns - 1 - r; each lane takes a
  // run of it, its sequences and FSE records loaded kSeqBatch at a time
  const u32 per = (ns + 63) / 64, r0 = lane * per < ns ? lane * per : ns;
  const u32 r1 = r0 + per < ns ? r0 + per : ns;
  struct SeqRec {
    Seq q;
    u32 ov;
    u32 f[3];  // FSE records (bits << 10 | value) of LL, OF, ML; unused for the last sequence
  };
  auto load = [&](u32 rb, SeqRec* v) {
#pragma unroll
    for (u32 b = 0; b < kSeqBatch; ++b) {
      const u32 i = ns - 1 - (rb + b < r1 ? rb + b : r1 - 1);
      v[b].q = S.cseq[i];
      v[b].ov = S.ov[i];
#pragma unroll
      for (u32 k = 0; k < 3; ++k) v[b].f[k] = S.fse[k][i];
    }
  };
  u32 bits = 0;
  for (u32 rb = r0; rb < r1; rb += kSeqBatch) {
    SeqRec v[kSeqBatch];
    load(rb, v);
#pragma unroll
    for (u32 b = 0; b < kSeqBatch; ++b) {
      if (rb + b >= r1) break;
      const u32 i = ns - 1 - (rb + b);
      const Seq& q = v[b].q;
      bits += xz::ll_bits(xz::ll_code(q.ll)) + xz::ml_bits(xz::ml_code(q.ml)) + xz::highbit(v[b].ov);
      if (i + 1 < ns) bits += (v[b].f[0] >> 10) + (v[b].f[1] >> 10) + (v[b].f[2] >> 10);
    }
  }
  const u32 inc = wave_incl_scan(bits);
  const u32 total = __shfl((int)inc, 63, 64);
  const u32 tl = (u32)L.sw.t[0].log, to = (u32)L.sw.t[1].log, tm = (u32)L.sw.t[2].log;
  const u32 nbytes = (total + tm + to + tl + 1 + 7) / 8;
  for (u32 i = lane; i < nbytes; i += 64) slot[base + i] = 0;
  __threadfence_block();
  __syncthreads();
  u32* w = reinterpret_cast<u32*>(slot);
  u64 pos = (u64)base * 8 + (inc - bits);
  u64 acc = 0;
  u32 nb = (u32)(pos & 31);
  u64 word = pos >> 5;
  bool first = true;
  auto put = [&](u32 v, u32 n) {
    acc |= (u64)(v & (n < 32 ? (1u << n) - 1u : 0xFFFFFFFFu)) << nb;
    nb += n;
    if (nb >= 32) {
      if (first) atomicOr(&w[word], (u32)acc);
      else w[word] = (u32)acc;
      first = false;
      acc >>= 32;
      nb -= 32;
      ++word;
    }
  };
  for (u32 rb = r0; rb < r1; rb += kSeqBatch) {
    SeqRec v[kSeqBatch];
    load(rb, v);
#pragma unroll
    for (u32 b = 0; b < kSeqBatch; ++b) {
      if (rb + b >= r1) break;
      const u32 i = ns - 1 - (rb + b);
      const Seq& q = v[b].q;
      const SeqCodes c = codes_of(q, v[b].ov);
      if (i + 1 < ns) {
        const u32 fo = v[b].f[1], fm = v[b].f[2], fl = v[b].f[0];
        put(fo & 1023u, fo >> 10);
        put(fm & 1023u, fm >> 10);
        put(fl & 1023u, fl >> 10);
      }
      put(q.ll - xz::ll_base(c.ll), xz::ll_bits(c.ll));
      put(q.ml - xz::ml_base(c.ml), xz::ml_bits(c.ml));
      put(c.ov - (1u << c.of), c.of);
    }
  }
  if (nb) atomicOr(&w[word], (u32)acc);
  __threadfence_block();
  __syncthreads();
  if (lane == 0) {  // final states (match lengths, offsets, literal lengths), end mark
    pos = (u64)base * 8 + total;
    acc = 0;
    nb = (u32)(pos & 31);
    word = pos >> 5;
    first = true;
    put(L.fin[2], tm);
    put(L.fin[1], to);
    put(L.fin[0], tl);
    put(1, 1);
    if (nb) atomicOr(&w[word], (u32)acc);
  }
  __threadfence_block();
  __syncthreads();
  return nbytes;
}

constexpr u32 kLitBatch = 16;  // literal bytes each lane loads at once

// Huffman-coded literals as write_huf_lits lays them out, written at their
// final place: slot + base (slot 256-byte aligned; words are addressed from
// the slot so that every 32-bit store and atomic is aligned).  Returns 0
// when not representable or not smaller than raw.
__device__ u32 dev_huf_lits(u8* slot, u32 base, const u8* lit, u32 n, LitLds& L) {
  u8* out = slot + base;
  const u32 lane = threadIdx.x;
  const bool four = n > 1023;
  const u32 dl = xz::huf_write_desc(L.desc, L.h, L.W);  // uniform
  if (!dl) return 0;
  const u32 nst = four ? 4 : 1, seg = four ? (n + 3) / 4 : n;
  u32 sbytes[4] = {0, 0, 0, 0};
  for (u32 k = 0; k < nst; ++k) {
    const u32 a = k * seg, e = k + 1 < nst ? a + seg : n;
    u32 bits = 0;
    for (u32 i0 = a; i0 < e; i0 += 64 * kLitBatch) {  // loads of a batch issued together
      u32 c[kLitBatch];
#pragma unroll
      for (u32 q = 0; q < kLitBatch; ++q) {
        const u32 i = i0 + q * 64 + lane;
        c[q] = lit[i < e ? i : e - 1];
      }
#pragma unroll
      for (u32 q = 0; q < kLitBatch; ++q)
        if (i0 + q * 64 + lane < e) bits += L.h.nbits[c[q]];
    }
    bits = wave_sum(bits);
    sbytes[k] = (bits + 8) / 8;  // with the end mark
    if (lane == 0) L.sbits[k] = bits;
  }
  u32 cs = dl + (four ? 6 : 0);
  for (u32 k = 0; k < nst; ++k) cs += sbytes[k];
  const u32 big = n > cs ? n : cs;
  u32 hdr;
  u64 v;
  if (!four) {
    if (big > 1023) return 0;
    hdr = 3;
    v = 2u | ((u64)n << 4) | ((u64)cs << 14);
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
  if (four)
    for (u32 k = 0; k < 3; ++k)
      if (sbytes[k] > 0xFFFF) return 0;
  // header, description, jump table: byte stores; the streams' bytes zeroed
  u32 at = hdr + dl + (four ? 6 : 0);
  for (u32 i = lane; i < total - at; i += 64) out[at + i] = 0;
  if (lane == 0) {
    for (u32 i = 0; i < hdr; ++i) out[i] = (u8)(v >> (8 * i));
    for (u32 i = 0; i < dl; ++i) out[hdr + i] = L.desc[i];
    if (four)
      for (u32 k = 0; k < 3; ++k) {
        out[hdr + dl + 2 * k] = (u8)sbytes[k];
        out[hdr + dl + 2 * k + 1] = (u8)(sbytes[k] >> 8);
      }
  }
  __threadfence_block();
  __syncthreads();
  // streams: literal i of a stream sits after the codes of the literals
  // following it; lane chunks of the reversed stream, offsets by scan
  u32* w = reinterpret_cast<u32*>(slot);
  for (u32 k = 0; k < nst; ++k) {
    const u32 a = k * seg, e = k + 1 < nst ? a + seg : n, m = e - a;
    const u32 per = (m + 63) / 64, r0 = lane * per < m ? lane * per : m;
    const u32 r1 = r0 + per < m ? r0 + per : m;
    // the lane's literals, read kLitBatch at a time (loads issued together)
    auto batch = [&](u32 rb, u32* c) {
#pragma unroll
      for (u32 q = 0; q < kLitBatch; ++q) c[q] = lit[e - 1 - (rb + q < r1 ? rb + q : r1 - 1)];
    };
    u32 bits = 0;
    for (u32 rb = r0; rb < r1; rb += kLitBatch) {
      u32 c[kLitBatch];
      batch(rb, c);
#pragma unroll
      for (u32 q = 0; q < kLitBatch; ++q)
        if (rb + q < r1) bits += L.h.nbits[c[q]];
    }
    const u32 inc = wave_incl_scan(bits);
    u64 pos = (u64)(base + at) * 8 + (inc - bits);  // bit position from the slot start
    u64 acc = 0;
    u32 nb = (u32)(pos & 31);
    u64 word = pos >> 5;
    bool first = true;
    for (u32 rb = r0; rb < r1; rb += kLitBatch) {
      u32 cb[kLitBatch];
      batch(rb, cb);
#pragma unroll
      for (u32 q = 0; q < kLitBatch; ++q) {
        if (rb + q >= r1) break;
        const u32 c = cb[q];
        acc |= (u64)L.h.code[c] << nb;
        nb += L.h.nbits[c];
        if (nb >= 32) {
          if (first) atomicOr(&w[word], (u32)acc);
          else w[word] = (u32)acc;
          first = false;
          acc >>= 32;
          nb -= 32;
          ++word;
        }
      }
    }
    if (nb) atomicOr(&w[word], (u32)acc);
    if (lane == 63) {  // the end mark after the stream's last code
      const u64 endb = (u64)(base + at) * 8 + inc;
      atomicOr(&w[endb >> 5], 1u << (endb & 31));
    }
    at += sbytes[k];
  }
  __threadfence_block();
  __syncthreads();
  return total;
}

// Word-wise reader of a block held in LDS (4-byte aligned, padded by kPad
// bytes): the same values as xz::ByteReader.
struct LdsReader {
  __device__ static u32 r32(const u8* d, u32 p) {
    const u32* w = reinterpret_cast<const u32*>(d);
    return __builtin_amdgcn_alignbyte(w[(p >> 2) + 1], w[p >> 2], p & 3u);
  }
  // 16 bytes per step: five aligned words of each side loaded together
  // (one LDS round trip), compared four bytes at a time.
  __device__ static u32 len(const u8* d, u32 p, u32 q, u32 end) {
    const u32* w = reinterpret_cast<const u32*>(d);
    u32 l = 0;
    while (p + l < end) {
      const u32 pa = (p + l) >> 2, ps = (p + l) & 3u, qa = (q + l) >> 2, qs = (q + l) & 3u;
      u32 a[5], b[5];
#pragma unroll
      for (u32 k = 0; k < 5; ++k) {
        a[k] = w[pa + k];
        b[k] = w[qa + k];
      }
      const u32 r = end - (p + l);
#pragma unroll
      for (u32 k = 0; k < 4; ++k) {
        const u32 x = __builtin_amdgcn_alignbyte(a[k + 1], a[k], ps) ^
                      __builtin_amdgcn_alignbyte(b[k + 1], b[k], qs);
        if (x) {
          const u32 e = 4 * k + ((u32)__builtin_ctz(x) >> 3);
          return l + (e < r ? e : r);
        }
        if (r <= 4 * (k + 1)) return l + r;
      }
      l += 16;
    }
    return l;
  }
  // out[c] = len(d, p, p - off[c], end) for every c (< 4) with bit c of
  // `want`, in shared rounds: one load of p's side and one LDS round trip
  // for all.
  __device__ static void lens(const u8* d, u32 p, const u32* off, u32 want, u32 end, u32* out) {
    const u32* w = reinterpret_cast<const u32*>(d);
    const u32 q[4] = {p - off[0], p - off[1], p - off[2], p - off[3]};
    for (u32 l = 0; want; l += 16) {
      const u32 r = end - (p + l);  // > 0: a measurement ends once l reaches end - p
      const u32 pa = (p + l) >> 2, ps = (p + l) & 3u;
      u32 a[5];
#pragma unroll
      for (u32 k = 0; k < 5; ++k) a[k] = w[pa + k];
#pragma unroll
      for (u32 c = 0; c < 4; ++c) {
        if (!(want >> c & 1u)) continue;
        const u32 qa = (q[c] + l) >> 2, qs = (q[c] + l) & 3u;
        u32 b[5];
#pragma unroll
        for (u32 k = 0; k < 5; ++k) b[k] = w[qa + k];
        u32 e = 16;
#pragma unroll
        for (u32 k = 0; k < 4; ++k) {
          const u32 x = __builtin_amdgcn_alignbyte(a[k + 1], a[k], ps) ^
                        __builtin_amdgcn_alignbyte(b[k + 1], b[k], qs);
          if (x && e == 16) e = 4 * k + ((u32)__builtin_ctz(x) >> 3);
        }
        if (e < 16 || r <= 16) {
          out[c] = l + (e < r ? e : r);
          want &= ~(1u << c);
        }
      }
    }
  }
  // len(d, p, q, end) given v0, v1 = the words at p and p + 4: the first
  // eight bytes cost three loads of q's side, and most hash hits end there
  // (a 4- or 8-byte hash collision usually differs in the first byte).
  __device__ static u32 len8(const u8* d, u32 p, u32 q, u32 end, u32 v0, u32 v1) {
    const u32* w = reinterpret_cast<const u32*>(d);
    const u32 qa = q >> 2, qs = q & 3u;
    const u32 a0 = w[qa], a1 = w[qa + 1], a2 = w[qa + 2];
    const u32 r = end - p;
    u32 e;
    u32 x = __builtin_amdgcn_alignbyte(a1, a0, qs) ^ v0;
    if (x) {
      e = (u32)__builtin_ctz(x) >> 3;
    } else {
      x = __builtin_amdgcn_alignbyte(a2, a1, qs) ^ v1;
      if (x) {
        e = 4 + ((u32)__builtin_ctz(x) >> 3);
      } else {
        if (r <= 8) return r;
        return 8 + len(d, p + 8, q + 8, end);
      }
    }
    return e < r ? e : r;
  }
};

constexpr u32 kPad = 32;  // LDS block copies: readable bytes past the end (match loads read 20)

// Copies block bytes [0, n) into LDS (16-byte aligned source and target:
// blocks start at multiples of kBlock in 256-byte aligned streams); the
// padding is zeroed so no load reads uninitialised memory.  Every lane issues
// all its 16-byte loads before its first LDS store, so the wave waits on
// global memory once per block, not once per load.
template <u32 kThreads>
__device__ __forceinline__ void load_block(u8* data, const u8* src, u32 n, u32 tid) {
  constexpr u32 kPer = kBlock / 16 / kThreads, kBat = kPer < 16 ? kPer : 16;
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* t = reinterpret_cast<uint4*>(data);
  const u32 nc = n / 16;
  if (nc) {
    // unconditional loads (clamped to the last chunk) keep v in registers
#pragma unroll
    for (u32 k0 = 0; k0 < kPer; k0 += kBat) {
      uint4 v[kBat];
#pragma unroll
      for (u32 k = 0; k < kBat; ++k) {
        const u32 i = tid + (k0 + k) * kThreads;
        v[k] = s[i < nc ? i : nc - 1];
      }
#pragma unroll
      for (u32 k = 0; k < kBat; ++k)
        if (tid + (k0 + k) * kThreads < nc) t[tid + (k0 + k) * kThreads] = v[k];
    }
  }
  for (u32 i = nc * 16 + tid; i < n + kPad; i += kThreads) data[i] = i < n ? src[i] : 0;
}

// Byte-equality of a block with itself shifted by 1, 8 and 16 bytes, one bit
// per position (bit x of offset o: d[x] == d[x - o]); the run of ones from
// bit p is the match length of offset o at p.
constexpr u32 kFixedOff[3] = {1, 8, 16};
constexpr u32 kEqWords = kBlock / 32 + 4;

// The 4-bit mask of the zero bytes of x (bit i: byte i of x is 0).
__device__ __forceinline__ u32 zero_bytes(u32 x) {
  const u32 z = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
  return ((z >> 7) & 1u) | ((z >> 14) & 2u) | ((z >> 21) & 4u) | ((z >> 28) & 8u);
}

// Length of the run of ones of eq from bit p, at most cap (<= kCap <= 35).
__device__ __forceinline__ u32 eq_run(const u32* eq, u32 p, u32 cap) {
  const u32 w = p >> 5, s = p & 31u;
  const u64 lo = ((u64)eq[w] | ((u64)eq[w + 1] << 32)) >> s;
  const u64 x = ~lo;
  u32 l = x ? (u32)__builtin_ctzll(x) : 64u;
  if (l == 64u - s && l < cap) {
    const u32 hi = ~eq[w + 2];
    l += hi ? (u32)__builtin_ctz(hi) : 32u;
  }
  return l < cap ? l : cap;
}

constexpr u32 kCandThreads = 512;  // k_zcand workgroup: eight wavefronts share a block

__global__ __launch_bounds__(kCandThreads) void k_zcand(const u8* __restrict__ pre,
                                               const ZBlk* __restrict__ blks,
                                               BlkScratch* __restrict__ scr) {
  // LDS (40 KiB: four workgroups per CU): the block, the 4-byte hash table
  // (u16), and a u32 area in which each table is built in turn with
  // atomicMin; the 8-byte table ends as u16 in its first half, the equality
  // masks take the second.  The block has no padding of its own: its copy's
  // zero padding and every read past its end (all clamped to the block's
  // bytes) fall in the 4-byte table.
  constexpr u32 kH = 1u << kHashLog;
  __shared__ __align__(16) u32 lds[(kBlock + 2 * kH + 4 * kH) / 4];
  static_assert(kPad <= 2 * kH, "padding falls in the 4-byte table");
  static_assert(3 * kEqWords * 4 <= 2 * kH, "masks fit the second half");
  u8* const data = reinterpret_cast<u8*>(lds);
  u16* const h4t = reinterpret_cast<u16*>(data + kBlock);
  u32* const bld = reinterpret_cast<u32*>(data + kBlock + 2 * kH);
  const ZBlk bk = blks[blockIdx.x];
  u32* cand = scr[blockIdx.x].cand;
  const u32 n = bk.n, tid = threadIdx.x, lane = tid & (kWave - 1);
  load_block<kCandThreads>(data, pre + bk.src, n, tid);
  for (u32 i = tid; i < kH; i += kCandThreads) bld[i] = 0xFFFFFFFFu;
  __syncthreads();
  // Runs of equal hashes on consecutive lanes (runs of equal bytes) leave
  // the update to the run's first lane, whose position is the smallest:
  // atomics on one address from a whole wavefront would serialise.
  for (u32 p0 = 0; p0 < n; p0 += kCandThreads) {
    const u32 p = p0 + tid;
    const bool ok = p + 4 <= n;
    const u32 h = ok ? xz::hash4(LdsReader::r32(data, p)) : 0xFFFFFFFFu;
    const u32 up = (u32)__shfl_up((int)h, 1, kWave);
    if (ok && (lane == 0 || up != h)) atomicMin(&bld[h], p);
  }
  __syncthreads();
#pragma unroll
  for (u32 k = 0; k < kH / kCandThreads; ++k) {
    const u32 t = bld[tid + kCandThreads * k];
    h4t[tid + kCandThreads * k] = (u16)(t < 0xFFFFu ? t : 0xFFFFu);
    bld[tid + kCandThreads * k] = 0xFFFFFFFFu;
  }
  __syncthreads();
  for (u32 p0 = 0; p0 < n; p0 += kCandThreads) {
    const u32 p = p0 + tid;
    const bool ok = p + 8 <= n;
    const u32 h = ok ? xz::hash8((u64)LdsReader::r32(data, p) | ((u64)LdsReader::r32(data, p + 4) << 32))
                     : 0xFFFFFFFFu;
    const u32 up = (u32)__shfl_up((int)h, 1, kWave);
    if (ok && (lane == 0 || up != h)) atomicMin(&bld[h], p);
  }
  __syncthreads();
  {
    u32 t[kH / kCandThreads];
#pragma unroll
    for (u32 k = 0; k < kH / kCandThreads; ++k) t[k] = bld[tid + kCandThreads * k];
    __syncthreads();
    u16* h16 = reinterpret_cast<u16*>(bld);
#pragma unroll
    for (u32 k = 0; k < kH / kCandThreads; ++k) h16[tid + kCandThreads * k] = (u16)(t[k] < 0xFFFFu ? t[k] : 0xFFFFu);
  }
  u32* eq = bld + kH / 2;  // [3][kEqWords]
  {
    const u32* dw = reinterpret_cast<const u32*>(data);
    const u32 rot = (tid >> 3) & 7u;  // conflict-free LDS banks across the wavefront
    for (u32 w = tid; w < kEqWords; w += kCandThreads) {
      u32 m1 = 0, m8 = 0, m16 = 0;
      if (32 * w < n) {
#pragma unroll
        for (u32 i = 0; i < 8; ++i) {
          const u32 ii = (i + rot) & 7u, j = 8 * w + ii;
          const u32 a = dw[j];
          const u32 p1 = j >= 1 ? dw[j - 1] : 0u;
          const u32 p2 = j >= 2 ? dw[j - 2] : 0u;
          const u32 p4 = j >= 4 ? dw[j - 4] : 0u;
          m1 |= zero_bytes(a ^ __builtin_amdgcn_alignbyte(a, p1, 3u)) << (4 * ii);
          m8 |= zero_bytes(a ^ p2) << (4 * ii);
          m16 |= zero_bytes(a ^ p4) << (4 * ii);
        }
      }
      eq[w] = m1;
      eq[kEqWords + w] = m8;
      eq[2 * kEqWords + w] = m16;
    }
  }
  __syncthreads();
  // xz::best_fixed: offsets 1, 8, 16 from the masks, then the hash hits
  const u16* h4 = h4t;
  const u16* h8 = reinterpret_cast<const u16*>(bld);
  for (u32 p = tid; p < n; p += kCandThreads) {
    const u32 e = xz::seg_end(p, n);
    u32 c = 0;
    if (p + kMinMatch <= e) {
      const u32 cap = p + xz::kCap < e ? xz::kCap : e - p;
      u32 best = 0, boff = 0;
      i32 bs = -(1 << 30);
      auto consider = [&](u32 off, u32 l) {
        if (l < kMinMatch) return;
        const i32 sc = (i32)l * xz::kLenWeight - ((i32)xz::highbit(off + 3) + xz::kOffCost);
        if (sc > bs) {
          bs = sc;
          best = l;
          boff = off;
        }
      };
#pragma unroll
      for (u32 k = 0; k < 3; ++k)
        if (kFixedOff[k] <= p) consider(kFixedOff[k], eq_run(eq + k * kEqWords, p, cap));
      // A hash hit is measured only if even a full-length match at its
      // offset could beat the best so far (ties keep the earlier candidate,
      // so skipping it cannot change the choice).
      auto can_win = [&](u32 off) {
        return (i32)cap * xz::kLenWeight - ((i32)xz::highbit(off + 3) + xz::kOffCost) > bs;
      };
      if (p + 4 <= n) {
        // p + 8 may pass n: v1 then holds zero padding, never compared
        // past the block end (len8 stops at p + cap <= n)
        const u32 v = LdsReader::r32(data, p), v1 = LdsReader::r32(data, p + 4);
        const u32 q4 = h4[xz::hash4(v)];
        if (q4 < p && can_win(p - q4)) consider(p - q4, LdsReader::len8(data, p, q4, p + cap, v, v1));
        if (p + 8 <= n) {
          const u32 q8 = h8[xz::hash8((u64)v | ((u64)v1 << 32))];
          // the same earliest position as the 4-byte hit measures the same
          if (q8 < p && q8 != q4 && can_win(p - q8))
            consider(p - q8, LdsReader::len8(data, p, q8, p + cap, v, v1));
        }
      }
      if (best) c = (best << 16) | boff;
    }
    cand[p] = c;
  }
}

constexpr u32 kParseBlocks = 1;  // blocks per k_zparse workgroup (32 segments each)
constexpr u32 kParseLanes = kParseBlocks * kSegs;  // one lane per segment
constexpr u32 kWin = 20;         // candidates per k_zparse window fetch
constexpr u32 kRefill = 4;       // a wave-wide refill reloads lanes this far into their window
// A window read starting at a position below kBlock stays inside the block's
// scratch (cand is its first member; the words past it are never used).
static_assert(offsetof(BlkScratch, cand) == 0 && sizeof(BlkScratch) >= sizeof(u32) * (kBlock + kWin),
              "window past scratch");

__global__ __launch_bounds__(kParseLanes) void k_zparse(const u8* __restrict__ pre,
                                                        const ZBlk* __restrict__ blks, u32 nblk,
                                                        BlkScratch* __restrict__ scr) {
  __shared__ __align__(16) u8 data[kParseBlocks][kBlock + kPad];
  __shared__ u32 hist[kParseBlocks][256];
  __shared__ u32 win[kParseLanes][kWin + 1];  // per lane: candidates of positions [wb, wb + kWin)
  const u32 lane = threadIdx.x;
  const u32 b0 = blockIdx.x * kParseBlocks;
  for (u32 k = 0; k < kParseBlocks; ++k) {
    if (b0 + k < nblk) load_block<kParseLanes>(data[k], pre + blks[b0 + k].src, blks[b0 + k].n, lane);
    for (u32 i = lane; i < 256; i += kParseLanes) hist[k][i] = 0;
  }
  __syncthreads();
  const u32 k = lane / kSegs, seg = lane % kSegs, b = b0 + k;
  if (b < nblk) {
    const u32 n = blks[b].n, nseg = (n + kSeg - 1) / kSeg;
    BlkScratch& S = scr[b];
    if (seg < nseg) {
      const u8* d = data[k];
      const u32* cand = S.cand;
      const u32 a = seg * kSeg, e = a + kSeg < n ? a + kSeg : n;
      Seq* sq = S.seqs + seg * kSegSeqs;
      u8* lit = S.lits + seg * kSeg;
      u32* h = hist[k];
      u32 ns = 0, nl = 0;
      // xz::cand_match on the resolved candidate (the capped length;
      // parse_segment extends it).  Candidates are fetched kWin at a time
      // into the lane's LDS window.  Lanes move through their segments at
      // different speeds, so some lane runs out of its window on most steps;
      // the refill is therefore wave-wide: when any active lane has left its
      // window, every active lane at least kRefill positions into its own
      // reloads as well, and the wave waits on global memory once for all.
      u32 wb = 0x80000000u;  // p - wb >= kWin for every p: the first call fetches
      u32* wr = win[lane];
      auto best_at = [&](u32 p) -> u32 {
        if (__any(p - wb >= kWin) && p - wb >= kRefill) {
          wb = p & ~3u;
          const uint4* src = reinterpret_cast<const uint4*>(cand + wb);
          uint4 v[kWin / 4];
#pragma unroll
          for (u32 q = 0; q < kWin / 4; ++q) v[q] = src[q];
#pragma unroll
          for (u32 q = 0; q < kWin / 4; ++q) {
            wr[4 * q] = v[q].x;
            wr[4 * q + 1] = v[q].y;
            wr[4 * q + 2] = v[q].z;
            wr[4 * q + 3] = v[q].w;
          }
        }
        return wr[p - wb];
      };
      const u32 tail = xz::parse_segment<LdsReader>(d, a, e, best_at,
                                                    [&](u32 l0, u32 p, u32 len, u32 off) {
        for (u32 i = l0; i < p; ++i) {
          lit[nl++] = d[i];
          atomicAdd(&h[d[i]], 1u);
        }
        sq[ns++] = Seq{p - l0, len, off};
      });
      for (u32 i = tail; i < e; ++i) {
        lit[nl++] = d[i];
        atomicAdd(&h[d[i]], 1u);
      }
      S.nseq[seg] = ns;
      S.nlit[seg] = nl;
      S.ntail[seg] = e - tail;
    }
  }
  __syncthreads();
  for (u32 kk = 0; kk < kParseBlocks; ++kk)
    if (b0 + kk < nblk)
      for (u32 i = lane; i < 256; i += kParseLanes) scr[b0 + kk].hist[i] = hist[kk][i];
}

__global__ __launch_bounds__(64) void k_zlits(const ZBlk* __restrict__ blks,
                                              BlkScratch* __restrict__ scr,
                                              u8* __restrict__ slots) {
  __shared__ LitLds L;
  const u32 b = blockIdx.x, lane = threadIdx.x;
  const ZBlk bk = blks[b];
  BlkScratch& S = scr[b];
  const u32 n = bk.n, nseg = (n + kSeg - 1) / kSeg;
  u8* out = slots + (u64)b * kSlot;
  // 1. literals compacted in segment order
  if (lane < kSegs) L.segl[lane] = lane < nseg ? S.nlit[lane] : 0u;
  for (u32 i = lane; i < 256; i += 64) L.hist[i] = S.hist[i];
  __syncthreads();
  {
    const u32 v = lane < kSegs ? L.segl[lane] : 0u;
    const u32 inc = wave_incl_scan(v);
    if (lane <= kSegs) L.loff[lane] = inc - v;
  }
  __syncthreads();
  const u32 nl = L.loff[nseg];
  // Four segments at a time, every lane's loads (up to 8 bytes of each)
  // issued before its stores, so the wave waits on memory once per group.
  for (u32 t0 = 0; t0 < nseg; t0 += 4) {
    u32 v[4][kSeg / 64];
#pragma unroll
    for (u32 tt = 0; tt < 4; ++tt) {
      // indices clamped to the segment's literals (slots are always readable)
      const u32 t = t0 + tt < kSegs ? t0 + tt : kSegs - 1;
      const u32 ln = t0 + tt < nseg ? L.segl[t] : 0u, top = ln ? ln - 1 : 0u;
#pragma unroll
      for (u32 j = 0; j < kSeg / 64; ++j)
        v[tt][j] = S.lits[t * kSeg + (j * 64 + lane < top ? j * 64 + lane : top)];
    }
#pragma unroll
    for (u32 tt = 0; tt < 4; ++tt) {
      const u32 t = t0 + tt;
      const u32 ln = t < nseg ? L.segl[t] : 0u, lo = t < nseg ? L.loff[t] : 0u;
#pragma unroll
      for (u32 j = 0; j < kSeg / 64; ++j)
        if (j * 64 + lane < ln) S.clit[lo + j * 64 + lane] = (u8)v[tt][j];
    }
  }
  __threadfence_block();
  __syncthreads();
  // 2. the literals section (xz::finish_block's choice)
  u8* body = out + 3;
  u32 ls = 0;
  u32 other = 0;  // a symbol count other than 0 and nl
  for (u32 i = lane; i < 256; i += 64) other |= (L.hist[i] != 0 && L.hist[i] != nl) ? 1u : 0u;
  const bool rle = nl > 1 && wave_max(other) == 0;
  if (rle) {
    if (lane == 0) ls = xz::write_raw_lits(body, S.clit, nl, 1);
    ls = __shfl((int)ls, 0, 64);
  } else {
    if (nl >= 64 && dev_huf_build(L.hist, L.h, L.W)) {
      u64 bits = 0;
      for (u32 c = lane; c < 256; c += 64) bits += (u64)L.hist[c] * L.h.nbits[c];
      bits = wave_sum(bits);
      if (bits / 8 + 16 < nl) ls = dev_huf_lits(out, 3, S.clit, nl, L);
    }
    if (!ls) {  // raw literals: header, then the bytes
      const u32 h = nl < 32 ? 1 : nl < 4096 ? 2 : 3;
      if (lane == 0) {
        if (h == 1) {
          body[0] = (u8)(nl << 3);
        } else {
          const u32 v = (nl << 4) | ((h == 2 ? 1u : 3u) << 2);
          for (u32 i = 0; i < h; ++i) body[i] = (u8)(v >> (8 * i));
        }
      }
      for (u32 i = lane; i < nl; i += 64) body[h + i] = S.clit[i];
      ls = h + nl;
    }
  }
  if (lane == 0) S.ls = ls;
}

__global__ __launch_bounds__(64) void k_zseqs(const ZBlk* __restrict__ blks,
                                              BlkScratch* __restrict__ scr,
                                              u8* __restrict__ slots,
                                              u64* __restrict__ slot_len) {
  __shared__ SeqLds L;
  const u32 b = blockIdx.x, lane = threadIdx.x;
  const ZBlk bk = blks[b];
  BlkScratch& S = scr[b];
  const u32 n = bk.n, nseg = (n + kSeg - 1) / kSeg;
  u8* out = slots + (u64)b * kSlot;
  // 1. per-segment counts, offsets and literal carries; sequences compacted
  //    (a segment's first sequence takes the literals left pending before it)
  if (lane < kSegs) {
    const bool in = lane < nseg;
    L.segn[lane] = in ? S.nseq[lane] : 0u;
    L.segl[lane] = in ? S.nlit[lane] : 0u;
    L.segt[lane] = in ? S.ntail[lane] : 0u;
  }
  __syncthreads();
  if (lane == 0) {
    u32 so = 0, carry = 0;
    for (u32 t = 0; t < nseg; ++t) {
      L.soff[t] = so;
      L.carry[t] = carry;
      so += L.segn[t];
      carry = L.segn[t] ? L.segt[t] : carry + L.segl[t];
    }
    L.soff[nseg] = so;
  }
  __syncthreads();
  const u32 ns = L.soff[nseg];
  // Four segments at a time, every lane's loads issued before its stores.
  for (u32 t0 = 0; t0 < nseg; t0 += 4) {
    Seq v[4][kSegSeqs / 64];
#pragma unroll
    for (u32 tt = 0; tt < 4; ++tt) {
      // indices clamped to the segment's sequences (slots are always readable)
      const u32 t = t0 + tt < kSegs ? t0 + tt : kSegs - 1;
      const u32 sn = t0 + tt < nseg ? L.segn[t] : 0u, top = sn ? sn - 1 : 0u;
#pragma unroll
      for (u32 j = 0; j < kSegSeqs / 64; ++j)
        v[tt][j] = S.seqs[t * kSegSeqs + (j * 64 + lane < top ? j * 64 + lane : top)];
    }
#pragma unroll
    for (u32 tt = 0; tt < 4; ++tt) {
      const u32 t = t0 + tt;
      const u32 sn = t < nseg ? L.segn[t] : 0u, so = t < nseg ? L.soff[t] : 0u;
#pragma unroll
      for (u32 j = 0; j < kSegSeqs / 64; ++j) {
        const u32 k = j * 64 + lane;
        if (k < sn) {
          Seq q = v[tt][j];
          if (k == 0) q.ll += L.carry[t];
          S.cseq[so + k] = q;
        }
      }
    }
  }
  __threadfence_block();
  __syncthreads();
  // 2. the sequences section, at kSeqOff in the slot
  u8* sq = out + kSeqOff;
  u32 ss;
  if (lane == 0) ss = xz::seq_count_header(sq, ns);
  ss = __shfl((int)ss, 0, 64);
  if (ns) {
    dev_offset_values(S.cseq, ns, S.ov);
    for (u32 i = lane; i < 3 * 64; i += 64) (&L.sw.cnt[0][0])[i] = 0;
    __syncthreads();
    for (u32 i0 = 0; i0 < ns; i0 += 64 * kSeqBatch) {
      Seq q[kSeqBatch];
      u32 ovb[kSeqBatch];
#pragma unroll
      for (u32 b = 0; b < kSeqBatch; ++b) {
        const u32 i = i0 + b * 64 + lane;
        q[b] = S.cseq[i < ns ? i : ns - 1];
        ovb[b] = S.ov[i < ns ? i : ns - 1];
      }
#pragma unroll
      for (u32 b = 0; b < kSeqBatch; ++b) {
        const u32 i = i0 + b * 64 + lane;
        if (i >= ns) break;
        const SeqCodes c = codes_of(q[b], ovb[b]);
        atomicAdd(&L.sw.cnt[0][c.ll], 1u);
        atomicAdd(&L.sw.cnt[1][c.of], 1u);
        atomicAdd(&L.sw.cnt[2][c.ml], 1u);
        S.sym[0][i] = c.ll;
        S.sym[1][i] = c.of;
        S.sym[2][i] = c.ml;
      }
    }
    __threadfence_block();
    __syncthreads();
    ss += dev_seq_write_tables(sq + ss, ns, L.sw, L.F);
    __threadfence_block();
    __syncthreads();
    ss += dev_seq_bits(out, kSeqOff + ss, S, ns, L);
  }
  // 3. the block: compressed, or raw when that is not smaller
  if (lane == 0) {
    const u32 size = S.ls + ss;
    const bool raw = size >= n;
    S.ss = ss;
    S.raw = raw ? 1u : 0u;
    slot_len[b] = 3 + (raw ? n : size);
  }
}

__global__ __launch_bounds__(256) void k_zgather(const u8* __restrict__ pre,
                                                 const ZBlk* __restrict__ blks,
                                                 const ZFrame* __restrict__ frames,
                                                 const BlkScratch* __restrict__ scr,
                                                 const u8* __restrict__ slots,
                                                 const u64* __restrict__ pos,
                                                 const u64* __restrict__ slot_len,
                                                 u8* __restrict__ out) {
  const u32 b = blockIdx.x;
  const ZBlk bk = blks[b];
  const ZFrame fr = frames[bk.frame];
  const BlkScratch& S = scr[b];
  u8* dst = out + pos[b] + 13ull * (bk.frame + 1);
  const u8* src = slots + (u64)b * kSlot;
  const u32 len = (u32)slot_len[b], ls = S.ls, raw = S.raw;
  const u32 last = b == fr.first + fr.nblk - 1 ? 1u : 0u;  // last block of the frame
  const u32 h = raw ? (bk.n << 3) | last : (2u << 1) | ((len - 3) << 3) | last;
  for (u32 i = threadIdx.x; i < len; i += 256) {
    u8 v;
    if (i < 3) v = (u8)(h >> (8 * i));
    else if (raw) v = pre[bk.src + i - 3];
    else if (i - 3 < ls) v = src[i];
    else v = src[kSeqOff + (i - 3 - ls)];
    dst[i] = v;
  }
  if (b == fr.first && threadIdx.x < 13) {
    u8* fh = out + pos[b] + 13ull * bk.frame;
    const u32 i = threadIdx.x;
    const u64 magic = 0xFD2FB528ull;
    fh[i] = i < 4 ? (u8)(magic >> (8 * i)) : i == 4 ? (u8)0xE0 : (u8)(fr.len >> (8 * (i - 5)));
  }
}

}  // namespace

size_t zframe_scratch_bytes(u32 nblk) {
  return (size_t)nblk * (sizeof(BlkScratch) + kSlot + 2 * sizeof(u64)) + 4096;
}

void zframe_encode(const u8* pre, const ZBlk* blks, u32 nblk, const ZFrame* frames,
                   void* scratch, u8* out, u64* pos, Arena& ar, hipStream_t s, u64* h_pos,
                   hipEvent_t sizes_ready) {
  auto* scr = reinterpret_cast<BlkScratch*>(scratch);
  u8* slots = reinterpret_cast<u8*>(scr + nblk);
  u64* slot_len = reinterpret_cast<u64*>(slots + (size_t)nblk * kSlot);
  k_zcand<<<nblk, kCandThreads, 0, s>>>(pre, blks, scr);
  check_hip(hipGetLastError(), "k_zcand");
  if (g_profiler) g_profiler->mark("k_zcand", s);
  k_zparse<<<(nblk + kParseBlocks - 1) / kParseBlocks, kParseLanes, 0, s>>>(pre, blks, nblk, scr);
  check_hip(hipGetLastError(), "k_zparse");
  if (g_profiler) g_profiler->mark("k_zparse", s);
  k_zlits<<<nblk, 64, 0, s>>>(blks, scr, slots);
  check_hip(hipGetLastError(), "k_zlits");
  if (g_profiler) g_profiler->mark("k_zlits", s);
  k_zseqs<<<nblk, 64, 0, s>>>(blks, scr, slots, slot_len);
  check_hip(hipGetLastError(), "k_zseqs");
  if (g_profiler) g_profiler->mark("k_zseqs", s);
  scan_excl_u64(slot_len, pos, nblk, pos + nblk, ar, s);
  if (h_pos) {
    check_hip(hipMemcpyAsync(h_pos, pos, (nblk + 1) * sizeof(u64), hipMemcpyDeviceToHost, s),
              "block positions D2H");
    if (sizes_ready) check_hip(hipEventRecord(sizes_ready, s), "hipEventRecord");
    if (g_profiler) g_profiler->mark(nullptr, s);  // the copy is not k_zgather's time
  }
  k_zgather<<<nblk, 256, 0, s>>>(pre, blks, frames, scr, slots, pos, slot_len, out);
  check_hip(hipGetLastError(), "k_zgather");
  if (g_profiler) g_profiler->mark("k_zgather", s);
}

}  // namespace xnyz
