// zframe.hip -- zstd-format frames built on the device from the assembled
// streams in HBM (format and parse rules in zframe.hpp).
//
//   k_zcand    one workgroup per 16 KiB block: the block and its two hash
//              tables in LDS (48 KiB); all 256 threads evaluate every
//              position's best fixed candidate and write it, resolved to
//              (capped length, offset), to global scratch.
//   k_zparse   one wavefront per two blocks, one lane per 512-byte segment:
//              only the blocks' bytes sit in LDS (33 KiB, so four such
//              wavefronts share a CU), the candidates stream from scratch;
//              each lane runs the lazy parse of its segment, writing its
//              sequences and literals into per-segment slots and the
//              block's literal histogram.
//   Both read the block with aligned 32-bit LDS loads and byte funnel
//   shifts (match lengths four bytes per step); results equal the
//   byte-wise definitions in zframe.hpp.
//   k_zfinish  one wavefront per block: slots compacted in order (the first
//              sequence of a segment takes the literals left pending by the
//              segments before it), then the block as zframe.hpp's
//              finish_block writes it -- the serial steps (Huffman and FSE
//              table builds, repeat-offset history, the FSE chain) executed
//              uniformly by all 64 lanes on LDS working storage, with
//              sequences fetched 64 at a time and broadcast by lane
//              shuffles; Huffman streams bit-packed by all lanes at offsets
//              from a wavefront scan.
//   k_zgather  blocks copied behind their frame headers, positions from a
//              device-wide scan of the block sizes.
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

constexpr u32 kSegSeqs = kSeg / kMinMatch;  // most sequences one segment can hold

struct BlkScratch {  // per block, in global memory
  u32 cand[kBlock];            // (capped length << 16) | offset, 0 = none (k_zcand)
  Seq seqs[kSegs * kSegSeqs];  // per-segment slots (k_zparse)
  Seq cseq[kSegs * kSegSeqs];  // compacted (k_zfinish)
  u8 lits[kBlock];
  u8 clit[kBlock];
  u32 hist[256];
  u32 nseq[kSegs], nlit[kSegs];
};

constexpr u32 kMaxSeqs = kBlock / kMinMatch;

struct FinLds {
  xz::Work W;
  xz::SeqWork sw;
  xz::HufTable h;
  u32 hist[256];
  u16 ov[kMaxSeqs];  // offset values (offset + 3 or a repeat code), < 2^15 in a 16 KiB block
  u32 segn[kSegs], segl[kSegs], segll[kSegs], carry[kSegs], soff[kSegs + 1], loff[kSegs + 1];
  u8 desc[160];
  u32 sbits[4];
};

__device__ __forceinline__ Seq shfl_seq(const Seq& q, u32 j) {
  return Seq{(u32)__shfl((int)q.ll, (int)j, 64), (u32)__shfl((int)q.ml, (int)j, 64),
             (u32)__shfl((int)q.off, (int)j, 64)};
}

// Repeat-offset history over the sequences (zframe.hpp seq_codes), one
// step per sequence, into L.ov.
__device__ void dev_offset_values(const Seq* cs, u32 ns, FinLds& L) {
  const u32 lane = threadIdx.x;
  u32 rep[3] = {0, 0, 0}, known = 0;
  for (u32 base = 0; base < ns; base += 64) {
    const Seq mine = base + lane < ns ? cs[base + lane] : Seq{0, 0, 0};
    const u32 m = ns - base < 64 ? ns - base : 64;
    for (u32 j = 0; j < m; ++j) {
      const Seq q = shfl_seq(mine, j);
      const u32 o = q.off, ll = q.ll;
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
      if (lane == 0) L.ov[base + j] = (u16)ov;
    }
  }
}

__device__ __forceinline__ SeqCodes codes_of(const Seq& q, u32 ov) {
  SeqCodes c;
  c.ll = (u8)xz::ll_code(q.ll);
  c.ml = (u8)xz::ml_code(q.ml);
  c.ov = ov;
  c.of = (u8)xz::highbit(ov);
  return c;
}

// Huffman-coded literals as write_huf_lits lays them out, written at their
// final place: slot + base (slot 256-byte aligned; words are addressed from
// the slot so that every 32-bit store and atomic is aligned).  Returns 0
// when not representable or not smaller than raw.
__device__ u32 dev_huf_lits(u8* slot, u32 base, const u8* lit, u32 n, FinLds& L) {
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
    for (u32 i = a + lane; i < e; i += 64) bits += L.h.nbits[lit[i]];
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
    const u32 per = (m + 63) / 64, r0 = lane * per, r1 = r0 + per < m ? r0 + per : m;
    u32 bits = 0;
    for (u32 r = r0; r < r1; ++r) bits += L.h.nbits[lit[e - 1 - r]];
    const u32 inc = wave_incl_scan(bits);
    u64 pos = (u64)(base + at) * 8 + (inc - bits);  // bit position from the slot start
    u64 acc = 0;
    u32 nb = (u32)(pos & 31);
    u64 word = pos >> 5;
    bool first = true;
    for (u32 r = r0; r < r1; ++r) {
      const u8 c = lit[e - 1 - r];
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

// Word-wise reader of a block held in LDS (4-byte aligned, padded by 8
// bytes): the same values as xz::ByteReader.
struct LdsReader {
  __device__ static u32 r32(const u8* d, u32 p) {
    const u32* w = reinterpret_cast<const u32*>(d);
    return __builtin_amdgcn_alignbyte(w[(p >> 2) + 1], w[p >> 2], p & 3u);
  }
  __device__ static u32 len(const u8* d, u32 p, u32 q, u32 end) {
    u32 l = 0;
    while (p + l < end) {
      const u32 x = r32(d, p + l) ^ r32(d, q + l);
      const u32 r = end - (p + l);
      if (x) {
        const u32 e = (u32)__builtin_ctz(x) >> 3;
        return l + (e < r ? e : r);
      }
      if (r <= 4) return l + r;
      l += 4;
    }
    return l;
  }
};

constexpr u32 kPad = 16;  // LDS block copies: readable bytes past the end

// Copies block bytes [0, n) into LDS (4-byte aligned source and target);
// the padding is zeroed so no load reads uninitialised memory.
__device__ __forceinline__ void load_block(u8* data, const u8* src, u32 n, u32 tid, u32 nthr) {
  for (u32 i = tid * 4; i + 4 <= n; i += nthr * 4)
    *reinterpret_cast<u32*>(data + i) = *reinterpret_cast<const u32*>(src + i);
  for (u32 i = (n & ~3u) + tid; i < n + kPad; i += nthr) data[i] = i < n ? src[i] : 0;
}

__global__ __launch_bounds__(256) void k_zcand(const u8* __restrict__ pre,
                                               const ZBlk* __restrict__ blks,
                                               BlkScratch* __restrict__ scr) {
  __shared__ __align__(16) u8 data[kBlock + kPad];
  __shared__ u32 ht4[1u << kHashLog], ht8[1u << kHashLog];
  const ZBlk bk = blks[blockIdx.x];
  u32* cand = scr[blockIdx.x].cand;
  const u32 n = bk.n, tid = threadIdx.x;
  load_block(data, pre + bk.src, n, tid, 256);
  for (u32 i = tid; i < (1u << kHashLog); i += 256) ht4[i] = ht8[i] = 0xFFFFFFFFu;
  __syncthreads();
  for (u32 p = tid; p + 4 <= n; p += 256) {
    const u32 v = LdsReader::r32(data, p);
    atomicMin(&ht4[xz::hash4(v)], p);
    if (p + 8 <= n)
      atomicMin(&ht8[xz::hash8((u64)v | ((u64)LdsReader::r32(data, p + 4) << 32))], p);
  }
  __syncthreads();
  for (u32 p = tid; p < n; p += 256) {
    const u8 info = xz::best_fixed<LdsReader>(data, n, ht4, ht8, p);
    u32 c = 0;
    if (info)
      c = (((info & 31u) + kMinMatch) << 16) |
          xz::cand_offset<LdsReader>(data, n, ht4, ht8, p, info >> 5);
    cand[p] = c;
  }
}

constexpr u32 kParseBlocks = 2;  // blocks per k_zparse wavefront (32 segments each)

__global__ __launch_bounds__(64) void k_zparse(const u8* __restrict__ pre,
                                               const ZBlk* __restrict__ blks, u32 nblk,
                                               BlkScratch* __restrict__ scr) {
  __shared__ __align__(16) u8 data[kParseBlocks][kBlock + kPad];
  __shared__ u32 hist[kParseBlocks][256];
  const u32 lane = threadIdx.x;
  const u32 b0 = blockIdx.x * kParseBlocks;
  for (u32 k = 0; k < kParseBlocks; ++k) {
    if (b0 + k < nblk) load_block(data[k], pre + blks[b0 + k].src, blks[b0 + k].n, lane, 64);
    for (u32 i = lane; i < 256; i += 64) hist[k][i] = 0;
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
      // xz::cand_match on the resolved candidate: a capped length is
      // extended to the segment end.
      auto best_at = [&](u32 p) -> u32 {
        const u32 c = cand[p];
        u32 len = c >> 16;
        if (len == xz::kCap) len = LdsReader::len(d, p, p - (c & 0xFFFFu), xz::seg_end(p, n));
        return (len << 16) | (c & 0xFFFFu);
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
    }
  }
  __syncthreads();
  for (u32 kk = 0; kk < kParseBlocks; ++kk)
    if (b0 + kk < nblk)
      for (u32 i = lane; i < 256; i += 64) scr[b0 + kk].hist[i] = hist[kk][i];
}

__global__ __launch_bounds__(64) void k_zfinish(const u8* __restrict__ pre,
                                                const ZBlk* __restrict__ blks,
                                                BlkScratch* __restrict__ scr,
                                                u8* __restrict__ slots,
                                                u64* __restrict__ slot_len) {
  __shared__ FinLds L;
  const u32 b = blockIdx.x, lane = threadIdx.x;
  const ZBlk bk = blks[b];
  BlkScratch& S = scr[b];
  const u32 n = bk.n, nseg = (n + kSeg - 1) / kSeg;
  u8* out = slots + (u64)b * kSlot;
  // 1. per-segment counts, offsets and literal carries
  if (lane < kSegs) {
    u32 ns = 0, nl = 0, llsum = 0;
    if (lane < nseg) {
      ns = S.nseq[lane];
      nl = S.nlit[lane];
      for (u32 k = 0; k < ns; ++k) llsum += S.seqs[lane * kSegSeqs + k].ll;
    }
    L.segn[lane] = ns;
    L.segl[lane] = nl;
    L.segll[lane] = llsum;
  }
  for (u32 i = lane; i < 256; i += 64) L.hist[i] = S.hist[i];
  __syncthreads();
  if (lane == 0) {
    u32 so = 0, lo = 0, carry = 0;
    for (u32 t = 0; t < nseg; ++t) {
      L.soff[t] = so;
      L.loff[t] = lo;
      L.carry[t] = carry;
      so += L.segn[t];
      lo += L.segl[t];
      carry = L.segn[t] ? L.segl[t] - L.segll[t] : carry + L.segl[t];
    }
    L.soff[nseg] = so;
    L.loff[nseg] = lo;
  }
  __syncthreads();
  const u32 ns = L.soff[nseg], nl = L.loff[nseg];
  // 2. compaction (separate arrays: no overlap)
  for (u32 t = 0; t < nseg; ++t) {
    const u32 sn = L.segn[t], so = L.soff[t], ln = L.segl[t], lo = L.loff[t];
    for (u32 k = lane; k < sn; k += 64) {
      Seq q = S.seqs[t * kSegSeqs + k];
      if (k == 0) q.ll += L.carry[t];
      S.cseq[so + k] = q;
    }
    for (u32 i = lane; i < ln; i += 64) S.clit[lo + i] = S.lits[t * kSeg + i];
  }
  __threadfence_block();
  __syncthreads();
  // 3. literals section (finish_block's choice)
  u8* body = out + 3;
  u32 ls = 0;
  bool rle = nl > 1;
  for (u32 i = 0; rle && i < 256; ++i) rle = L.hist[i] == 0 || L.hist[i] == nl;
  if (rle) {
    if (lane == 0) ls = xz::write_raw_lits(body, S.clit, nl, 1);
    ls = __shfl((int)ls, 0, 64);
  } else {
    if (nl >= 64 && xz::huf_build(L.hist, L.h, L.W)) {  // uniform
      u64 bits = 0;
      for (u32 c = 0; c < 256; ++c) bits += (u64)L.hist[c] * L.h.nbits[c];
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
  __threadfence_block();
  __syncthreads();
  // 4. sequences section
  u8* sq = body + ls;
  u32 ss;
  if (lane == 0) ss = xz::seq_count_header(sq, ns);
  ss = __shfl((int)ss, 0, 64);
  if (ns) {
    dev_offset_values(S.cseq, ns, L);
    for (u32 i = lane; i < 3 * 64; i += 64) (&L.sw.cnt[0][0])[i] = 0;
    __syncthreads();
    for (u32 i = lane; i < ns; i += 64) {
      const SeqCodes c = codes_of(S.cseq[i], L.ov[i]);
      atomicAdd(&L.sw.cnt[0][c.ll], 1u);
      atomicAdd(&L.sw.cnt[1][c.of], 1u);
      atomicAdd(&L.sw.cnt[2][c.ml], 1u);
    }
    __syncthreads();
    ss += xz::seq_write_tables(sq + ss, ns, L.sw, L.W);  // uniform
    xz::BitW bw(sq + ss);
    xz::FseState st[3];
    {
      const Seq q = S.cseq[ns - 1];
      xz::seq_encode_last(bw, st, L.sw, q, codes_of(q, L.ov[ns - 1]));
    }
    for (u32 top = ns - 1; top > 0;) {  // sequences top-1 down to 0, 64 at a time
      const u32 m = top < 64 ? top : 64;
      const u32 idx = top - 1 - lane;
      const Seq mine = lane < m ? S.cseq[idx] : Seq{0, 0, 0};
      const u32 myov = lane < m ? L.ov[idx] : 0u;
      for (u32 j = 0; j < m; ++j) {
        const Seq q = shfl_seq(mine, j);
        const u32 ov = (u32)__shfl((int)myov, (int)j, 64);
        xz::seq_encode_one(bw, st, L.sw, q, codes_of(q, ov));
      }
      top -= m;
    }
    ss += xz::seq_encode_close(bw, st, L.sw);
  }
  __threadfence_block();
  __syncthreads();
  // 5. the block: compressed, or raw when that is not smaller
  const u32 size = ls + ss;
  if (size >= n) {
    for (u32 i = lane; i < n; i += 64) out[3 + i] = pre[bk.src + i];
    if (lane == 0) {
      out[0] = (u8)(n << 3);
      out[1] = (u8)(n >> 5);
      out[2] = (u8)(n >> 13);
      slot_len[b] = 3 + n;
    }
  } else if (lane == 0) {
    const u32 h = (2u << 1) | (size << 3);
    out[0] = (u8)h;
    out[1] = (u8)(h >> 8);
    out[2] = (u8)(h >> 16);
    slot_len[b] = 3 + size;
  }
}

// Frame f at out + pos[frames[f].first] + 13 f: magic, descriptor (single
// segment, 8-byte content size), content size, then its blocks.
__global__ __launch_bounds__(256) void k_zgather(const ZBlk* __restrict__ blks,
                                                 const ZFrame* __restrict__ frames,
                                                 const u8* __restrict__ slots,
                                                 const u64* __restrict__ pos,
                                                 const u64* __restrict__ slot_len,
                                                 u8* __restrict__ out) {
  const u32 b = blockIdx.x;
  const ZBlk bk = blks[b];
  const ZFrame fr = frames[bk.frame];
  u8* dst = out + pos[b] + 13ull * (bk.frame + 1);
  const u8* src = slots + (u64)b * kSlot;
  const u32 len = (u32)slot_len[b];
  for (u32 i = threadIdx.x; i < len; i += 256) {
    u8 v = src[i];
    if (i == 0 && b == fr.first + fr.nblk - 1) v |= 1;  // last block of the frame
    dst[i] = v;
  }
  if (b == fr.first && threadIdx.x < 13) {
    u8* h = out + pos[b] + 13ull * bk.frame;
    const u32 i = threadIdx.x;
    const u64 magic = 0xFD2FB528ull;
    h[i] = i < 4 ? (u8)(magic >> (8 * i)) : i == 4 ? (u8)0xE0 : (u8)(fr.len >> (8 * (i - 5)));
  }
}

}  // namespace

size_t zframe_scratch_bytes(u32 nblk) {
  return (size_t)nblk * (sizeof(BlkScratch) + kSlot + 2 * sizeof(u64)) + 4096;
}

void zframe_encode(const u8* pre, const ZBlk* blks, u32 nblk, const ZFrame* frames,
                   void* scratch, u8* out, u64* pos, Arena& ar, hipStream_t s) {
  auto* scr = reinterpret_cast<BlkScratch*>(scratch);
  u8* slots = reinterpret_cast<u8*>(scr + nblk);
  u64* slot_len = reinterpret_cast<u64*>(slots + (size_t)nblk * kSlot);
  k_zcand<<<nblk, 256, 0, s>>>(pre, blks, scr);
  check_hip(hipGetLastError(), "k_zcand");
  if (g_profiler) g_profiler->mark("k_zcand", s);
  k_zparse<<<(nblk + kParseBlocks - 1) / kParseBlocks, 64, 0, s>>>(pre, blks, nblk, scr);
  check_hip(hipGetLastError(), "k_zparse");
  if (g_profiler) g_profiler->mark("k_zparse", s);
  k_zfinish<<<nblk, 64, 0, s>>>(pre, blks, scr, slots, slot_len);
  check_hip(hipGetLastError(), "k_zfinish");
  if (g_profiler) g_profiler->mark("k_zfinish", s);
  scan_excl_u64(slot_len, pos, nblk, pos + nblk, ar, s);
  k_zgather<<<nblk, 256, 0, s>>>(blks, frames, slots, pos, slot_len, out);
  check_hip(hipGetLastError(), "k_zgather");
  if (g_profiler) g_profiler->mark("k_zgather", s);
}

}  // namespace xnyz
