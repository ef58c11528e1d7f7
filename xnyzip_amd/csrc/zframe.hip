// zframe.hip -- zstd-format frames built on the device from the assembled
// streams in HBM (format and parse rules in zframe.hpp).
//
//   k_zparse   one workgroup per 64 KiB block: the block and its two hash
//              tables in LDS (128 KiB); the best fixed candidate of every
//              position, evaluated by all 256 threads; then one thread per
//              1 KiB segment runs the lazy parse over those candidates,
//              writing its sequences and literals into per-segment slots
//              and the literal histogram.
//   k_zfinish  one lane per block: slots compacted in order (the first
//              sequence of a segment takes the literals left pending by the
//              segments before it), then the literals and sequences
//              sections (finish_block) into the block's output slot.
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

constexpr u32 kSegSeqs = kSeg / kMinMatch;  // most sequences one segment can hold

struct BlkScratch {  // per block, in global memory
  u32 cand[kBlock];
  Seq seqs[kSegs * kSegSeqs];
  SeqCodes codes[kSegs * kSegSeqs];
  u8 lits[kBlock];
  u32 hist[256];
  u32 nseq[kSegs], nlit[kSegs];
};

__global__ __launch_bounds__(256) void k_zparse(const u8* __restrict__ pre,
                                                const ZBlk* __restrict__ blks,
                                                BlkScratch* __restrict__ scr) {
  __shared__ u8 data[kBlock];
  __shared__ u32 ht4[1u << kHashLog], ht8[1u << kHashLog];
  __shared__ u32 hist[256];
  const ZBlk bk = blks[blockIdx.x];
  BlkScratch& S = scr[blockIdx.x];
  const u32 n = bk.n, tid = threadIdx.x;
  const u8* src = pre + bk.src;
  for (u32 i = tid * 4; i + 4 <= n; i += kBlock / 64)  // 4-byte aligned source
    *reinterpret_cast<u32*>(data + i) = *reinterpret_cast<const u32*>(src + i);
  for (u32 i = (n & ~3u) + tid; i < n; i += 256) data[i] = src[i];
  for (u32 i = tid; i < (1u << kHashLog); i += 256) ht4[i] = ht8[i] = 0xFFFFFFFFu;
  hist[tid] = 0;
  __syncthreads();
  for (u32 p = tid; p + 4 <= n; p += 256) {
    atomicMin(&ht4[xz::hash4(xz::read32(data + p))], p);
    if (p + 8 <= n) atomicMin(&ht8[xz::hash8(xz::read64(data + p))], p);
  }
  __syncthreads();
  for (u32 p = tid; p < n; p += 256) S.cand[p] = xz::best_fixed(data, n, ht4, ht8, p);
  __syncthreads();  // cand[] of every position visible to the parse threads
  const u32 nseg = (n + kSeg - 1) / kSeg;
  if (tid < nseg) {
    const u32 a = tid * kSeg, b = a + kSeg < n ? a + kSeg : n;
    Seq* sq = S.seqs + tid * kSegSeqs;
    u8* lit = S.lits + tid * kSeg;
    u32 ns = 0, nl = 0;
    auto best_at = [&](u32 p) { return S.cand[p]; };
    const u32 tail = xz::parse_segment(data, a, b, best_at, [&](u32 l0, u32 p, u32 len, u32 off) {
      for (u32 i = l0; i < p; ++i) {
        lit[nl++] = data[i];
        atomicAdd(&hist[data[i]], 1u);
      }
      sq[ns++] = Seq{p - l0, len, off};
    });
    for (u32 i = tail; i < b; ++i) {
      lit[nl++] = data[i];
      atomicAdd(&hist[data[i]], 1u);
    }
    S.nseq[tid] = ns;
    S.nlit[tid] = nl;
  }
  __syncthreads();
  S.hist[tid] = hist[tid];
}

__global__ __launch_bounds__(64) void k_zfinish(const u8* __restrict__ pre,
                                                const ZBlk* __restrict__ blks, u32 nblk,
                                                BlkScratch* __restrict__ scr,
                                                u8* __restrict__ slots,
                                                u64* __restrict__ slot_len) {
  const u32 b = blockIdx.x * 64 + threadIdx.x;
  if (b >= nblk) return;
  const ZBlk bk = blks[b];
  BlkScratch& S = scr[b];
  const u32 n = bk.n, nseg = (n + kSeg - 1) / kSeg;
  // compact the per-segment slots in place (destinations never pass sources)
  u32 ns = 0, nl = 0, carry = 0;
  for (u32 t = 0; t < nseg; ++t) {
    const Seq* sq = S.seqs + t * kSegSeqs;
    u32 in_seqs = 0;
    for (u32 k = 0; k < S.nseq[t]; ++k) {
      Seq q = sq[k];
      in_seqs += q.ll;
      if (k == 0) q.ll += carry;
      S.seqs[ns++] = q;
    }
    const u8* lt = S.lits + t * kSeg;
    const u32 m = S.nlit[t];
    for (u32 i = 0; i < m; ++i) S.lits[nl + i] = lt[i];
    nl += m;
    carry = S.nseq[t] ? m - in_seqs : carry + m;
  }
  xz::HufTable h;
  slot_len[b] = xz::finish_block(slots + (u64)b * kSlot, pre + bk.src, n, S.lits, nl, S.hist,
                                 S.seqs, ns, S.codes, h);
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
  k_zparse<<<nblk, 256, 0, s>>>(pre, blks, scr);
  check_hip(hipGetLastError(), "k_zparse");
  if (g_profiler) g_profiler->mark("k_zparse", s);
  k_zfinish<<<(nblk + 63) / 64, 64, 0, s>>>(pre, blks, nblk, scr, slots, slot_len);
  check_hip(hipGetLastError(), "k_zfinish");
  if (g_profiler) g_profiler->mark("k_zfinish", s);
  scan_excl_u64(slot_len, pos, nblk, pos + nblk, ar, s);
  k_zgather<<<nblk, 256, 0, s>>>(blks, frames, slots, pos, slot_len, out);
  check_hip(hipGetLastError(), "k_zgather");
  if (g_profiler) g_profiler->mark("k_zgather", s);
}

}  // namespace xnyz
