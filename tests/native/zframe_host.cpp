// Host driver of the zstd-format block encoder (xnyzip_amd/csrc/zframe.hpp),
// test infrastructure only: runs the encoder's steps in the order the device
// kernel does (earliest-position hash tables, one candidate byte per
// position, segments parsed independently, literal-length carry across
// segments), so the device
// frames can be compared with these byte for byte and these checked against
// libzstd's decoder.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../xnyzip_amd/csrc/zframe.hpp"

using namespace xz;

static u32 encode_block(const u8* d, u32 n, u8* out) {
  std::vector<u32> ht4(1u << kHashLog, 0xFFFFFFFFu), ht8(1u << kHashLog, 0xFFFFFFFFu);
  for (u32 p = 0; p + 4 <= n; ++p) {
    u32& e = ht4[hash4(read32(d + p))];
    if (p < e) e = p;
    if (p + 8 <= n) {
      u32& f = ht8[hash8(read64(d + p))];
      if (p < f) f = p;
    }
  }
  std::vector<u8> info(n);
  for (u32 p = 0; p < n; ++p) info[p] = best_fixed(d, n, ht4.data(), ht8.data(), p);
  std::vector<Seq> seqs;
  std::vector<u8> lits;
  u32 carry = 0;  // literals pending from earlier segments
  for (u32 a = 0; a < n; a += kSeg) {
    const u32 b = a + kSeg < n ? a + kSeg : n;
    bool first = true;
    auto best_at = [&](u32 p) { return cand_match(d, n, ht4.data(), ht8.data(), p, info[p]); };
    const u32 tail = parse_segment(d, a, b, best_at, [&](u32 lit0, u32 p, u32 len, u32 off) {
      for (u32 i = lit0; i < p; ++i) lits.push_back(d[i]);
      seqs.push_back(Seq{(p - lit0) + (first ? carry : 0), len, off});
      if (first) carry = 0;
      first = false;
    });
    for (u32 i = tail; i < b; ++i) lits.push_back(d[i]);
    carry += b - tail;
  }
  u32 hist[256] = {0};
  for (u8 c : lits) hist[c]++;
  std::vector<SeqCodes> codes(seqs.size() + 1);
  HufTable h;
  static Work W;
  static SeqWork sw;
  return finish_block(out, d, n, lits.data(), (u32)lits.size(), hist, seqs.data(),
                      (u32)seqs.size(), codes.data(), h, W, sw);
}

extern "C" uint64_t zf_frame_bound(uint64_t n) { return 13 + n + 3 * (n / kBlock + 1) + 64; }

// Whole frame: header (single segment, 8-byte content size) and blocks.
extern "C" uint64_t zf_encode_frame(const uint8_t* src, uint64_t n, uint8_t* dst) {
  const uint32_t magic = 0xFD2FB528u;
  std::memcpy(dst, &magic, 4);
  dst[4] = 0xE0;
  std::memcpy(dst + 5, &n, 8);
  uint64_t at = 13;
  if (n == 0) {  // one empty raw block
    dst[at] = 1;
    dst[at + 1] = dst[at + 2] = 0;
    return at + 3;
  }
  for (uint64_t b = 0; b < n; b += kBlock) {
    const u32 m = (u32)(n - b < kBlock ? n - b : kBlock);
    const u32 w = encode_block(src + b, m, dst + at);
    if (b + m == n) dst[at] |= 1;  // last block
    at += w;
  }
  return at;
}
