// ref_harness.cpp -- C entry points over the reference's own encoder headers
// and Hilbert source, compiled as they lie in the reference checkout.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ when
// the reference checkout is present; used by tests/ to pin the oracle's
// restatement (oracle/xnyz_oracle.cpp) against the reference code itself and
// by tests/golden/make_golden.py to write fixtures.  The parts of the
// reference that need Eigen (quantizers, block_quantizer, composers) cannot
// be built here and are not included.
//
// Reference files compiled: source/encoder/{bitpacking,p_for_delta,
// golomb_rice,huffman}_encoder.hpp and
// source/quantizer/hilbert-curves-cpp/hilbert.cpp.

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "encoder/bitpacking_encoder.hpp"
#include "encoder/golomb_rice_encoder.hpp"
#include "encoder/huffman_encoder.hpp"
#include "encoder/p_for_delta_encoder.hpp"
#include "quantizer/hilbert-curves-cpp/hilbert.hpp"

namespace {
uint8_t* dup(const std::vector<uint8_t>& v) {
  uint8_t* p = static_cast<uint8_t*>(std::malloc(v.size() ? v.size() : 1));
  if (!v.empty()) std::memcpy(p, v.data(), v.size());
  return p;
}
void out(const std::vector<uint8_t>& m, const std::vector<uint8_t>& d, uint8_t** meta,
         uint64_t* ml, uint8_t** data, uint64_t* dl) {
  *meta = dup(m);
  *ml = m.size();
  *data = dup(d);
  *dl = d.size();
}

// mode 0: one build; mode 1: the RLE counts double build
// (block_compressor_rle.hpp:290-303); mode 2: exp-golomb preset; mode 3:
// build_complete over 0..255 (fit_profile, block_compressor.hpp:364-370);
// mode 4: that meta loaded into a fresh encoder and written again
// (compress_with_profile, :362-363).
template <typename T>
int huff(const T* v, uint64_t n, int mode, uint8_t** meta, uint64_t* ml, uint8_t** data,
         uint64_t* dl) {
  std::vector<T> d(v, v + n);
  XnYZip::HuffmanEncoder<T> enc;
  if (mode == 3 || mode == 4) {  // build_complete(d, 255); 4: a fresh encoder loads its meta
    enc.build_complete(d, T(255));
    if (mode == 4) {
      XnYZip::HuffmanEncoder<T> e2;
      e2.load_meta(enc.get_meta());
      out(e2.get_meta(), e2.encode(d), meta, ml, data, dl);
      return 0;
    }
    out(enc.get_meta(), enc.encode(d), meta, ml, data, dl);
    return 0;
  }
  if (mode == 2) enc.set_use_exp_golomb(true);
  enc.build(d);
  auto m = enc.get_meta();
  auto e = enc.encode(d);
  if (mode == 1) {
    if (static_cast<double>(m.size()) > static_cast<double>(e.size()) * 0.25)
      enc.set_use_exp_golomb(true);
    enc.build(d);
    m = enc.get_meta();
    e = enc.encode(d);
  }
  out(m, e, meta, ml, data, dl);
  return 0;
}
}  // namespace

extern "C" {

int ref_bitpack_u64(const uint64_t* v, uint64_t n, uint8_t** meta, uint64_t* ml,
                    uint8_t** data, uint64_t* dl) {
  try {
    auto r = BitPacker<uint64_t>::pack(std::vector<uint64_t>(v, v + n));
    out(r.meta, r.data, meta, ml, data, dl);
    return 0;
  } catch (...) {
    return -1;
  }
}

int ref_gr_pack_u64(const uint64_t* v, uint64_t n, uint8_t** meta, uint64_t* ml,
                    uint8_t** data, uint64_t* dl) {
  auto r = GolombRiceCoder<uint64_t>::pack(std::vector<uint64_t>(v, v + n));
  out(r.meta, r.data, meta, ml, data, dl);
  return 0;
}

int ref_gr_pack_u8(const uint8_t* v, uint64_t n, uint8_t** meta, uint64_t* ml,
                   uint8_t** data, uint64_t* dl) {
  auto r = GolombRiceCoder<uint8_t>::pack(std::vector<uint8_t>(v, v + n));
  out(r.meta, r.data, meta, ml, data, dl);
  return 0;
}

int ref_pfd_u64(const uint64_t* v, uint64_t n, uint64_t* deltas, uint64_t* patches,
                uint64_t* npatches) {
  XnYZip::PForDeltaEncoder<uint64_t> e;
  auto r = e.encode(std::vector<uint64_t>(v, v + n));
  std::memcpy(deltas, r.deltas.data(), r.deltas.size() * 8);
  std::memcpy(patches, r.patches.data(), r.patches.size() * 8);
  *npatches = r.patches.size();
  return 0;
}

int ref_huffman(const void* v, uint64_t n, int sym_type, int mode, uint8_t** meta,
                uint64_t* ml, uint8_t** data, uint64_t* dl) {
  if (sym_type == 0) return huff(static_cast<const uint64_t*>(v), n, mode, meta, ml, data, dl);
  if (sym_type == 1) return huff(static_cast<const uint8_t*>(v), n, mode, meta, ml, data, dl);
  return huff(static_cast<const int*>(v), n, mode, meta, ml, data, dl);
}

uint64_t ref_morton_to_hilbert3(uint64_t m, unsigned it) { return mve::mortonToHilbert3(m, it); }
uint64_t ref_hilbert_to_morton3(uint64_t h, unsigned it) { return mve::hilbertToMorton3(h, it); }

void ref_free(void* p) { std::free(p); }

}  // extern "C"
