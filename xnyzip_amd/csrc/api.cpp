// api.cpp -- C ABI of libxnyz.so: context, the compression pipeline on one
// HIP stream, stream assembly and the zstd frame stage.
//
// Pipeline (reference: BlockCompressor[RLE]<float>::compress,
// source/composer/block_compressor.hpp:39-222, block_compressor_rle.hpp:38-321):
//   device  bbox minimum -> offsets -> quantize + range (+ RLE diagnostic)
//   host    sync 1: ranges decide block vs direct mode and the key layout
//   device  keygen -> LSD radix sort -> [RLE cells + counts] -> block runs
//           -> symbol histograms / first occurrences, patch count and sum
//   host    sync 2: Huffman / Exp-Golomb codebooks, Golomb-Rice k, widths
//   device  per-tile bit counts of the variable-length sections
//   host    sync 3: section sizes -> byte layout of the assembled stream
//   device  zero fill, bit packers write every data section in place
//   host    sync 4: copy back, write header / size fields / metas / trailer,
//           zstd level 3 (source/composer/utils.hpp:35-53)
#include <hip/hip_runtime.h>
#include <zstd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xnyz.h"
#include "codebook.hpp"
#include "common.hpp"
#include "kernels.hpp"

namespace xnyz {
HilbertSM build_hilbert_sm();
}

using namespace xnyz;

namespace {

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (p) check_hip(hipFree(p), "hipFree");
    p = nullptr;
    cap = 0;
    size_t want = bytes + bytes / 8 + 4096;
    if (hipMalloc(&p, want) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw Error(XNYZ_ENOMEM, "hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    cap = want;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct HBuf {  // pinned host memory
  void* p = nullptr;
  size_t cap = 0;
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (p) check_hip(hipHostFree(p), "hipHostFree");
    p = nullptr;
    cap = 0;
    size_t want = bytes + bytes / 8 + 4096;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw Error(XNYZ_ENOMEM, "hipHostMalloc of " + std::to_string(want) + " bytes failed");
    }
    cap = want;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Codebook tables uploaded to the device after sync 2.
struct LutBlock {
  HufEntry sym[kSymBins];
  HufEntry quad[kSymBins];
  HufEntry cnt[kCntBins];
};

u32 bitlen(u64 v) { return v ? 64u - (u32)__builtin_clzll(v) : 0u; }

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Per-kernel device time from HIP events recorded on the library's stream.
// The launch wrappers mark the end of each kernel; the host marks a fresh
// start after every synchronisation or copy it enqueues, so an interval
// between two consecutive marks covers exactly one kernel.
struct EventProfiler : Profiler {
  struct Mark {
    const char* name;
    hipEvent_t ev;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<std::string, std::pair<u64, double>>> acc;  // name -> calls, ms

  void mark(const char* name, hipStream_t s) override {
    if (used == pool.size()) {
      hipEvent_t e;
      check_hip(hipEventCreate(&e), "hipEventCreate");
      pool.push_back(e);
    }
    hipEvent_t e = pool[used++];
    check_hip(hipEventRecord(e, s), "hipEventRecord");
    marks.push_back(Mark{name, e});
  }
  void reset(hipStream_t s) { mark(nullptr, s); }
  void collect() {  // after a stream synchronisation
    for (size_t i = 1; i < marks.size(); ++i) {
      if (!marks[i].name) continue;
      float ms = 0;
      check_hip(hipEventElapsedTime(&ms, marks[i - 1].ev, marks[i].ev), "hipEventElapsedTime");
      auto it = std::find_if(acc.begin(), acc.end(),
                             [&](const auto& p) { return p.first == marks[i].name; });
      if (it == acc.end()) {
        acc.push_back({marks[i].name, {0, 0.0}});
        it = acc.end() - 1;
      }
      it->second.first += 1;
      it->second.second += ms;
    }
    marks.clear();
    used = 0;
  }
  void destroy() {
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }
};

}  // namespace

struct xnyz_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  HilbertSM hs{};
  DBuf ctl, keys[2], lo[2], pos, pos2, large, lut, bigv, bigc, out, arena, upload, tmp;
  DBuf idx[2], lo64[2];  // permutation payload
  HBuf h_ctl, h_out, h_lut;
  bool prof_on = false;
  EventProfiler prof;

  void sync() {
    check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (g_profiler) {
      prof.collect();
      prof.reset(stream);
    }
  }
  void copied() {  // after enqueueing a copy: the next kernel starts a new interval
    if (g_profiler) prof.reset(stream);
  }
  void fetch_ctl(size_t bytes) {
    check_hip(hipMemcpyAsync(h_ctl.p, ctl.p, bytes, hipMemcpyDeviceToHost, stream), "ctl D2H");
    sync();
  }
  const DevCtl& hc() const { return *h_ctl.as<DevCtl>(); }
};

namespace {

struct Section {  // one [u64 data_size][u64 meta_size][meta][data] record
  std::vector<uint8_t> meta;
  u64 data_bytes = 0;
  u64 at = 0;       // byte offset of the size fields
  u64 data_at = 0;  // byte offset of the data
};

void put_bytes(uint8_t* dst, const void* src, size_t n) { std::memcpy(dst, src, n); }

// zstd level 3 over the assembled stream (source/composer/utils.hpp:35-53).
void zstd_frame(const uint8_t* src, u64 len, int mode, int workers, xnyz_buf* out) {
  const size_t bound = ZSTD_compressBound(len);
  uint8_t* dst = static_cast<uint8_t*>(std::malloc(bound ? bound : 1));
  if (!dst) throw Error(XNYZ_ENOMEM, "malloc of the frame buffer failed");
  size_t r;
  if (mode == XNYZ_FRAME_ZSTD3) {
    r = ZSTD_compress(dst, bound, src, len, 3);
  } else {
    ZSTD_CCtx* cc = ZSTD_createCCtx();
    if (!cc) {
      std::free(dst);
      throw Error(XNYZ_ENOMEM, "ZSTD_createCCtx failed");
    }
    int nw = workers > 0 ? workers : (int)std::max(1u, std::thread::hardware_concurrency());
    ZSTD_CCtx_setParameter(cc, ZSTD_c_compressionLevel, 3);
    ZSTD_CCtx_setParameter(cc, ZSTD_c_contentSizeFlag, 1);
    size_t e = ZSTD_CCtx_setParameter(cc, ZSTD_c_nbWorkers, nw);
    if (ZSTD_isError(e)) {
      ZSTD_freeCCtx(cc);
      std::free(dst);
      throw Error(XNYZ_EZSTD, std::string("zstd workers unavailable: ") + ZSTD_getErrorName(e));
    }
    r = ZSTD_compress2(cc, dst, bound, src, len);
    ZSTD_freeCCtx(cc);
  }
  if (ZSTD_isError(r)) {
    std::free(dst);
    throw Error(XNYZ_EZSTD, ZSTD_getErrorName(r));
  }
  out->data = dst;
  out->len = r;
}

void validate(const xnyz_opts* o, u64 n) {
  if (!o) throw Error(XNYZ_EINVAL, "opts is null");
  if (o->quantizer != XNYZ_CUBE && o->quantizer != XNYZ_TRUNCATED_OCTAHEDRON)
    throw Error(XNYZ_EINVAL, "quantizer must be XNYZ_CUBE or XNYZ_TRUNCATED_OCTAHEDRON");
  if (!(o->l2_bound > 0.0f) || !std::isfinite(o->l2_bound))
    throw Error(XNYZ_EINVAL, "l2_bound must be a positive finite float");
  if (o->frame_mode < XNYZ_FRAME_ZSTD3 || o->frame_mode > XNYZ_FRAME_NONE)
    throw Error(XNYZ_EINVAL, "bad frame_mode");
  if (n == 0) throw Error(XNYZ_EINVAL, "empty input (the reference indexes an empty vector)");
  if (n >= 0xFFFFFFFFull) throw Error(XNYZ_EUNSUPPORTED, "more than 2^32-2 points in one call");
}

QuantParams quant_params(const xnyz_opts* o) {
  QuantParams qp{};
  qp.to = o->quantizer == XNYZ_TRUNCATED_OCTAHEDRON;
  const float two_b = 2 * o->l2_bound;  // int * float: float
  const double root = std::sqrt(qp.to ? 5.0 : 3.0);
  qp.div = static_cast<float>(static_cast<double>(two_b) / root);
  qp.b = o->l2_bound;
  qp.rec_div = static_cast<float>(root);
  return qp;
}

std::vector<SymCount> sym_counts(const u32* hist, const u32* first, u32 bins) {
  std::vector<SymCount> v;
  for (u32 s = 0; s < bins; ++s)
    if (hist[s]) v.push_back(SymCount{s, hist[s], first[s]});
  return v;
}

template <typename T>
void fill_lut(const Codebook<T>& cb, HufEntry* lut, u32 bins) {
  for (u32 s = 0; s < bins; ++s) lut[s] = HufEntry{0, 0};
  for (auto& kv : cb.code_table) {
    const u64 s = static_cast<u64>(kv.first);
    if (s < bins) lut[s] = HufEntry{kv.second.bits, kv.second.length};
  }
}

struct ProfGuard {  // installs the context's profiler for one call
  xnyz_ctx* c;
  explicit ProfGuard(xnyz_ctx* ctx) : c(ctx) {
    if (c->prof_on) {
      g_profiler = &c->prof;
      c->prof.reset(c->stream);
    }
  }
  ~ProfGuard() {
    if (g_profiler == &c->prof) {
      c->prof.marks.clear();
      c->prof.used = 0;
      g_profiler = nullptr;
    }
  }
};

int compress_impl(xnyz_ctx* ctx, const float* d_xyz, u64 n64, const xnyz_opts* o,
                  xnyz_buf* out_frame, xnyz_buf* out_pre, uint32_t* perm, xnyz_stats* st,
                  double t_start) {
  validate(o, n64);
  const u32 N = (u32)n64;
  const bool rle = o->rle != 0;
  const bool hil = o->is_hilbert != 0;
  hipStream_t s = ctx->stream;
  ProfGuard prof_guard(ctx);
  const QuantParams qp = quant_params(o);

  ctx->ctl.ensure(sizeof(DevCtl));
  ctx->h_ctl.ensure(sizeof(DevCtl));
  ctx->arena.ensure((1u << 20) + (size_t)N / 2 + 65536);
  Arena ar{ctx->arena.as<char>(), ctx->arena.cap, 0};
  DevCtl* dctl = ctx->ctl.as<DevCtl>();

  // ---- stage 1: bbox minimum, quantize, range scan ------------------------
  launch_init_ctl(dctl, s);
  launch_bbox(d_xyz, N, dctl, s);
  launch_quant_stats(d_xyz, N, qp, rle, dctl, s);
  ctx->fetch_ctl(offsetof(DevCtl, hist_sym));
  const DevCtl& h = ctx->hc();
  const int rx = h.range[0], ry = h.range[1], rz = h.range[2];
  const u32 W = h.wrapped;
  const float off[3] = {h.offset[0], h.offset[1], h.offset[2]};
  const int max_range = std::max({rx, ry, rz});
  const bool direct = max_range < o->direct_threshold;
  double max_l2 = 0, mse = 0;
  if (rle) {
    double m2;
    std::memcpy(&m2, &h.max_err2_bits, 8);
    max_l2 = std::sqrt(m2);
    mse = h.sum_err2 / N;
  }

  // ---- key layout ------------------------------------------------------------
  KeyParams kp{};
  kp.direct = direct;
  kp.hilbert = hil;
  u32 key_bits = 0, hi_bits = 0;
  if (direct) {
    if (max_range >= (1 << 21) - 1)
      throw Error(XNYZ_EUNSUPPORTED, "direct mode with a quantized range >= 2^21-1");
    u32 L = bitlen((u64)max_range);
    kp.L = L ? L : 1;
    key_bits = W ? 64 : 3 * kp.L;
  } else {
    const u64 nx = (u64)rx / 128 + 1, ny = (u64)ry / 128 + 1, nz = (u64)rz / 128 + 1;
    const unsigned __int128 nb = (unsigned __int128)nx * ny * nz;
    if (nb > ((unsigned __int128)1 << 57))
      throw Error(XNYZ_EUNSUPPORTED, "block grid larger than 2^57 blocks");
    kp.nx = nx;
    kp.nxny = nx * ny;
    kp.C = ((1ull << 57) - 1) * kp.nxny;  // wraps mod 2^64 as the reference's size_t
    kp.idb = bitlen((u64)nb - 1);
    if (W && (u64)nb > kp.C)
      throw Error(XNYZ_EUNSUPPORTED, "wrapped block ids overlap the regular ids");
    hi_bits = kp.idb + (W ? 1 : 0);
    key_bits = hi_bits + 21;
    kp.wide = key_bits > 64;
  }

  // ---- stage 2: keys, sort, runs, statistics -----------------------------------
  for (int i = 0; i < 2; ++i) ctx->keys[i].ensure((size_t)N * 8);
  if (kp.wide)
    for (int i = 0; i < 2; ++i) ctx->lo[i].ensure((size_t)N * 4);
  ctx->pos.ensure((size_t)N * 4 + 16);
  ctx->pos2.ensure((size_t)N * 4 + 16);
  const u32 large_cap = N / kCntBins + 2;
  ctx->large.ensure((size_t)large_cap * 8);
  u64* ka = ctx->keys[0].as<u64>();
  u64* kb = ctx->keys[1].as<u64>();
  u32* la = kp.wide ? ctx->lo[0].as<u32>() : nullptr;
  u32* lb = kp.wide ? ctx->lo[1].as<u32>() : nullptr;

  // The reference reorders the caller's points by the sort (ords,
  // block_compressor.hpp:141-144); here that order is an index payload of the
  // sort, returned through `perm` (non-RLE only, as the reference's RLE
  // reorder indexes N points with U indices).
  const bool want_perm = perm && !rle;
  u32* ia = nullptr;
  u32* ib = nullptr;
  u64* l64a = nullptr;
  u64* l64b = nullptr;
  if (want_perm) {
    for (int i = 0; i < 2; ++i) ctx->idx[i].ensure((size_t)N * 4);
    ia = ctx->idx[0].as<u32>();
    ib = ctx->idx[1].as<u32>();
    if (kp.wide) {
      for (int i = 0; i < 2; ++i) ctx->lo64[i].ensure((size_t)N * 8);
      l64a = ctx->lo64[0].as<u64>();
      l64b = ctx->lo64[1].as<u64>();
    }
  }
  launch_keygen(d_xyz, N, qp, kp, ctx->hs, dctl, ka, l64a ? nullptr : la, ia, l64a, s);
  int cur;
  const u32* perm_dev = nullptr;
  if (want_perm && kp.wide) {
    cur = radix_sort_wide_v(ka, kb, l64a, l64b, N, hi_bits, ar, s);
    launch_split_lo64(cur ? l64b : l64a, ctx->lo[cur].as<u32>(), ia, N, s);
    perm_dev = ia;
  } else if (want_perm) {
    cur = radix_sort_narrow_v(ka, kb, ia, ib, N, key_bits, ar, s);
    perm_dev = cur ? ib : ia;
  } else if (kp.wide) {
    cur = radix_sort_wide(ka, kb, la, lb, N, 21, hi_bits, ar, s);
  } else {
    cur = radix_sort_narrow(ka, kb, N, key_bits, ar, s);
  }
  if (perm_dev) {
    check_hip(hipMemcpyAsync(perm, perm_dev, (size_t)N * 4, hipMemcpyDeviceToHost, s),
              "perm D2H");
    ctx->copied();
  }
  KeyView kv{ctx->keys[cur].as<u64>(), kp.wide ? ctx->lo[cur].as<u32>() : nullptr, kp.idb, kp.C};
  KeyView kfree{ctx->keys[cur ^ 1].as<u64>(), kp.wide ? ctx->lo[cur ^ 1].as<u32>() : nullptr,
                kp.idb, kp.C};
  const u32* n_dev = nullptr;  // element count after RLE (device)
  u32* cnt = ctx->pos2.as<u32>();
  if (rle) {
    runs_unique(kv, kp.wide, N, ctx->pos.as<u32>(), &dctl->n_unique, ar, s);
    launch_rle_finish(kv, kp.wide, ctx->pos.as<u32>(), &dctl->n_unique, N,
                      ctx->keys[cur ^ 1].as<u64>(), kp.wide ? ctx->lo[cur ^ 1].as<u32>() : nullptr,
                      cnt, dctl, ctx->large.as<u32>(), large_cap, s);
    kv = kfree;
    n_dev = &dctl->n_unique;
  }
  if (direct) {
    launch_stats_direct(kv.a, N, n_dev, dctl, s);
  } else {
    runs_blocks(kv, kp.wide, N, n_dev, ctx->pos.as<u32>(), &dctl->n_blocks, ar, s);
    launch_stats_block(kv, kp.wide, ctx->pos.as<u32>(), &dctl->n_blocks, N, n_dev, dctl, s);
  }
  ctx->fetch_ctl(sizeof(DevCtl));
  const u32 n = rle ? h.n_unique : N;  // elements of the coded streams
  const u32 B = direct ? 0 : h.n_blocks;
  if (h.n_qpatches) throw Error(XNYZ_EINTERNAL, "quads delta above 254");
  if (rle && h.n_large > large_cap) throw Error(XNYZ_EINTERNAL, "large-count list overflow");

  // ---- host: codebooks --------------------------------------------------------
  ctx->h_lut.ensure(sizeof(LutBlock));
  LutBlock& L = *ctx->h_lut.as<LutBlock>();
  Codebook<uint64_t> cb_sym;
  cb_sym.build(sym_counts(h.hist_sym, h.first_sym, kSymBins), n);
  fill_lut(cb_sym, L.sym, kSymBins);
  Codebook<uint8_t> cb_quad;
  if (!direct) {
    cb_quad.build(sym_counts(h.hist_quad, h.first_quad, kSymBins), n);
    fill_lut(cb_quad, L.quad, kSymBins);
  }
  const u64 P = h.n_patches;
  const uint8_t grk = golomb_rice_k(P, h.patch_sum, false);
  Codebook<int> cb_cnt;
  std::vector<u32> big_vals;
  std::vector<HufEntry> big_codes;
  if (rle) {
    std::vector<SymCount> cs = sym_counts(h.hist_cnt, h.first_cnt, kCntBins);
    if (h.n_large) {
      std::vector<u32> raw((size_t)h.n_large * 2);
      check_hip(hipMemcpyAsync(raw.data(), ctx->large.p, raw.size() * 4, hipMemcpyDeviceToHost, s),
                "large D2H");
      ctx->sync();
      std::vector<std::pair<u32, u32>> lv;
      for (u32 i = 0; i < h.n_large; ++i) lv.emplace_back(raw[2 * i], raw[2 * i + 1]);
      std::sort(lv.begin(), lv.end());
      for (size_t i = 0; i < lv.size();) {
        size_t j = i;
        while (j < lv.size() && lv[j].first == lv[i].first) ++j;
        cs.push_back(SymCount{lv[i].first, (u64)(j - i), lv[i].second});
        big_vals.push_back(lv[i].first);
        i = j;
      }
    }
    // block_compressor_rle.hpp:290-303: build, compare the meta with a
    // quarter of the data, then build again on the same object.
    cb_cnt.build(cs, n);
    const double meta1 = (double)cb_cnt.meta().size();
    const double data1 = (double)((cb_cnt.data_bits() + 7) / 8);
    if (meta1 > data1 * 0.25) cb_cnt.use_eg = true;
    cb_cnt.build(cs, n);
    if (!cb_cnt.use_eg) {
      fill_lut(cb_cnt, L.cnt, kCntBins);
      for (u32 v : big_vals) {
        const auto& c = cb_cnt.code_table.at((int)v);
        big_codes.push_back(HufEntry{c.bits, c.length});
      }
    }
  }
  if (cb_sym.max_length() > 56 || cb_quad.max_length() > 56 || cb_cnt.max_length() > 56)
    throw Error(XNYZ_EUNSUPPORTED, "Huffman code longer than 56 bits");
  ctx->lut.ensure(sizeof(LutBlock));
  check_hip(hipMemcpyAsync(ctx->lut.p, &L, sizeof(LutBlock), hipMemcpyHostToDevice, s), "lut H2D");
  if (!big_vals.empty()) {
    ctx->bigv.ensure(big_vals.size() * 4);
    ctx->bigc.ensure(big_codes.size() * sizeof(HufEntry) + 8);
    check_hip(hipMemcpyAsync(ctx->bigv.p, big_vals.data(), big_vals.size() * 4,
                             hipMemcpyHostToDevice, s), "big H2D");
    if (!big_codes.empty())
      check_hip(hipMemcpyAsync(ctx->bigc.p, big_codes.data(), big_codes.size() * sizeof(HufEntry),
                               hipMemcpyHostToDevice, s), "big H2D");
    // pageable sources: finish before the vectors go out of scope
    ctx->sync();
  }
  ctx->copied();
  const LutBlock* dl = ctx->lut.as<LutBlock>();

  // ---- stage 3: variable-length section sizes ---------------------------------
  CoderArgs ca{};
  ca.kv = kv;
  ca.wide = kp.wide;
  ca.sfc = kv.a;
  ca.cnt = cnt;
  VlPlan plan_gr{}, plan_sym{}, plan_quad{}, plan_cnt{};
  const CoderKind k_gr = direct ? kCoderGRDirect : kCoderGRBlock;
  const CoderKind k_sym = direct ? kCoderHufDirect : kCoderHufBlock;
  const CoderKind k_cnt = cb_cnt.use_eg ? kCoderCntEG : kCoderCntHuf;
  CoderArgs a_gr = ca, a_sym = ca, a_quad = ca, a_cnt = ca;
  a_gr.k = grk;
  a_sym.lut = dl->sym;
  a_quad.lut = dl->quad;
  a_cnt.lut = dl->cnt;
  a_cnt.k = cb_cnt.eg_k;
  a_cnt.big_vals = ctx->bigv.as<u32>();
  a_cnt.big_codes = ctx->bigc.as<HufEntry>();
  a_cnt.n_big = (u32)big_vals.size();
  if (P) {
    plan_gr = vl_plan(n, ar);
    vl_count(k_gr, a_gr, n, plan_gr, &dctl->sec_bits[kSecGR], ar, s);
  }
  plan_sym = vl_plan(n, ar);
  vl_count(k_sym, a_sym, n, plan_sym, &dctl->sec_bits[kSecHuf], ar, s);
  if (!direct) {
    plan_quad = vl_plan(n, ar);
    vl_count(kCoderQuadHuf, a_quad, n, plan_quad, &dctl->sec_bits[kSecQuadHuf], ar, s);
  }
  if (rle) {
    plan_cnt = vl_plan(n, ar);
    vl_count(k_cnt, a_cnt, n, plan_cnt, &dctl->sec_bits[kSecCnt], ar, s);
  }
  check_hip(hipMemcpyAsync(ctx->h_ctl.as<DevCtl>()->sec_bits, &dctl->sec_bits,
                           sizeof(h.sec_bits), hipMemcpyDeviceToHost, s), "sec D2H");
  ctx->sync();
  const u64 gr_bits = P ? h.sec_bits[kSecGR] : 0;
  if (h.sec_bits[kSecHuf] != cb_sym.data_bits() ||
      (!direct && h.sec_bits[kSecQuadHuf] != cb_quad.data_bits()) ||
      (rle && h.sec_bits[kSecCnt] != cb_cnt.data_bits()))
    throw Error(XNYZ_EINTERNAL, "device and host disagree on a section size");

  // ---- layout (Appendix: header, sections, trailer) ---------------------------
  std::vector<Section> secs;
  auto add = [&](std::vector<uint8_t> meta, u64 data_bytes) {
    secs.push_back(Section{std::move(meta), data_bytes, 0, 0});
    return secs.size() - 1;
  };
  auto le_meta = [](u32 count, uint8_t last) {  // BitPacker / GR meta: u32 LE + u8
    return std::vector<uint8_t>{uint8_t(count), uint8_t(count >> 8), uint8_t(count >> 16),
                                uint8_t(count >> 24), last};
  };
  size_t s1 = 0, s2 = 0, s3, s4, s6 = 0, s7 = 0;
  u32 w1 = 0, w2 = 0;
  if (!direct) {
    w1 = h.max_blkd ? bitlen(h.max_blkd) : 1;
    w2 = h.max_cntd ? bitlen(h.max_cntd) : 1;
    s1 = add(le_meta(B, (uint8_t)w1), ((u64)B * w1 + 7) / 8);
    s2 = add(le_meta(B, (uint8_t)w2), ((u64)B * w2 + 7) / 8);
  }
  s3 = add(P ? le_meta((u32)P, grk) : std::vector<uint8_t>{}, (gr_bits + 7) / 8);
  s4 = add(cb_sym.meta(), (cb_sym.data_bits() + 7) / 8);
  if (!direct) {
    add({}, 0);  // quads patches: none (deltas are at most 7)
    s6 = add(cb_quad.meta(), (cb_quad.data_bits() + 7) / 8);
  }
  if (rle) s7 = add(cb_cnt.meta(), (cb_cnt.data_bits() + 7) / 8);
  u64 at = 30;
  for (auto& sc : secs) {
    sc.at = at;
    sc.data_at = at + 16 + sc.meta.size();
    at = sc.data_at + sc.data_bytes;
  }
  const u64 total = at + 12;

  // ---- stage 4: write the data sections in place ------------------------------
  ctx->out.ensure(total + 64);
  u32* ow = ctx->out.as<u32>();
  launch_fill_u32(ow, (total + 64) / 4, 0, s);
  if (!direct) {
    pack_block_deltas(0, kv, kp.wide, ctx->pos.as<u32>(), B, n, w1, secs[s1].data_at, ow, s);
    pack_block_deltas(1, kv, kp.wide, ctx->pos.as<u32>(), B, n, w2, secs[s2].data_at, ow, s);
  }
  if (P) vl_write(k_gr, a_gr, n, plan_gr, secs[s3].data_at, ow, s);
  vl_write(k_sym, a_sym, n, plan_sym, secs[s4].data_at, ow, s);
  if (!direct) vl_write(kCoderQuadHuf, a_quad, n, plan_quad, secs[s6].data_at, ow, s);
  if (rle) vl_write(k_cnt, a_cnt, n, plan_cnt, secs[s7].data_at, ow, s);
  ctx->h_out.ensure(total);
  uint8_t* hb = ctx->h_out.as<uint8_t>();
  check_hip(hipMemcpyAsync(hb, ow, total, hipMemcpyDeviceToHost, s), "stream D2H");
  ctx->sync();

  // header (block_compressor.hpp:61-87): u64 points, u64 params.size() = 0,
  // i32 ranges, bool curve, bool direct; per section u64 data size, u64 meta
  // size, meta; trailer: the float offset.
  const u64 np = n;
  const u64 zero = 0;
  put_bytes(hb, &np, 8);
  put_bytes(hb + 8, &zero, 8);
  const int32_t rr[3] = {rx, ry, rz};
  put_bytes(hb + 16, rr, 12);
  hb[28] = hil ? 1 : 0;
  hb[29] = direct ? 1 : 0;
  for (auto& sc : secs) {
    const u64 ds = sc.data_bytes, ms = sc.meta.size();
    put_bytes(hb + sc.at, &ds, 8);
    put_bytes(hb + sc.at + 8, &ms, 8);
    if (ms) put_bytes(hb + sc.at + 16, sc.meta.data(), ms);
  }
  put_bytes(hb + total - 12, off, 12);
  const double t_dev = now_ms();

  if (out_pre) {
    out_pre->data = static_cast<uint8_t*>(std::malloc(total));
    if (!out_pre->data) throw Error(XNYZ_ENOMEM, "malloc failed");
    std::memcpy(out_pre->data, hb, total);
    out_pre->len = total;
  }
  u64 frame_len = 0;
  if (out_frame) {
    if (o->frame_mode == XNYZ_FRAME_NONE) {
      out_frame->data = static_cast<uint8_t*>(std::malloc(total));
      if (!out_frame->data) throw Error(XNYZ_ENOMEM, "malloc failed");
      std::memcpy(out_frame->data, hb, total);
      out_frame->len = total;
    } else {
      zstd_frame(hb, total, o->frame_mode, o->zstd_workers, out_frame);
    }
    frame_len = out_frame->len;
  }
  const double t_end = now_ms();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->input_points = N;
    st->num_points = n;
    st->range[0] = rx;
    st->range[1] = ry;
    st->range[2] = rz;
    st->is_direct = direct;
    st->blocks = B;
    st->patches = P;
    st->quad_patches = 0;
    st->wrapped_points = W;
    st->counts_exp_golomb = rle ? cb_cnt.use_eg : 0;
    st->counts_eg_k = rle ? cb_cnt.eg_k : 0;
    st->entropy_repos = cb_sym.entropy;
    st->entropy_quads = cb_quad.entropy;
    st->entropy_counts = cb_cnt.entropy;
    st->max_l2 = max_l2;
    st->mse = mse;
    st->prezstd_bytes = total;
    st->frame_bytes = frame_len;
    st->ms_device = t_dev - t_start;
    st->ms_zstd = t_end - t_dev;
    st->ms_total = t_end - t_start;
  }
  return XNYZ_OK;
}

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    return set_error(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return set_error(XNYZ_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return set_error(XNYZ_EINTERNAL, e.what());
  }
}

void use_device(xnyz_ctx* ctx) { check_hip(hipSetDevice(ctx->device), "hipSetDevice"); }

}  // namespace

extern "C" {

int xnyz_abi_version(void) { return XNYZ_ABI_VERSION; }

const char* xnyz_last_error(void) { return g_last_error.c_str(); }

int xnyz_create(int device, xnyz_ctx** out) {
  return guarded([&] {
    if (!out) throw Error(XNYZ_EINVAL, "out is null");
    int count = 0;
    check_hip(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count) throw Error(XNYZ_EINVAL, "no such HIP device");
    auto* c = new xnyz_ctx();
    c->device = device;
    try {
      use_device(c);
      check_hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
      c->hs = build_hilbert_sm();
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
    return XNYZ_OK;
  });
}

int xnyz_destroy(xnyz_ctx* ctx) {
  if (!ctx) return XNYZ_OK;
  return guarded([&] {
    use_device(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    for (DBuf* b : {&ctx->ctl, &ctx->keys[0], &ctx->keys[1], &ctx->lo[0], &ctx->lo[1],
                    &ctx->pos, &ctx->pos2, &ctx->large, &ctx->lut, &ctx->bigv, &ctx->bigc,
                    &ctx->out, &ctx->arena, &ctx->upload, &ctx->tmp, &ctx->idx[0],
                    &ctx->idx[1], &ctx->lo64[0], &ctx->lo64[1]})
      b->release();
    ctx->h_ctl.release();
    ctx->h_out.release();
    ctx->h_lut.release();
    ctx->prof.destroy();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return XNYZ_OK;
  });
}

int xnyz_compress_device(xnyz_ctx* ctx, const float* d_xyz, uint64_t n, const xnyz_opts* opts,
                         xnyz_buf* out_frame, xnyz_buf* out_prezstd, uint32_t* perm,
                         xnyz_stats* stats) {
  return guarded([&] {
    if (!ctx || !d_xyz) throw Error(XNYZ_EINVAL, "null context or input");
    use_device(ctx);
    return compress_impl(ctx, d_xyz, n, opts, out_frame, out_prezstd, perm, stats, now_ms());
  });
}

int xnyz_compress(xnyz_ctx* ctx, const float* xyz, uint64_t n, const xnyz_opts* opts,
                  xnyz_buf* out_frame, xnyz_buf* out_prezstd, uint32_t* perm,
                  xnyz_stats* stats) {
  return guarded([&] {
    if (!ctx || !xyz) throw Error(XNYZ_EINVAL, "null context or input");
    validate(opts, n);
    use_device(ctx);
    const double t0 = now_ms();
    ctx->upload.ensure((size_t)n * 12);
    check_hip(hipMemcpyAsync(ctx->upload.p, xyz, (size_t)n * 12, hipMemcpyHostToDevice,
                             ctx->stream), "points H2D");
    return compress_impl(ctx, ctx->upload.as<float>(), n, opts, out_frame, out_prezstd, perm,
                         stats, t0);
  });
}

int xnyz_quantize_device(xnyz_ctx* ctx, const float* d_xyz, uint64_t n, int32_t quantizer,
                         float l2_bound, int32_t* nodes, float* offset) {
  return guarded([&] {
    if (!ctx || !d_xyz || !nodes) throw Error(XNYZ_EINVAL, "null argument");
    xnyz_opts o{};
    o.quantizer = quantizer;
    o.l2_bound = l2_bound;
    validate(&o, n);
    use_device(ctx);
    hipStream_t s = ctx->stream;
    ctx->ctl.ensure(sizeof(DevCtl));
    ctx->h_ctl.ensure(sizeof(DevCtl));
    ctx->tmp.ensure((size_t)n * 12);
    DevCtl* dctl = ctx->ctl.as<DevCtl>();
    const QuantParams qp = quant_params(&o);
    launch_init_ctl(dctl, s);
    launch_bbox(d_xyz, (u32)n, dctl, s);
    launch_quantize_nodes(d_xyz, (u32)n, qp, dctl, ctx->tmp.as<int32_t>(), s);
    check_hip(hipMemcpyAsync(nodes, ctx->tmp.p, (size_t)n * 12, hipMemcpyDeviceToHost, s),
              "nodes D2H");
    ctx->fetch_ctl(offsetof(DevCtl, hist_sym));
    if (offset)
      for (int a = 0; a < 3; ++a) offset[a] = ctx->hc().offset[a];
    return XNYZ_OK;
  });
}

int xnyz_morton_to_hilbert(xnyz_ctx* ctx, const uint64_t* in, uint64_t n, uint32_t iteration,
                           uint64_t* out) {
  return guarded([&] {
    if (!ctx || !in || !out) throw Error(XNYZ_EINVAL, "null argument");
    if (iteration > 21) throw Error(XNYZ_EUNSUPPORTED, "iteration above 21");
    if (n == 0) return XNYZ_OK;
    if (n >= 0xFFFFFFFFull) throw Error(XNYZ_EUNSUPPORTED, "too many codes");
    use_device(ctx);
    hipStream_t s = ctx->stream;
    ctx->tmp.ensure((size_t)n * 16);
    u64* d = ctx->tmp.as<u64>();
    check_hip(hipMemcpyAsync(d, in, n * 8, hipMemcpyHostToDevice, s), "codes H2D");
    launch_m2h(d, (u32)n, iteration, ctx->hs, d + n, s);
    check_hip(hipMemcpyAsync(out, d + n, n * 8, hipMemcpyDeviceToHost, s), "codes D2H");
    ctx->sync();
    return XNYZ_OK;
  });
}

int xnyz_profile_enable(xnyz_ctx* ctx, int on) {
  return guarded([&] {
    if (!ctx) throw Error(XNYZ_EINVAL, "null context");
    ctx->prof_on = on != 0;
    ctx->prof.acc.clear();
    return XNYZ_OK;
  });
}

int xnyz_profile_report(xnyz_ctx* ctx, char* buf, uint64_t cap, uint64_t* needed) {
  return guarded([&] {
    if (!ctx) throw Error(XNYZ_EINVAL, "null context");
    std::string js = "{";
    char line[256];
    for (size_t i = 0; i < ctx->prof.acc.size(); ++i) {
      const auto& a = ctx->prof.acc[i];
      std::snprintf(line, sizeof(line), "%s\"%s\":{\"calls\":%llu,\"ms\":%.6f}", i ? "," : "",
                    a.first.c_str(), (unsigned long long)a.second.first, a.second.second);
      js += line;
    }
    js += "}";
    if (needed) *needed = js.size() + 1;
    if (buf && cap) {
      const size_t n = std::min<size_t>(cap - 1, js.size());
      std::memcpy(buf, js.data(), n);
      buf[n] = 0;
    }
    return XNYZ_OK;
  });
}

void xnyz_free(xnyz_buf* buf) {
  if (!buf) return;
  std::free(buf->data);
  buf->data = nullptr;
  buf->len = 0;
}

}  // extern "C"
