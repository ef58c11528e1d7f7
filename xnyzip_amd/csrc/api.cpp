// api.cpp -- C ABI of libxnyz.so: context, the segmented compression pipeline
// on one HIP stream, stream assembly and the zstd frame stage.
//
// One call compresses S clouds ("segments"; S = 1 for xnyz_compress) with one
// launch set per stage.  Pipeline (reference: BlockCompressor[RLE]<float>::
// compress, source/composer/block_compressor.hpp:39-222,
// block_compressor_rle.hpp:38-321):
//   device  bbox minimum -> offsets -> quantize + range (+ RLE diagnostic)
//   host    sync 1: ranges decide block vs direct mode and the key layout
//   device  keygen -> segmented LSD radix sort -> RLE cells + counts ->
//           block runs -> symbol histograms / first occurrences, patches
//   host    sync 2: Huffman / Exp-Golomb codebooks, Golomb-Rice k, widths
//   device  per-tile bit counts of every section -> per-segment sizes
//   host    sync 3: byte layout of every segment's assembled stream
//   device  zero fill, every section written in place
//   host    sync 4: copy back, header / size fields / metas / trailer,
//           zstd level 3 (source/composer/utils.hpp:35-53)
#include <hip/hip_runtime.h>
#define ZSTD_STATIC_LINKING_ONLY  // block-level API (ZSTD_compressBlock)
#include <zstd.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xnyz.h"
#include "codebook.hpp"
#include "common.hpp"
#include "decode_dev.hpp"
#include "kernels.hpp"
#include "mlp.hpp"
#include "modes.hpp"
#include "zframe.hpp"
#include "zframe_dev.hpp"

// Exported by libzstd 1.4 (zstd_compress.c) though absent from its public
// header: makes the next block start without repeat offsets, so a chunk
// compressed on its own may follow any other chunk in one frame (zstd's own
// multithreaded compressor does the same between jobs).
extern "C" void ZSTD_invalidateRepCodes(ZSTD_CCtx* cctx);

namespace xnyz {
HilbertSM build_hilbert_sm();
int decompress_impl(const uint8_t* frame, uint64_t len, const xnyz_opts* o, xnyz_buf* out);
// container.cpp
std::vector<xnyz_chunk> read_manifest(const char* path);
void write_container(const char* part_path, const char* manifest_path, const uint8_t* out,
                     const uint64_t* out_offsets, uint32_t count, xnyz_chunk* chunks);
std::vector<uint8_t> read_part(const char* part_path, const xnyz_chunk& c);
}  // namespace xnyz

using namespace xnyz;

namespace {

// Host staging of a no-allocation call (per-segment parameters, statistics
// and codebook tables for up to XNYZ_NOALLOC_MAX_CLOUDS clouds).
constexpr size_t kNoAllocHostBytes = size_t(4) << 20;

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// Fixed regions that DBuf / HBuf carve from instead of allocating, while a
// no-allocation call runs on this thread (xnyz_compress_batch_device_noalloc):
// the device scratch array and the context's host staging block.  Each
// buffer the call asks for gets a fresh 256-byte aligned slice; running out
// is an error, not an allocation.
struct Bump {
  char* dev;
  size_t dcap, doff;
  char* host;
  size_t hcap, hoff;
  static void* take(char* base, size_t cap, size_t& off, size_t bytes, const char* name,
                    const char* what) {
    const size_t b = (bytes + 255) & ~size_t(255);
    if (off + b > cap)
      throw Error(XNYZ_EUNSUPPORTED, std::string("no-allocation ") + what + " of " +
                                         std::to_string(cap >> 20) + " MiB exhausted by " +
                                         name + " (" + std::to_string(bytes) + " bytes)");
    void* p = base + off;
    off += b;
    return p;
  }
};
thread_local Bump* g_bump = nullptr;

struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
  const char* name = "device buffer";
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (g_bump) {
      p = Bump::take(g_bump->dev, g_bump->dcap, g_bump->doff, bytes, name, "device scratch");
      cap = bytes;
      return;
    }
    if (p) check_hip(hipFree(p), "hipFree");
    p = nullptr;
    cap = 0;
    size_t want = bytes + bytes / 8 + 4096;
    if (hipMalloc(&p, want) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw Error(XNYZ_ENOMEM, std::string("hipMalloc of ") + std::to_string(want) +
                                   " bytes for " + name + " (" + std::to_string(bytes) +
                                   " requested) failed");
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
  const char* name = "host buffer";
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (g_bump) {
      p = Bump::take(g_bump->host, g_bump->hcap, g_bump->hoff, bytes, name, "host staging");
      cap = bytes;
      return;
    }
    if (p) check_hip(hipHostFree(p), "hipHostFree");
    p = nullptr;
    cap = 0;
    size_t want = bytes + bytes / 8 + 4096;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw Error(XNYZ_ENOMEM, std::string("hipHostMalloc of ") + std::to_string(want) +
                                   " bytes for " + name + " (" + std::to_string(bytes) +
                                   " requested) failed");
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

// Pinned host buffers handed out by xnyz_host_alloc: base -> size.  Output
// inside one of them takes device-built frame bytes straight from the DMA
// engine, with no staging copy.
std::mutex g_pinned_mu;
std::map<uintptr_t, u64> g_pinned;

bool pinned_range(const void* p, u64 n) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  std::lock_guard<std::mutex> lk(g_pinned_mu);
  auto it = g_pinned.upper_bound(a);
  if (it == g_pinned.begin()) return false;
  --it;
  return a - it->first <= it->second && n <= it->second - (a - it->first);
}

void* pinned_alloc(u64 bytes) {
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    throw Error(XNYZ_ENOMEM, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  std::lock_guard<std::mutex> lk(g_pinned_mu);
  g_pinned[reinterpret_cast<uintptr_t>(p)] = bytes;
  return p;
}

void pinned_free(void* p) {
  {
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    if (!g_pinned.erase(reinterpret_cast<uintptr_t>(p)))
      throw Error(XNYZ_EINVAL, "pointer did not come from xnyz_host_alloc");
  }
  check_hip(hipHostFree(p), "hipHostFree");
}

u32 bitlen(u64 v) { return v ? 64u - (u32)__builtin_clzll(v) : 0u; }

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int host_threads(int want) {
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  return want > 0 ? std::min(want, 256) : hw;
}

// Runs f(i) for i in [0, n) on up to `threads` host threads.
void parallel_for(size_t n, int threads, const std::function<void(size_t)>& f) {
  if (n == 0) return;
  const size_t t = std::min<size_t>(n, (size_t)std::max(1, threads));
  if (t == 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::exception_ptr> errs(t);
  std::vector<std::thread> pool;
  for (size_t w = 0; w < t; ++w)
    pool.emplace_back([&, w] {
      try {
        for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
      } catch (...) {
        errs[w] = std::current_exception();
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errs)
    if (e) std::rethrow_exception(e);
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

struct FrameState;

// Bytes that determine a launch sequence.
struct GraphKey {
  std::vector<uint8_t> b;
  template <typename T>
  GraphKey& add(const T& v) {
    const auto* p = reinterpret_cast<const uint8_t*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
    return *this;
  }
  bool operator==(const GraphKey& o) const { return b == o.b; }
};

// One captured launch sequence, replayed while its key repeats and recaptured
// when it changes.  `cur` / `perm` are host-side results of the enqueueing
// code, stored with the graph.
struct CachedGraph {
  GraphKey key;
  hipGraphExec_t exec = nullptr;
  int cur = 0;
  const uint32_t* perm = nullptr;
  void clear() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.b.clear();
  }
  template <typename F>
  void run(const GraphKey& k, hipStream_t st, F&& enqueue, int& cur_out, const uint32_t*& perm_out) {
    if (!exec || !(key == k)) {
      clear();
      xnyz::check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture");
      try {
        enqueue();
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hipGraph_t g = nullptr;
      xnyz::check_hip(hipStreamEndCapture(st, &g), "end capture");
      const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      xnyz::check_hip(e, "graph instantiate");
      key = k;
      cur = cur_out;
      perm = perm_out;
    }
    xnyz::check_hip(hipGraphLaunch(exec, st), "graph launch");
    cur_out = cur;
    perm_out = perm;
  }
};

struct xnyz_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  HilbertSM hs{};
  HilbertInv hinv{};
  // device buffers
  DBuf tiles, sp, sc, sh, lut, bigv, bigc, large, nlarge;
  DBuf keys[2], lo[2], cells, cells_lo, pos, cnt, out, arena, upload, tmp;
  DBuf idx[2], lo64[2];  // permutation payload
  DBuf feat;             // selector statistics
  DBuf patch;            // stream header / meta bytes for k_patch
  DBuf ztab, zscr, zout, zpos;  // device-built frames
  // Packed outputs of a batch (k_gather) and its segment table; the
  // asynchronous form keeps two stream buffers and two packed buffers, one
  // of each possibly still being copied to the host.
  DBuf pk, gtab, aout[2], apk[2], agtab[2];
  HBuf h_gtab;
  hipEvent_t aev[2] = {nullptr, nullptr};  // copy of slot i done
  bool abusy[2] = {false, false};
  int aslot = 0;
  // device decoder: stream, symbols, quads, block flags, patches and their
  // prefix sums, block arrays, RLE counts and offsets, points, tables, scratch
  DBuf dstream, dsym, dquad, dhead, dpatch, dps, dblk, dcnt, dpts, dtab, darena;
  HBuf h_dflag;
  // selector: parameters, FP16 weights, optimiser state, samples
  DBuf mP, mWh, mG, mM, mV, mX, mY, mIdx, mScratch, mLoss, mOut, mConf;
  // pinned host mirrors
  HBuf h_tiles, h_sp, h_sc, h_sh, h_lut, h_out, h_feat, h_zout, h_ztab, h_zpos;
  // pinned staging of the small copies inside a call (counts, large-count
  // lists, patches, big code tables, selector inputs and outputs): a copy
  // from or to pageable memory can wait behind another stream's copy in
  // flight (an earlier call's stream crossing PCIe)
  HBuf h_small, h_patch, h_big, h_mlp;
  bool prof_on = false;
  EventProfiler prof;
  FrameState* fs = nullptr;  // frame stage (created on first use)
  bool graphs = true;        // replay stage 2 from a captured graph
  std::vector<hipEvent_t> slice_ev;  // copy-back slices (frame overlap)
  hipStream_t copy_stream = nullptr;  // copy-back beside device-built frame blocks
  hipEvent_t ev_pipe = nullptr;       // end of the stream-building work
  hipEvent_t ev_fixed = nullptr;      // BitPacker sections written (early copy-back)
  hipEvent_t dec_ev[5] = {};          // decoding: upload pieces landed; [4] prior work done
  hipEvent_t ev_sizes = nullptr;      // device-built block sizes in host memory
  xnyz_tuning tuning[XNYZ_SIZE_CLASSES];
  CachedGraph stage2;
  // No-allocation calls: the device address of this GPU's scratch array and a
  // pinned host staging block, both set up by xnyz_create.
  void* na_dev = nullptr;
  HBuf na_host;

  std::vector<DBuf*> dbufs() {
    return {&tiles, &sp, &sc, &sh, &lut, &bigv, &bigc, &large, &nlarge, &keys[0], &keys[1],
            &lo[0], &lo[1], &cells, &cells_lo, &pos, &cnt, &out, &arena, &upload, &tmp,
            &idx[0], &idx[1], &lo64[0], &lo64[1], &feat, &patch, &ztab, &zscr, &zout, &zpos,
            &mP, &mWh, &mG, &mM, &mV, &mX, &mY, &mIdx, &mScratch, &mLoss, &mOut, &mConf,
            &dstream, &dsym, &dquad, &dhead, &dpatch, &dps, &dblk, &dcnt, &dpts, &dtab, &darena,
            &pk, &gtab, &aout[0], &aout[1], &apk[0], &apk[1], &agtab[0], &agtab[1]};
  }
  std::vector<HBuf*> hbufs() {
    return {&h_tiles, &h_sp, &h_sc, &h_sh, &h_lut, &h_out, &h_feat, &h_zout, &h_ztab, &h_zpos,
            &h_dflag, &h_gtab, &h_small, &h_patch, &h_big, &h_mlp};
  }

  // Every buffer's member name, for allocation errors.
  void name_buffers() {
#define XNYZ_NAME(b) b.name = #b
    XNYZ_NAME(tiles); XNYZ_NAME(sp); XNYZ_NAME(sc); XNYZ_NAME(sh); XNYZ_NAME(lut);
    XNYZ_NAME(bigv); XNYZ_NAME(bigc); XNYZ_NAME(large); XNYZ_NAME(nlarge);
    XNYZ_NAME(keys[0]); XNYZ_NAME(keys[1]); XNYZ_NAME(lo[0]); XNYZ_NAME(lo[1]);
    XNYZ_NAME(cells); XNYZ_NAME(cells_lo); XNYZ_NAME(pos); XNYZ_NAME(cnt); XNYZ_NAME(out);
    XNYZ_NAME(arena); XNYZ_NAME(upload); XNYZ_NAME(tmp); XNYZ_NAME(idx[0]); XNYZ_NAME(idx[1]);
    XNYZ_NAME(lo64[0]); XNYZ_NAME(lo64[1]); XNYZ_NAME(feat); XNYZ_NAME(patch); XNYZ_NAME(ztab);
    XNYZ_NAME(zscr); XNYZ_NAME(zout); XNYZ_NAME(zpos); XNYZ_NAME(mP); XNYZ_NAME(mWh);
    XNYZ_NAME(mG); XNYZ_NAME(mM); XNYZ_NAME(mV); XNYZ_NAME(mX); XNYZ_NAME(mY);
    XNYZ_NAME(mIdx); XNYZ_NAME(mScratch); XNYZ_NAME(mLoss); XNYZ_NAME(mOut); XNYZ_NAME(mConf);
    XNYZ_NAME(h_tiles); XNYZ_NAME(h_sp); XNYZ_NAME(h_sc); XNYZ_NAME(h_sh); XNYZ_NAME(h_lut);
    XNYZ_NAME(h_out); XNYZ_NAME(h_feat); XNYZ_NAME(h_zout); XNYZ_NAME(h_ztab); XNYZ_NAME(h_zpos);
    XNYZ_NAME(na_host); XNYZ_NAME(dstream); XNYZ_NAME(dsym); XNYZ_NAME(dquad);
    XNYZ_NAME(dhead); XNYZ_NAME(dpatch); XNYZ_NAME(dps); XNYZ_NAME(dblk); XNYZ_NAME(dcnt);
    XNYZ_NAME(dpts); XNYZ_NAME(dtab); XNYZ_NAME(darena); XNYZ_NAME(h_dflag);
    XNYZ_NAME(h_small); XNYZ_NAME(h_patch); XNYZ_NAME(h_big); XNYZ_NAME(h_mlp);
    XNYZ_NAME(pk); XNYZ_NAME(gtab); XNYZ_NAME(aout[0]); XNYZ_NAME(aout[1]); XNYZ_NAME(apk[0]);
    XNYZ_NAME(apk[1]); XNYZ_NAME(agtab[0]); XNYZ_NAME(agtab[1]); XNYZ_NAME(h_gtab);
#undef XNYZ_NAME
  }

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
  // Tables between pinned host memory and the device: a copy kernel (see
  // launch_copy_words), never queued behind another stream's copy.
  void h2d(DBuf& d, const HBuf& h, size_t bytes) { launch_copy_words(h.p, d.p, bytes, stream); }
  void d2h(HBuf& h, const DBuf& d, size_t bytes) { launch_copy_words(d.p, h.p, bytes, stream); }
};

namespace {

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

// ---------------------------------------------------------------------------
// Frame stage (source/composer/utils.hpp:35-53: ZSTD_compress at level 3).

uint8_t* xmalloc(size_t n) {
  uint8_t* p = static_cast<uint8_t*>(std::malloc(n ? n : 1));
  if (!p) throw Error(XNYZ_ENOMEM, "malloc failed");
  return p;
}

// One output as byte pieces in order (a frame header and the chunks of a
// host-threaded frame, or a single piece).
struct Piece {
  const uint8_t* p;
  u64 n;
};
using Pieces = std::vector<Piece>;

u64 pieces_size(const Pieces& v) {
  u64 t = 0;
  for (const Piece& x : v) t += x.n;
  return t;
}

// Runs f(i, worker) for i in [0, n) on up to `threads` host threads.
void parallel_for_w(size_t n, int threads, const std::function<void(size_t, size_t)>& f) {
  if (n == 0) return;
  const size_t t = std::min<size_t>(n, (size_t)std::max(1, threads));
  if (t == 1) {
    for (size_t i = 0; i < n; ++i) f(i, 0);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::exception_ptr> errs(t);
  std::vector<std::thread> pool;
  for (size_t w = 0; w < t; ++w)
    pool.emplace_back([&, w] {
      try {
        for (size_t i; (i = next.fetch_add(1)) < n;) f(i, w);
      } catch (...) {
        errs[w] = std::current_exception();
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errs)
    if (e) std::rethrow_exception(e);
}

// Host-threaded frames: bound of one compressed chunk of `chunk` bytes
// (its blocks, each with a 3-byte header).
size_t zchunk_bound(size_t chunk) { return ZSTD_compressBound(chunk) + 3 * (chunk / 1024 + 2); }

// Scratch bytes the frame of `len` bytes needs in `mode` with `threads`.
u64 frame_scratch(u64 len, int mode, int threads, u64 chunk) {
  if (mode == XNYZ_FRAME_NONE || mode == XNYZ_FRAME_ZSTD_GPU) return 0;
  const u64 nchunks = (len + chunk - 1) / chunk;
  if (mode == XNYZ_FRAME_ZSTD3_MT && nchunks > 1 && threads > 1)
    return 16 + nchunks * zchunk_bound(chunk);
  return ZSTD_compressBound(len);
}

// ---- run-time parameters (xnyz_tuning) ---------------------------------------

xnyz_tuning default_tuning() { return xnyz_tuning{20, 24, 0.f, 2048, 1, 8, 4096}; }

// Size class of a call: floor(log2(total points)), 0 for 0 or 1 point.
int size_class(u64 n) { return n <= 1 ? 0 : 63 - __builtin_clzll(n); }

// Split frames: bytes of a stream of `total` bytes whose blocks the device
// encodes (the head), or 0 when the stream is too short to split: the head
// is the device's share rounded up to whole chunks, at least 1 MiB, and
// leaves the host at least one chunk.
u64 split_head(u64 total, const xnyz_tuning& t, u64 chunk) {
  u64 h = (u64)((double)total * (double)t.gpu_frame_share);
  h = (h + chunk - 1) / chunk * chunk;
  if (h < (1u << 20) || h + chunk > total) return 0;
  return h;
}

void check_tuning(const xnyz_tuning& t) {
  if (t.zstd_chunk_log < 16 || t.zstd_chunk_log > 24)
    throw Error(XNYZ_EINVAL, "zstd_chunk_log must be in [16, 24]");
  if (t.copy_slice_log < 20 || t.copy_slice_log > 30)
    throw Error(XNYZ_EINVAL, "copy_slice_log must be in [20, 30]");
  if (!(t.gpu_frame_share >= 0.f && t.gpu_frame_share <= 0.9f))
    throw Error(XNYZ_EINVAL, "gpu_frame_share must be in [0, 0.9]");
  if (t.persist_blocks < 64 || t.persist_blocks > 16384)
    throw Error(XNYZ_EINVAL, "persist_blocks must be in [64, 16384]");
  if (t.copy_first > 1) throw Error(XNYZ_EINVAL, "copy_first must be 0 or 1");
  if (t.radix_bits < 6 || t.radix_bits > 8) throw Error(XNYZ_EINVAL, "radix_bits must be in [6, 8]");
  if (t.radix_tile != 2048 && t.radix_tile != 4096)
    throw Error(XNYZ_EINVAL, "radix_tile must be 2048 or 4096");
}

}  // namespace

// Frame-stage state kept by a context across calls: one libzstd compression
// context per host worker and the scratch the compressed bytes land in.
struct FrameState {
  std::vector<uint8_t> buf;
  std::vector<ZSTD_CCtx*> cctx;
  ZSTD_CCtx* get(size_t w) {
    if (cctx.size() <= w) cctx.resize(w + 1, nullptr);
    if (!cctx[w] && !(cctx[w] = ZSTD_createCCtx()))
      throw xnyz::Error(XNYZ_ENOMEM, "ZSTD_createCCtx failed");
    return cctx[w];
  }
  void ensure_workers(size_t n) {
    for (size_t w = 0; w < n; ++w) get(w);
  }
  ~FrameState() {
    for (auto c : cctx) ZSTD_freeCCtx(c);
  }
};

namespace {

// source/composer/utils.hpp:35-53: ZSTD_compress at level 3 (byte-identical:
// ZSTD_compress is ZSTD_compressCCtx on a fresh context).
Pieces zstd_single(ZSTD_CCtx* cc, const uint8_t* src, u64 len, uint8_t* dst, u64 cap) {
  const size_t r = ZSTD_compressCCtx(cc, dst, cap, src, len, 3);
  if (ZSTD_isError(r)) throw Error(XNYZ_EZSTD, ZSTD_getErrorName(r));
  return {Piece{dst, r}};
}

// One zstd frame, level 3, whose chunks (`chunk` bytes each) are compressed
// on host threads with the block-level API: single-segment frame header with
// the content size `fcs`, then each chunk's blocks (compressed, or raw when
// zstd declines).  Chunks after the first start without repeat offsets, so
// every chunk decodes after any other.  Valid for every zstd decoder; not
// byte-identical to the single-shot frame.
// `own` compresses the frame when it is not split (one chunk or one thread).
// `ready` (nullable) is called with a chunk's end offset before the chunk is
// read: the source may still be arriving from the device.
// With `header` false, only the blocks of [0, len) are written (the last one
// flagged): the tail of a frame whose header and first blocks come from
// elsewhere.  `place` (nullable, with `header` false) is where those blocks
// go in the end: the workers copy every chunk there as soon as all chunks
// before it are done, while later chunks are still being compressed.
Pieces zstd_chunked(FrameState& fs, ZSTD_CCtx* own, const uint8_t* src, u64 len, int threads,
                    uint8_t* dst, u64 chunk, const std::function<void(u64)>* ready, u64 fcs,
                    bool header, uint8_t* place = nullptr) {
  const size_t nchunks = (len + chunk - 1) / chunk;
  if (place && header) throw Error(XNYZ_EINTERNAL, "placed chunks carry no frame header");
  if (header && (nchunks <= 1 || threads <= 1)) {
    if (ready) (*ready)(len);
    return zstd_single(own, src, len, dst, ZSTD_compressBound(len));
  }
  const int t = std::max(1, std::min<int>(threads, (int)nchunks));
  fs.ensure_workers(t);  // contexts are created before the workers start
  const size_t cb = zchunk_bound(chunk);
  uint8_t* hdr = dst;  // magic, descriptor (single segment, 8-byte content size), size
  const uint32_t magic = 0xFD2FB528u;
  std::memcpy(hdr, &magic, 4);
  hdr[4] = 0xE0;
  const uint64_t fcs64 = fcs;
  std::memcpy(hdr + 5, &fcs64, 8);
  Pieces out(nchunks + 1);
  out[0] = Piece{hdr, header ? 13u : 0u};
  uint8_t* parts = dst + 16;
  // in-order placement: chunks [0, placed) have their final offsets
  std::mutex mu;
  std::vector<char> done(place ? nchunks : 0, 0);
  size_t placed = 0;
  u64 place_at = 0;
  parallel_for_w(nchunks, t, [&](size_t c, size_t w) {
    ZSTD_CCtx* cc = fs.cctx[w];
    const size_t off = c * chunk, n = std::min<size_t>(chunk, len - off);
    if (ready) (*ready)(off + n);
    uint8_t* o = parts + c * cb;
    size_t r = ZSTD_compressBegin(cc, 3);
    if (ZSTD_isError(r)) throw Error(XNYZ_EZSTD, ZSTD_getErrorName(r));
    if (c || !header) ZSTD_invalidateRepCodes(cc);
    const size_t bs = ZSTD_getBlockSize(cc);
    size_t at = 0;
    for (size_t b = 0; b < n; b += bs) {
      const size_t m = std::min(bs, n - b);
      const bool last = (c + 1 == nchunks) && (b + m == n);
      r = ZSTD_compressBlock(cc, o + at + 3, cb - at - 3, src + off + b, m);
      if (ZSTD_isError(r)) throw Error(XNYZ_EZSTD, ZSTD_getErrorName(r));
      u32 type = 2, size = (u32)r;
      if (r == 0 || r >= m) {  // stored as a raw block
        type = 0;
        size = (u32)m;
        std::memcpy(o + at + 3, src + off + b, m);
      }
      const u32 h = (last ? 1u : 0u) | (type << 1) | (size << 3);
      o[at] = uint8_t(h);
      o[at + 1] = uint8_t(h >> 8);
      o[at + 2] = uint8_t(h >> 16);
      at += 3 + size;
    }
    out[c + 1] = Piece{o, at};
    if (place) {
      size_t first, last;
      u64 to;
      {
        std::lock_guard<std::mutex> lk(mu);
        done[c] = 1;
        first = placed;
        to = place_at;
        while (placed < nchunks && done[placed]) place_at += out[++placed].n;
        last = placed;
      }
      for (size_t k = first; k < last; ++k) {  // this worker's share of the prefix
        std::memcpy(place + to, out[k + 1].p, out[k + 1].n);
        to += out[k + 1].n;
      }
    }
  });
  return out;
}

// Copies pieces to dst, splitting large copies over host threads.
void copy_pieces(const Pieces& v, uint8_t* dst, int threads) {
  constexpr u64 kSlice = 8u << 20;
  std::vector<std::pair<u64, u64>> jobs;  // (piece, offset in piece)
  std::vector<u64> at(v.size());
  u64 t = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    at[i] = t;
    t += v[i].n;
    for (u64 o = 0; o < v[i].n; o += kSlice) jobs.emplace_back(i, o);
  }
  parallel_for(jobs.size(), t >= (32u << 20) ? threads : 1, [&](size_t j) {
    const Piece& x = v[jobs[j].first];
    const u64 o = jobs[j].second, n = std::min(kSlice, x.n - o);
    std::memcpy(dst + at[jobs[j].first] + o, x.p + o, n);
  });
}

// ---------------------------------------------------------------------------
// Validation and per-segment setup

void validate(const xnyz_opts* o, u64 n) {
  if (!o) throw Error(XNYZ_EINVAL, "opts is null");
  if (o->quantizer != XNYZ_CUBE && o->quantizer != XNYZ_TRUNCATED_OCTAHEDRON)
    throw Error(XNYZ_EINVAL, "quantizer must be XNYZ_CUBE or XNYZ_TRUNCATED_OCTAHEDRON");
  if (!(o->l2_bound > 0.0f) || !std::isfinite(o->l2_bound))
    throw Error(XNYZ_EINVAL, "l2_bound must be a positive finite float");
  if (o->frame_mode < XNYZ_FRAME_ZSTD3 || o->frame_mode > XNYZ_FRAME_ZSTD_GPU)
    throw Error(XNYZ_EINVAL, "bad frame_mode");
  if (o->mode_flags & ~(XNYZ_MODE_GR_K | XNYZ_MODE_PFD_WIDTH))
    throw Error(XNYZ_EINVAL, "unknown mode_flags bits");
  if ((o->mode_flags & XNYZ_MODE_GR_K) && (o->gr_k < 0 || o->gr_k > 62))
    throw Error(XNYZ_EINVAL, "gr_k must be in 0..62");
  if ((o->mode_flags & XNYZ_MODE_PFD_WIDTH) && (o->pfd_width < 1 || o->pfd_width > 8))
    throw Error(XNYZ_EINVAL, "pfd_width must be in 1..8");
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
  qp.inv = 1.0f / qp.div;
  qp.rec_inv = 1.0f / qp.rec_div;
  qp.fast = div_const_ok(qp.div) ? 1 : 0;
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

// Host-side state of one segment across the stages.
struct SegHost {
  u64 n = 0;
  bool rle = false, hil = false, direct = false;
  int rx = 0, ry = 0, rz = 0;
  u32 W = 0;
  float off[3] = {0, 0, 0};
  u32 key_bits = 0, hi_bits = 0;
  bool wide_needed = false;
  double max_l2 = 0, mse = 0;
  u64 ncoded = 0, B = 0, P = 0;
  Codebook<uint64_t> cb_sym;
  Codebook<uint8_t> cb_quad;
  Codebook<int> cb_cnt;
  std::vector<u32> big_vals;
  std::vector<HufEntry> big_codes;
  std::vector<std::pair<u32, u32>> large;  // (count, first cell index)
  uint8_t grk = 0;
  u32 w1 = 0, w2 = 0;
  std::vector<std::vector<uint8_t>> metas;  // per section in stream order
  u64 base = 0, total = 0;                  // stream bytes in the output buffer
  u64 cells_at = ~0ull;  // first byte (from base) of the cell-space sections' data
};

// Layout of one stream: section records in stream order.
struct Rec {
  int kind;  // SecKind, or -1 for the always-empty quads patch section
  u64 data_bytes;
};

std::vector<Rec> records(const SegHost& g, const SegCtl& c) {
  std::vector<Rec> r;
  auto bytes = [](u64 bits) { return (bits + 7) / 8; };
  if (!g.direct) {
    r.push_back({kSecBlk, bytes(c.sec_bits[kSecBlk])});
    r.push_back({kSecCnt, bytes(c.sec_bits[kSecCnt])});
  }
  r.push_back({kSecGR, bytes(c.sec_bits[kSecGR])});
  r.push_back({kSecHuf, bytes(c.sec_bits[kSecHuf])});
  if (!g.direct) {
    r.push_back({-1, 0});
    r.push_back({kSecQuad, bytes(c.sec_bits[kSecQuad])});
  }
  if (g.rle) r.push_back({kSecRle, bytes(c.sec_bits[kSecRle])});
  return r;
}

std::vector<uint8_t> le_meta(u64 count, uint8_t last) {  // BitPacker / GR meta
  return {uint8_t(count), uint8_t(count >> 8), uint8_t(count >> 16), uint8_t(count >> 24), last};
}

std::vector<uint8_t> meta_of(const SegHost& g, int kind) {
  switch (kind) {
    case kSecBlk: return le_meta(g.B, (uint8_t)g.w1);
    case kSecCnt: return le_meta(g.B, (uint8_t)g.w2);
    case kSecGR: return g.P ? le_meta(g.P, g.grk) : std::vector<uint8_t>{};
    case kSecHuf: return g.cb_sym.meta();
    case kSecQuad: return g.cb_quad.meta();
    case kSecRle: return g.cb_cnt.meta();
    default: return {};
  }
}

// ---------------------------------------------------------------------------
// The pipeline

struct CallOut {
  xnyz_buf* frames;
  xnyz_buf* pre;
  xnyz_stats* stats;
  uint32_t* perm;  // single segment, non-RLE only
  // Mode search: stop after the statistics and return each segment's size
  // model instead of a stream.
  std::vector<ModeStats>* search = nullptr;
  // Selector features: stop after the range scan and return kFeat floats
  // per segment.
  std::vector<float>* features = nullptr;
  // Outputs into caller memory instead of `frames`: segment s in
  // [dst + dst_off[s], dst + dst_off[s + 1]), back to back; dst_off[0] is 0
  // except for an anchored split frame.  When cap is too small, dst_off[S]
  // is the size needed.
  uint8_t* dst = nullptr;
  u64 cap = 0;
  uint64_t* dst_off = nullptr;
  // Per segment (nullable): settle the Golomb-Rice k and PFD width of the
  // segment's curve x RLE variant from its own statistics with the exact
  // size model (modes.hpp) -- the statistics pass runs at the reference limit
  // 254 and also collects the patch bit populations.  The opts' k and width
  // stay when they cost no more than the optimum, else the optimum replaces
  // them.  The mode used lands in modes_out[s] (nullable), and
  // corrected[s] (nullable) says whether the opts' k / width were replaced.
  const char* refine = nullptr;
  xnyz_mode* modes_out = nullptr;
  char* corrected = nullptr;
  // Shared codebooks (block_compressor.hpp:241-431, compress_internal): 1 =
  // fit_profile (the repos and quads codebooks cover every symbol 0..255),
  // 2 = compress_with_profile (both loaded from profile_metas[0], [1]).
  // Every segment is coded in block mode; profile_out (fit) gets segment 0's
  // serialized profile (block_encoding_profile.hpp:34-44).
  int shared = 0;
  // The MPI driver's mapreduce mode on one GPU (mpi_driver.cpp:365-648 with
  // encode_cells_with_counts, block_compressor_rle.hpp:341-470): RLE, block
  // mode always, the largest quantized coordinate as the range of all three
  // axes.
  bool global_rle = false;
  // Assembled streams into pinned caller memory, returning once the copy is
  // queued on the copy stream (xnyz_wait ends it); the stream buffer
  // alternates between two slots, a slot's previous copy waited for first.
  bool async_out = false;
  const std::vector<uint8_t>* profile_metas = nullptr;
  std::vector<uint8_t>* profile_out = nullptr;
  // No-allocation call: every buffer comes from the fixed regions (g_bump),
  // stage 2 runs without a graph, and each segment's stream is copied from
  // device memory straight into dst (XNYZ_FRAME_NONE only).
  bool noalloc = false;
};

// HyperLogLog estimate from 256 registers (small-range correction by linear
// counting).
double hll_estimate(const u32* reg) {
  const double m = kHll, alpha = 0.7213 / (1.0 + 1.079 / m);
  double sum = 0.0;
  u32 zeros = 0;
  for (u32 j = 0; j < kHll; ++j) {
    sum += std::ldexp(1.0, -(int)reg[j]);
    zeros += reg[j] == 0;
  }
  double e = alpha * m * m / sum;
  if (e <= 2.5 * m && zeros) e = m * std::log(m / zeros);
  return e;
}

// The selector's 32 inputs for one segment.
void segment_features(const SegFeat& f, const SegCtl& c, const xnyz_opts& o, u64 n, float* x) {
  const double nn = (double)n;
  const double rx = c.range[0], ry = c.range[1], rz = c.range[2];
  const double U = std::min(nn, std::max(1.0, hll_estimate(f.hll_cell)));
  const double Sb = std::min(U, std::max(1.0, hll_estimate(f.hll_sub)));
  const double B = std::min(Sb, std::max(1.0, hll_estimate(f.hll_blk)));
  const double pairs = std::max(1.0, (double)f.pairs);
  const int max_range = std::max({c.range[0], c.range[1], c.range[2]});
  x[0] = (float)std::log2(nn);
  x[1] = (float)std::log2(1 + rx);
  x[2] = (float)std::log2(1 + ry);
  x[3] = (float)std::log2(1 + rz);
  x[4] = (float)(std::log2(nn) - std::log2((1 + rx) * (1 + ry) * (1 + rz)));
  x[5] = (float)std::log2(U);
  x[6] = (float)std::log2(nn / U);
  x[7] = (float)std::log2(Sb);
  x[8] = (float)std::log2(B);
  x[9] = (float)std::log2(U / Sb);
  x[10] = (float)std::log2(U / B);
  x[11] = o.quantizer == XNYZ_TRUNCATED_OCTAHEDRON ? 1.f : 0.f;
  x[12] = (float)std::log10((double)o.l2_bound);
  x[13] = max_range < o.direct_threshold ? 1.f : 0.f;
  x[14] = (float)(c.wrapped / nn);
  x[15] = (float)(f.same_cell / pairs);
  x[16] = (float)(f.same_sub / pairs);
  x[17] = (float)(f.same_blk / pairs);
  for (int b = 1; b <= 14; ++b) x[17 + b] = (float)(f.xor_hist[b] / pairs);
}

int compress_impl(xnyz_ctx* ctx, const float* d_xyz, const std::vector<u64>& off,
                  const xnyz_opts* opts, const CallOut& co, double t_start) {
  const u32 S = (u32)(off.size() - 1);
  const u64 N64 = off[S];
  if (N64 >= 0xFFFFFFFFull) throw Error(XNYZ_EUNSUPPORTED, "more than 2^32-2 points in one call");
  for (u32 s = 0; s < S; ++s) validate(&opts[s], off[s + 1] - off[s]);
  const u32 N = (u32)N64;
  hipStream_t st = ctx->stream;
  ProfGuard prof_guard(ctx);
  const int threads = host_threads(0);
  const xnyz_tuning tu = ctx->tuning[size_class(N64)];
  RadixShape rshape;
  rshape.bits = tu.radix_bits;
  rshape.sub = tu.radix_tile == 2048 ? 2 : 1;

  // ---- tiles and segment parameters ----------------------------------------
  u32 ntiles = 0;
  for (u32 s = 0; s < S; ++s) ntiles += (u32)((off[s + 1] - off[s] + kTile - 1) / kTile);
  ctx->h_tiles.ensure(sizeof(Tile) * ntiles);
  ctx->h_sp.ensure(sizeof(SegParam) * S);
  ctx->h_sc.ensure(sizeof(SegCtl) * S);
  ctx->h_sh.ensure(sizeof(SegHist) * S);
  ctx->h_lut.ensure(sizeof(SegLut) * S);
  Tile* tiles = ctx->h_tiles.as<Tile>();
  SegParam* sp = ctx->h_sp.as<SegParam>();
  SegCtl* sc = ctx->h_sc.as<SegCtl>();
  SegHist* shh = ctx->h_sh.as<SegHist>();
  SegLut* lut = ctx->h_lut.as<SegLut>();
  std::vector<SegHost> g(S);
  bool any_rle = false, want_perm = co.perm && S == 1 && !opts[0].rle;
  {
    u32 t = 0;
    for (u32 s = 0; s < S; ++s) {
      std::memset(&sp[s], 0, sizeof(SegParam));
      const u32 n = (u32)(off[s + 1] - off[s]);
      sp[s].p0 = (u32)off[s];
      sp[s].n = n;
      sp[s].tile0 = t;
      sp[s].ntiles = (n + kTile - 1) / kTile;
      sp[s].qp = quant_params(&opts[s]);
      sp[s].rle = opts[s].rle != 0;
      sp[s].diag = sp[s].rle && !co.search;
      const bool refine = co.refine && co.refine[s];
      sp[s].lim = (opts[s].mode_flags & XNYZ_MODE_PFD_WIDTH) && !refine
                      ? pfd_limit(opts[s].pfd_width)
                      : 254;
      sp[s].search = co.search != nullptr || refine;
      for (u32 l = 0; l < sp[s].ntiles; ++l) tiles[t++] = Tile{s, l};
      g[s].n = n;
      g[s].rle = sp[s].rle;
      g[s].hil = opts[s].is_hilbert != 0;
      any_rle |= g[s].rle;
    }
  }
  ctx->tiles.ensure(sizeof(Tile) * ntiles);
  ctx->sp.ensure(sizeof(SegParam) * S);
  ctx->sc.ensure(sizeof(SegCtl) * S);
  ctx->sh.ensure(sizeof(SegHist) * S);
  ctx->lut.ensure(sizeof(SegLut) * S);
  const u32 large_cap = N / kCntBinsSeg + S + 2;
  ctx->large.ensure((size_t)large_cap * 12);
  ctx->nlarge.ensure(8);
  // radix digit counts (up to 256 digits x 2 radix tiles per tile), section
  // sizing, scan partials
  ctx->arena.ensure((size_t)ntiles * (512 * 4 + kNumSec * 16 + 64) + (4u << 20));
  Arena ar{ctx->arena.as<char>(), ctx->arena.cap, 0};
  launch_copy_ranges(CopyRanges{{ctx->h_tiles.p, ctx->h_sp.p}, {ctx->tiles.p, ctx->sp.p},
                                {sizeof(Tile) * ntiles, sizeof(SegParam) * S}, 2}, st);

  Batch bt;
  std::memset(&bt, 0, sizeof(bt));  // padding included: bt is part of a graph key
  bt.S = S;
  bt.ntiles = ntiles;
  bt.N = N;
  bt.tiles = ctx->tiles.as<Tile>();
  bt.sp = ctx->sp.as<SegParam>();
  bt.sc = ctx->sc.as<SegCtl>();
  bt.sh = ctx->sh.as<SegHist>();
  bt.lut = ctx->lut.as<SegLut>();
  bt.large = ctx->large.as<u32>();
  bt.n_large = ctx->nlarge.as<u32>();
  bt.large_cap = large_cap;
  bt.grid_cap = tu.persist_blocks;

  // ---- stage 1: bbox minimum, quantize, range scan ----------------------------
  launch_init(bt, st);
  launch_bbox(bt, d_xyz, st);
  launch_quant_stats(bt, d_xyz, st);
  if (co.features) {
    ctx->feat.ensure(sizeof(SegFeat) * S);
    ctx->h_feat.ensure(sizeof(SegFeat) * S);
    launch_features(bt, d_xyz, ctx->feat.as<SegFeat>(), st);
    ctx->d2h(ctx->h_feat, ctx->feat, sizeof(SegFeat) * S);
  }
  ctx->d2h(ctx->h_sc, ctx->sc, sizeof(SegCtl) * S);
  ctx->sync();
  if (co.features) {
    co.features->resize((size_t)S * kFeat);
    const SegFeat* hf = ctx->h_feat.as<SegFeat>();
    for (u32 s = 0; s < S; ++s)
      segment_features(hf[s], sc[s], opts[s], g[s].n, co.features->data() + (size_t)s * kFeat);
    return XNYZ_OK;
  }

  // ---- key layout per segment (block_compressor.hpp:69-139) -------------------
  bool wide = false, any_block = false, any_direct = false;
  u32 narrow_bits = 0, hi_bits = 0;
  for (u32 s = 0; s < S; ++s) {
    SegHost& h = g[s];
    const SegCtl& c = sc[s];
    h.rx = c.range[0];
    h.ry = c.range[1];
    h.rz = c.range[2];
    h.W = c.wrapped;
    for (int a = 0; a < 3; ++a) h.off[a] = c.offset[a];
    if (h.rle) {
      double m2;
      std::memcpy(&m2, &c.max_err2_bits, 8);
      h.max_l2 = std::sqrt(m2);
      h.mse = c.sum_err2 / h.n;
    }
    const int max_range = std::max({h.rx, h.ry, h.rz});
    h.direct = !co.shared && !co.global_rle && max_range < opts[s].direct_threshold;
    if (co.global_rle) h.rx = h.ry = h.rz = max_range;
    KeyParams& kp = sp[s].kp;
    kp = KeyParams{};
    kp.direct = h.direct;
    kp.hilbert = h.hil;
    if (h.direct) {
      if (max_range >= (1 << 21) - 1)
        throw Error(XNYZ_EUNSUPPORTED, "direct mode with a quantized range >= 2^21-1");
      const u32 L = bitlen((u64)max_range);
      kp.L = L ? L : 1;
      h.key_bits = h.W ? 64 : 3 * kp.L;
      h.hi_bits = h.key_bits;  // wide layout: the code sits in the high word
      any_direct = true;
    } else {
      const u64 nx = (u64)h.rx / 128 + 1, ny = (u64)h.ry / 128 + 1, nz = (u64)h.rz / 128 + 1;
      const unsigned __int128 nb = (unsigned __int128)nx * ny * nz;
      if (nb > ((unsigned __int128)1 << 57))
        throw Error(XNYZ_EUNSUPPORTED, "block grid larger than 2^57 blocks");
      kp.nx = nx;
      kp.nxny = nx * ny;
      kp.C = ((1ull << 57) - 1) * kp.nxny;  // wraps mod 2^64 as the reference's size_t
      kp.idb = bitlen((u64)nb - 1);
      if (h.W && (u64)nb > kp.C)
        throw Error(XNYZ_EUNSUPPORTED, "wrapped block ids overlap the regular ids");
      h.hi_bits = kp.idb + (h.W ? 1 : 0);
      h.key_bits = h.hi_bits + 21;
      h.wide_needed = h.key_bits > 64;
      any_block = true;
    }
    wide |= h.wide_needed;
    narrow_bits = std::max(narrow_bits, h.key_bits);
    hi_bits = std::max(hi_bits, h.hi_bits);
  }
  bt.wide = wide;
  // Passes each segment's keys need: the batch sorts for its widest keys, and
  // a segment with narrower ones keeps its keys in place after its own.
  for (u32 s = 0; s < S; ++s)
    sp[s].passes = wide ? radix_passes(0, g[s].hi_bits, true, rshape)
                        : radix_passes(g[s].key_bits, 0, false, rshape);
  ctx->h2d(ctx->sp, ctx->h_sp, sizeof(SegParam) * S);

  // ---- stage 2: keys, sort, cells, runs, statistics ---------------------------
  // Every buffer is sized before the launches, which may be captured.
  for (int i = 0; i < 2; ++i) ctx->keys[i].ensure((size_t)N * 8);
  if (wide)
    for (int i = 0; i < 2; ++i) ctx->lo[i].ensure((size_t)N * 4);
  ctx->pos.ensure((size_t)N * 4 + 16);
  if (want_perm) {
    for (int i = 0; i < 2; ++i) ctx->idx[i].ensure((size_t)N * 4);
    if (wide)
      for (int i = 0; i < 2; ++i) ctx->lo64[i].ensure((size_t)N * 8);
  }
  if (any_rle) {
    ctx->cells.ensure((size_t)N * 8);
    if (wide) ctx->cells_lo.ensure((size_t)N * 4);
    ctx->cnt.ensure((size_t)N * 4 + 16);
  }
  u64* ka = ctx->keys[0].as<u64>();
  u64* kb = ctx->keys[1].as<u64>();
  u32* la = wide ? ctx->lo[0].as<u32>() : nullptr;
  u32* lb = wide ? ctx->lo[1].as<u32>() : nullptr;
  // The reference reorders the caller's points by the sort (ords,
  // block_compressor.hpp:141-144); here that order is an index payload of the
  // sort, returned through `perm`.
  u32 *ia = nullptr, *ib = nullptr;
  u64 *l64a = nullptr, *l64b = nullptr;
  if (want_perm) {
    ia = ctx->idx[0].as<u32>();
    ib = ctx->idx[1].as<u32>();
    if (wide) {
      l64a = ctx->lo64[0].as<u64>();
      l64b = ctx->lo64[1].as<u64>();
    }
  }
  u32* cnt = any_rle ? ctx->cnt.as<u32>() : nullptr;
  // Block starts (each block's first cell): with RLE cells, found while the
  // cells are and kept in the sort's spare key buffer (free once the keys
  // are sorted), since the cell starts still occupy `pos`; otherwise in `pos`.
  auto block_starts = [&](int sorted) {
    return any_rle ? ctx->keys[sorted ^ 1].as<u32>() : ctx->pos.as<u32>();
  };
  int cur = 0;
  const u32* perm_dev = nullptr;
  auto stage2 = [&] {
    launch_keygen(bt, d_xyz, ctx->hs, ka, l64a ? nullptr : la, ia, l64a, st);
    if (want_perm && wide) {
      cur = radix_sort_wide_v(bt, ka, kb, l64a, l64b, hi_bits, rshape, ar, st);
      launch_split_lo64(cur ? l64b : l64a, ctx->lo[cur].as<u32>(), ia, N, st);
      perm_dev = ia;
    } else if (wide) {
      cur = radix_sort_wide(bt, ka, kb, la, lb, hi_bits, rshape, ar, st);
    } else {
      cur = radix_sort(bt, ka, kb, ia, ib, narrow_bits, rshape, ar, st);
      perm_dev = want_perm ? (cur ? ib : ia) : nullptr;
    }
    KeyView k{ctx->keys[cur].as<u64>(), wide ? ctx->lo[cur].as<u32>() : nullptr};
    if (any_rle)
      rle_cells(bt, k, ctx->pos.as<u32>(), block_starts(cur), ctx->cells.as<u64>(),
                wide ? ctx->cells_lo.as<u32>() : nullptr, cnt, ar, st);
    if (any_rle) k = KeyView{ctx->cells.as<u64>(), wide ? ctx->cells_lo.as<u32>() : nullptr};
    block_runs_and_stats(bt, k, any_direct, any_block, any_rle, block_starts(cur), ar, st);
  };
  if (ctx->graphs && !g_profiler && !co.noalloc) {
    // The launch sequence is a function of this signature alone (kernel
    // parameters are device-resident), so an instantiated graph replays it.
    GraphKey key;
    key.add(d_xyz).add(bt).add(ka).add(kb).add(la).add(lb).add(ia).add(ib).add(l64a).add(l64b)
        .add(cnt).add(ctx->pos.p).add(ctx->cells.p).add(ctx->cells_lo.p).add(ar.base).add(ar.off)
        .add(narrow_bits).add(hi_bits).add(any_rle).add(any_block).add(any_direct).add(want_perm)
        .add(rshape.bits).add(rshape.sub);
    ctx->stage2.run(key, st, stage2, cur, perm_dev);
  } else {
    stage2();
  }
  if (perm_dev) {
    check_hip(hipMemcpyAsync(co.perm, perm_dev, (size_t)N * 4, hipMemcpyDeviceToHost, st),
              "perm D2H");
    ctx->copied();
  }
  KeyView kv{ctx->keys[cur].as<u64>(), wide ? ctx->lo[cur].as<u32>() : nullptr};
  if (any_rle) kv = KeyView{ctx->cells.as<u64>(), wide ? ctx->cells_lo.as<u32>() : nullptr};
  ctx->h_small.ensure(64);
  launch_copy_ranges(CopyRanges{{ctx->sc.p, ctx->sh.p, ctx->nlarge.p},
                                {ctx->h_sc.p, ctx->h_sh.p, ctx->h_small.p},
                                {sizeof(SegCtl) * S, sizeof(SegHist) * S, 4}, 3}, st);
  ctx->sync();
  const u32 n_large = *ctx->h_small.as<u32>();
  if (n_large > large_cap) throw Error(XNYZ_EINTERNAL, "large-count list overflow");
  if (n_large) {
    ctx->h_small.ensure((size_t)n_large * 12);
    launch_copy_words(ctx->large.p, ctx->h_small.p, (size_t)n_large * 12, st);
    ctx->sync();
    const u32* raw = ctx->h_small.as<u32>();
    for (u32 i = 0; i < n_large; ++i) g[raw[3 * i]].large.emplace_back(raw[3 * i + 1], raw[3 * i + 2]);
  }

  // ---- host: codebooks per segment (in parallel) ----------------------------
  // Every stream byte outside the repos / sfc-delta sections (header, trailer,
  // block, quads and RLE sections with their size fields), once the other
  // codebooks are built.
  auto fixed_bytes = [](const SegHost& h) {
    auto sect = [](u64 meta, u64 bits) { return 16 + meta + (bits + 7) / 8; };
    u64 fixed = 30 + 12;
    if (!h.direct) {
      fixed += sect(5, h.B * h.w1) + sect(5, h.B * h.w2) + 16 +
               sect(h.cb_quad.meta().size(), h.cb_quad.data_bits());
    }
    if (h.rle) fixed += sect(h.cb_cnt.meta().size(), h.cb_cnt.data_bits());
    return fixed;
  };
  std::vector<xnyz_mode> used(S);
  parallel_for(S, threads, [&](size_t s) {
    SegHost& h = g[s];
    const SegCtl& c = sc[s];
    const SegHist& hh = shh[s];
    if (c.n_qpatches) throw Error(XNYZ_EINTERNAL, "quads delta above 254");
    h.ncoded = c.ne;
    h.B = h.direct ? 0 : c.nb;
    h.P = c.n_patches;
    if (!h.direct) {
      std::vector<SymCount> qs = sym_counts(hh.quad, hh.first_quad, kSymBins);
      if (co.shared == 1) {
        h.cb_quad.build_complete(qs, h.ncoded, 255);
      } else if (co.shared == 2) {
        if (!h.cb_quad.load_meta(co.profile_metas[1], qs) || !h.cb_quad.covers_data())
          throw Error(XNYZ_EINVAL, "profile's quads codebook does not cover the cloud");
      } else {
        h.cb_quad.build(qs, h.ncoded);
      }
      fill_lut(h.cb_quad, lut[s].quad, kSymBins);
      h.w1 = c.max_blkd ? bitlen(c.max_blkd) : 1;
      h.w2 = c.max_cntd ? bitlen(c.max_cntd) : 1;
    }
    if (h.rle) {
      std::vector<SymCount> cs = sym_counts(hh.cnt, hh.first_cnt, kCntBinsSeg);
      std::sort(h.large.begin(), h.large.end());
      for (size_t i = 0; i < h.large.size();) {
        size_t j = i;
        u32 first = h.large[i].second;
        while (j < h.large.size() && h.large[j].first == h.large[i].first) {
          first = std::min(first, h.large[j].second);
          ++j;
        }
        cs.push_back(SymCount{h.large[i].first, (u64)(j - i), first});
        h.big_vals.push_back(h.large[i].first);
        i = j;
      }
      // block_compressor_rle.hpp:290-303: build, compare the meta with a
      // quarter of the data, then build again on the same object.
      h.cb_cnt.build(cs, h.ncoded);
      const double meta1 = (double)h.cb_cnt.meta().size();
      const double data1 = (double)((h.cb_cnt.data_bits() + 7) / 8);
      if (meta1 > data1 * 0.25) h.cb_cnt.use_eg = true;
      h.cb_cnt.build(cs, h.ncoded);
      if (!h.cb_cnt.use_eg) {
        fill_lut(h.cb_cnt, lut[s].cnt, kCntBinsSeg);
        for (u32 v : h.big_vals) {
          const auto& cd = h.cb_cnt.code_table.at((int)v);
          h.big_codes.push_back(HufEntry{cd.bits, cd.length});
        }
      }
    }
    // The repos / sfc-delta symbols and patches (p_for_delta_encoder.hpp:27-43).
    std::vector<SymCount> syms = sym_counts(hh.sym, hh.first_sym, kSymBins);
    h.grk = (opts[s].mode_flags & XNYZ_MODE_GR_K) ? (uint8_t)opts[s].gr_k
                                                   : golomb_rice_k(h.P, c.patch_sum, false);
    used[s] = xnyz_mode{h.hil, h.rle, h.grk, (opts[s].mode_flags & XNYZ_MODE_PFD_WIDTH)
                                                  ? opts[s].pfd_width : 8};
    if (co.refine && co.refine[s]) {
      // Statistics at limit 254: cost the asked (k, width) and the optimum
      // exactly, keep the asked one unless it costs more.
      ModeStats m;
      for (u32 i = 0; i < kSymBins; ++i) m.hist[i] = hh.sym[i];
      for (u32 b = 0; b < 64; ++b) m.bitpop[b] = hh.bitpop[b];
      m.patch_sum = c.patch_sum;
      m.fixed_bytes = fixed_bytes(h);
      ModeChoice ref{}, best{};
      search_modes(m, ref, best);
      const int ak = used[s].gr_k, aw = used[s].pfd_width;
      const bool keep = mode_cost(m, ak, aw).bytes <= best.bytes;
      const int k = keep ? ak : best.k, w = keep ? aw : best.w;
      if (co.corrected) co.corrected[s] = !keep;
      // Fold the values in (lim(w), 254] into the patch symbol 0: they become
      // patches, and symbol 0 first occurs where the first of them does.
      const u32 lim = pfd_limit(w);
      std::vector<SymCount> folded;
      SymCount zero{0, 0, ~0ull};
      for (const SymCount& x : syms) {
        if (x.sym == 0 || x.sym - 1 > lim) {
          zero.freq += x.freq;
          zero.first = std::min(zero.first, x.first);
          if (x.sym) h.P += x.freq;
        } else {
          folded.push_back(x);
        }
      }
      if (zero.freq) folded.insert(folded.begin(), zero);
      syms.swap(folded);
      sp[s].lim = lim;
      h.grk = (uint8_t)k;
      used[s].gr_k = k;
      used[s].pfd_width = w;
    }
    if (co.shared == 1) {
      h.cb_sym.build_complete(syms, h.ncoded, 255);
    } else if (co.shared == 2) {
      if (!h.cb_sym.load_meta(co.profile_metas[0], syms) || !h.cb_sym.covers_data())
        throw Error(XNYZ_EINVAL, "profile's repos codebook does not cover the cloud");
    } else {
      h.cb_sym.build(syms, h.ncoded);
    }
    fill_lut(h.cb_sym, lut[s].sym, kSymBins);
    if (h.cb_sym.max_length() > 56 || h.cb_quad.max_length() > 56 ||
        h.cb_cnt.max_length() > 56)
      throw Error(XNYZ_EUNSUPPORTED, "Huffman code longer than 56 bits");
  });
  if (co.modes_out)
    for (u32 s = 0; s < S; ++s) co.modes_out[s] = used[s];
  if (co.profile_out) {  // serialize_profile: u32 LE size + repos meta, u32 LE size + quads meta
    co.profile_out->clear();
    for (const auto& m : {g[0].cb_sym.meta(), g[0].cb_quad.meta()}) {
      const uint32_t n = (uint32_t)m.size();
      for (int k = 0; k < 4; ++k) co.profile_out->push_back(uint8_t(n >> (8 * k)));
      co.profile_out->insert(co.profile_out->end(), m.begin(), m.end());
    }
  }
  if (co.search) {  // size model per segment (modes.hpp)
    co.search->resize(S);
    for (u32 s = 0; s < S; ++s) {
      const SegHist& hh = shh[s];
      ModeStats& m = (*co.search)[s];
      for (u32 i = 0; i < kSymBins; ++i) m.hist[i] = hh.sym[i];
      for (u32 b = 0; b < 64; ++b) m.bitpop[b] = hh.bitpop[b];
      m.patch_sum = sc[s].patch_sum;
      m.fixed_bytes = fixed_bytes(g[s]);
    }
    return XNYZ_OK;
  }
  std::vector<u32> big_vals;
  std::vector<HufEntry> big_codes;
  for (u32 s = 0; s < S; ++s) {
    SegHost& h = g[s];
    sp[s].grk = h.grk;
    sp[s].cnt_eg = h.cb_cnt.use_eg;
    sp[s].egk = h.cb_cnt.eg_k;
    sp[s].w1 = h.w1;
    sp[s].w2 = h.w2;
    sp[s].big0 = (u32)big_vals.size();
    sp[s].nbig = (u32)h.big_codes.size();
    if (!h.big_codes.empty()) {
      big_vals.insert(big_vals.end(), h.big_vals.begin(), h.big_vals.end());
      big_codes.insert(big_codes.end(), h.big_codes.begin(), h.big_codes.end());
    }
  }
  ctx->bigv.ensure(big_vals.size() * 4 + 16);
  ctx->bigc.ensure(big_codes.size() * sizeof(HufEntry) + 16);
  CopyRanges cr{{ctx->h_sp.p, ctx->h_lut.p}, {ctx->sp.p, ctx->lut.p},
                {sizeof(SegParam) * S, sizeof(SegLut) * S}, 2};
  if (!big_vals.empty()) {
    const size_t vb = big_vals.size() * 4, cb = big_codes.size() * sizeof(HufEntry);
    const size_t cat = (vb + 15) & ~size_t(15);
    ctx->h_big.ensure(cat + cb);
    std::memcpy(ctx->h_big.p, big_vals.data(), vb);
    std::memcpy(ctx->h_big.as<char>() + cat, big_codes.data(), cb);
    cr.src[2] = ctx->h_big.p;
    cr.dst[2] = ctx->bigv.p;
    cr.bytes[2] = vb;
    cr.src[3] = ctx->h_big.as<char>() + cat;
    cr.dst[3] = ctx->bigc.p;
    cr.bytes[3] = cb;
    cr.n = 4;
  }
  launch_copy_ranges(cr, st);
  bt.big_vals = ctx->bigv.as<u32>();
  bt.big_codes = ctx->bigc.as<HufEntry>();

  // ---- stage 3: section sizes -------------------------------------------------
  const SecPlan plan = plan_sections(bt, ar);
  size_sections(bt, kv, cnt, block_starts(cur), plan, ar, st);
  ctx->d2h(ctx->h_sc, ctx->sc, sizeof(SegCtl) * S);
  ctx->sync();

  // ---- layout: each segment's stream at a 256-byte aligned base ---------------
  u64 out_total = 0;
  for (u32 s = 0; s < S; ++s) {
    SegHost& h = g[s];
    const SegCtl& c = sc[s];
    if (c.sec_bits[kSecHuf] != h.cb_sym.data_bits() ||
        (!h.direct && c.sec_bits[kSecQuad] != h.cb_quad.data_bits()) ||
        (h.rle && c.sec_bits[kSecRle] != h.cb_cnt.data_bits()) ||
        (!h.direct && (c.sec_bits[kSecBlk] != h.B * h.w1 || c.sec_bits[kSecCnt] != h.B * h.w2)))
      throw Error(XNYZ_EINTERNAL, "device and host disagree on a section size");
    h.base = out_total;
    u64 at = h.base + 30;
    for (const Rec& r : records(h, c)) {
      std::vector<uint8_t> m = meta_of(h, r.kind);
      const u64 data_at = at + 16 + m.size();
      if (r.kind >= 0) sp[s].sec_byte[r.kind] = data_at;
      if (r.kind >= kSecGR && r.data_bytes && data_at - h.base < h.cells_at) h.cells_at = data_at - h.base;
      at = data_at + r.data_bytes;
      h.metas.push_back(std::move(m));
    }
    h.total = at + 12 - h.base;
    out_total = (h.base + h.total + 255) & ~u64(255);
  }
  ctx->h2d(ctx->sp, ctx->h_sp, sizeof(SegParam) * S);

  // ---- stage 4: the whole stream in device memory -----------------------------
  // Header (block_compressor.hpp:61-87): u64 points, u64 params.size() = 0,
  // i32 ranges, bool curve, bool direct; per section u64 data size, u64 meta
  // size, meta; trailer: the float offset.  These bytes are copied in by
  // k_patch before the sections are ORed in around them.
  std::vector<uint8_t> pbytes;
  std::vector<Patch> patches;
  auto put = [&](u64 dst, const void* src, size_t len) {
    if (!len) return;
    patches.push_back(Patch{dst, (u32)pbytes.size(), (u32)len});
    const auto* q = static_cast<const uint8_t*>(src);
    pbytes.insert(pbytes.end(), q, q + len);
  };
  for (u32 s = 0; s < S; ++s) {
    const SegHost& h = g[s];
    uint8_t head[30];
    const u64 np = h.ncoded, zero = 0;
    std::memcpy(head, &np, 8);
    std::memcpy(head + 8, &zero, 8);
    const int32_t rr[3] = {h.rx, h.ry, h.rz};
    std::memcpy(head + 16, rr, 12);
    head[28] = h.hil ? 1 : 0;
    head[29] = h.direct ? 1 : 0;
    put(h.base, head, 30);
    u64 at = h.base + 30;
    size_t i = 0;
    for (const Rec& r : records(h, sc[s])) {
      const std::vector<uint8_t>& m = h.metas[i++];
      const u64 sizes[2] = {r.data_bytes, (u64)m.size()};
      put(at, sizes, 16);
      put(at + 16, m.data(), m.size());
      at += 16 + m.size() + r.data_bytes;
    }
    put(at, h.off, 12);
  }
  // Assembled streams into caller memory that the DMA engine can write
  // (pinned, or any memory for a no-allocation call) leave straight from
  // device memory: one copy for one cloud; a batch is packed on the device
  // first (k_gather), then copied at once.
  bool all_none = true;
  for (u32 s = 0; s < S; ++s) all_none &= opts[s].frame_mode == XNYZ_FRAME_NONE;
  const bool direct_out =
      co.noalloc || (all_none && co.dst && !co.pre && co.cap && pinned_range(co.dst, co.cap));
  if (co.async_out && (!direct_out || co.noalloc))
    throw Error(XNYZ_EINVAL, "asynchronous calls return assembled streams (XNYZ_FRAME_NONE) "
                             "into pinned memory (xnyz_host_alloc)");
  const int slot = ctx->aslot;
  if (co.async_out && ctx->abusy[slot]) {  // this slot's previous copy must have ended
    check_hip(hipEventSynchronize(ctx->aev[slot]), "hipEventSynchronize");
    ctx->abusy[slot] = false;
  }
  DBuf& obuf = co.async_out ? ctx->aout[slot] : ctx->out;
  obuf.ensure(out_total + 64);
  u32* ow = obuf.as<u32>();
  launch_fill_u32(ow, (out_total + 64) / 4, 0, st);
  const size_t prec = patches.size() * sizeof(Patch);
  ctx->patch.ensure(prec + pbytes.size() + 16);
  ctx->h_patch.ensure(prec + pbytes.size() + 16);
  std::memcpy(ctx->h_patch.p, patches.data(), prec);
  if (!pbytes.empty()) std::memcpy(ctx->h_patch.as<uint8_t>() + prec, pbytes.data(), pbytes.size());
  launch_copy_words(ctx->h_patch.p, ctx->patch.p, prec + pbytes.size(), st);
  launch_patch(ctx->patch.as<Patch>(), (u32)patches.size(), ctx->patch.as<uint8_t>() + prec,
               reinterpret_cast<uint8_t*>(ow), st);
  // One cloud copied out as it is written: the bytes before the first word
  // the cell-space writer touches are final once the BitPacker sections are,
  // so their copy to the host starts on the copy stream while the cell-space
  // sections are written.
  u64 early = 0;
  if (S == 1 && direct_out && !co.noalloc && !co.async_out && any_block && !any_direct &&
      g[0].cells_at != ~0ull)
    early = g[0].cells_at & ~u64(3);
  if (early) {
    if (!ctx->copy_stream)
      check_hip(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking),
                "hipStreamCreate");
    if (!ctx->ev_fixed)
      check_hip(hipEventCreateWithFlags(&ctx->ev_fixed, hipEventDisableTiming), "hipEventCreate");
  }
  write_sections(bt, kv, cnt, block_starts(cur), plan, ow, any_block, any_direct, st,
                 early ? ctx->ev_fixed : nullptr);

  // Fills co.stats (frame bytes from co.frames / co.dst_off) once the outputs
  // are in place.
  auto fill_stats = [&](double t_dev, double t_end) {
    if (!co.stats) return;
    for (u32 s = 0; s < S; ++s) {
      const SegHost& h = g[s];
      xnyz_stats& x = co.stats[s];
      std::memset(&x, 0, sizeof(x));
      x.input_points = h.n;
      x.num_points = h.ncoded;
      x.range[0] = h.rx;
      x.range[1] = h.ry;
      x.range[2] = h.rz;
      x.is_direct = h.direct;
      x.blocks = h.B;
      x.patches = h.P;
      x.wrapped_points = h.W;
      x.counts_exp_golomb = h.rle ? h.cb_cnt.use_eg : 0;
      x.counts_eg_k = h.rle ? h.cb_cnt.eg_k : 0;
      x.entropy_repos = h.cb_sym.entropy;
      x.entropy_quads = h.cb_quad.entropy;
      x.entropy_counts = h.cb_cnt.entropy;
      x.max_l2 = h.max_l2;
      x.mse = h.mse;
      x.prezstd_bytes = h.total;
      x.frame_bytes = co.frames ? co.frames[s].len
                     : co.dst   ? co.dst_off[s + 1] - co.dst_off[s]
                                : 0;
      x.ms_device = t_dev - t_start;
      x.ms_zstd = t_end - t_dev;
      x.ms_total = t_end - t_start;
    }
  };

  if (direct_out) {
    co.dst_off[0] = 0;
    for (u32 s = 0; s < S; ++s) co.dst_off[s + 1] = co.dst_off[s] + g[s].total;
    if (co.dst_off[S] > co.cap)
      throw Error(XNYZ_ENOSPC, "output buffer too small: " + std::to_string(co.dst_off[S]) +
                                   " bytes needed");
    if (co.noalloc) {  // one copy per cloud: no packed buffer to carve
      for (u32 s = 0; s < S; ++s)
        check_hip(hipMemcpyAsync(co.dst + co.dst_off[s],
                                 reinterpret_cast<const uint8_t*>(ow) + g[s].base, g[s].total,
                                 hipMemcpyDeviceToHost, st), "stream D2H");
      ctx->copied();
      ctx->sync();
      const double t = now_ms();
      fill_stats(t, t);
      return XNYZ_OK;
    }
    const u64 bytes = co.dst_off[S];
    const uint8_t* src = reinterpret_cast<const uint8_t*>(ow) + g[0].base;
    if (S > 1) {
      DBuf& pkb = co.async_out ? ctx->apk[slot] : ctx->pk;
      DBuf& tb = co.async_out ? ctx->agtab[slot] : ctx->gtab;
      const size_t tbytes = S * sizeof(GatherSeg);
      ctx->h_gtab.ensure(tbytes);
      GatherSeg* gt = ctx->h_gtab.as<GatherSeg>();
      for (u32 s = 0; s < S; ++s) gt[s] = GatherSeg{g[s].base, co.dst_off[s], g[s].total};
      tb.ensure(tbytes);
      pkb.ensure(bytes + 64);
      launch_copy_words(gt, tb.p, tbytes, st);
      launch_gather(reinterpret_cast<const uint8_t*>(ow), tb.as<GatherSeg>(), S, bytes,
                    pkb.as<uint8_t>(), st);
      src = pkb.as<uint8_t>();
    }
    if (co.async_out) {
      if (!ctx->copy_stream)
        check_hip(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking),
                  "hipStreamCreate");
      if (!ctx->ev_pipe)
        check_hip(hipEventCreateWithFlags(&ctx->ev_pipe, hipEventDisableTiming), "hipEventCreate");
      if (!ctx->aev[slot])
        check_hip(hipEventCreateWithFlags(&ctx->aev[slot], hipEventDisableTiming),
                  "hipEventCreate");
      check_hip(hipEventRecord(ctx->ev_pipe, st), "hipEventRecord");
      check_hip(hipStreamWaitEvent(ctx->copy_stream, ctx->ev_pipe, 0), "hipStreamWaitEvent");
      check_hip(hipMemcpyAsync(co.dst, src, bytes, hipMemcpyDeviceToHost, ctx->copy_stream),
                "stream D2H");
      check_hip(hipEventRecord(ctx->aev[slot], ctx->copy_stream), "hipEventRecord");
      ctx->abusy[slot] = true;
      ctx->aslot = slot ^ 1;
      const double t = now_ms();
      fill_stats(t, t);
      return XNYZ_OK;
    }
    if (early && early <= bytes) {
      check_hip(hipStreamWaitEvent(ctx->copy_stream, ctx->ev_fixed, 0), "hipStreamWaitEvent");
      check_hip(hipMemcpyAsync(co.dst, src, early, hipMemcpyDeviceToHost, ctx->copy_stream),
                "stream D2H");
      if (bytes > early)
        check_hip(hipMemcpyAsync(co.dst + early, src + early, bytes - early,
                                 hipMemcpyDeviceToHost, st), "stream D2H");
      ctx->copied();
      ctx->sync();
      check_hip(hipStreamSynchronize(ctx->copy_stream), "hipStreamSynchronize");
      const double t = now_ms();
      fill_stats(t, t);
      return XNYZ_OK;
    }
    check_hip(hipMemcpyAsync(co.dst, src, bytes, hipMemcpyDeviceToHost, st), "stream D2H");
    ctx->copied();
    ctx->sync();
    const double t = now_ms();
    fill_stats(t, t);
    return XNYZ_OK;
  }

  // ---- frame stage -------------------------------------------------------------
  const bool want_out = co.frames || co.dst;
  bool host_pre = co.pre != nullptr;
  std::vector<u32> gsegs;  // segments framed on the device, in order
  for (u32 s = 0; s < S; ++s) {
    if (want_out && opts[s].frame_mode == XNYZ_FRAME_ZSTD_GPU) gsegs.push_back(s);
    else if (want_out) host_pre = true;
  }
  const bool per_seg_threads = S > 1;
  const u64 chunk = 1ull << tu.zstd_chunk_log;
  // One cloud framed on host threads: its stream comes back in slices, each
  // followed by an event, and a chunk is compressed as soon as its slice has
  // landed -- the copy overlaps the frame stage.
  const bool overlap = host_pre && S == 1 && !co.pre && opts[0].frame_mode == XNYZ_FRAME_ZSTD3_MT;
  // Split frame: the device encodes the stream's head [0, head) -- the
  // sections zstd spends the most time per byte on -- as the frame's first
  // blocks while the host threads compress the tail; only the tail is copied
  // back.
  u64 head = 0;
  if (overlap && tu.gpu_frame_share > 0.f) head = split_head(g[0].total, tu, chunk);
  const bool split = head > 0;
  hipStream_t cst = st;  // stream of the copy-back
  if (split && !tu.copy_first) {
    if (!ctx->copy_stream)
      check_hip(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking), "hipStreamCreate");
    if (!ctx->ev_pipe)
      check_hip(hipEventCreateWithFlags(&ctx->ev_pipe, hipEventDisableTiming), "hipEventCreate");
    cst = ctx->copy_stream;
    check_hip(hipEventRecord(ctx->ev_pipe, st), "hipEventRecord");
    check_hip(hipStreamWaitEvent(cst, ctx->ev_pipe, 0), "hipStreamWaitEvent");
  }
  struct Drain {  // an error in the frame stage leaves no copy or kernel in flight
    hipStream_t a, b;
    bool on;
    ~Drain() {
      if (!on) return;
      (void)hipStreamSynchronize(a);
      if (b != a) (void)hipStreamSynchronize(b);
    }
  } drain{st, cst, overlap};
  uint8_t* hb = nullptr;
  std::function<void(u64)> ready;
  if (host_pre) {
    ctx->h_out.ensure(out_total + 64);
    hb = ctx->h_out.as<uint8_t>();
    if (overlap) {
      // Copied: [c0, c1) of the stream buffer; a chunk ending at `end`
      // (relative to c0) waits for the slice holding byte end - 1.
      const u64 slice = 1ull << tu.copy_slice_log;
      const u64 c0 = g[0].base + head, c1 = g[0].base + g[0].total;
      const u64 nsl = (c1 - c0 + slice - 1) / slice;
      while (ctx->slice_ev.size() < nsl) {
        hipEvent_t e;
        check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        ctx->slice_ev.push_back(e);
      }
      for (u64 i = 0; i < nsl; ++i) {
        const u64 a = c0 + i * slice, m = std::min(slice, c1 - a);
        check_hip(hipMemcpyAsync(hb + a, reinterpret_cast<const uint8_t*>(ow) + a, m,
                                 hipMemcpyDeviceToHost, cst), "stream D2H");
        check_hip(hipEventRecord(ctx->slice_ev[i], cst), "hipEventRecord");
      }
      ready = [ctx, slice](u64 end) {
        check_hip(hipEventSynchronize(ctx->slice_ev[(end - 1) / slice]), "hipEventSynchronize");
      };
    } else {
      check_hip(hipMemcpyAsync(hb, ow, out_total, hipMemcpyDeviceToHost, st), "stream D2H");
    }
    ctx->copied();
  }

  // Frames built on the device: whole XNYZ_FRAME_ZSTD_GPU segments, or the
  // head of a split frame (its header and last block are patched when
  // assembling).
  struct DevFrame {
    u64 src, len;
  };
  std::vector<DevFrame> dfr;
  for (u32 s : gsegs) dfr.push_back({g[s].base, g[s].total});
  if (split) dfr.push_back({g[0].base, head});
  std::vector<ZFrame> zf;
  const u64* zpos = nullptr;  // block positions in the device-built frames (pinned)
  uint8_t* zb = nullptr;
  u32 znblk = 0;
  if (!dfr.empty()) {
    // Block and frame tables staged in pinned memory: the copies stay
    // asynchronous, so the host is not held until the stream drains.
    u64 nb = 0;
    for (const DevFrame& d : dfr) nb += (d.len + xz::kBlock - 1) / xz::kBlock;
    znblk = (u32)nb;
    const size_t tb = (nb * sizeof(ZBlk) + 255) & ~size_t(255);
    const size_t tbytes = tb + dfr.size() * sizeof(ZFrame);
    ctx->h_ztab.ensure(tbytes);
    ZBlk* zk = ctx->h_ztab.as<ZBlk>();
    u64 bound = 0;
    u32 k = 0;
    for (u32 f = 0; f < (u32)dfr.size(); ++f) {
      const u32 first = k;
      // the encoder's kernels load blocks 16 bytes at a time
      if (dfr[f].src & 15u) throw Error(XNYZ_EINTERNAL, "device frame source not 16-byte aligned");
      for (u64 o = 0; o < dfr[f].len; o += xz::kBlock)
        zk[k++] = ZBlk{dfr[f].src + o, (u32)std::min<u64>(xz::kBlock, dfr[f].len - o), f};
      zf.push_back(ZFrame{dfr[f].len, first, k - first});
      bound += 13 + dfr[f].len + 3 * (k - first);
    }
    std::memcpy(ctx->h_ztab.as<uint8_t>() + tb, zf.data(), zf.size() * sizeof(ZFrame));
    ctx->ztab.ensure(tbytes + 256);
    check_hip(hipMemcpyAsync(ctx->ztab.p, ctx->h_ztab.p, tbytes, hipMemcpyHostToDevice, st),
              "H2D");
    ctx->copied();
    ctx->zscr.ensure(zframe_scratch_bytes(znblk));
    ctx->zout.ensure(bound + 64);
    ctx->zpos.ensure((znblk + 1) * sizeof(u64) + 64);
    ctx->h_zpos.ensure((znblk + 1) * sizeof(u64));
    zpos = ctx->h_zpos.as<u64>();
    if (!ctx->ev_sizes)
      check_hip(hipEventCreateWithFlags(&ctx->ev_sizes, hipEventDisableTiming), "hipEventCreate");
    // block positions come back (and ev_sizes fires) before the blocks are
    // gathered, so the host can place its own bytes meanwhile
    zframe_encode(reinterpret_cast<const uint8_t*>(ow), ctx->ztab.as<ZBlk>(), znblk,
                  reinterpret_cast<const ZFrame*>(ctx->ztab.as<uint8_t>() + tb), ctx->zscr.p,
                  ctx->zout.as<uint8_t>(), ctx->zpos.as<u64>(), ar, st, ctx->h_zpos.as<u64>(),
                  ctx->ev_sizes);
  }
  // Brings the device-built frames to host memory (after the host frames
  // when the frame is split, so the two run side by side).
  auto fetch_dev_frames = [&] {
    check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    const u64 total = zpos[znblk] + 13 * zf.size();
    ctx->h_zout.ensure(total + 64);
    zb = ctx->h_zout.as<uint8_t>();
    check_hip(hipMemcpyAsync(zb, ctx->zout.p, total, hipMemcpyDeviceToHost, st), "frames D2H");
    check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
  };
  if (!dfr.empty() && !split) fetch_dev_frames();
  if (!overlap) ctx->sync();

  const double t_dev = now_ms();
  if (co.pre) {
    parallel_for(S, per_seg_threads ? threads : 1, [&](size_t s) {
      co.pre[s].data = xmalloc(g[s].total);
      std::memcpy(co.pre[s].data, hb + g[s].base, g[s].total);
      co.pre[s].len = g[s].total;
    });
  }
  if (co.frames || co.dst) {
    if (!ctx->fs) ctx->fs = new FrameState();
    FrameState& fs = *ctx->fs;
    std::vector<u64> zoff(S + 1, 0);
    std::vector<int> zthreads(S);
    for (u32 s = 0; s < S; ++s) {
      const xnyz_opts& o = opts[s];
      zthreads[s] = per_seg_threads ? 1 : host_threads(o.zstd_workers > 0 ? o.zstd_workers : threads);
      // A split frame's tail always takes the chunked block path.
      const u64 need = (split && s == 0)
                           ? 16 + (g[s].total - head + chunk - 1) / chunk * zchunk_bound(chunk)
                           : frame_scratch(g[s].total, o.frame_mode, zthreads[s], chunk);
      zoff[s + 1] = zoff[s] + ((need + 63) & ~u64(63));
    }
    if (fs.buf.size() < zoff[S]) fs.buf.resize(zoff[S]);
    // Anchored split frame (caller memory with room for it): the host's
    // blocks are placed at out + anchor while they are compressed, anchor
    // being the bound of the device-built part, and that part lands right
    // in front of them once its size is known -- the frame then starts
    // past `out`, and no host block is copied after the device is done.
    u64 anchor = 0;
    if (split && co.dst) {
      const u64 dev_bound = 13 + head + 3 * ((head + xz::kBlock - 1) / xz::kBlock);
      const u64 host_bound = (g[0].total - head + chunk - 1) / chunk * zchunk_bound(chunk);
      if (co.cap >= dev_bound + host_bound) anchor = dev_bound;
    }
    // Every worker's context exists before the workers start (no resizing
    // while they read the table).
    fs.ensure_workers(per_seg_threads ? (size_t)std::min<u64>(S, (u64)threads) : 1);
    std::vector<Pieces> outs(S);
    parallel_for_w(S, per_seg_threads ? threads : 1, [&](size_t s, size_t w) {
      const uint8_t* b = hb ? hb + g[s].base : nullptr;
      const u64 len = g[s].total;
      uint8_t* z = fs.buf.data() + zoff[s];
      switch (opts[s].frame_mode) {
        case XNYZ_FRAME_ZSTD_GPU: {
          const u32 f = (u32)(std::lower_bound(gsegs.begin(), gsegs.end(), (u32)s) - gsegs.begin());
          const u64 st0 = zpos[zf[f].first] + 13ull * f;
          const u64 en = zpos[zf[f].first + zf[f].nblk] + 13ull * (f + 1);
          outs[s] = {Piece{zb + st0, en - st0}};
          break;
        }
        case XNYZ_FRAME_NONE: outs[s] = {Piece{b, len}}; break;
        case XNYZ_FRAME_ZSTD3_MT:
          outs[s] = split ? zstd_chunked(fs, fs.cctx[w], b + head, len - head, zthreads[s], z,
                                         chunk, &ready, 0, false,
                                         anchor ? co.dst + anchor : nullptr)
                          : zstd_chunked(fs, fs.cctx[w], b, len, zthreads[s], z, chunk,
                                         overlap ? &ready : nullptr, len, true);
          break;
        default: outs[s] = zstd_single(fs.cctx[w], b, len, z, zoff[s + 1] - zoff[s]);
      }
    });
    if (split) {
      // The device-built blocks open the frame.  Their total size is known
      // before the blocks themselves arrive, so the host's blocks are put in
      // place while the device gathers and copies its own.
      check_hip(hipEventSynchronize(ctx->ev_sizes), "hipEventSynchronize");
      const u64 head_bytes = 13 + zpos[znblk], host_bytes = pieces_size(outs[0]);
      const u64 n = head_bytes + host_bytes;
      uint8_t* dst;
      if (co.frames) {
        dst = xmalloc(n);
        co.frames[0].data = dst;
        co.frames[0].len = n;
      } else if (anchor) {
        if (head_bytes > anchor) throw Error(XNYZ_EINTERNAL, "device frame exceeds its bound");
        co.dst_off[0] = anchor - head_bytes;
        co.dst_off[1] = anchor + host_bytes;
        dst = co.dst + co.dst_off[0];
      } else {
        co.dst_off[0] = 0;
        co.dst_off[1] = n;
        if (n > co.cap)
          throw Error(XNYZ_ENOSPC, "output buffer too small: " + std::to_string(n) +
                                       " bytes needed");
        dst = co.dst;
      }
      // the device-built blocks come back (behind the gather) while the
      // host's blocks are copied into place: straight into the output when
      // it is pinned, else through the context's staging buffer
      const bool direct = co.dst && pinned_range(dst, n);
      if (!direct) ctx->h_zout.ensure(head_bytes + 64);
      zb = direct ? dst : ctx->h_zout.as<uint8_t>();
      check_hip(hipMemcpyAsync(zb, ctx->zout.p, head_bytes, hipMemcpyDeviceToHost, st), "frames D2H");
      ctx->copied();
      if (!anchor) copy_pieces(outs[0], dst + head_bytes, threads);
      check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
      const uint64_t fcs = g[0].total;  // the header covers the whole stream
      std::memcpy(zb + 5, &fcs, 8);
      zb[zpos[znblk - 1] + 13] &= 0xFE;  // the host's blocks follow the last one
      if (!direct) copy_pieces({Piece{zb, head_bytes}}, dst, threads);
    } else if (co.frames) {
      parallel_for(S, per_seg_threads ? threads : 1, [&](size_t s) {
        const u64 n = pieces_size(outs[s]);
        co.frames[s].data = xmalloc(n);
        co.frames[s].len = n;
        copy_pieces(outs[s], co.frames[s].data, per_seg_threads ? 1 : threads);
      });
    } else {
      co.dst_off[0] = 0;
      for (u32 s = 0; s < S; ++s) co.dst_off[s + 1] = co.dst_off[s] + pieces_size(outs[s]);
      if (co.dst_off[S] > co.cap)
        throw Error(XNYZ_ENOSPC, "output buffer too small: " + std::to_string(co.dst_off[S]) +
                                     " bytes needed");
      parallel_for(S, per_seg_threads ? threads : 1, [&](size_t s) {
        copy_pieces(outs[s], co.dst + co.dst_off[s], per_seg_threads ? 1 : threads);
      });
    }
  }
  if (overlap && cst != st) check_hip(hipStreamSynchronize(cst), "hipStreamSynchronize");
  if (overlap) ctx->sync();
  const double t_end = now_ms();
  fill_stats(t_dev, t_end);
  return XNYZ_OK;
}

// ---------------------------------------------------------------------------
// Device decoder (BlockDecompressor[RLE]<float>::decompress,
// source/composer/block_decompressor.hpp:31-192,
// block_decompressor_rle.hpp:32-207): the host reads the header, the section
// records and the metas; the device decodes every section and rebuilds the
// points (decode_dev.hip).

struct DevSec {  // one section of the assembled stream
  u64 meta_at, meta_len, data_at, data_len;
};

// Host-side code tables of one entropy-coded section, in the layout CodeDesc
// points into (uploaded into one table buffer).
struct HostCode {
  CodeDesc d{};
  std::vector<u32> lut, tree;
  std::vector<u64> values;
};

u32 rd_be32(const uint8_t* m) {
  return uint32_t(m[0]) << 24 | uint32_t(m[1]) << 16 | uint32_t(m[2]) << 8 | uint32_t(m[3]);
}

// huffman_encoder.hpp:126-146 (meta), :172-201 (load_meta): Exp-Golomb
// header, or u32 BE entry count and entries (symbol, u8 length, u32 BE
// bits).  sym_bytes: 8 (u64, big-endian), 1 (u8) or 4 (int, raw LE).
HostCode huffman_code(const uint8_t* m, u64 ml, u32 sym_bytes, bool is_signed, u64 max_symbol) {
  HostCode h;
  if (ml >= 4 && m[0] == 0xE1 && m[1] == 0xC1 && m[2] == 0x01) {
    if (m[3] != 0x01) throw Error(XNYZ_ECORRUPT, "unknown Huffman meta mode");
    h.d.kind = kCodeEG;
    h.d.k = ml > 4 ? m[4] : 0;
    if (h.d.k > 62) throw Error(XNYZ_ECORRUPT, "Exp-Golomb k too large");
    h.d.zigzag = is_signed;
    return h;
  }
  if (ml < 4) throw Error(XNYZ_ECORRUPT, "Huffman meta too short");
  const u32 entries = rd_be32(m);
  const u64 esz = sym_bytes + 5;
  if (4 + (u64)entries * esz > ml) throw Error(XNYZ_ECORRUPT, "Huffman meta truncated");
  if (!entries) throw Error(XNYZ_ECORRUPT, "Huffman meta without entries");
  struct E {
    u64 sym;
    u32 len, bits;
  };
  std::vector<E> es(entries);
  u32 maxlen = 0;
  for (u32 i = 0; i < entries; ++i) {
    const uint8_t* e = m + 4 + i * esz;
    u64 v = 0;
    if (sym_bytes == 8) {
      for (int k = 0; k < 8; ++k) v = (v << 8) | e[k];
    } else if (sym_bytes == 4) {
      int32_t x;
      std::memcpy(&x, e, 4);
      v = (u64)(int64_t)x;
    } else {
      v = e[0];
    }
    const u32 len = e[sym_bytes];
    const u32 bits = rd_be32(e + sym_bytes + 1);
    if (!is_signed && v > max_symbol) throw Error(XNYZ_EUNSUPPORTED, "Huffman symbol out of range");
    if (len == 0 || len > 32) throw Error(XNYZ_EUNSUPPORTED, "Huffman code length outside 1..32");
    es[i] = E{v, len, bits};
    maxlen = std::max(maxlen, len);
  }
  if (entries == 1) {  // huffman_encoder.hpp:211-215: every symbol is this one
    h.d.kind = kCodeConst;
    h.d.const_value = es[0].sym;
    return h;
  }
  h.d.kind = kCodeHuf;
  // Symbols below 2^24 go into the table entries themselves (the decoder
  // then needs no second lookup per value).
  bool dv = true;
  for (const E& e : es) dv &= e.sym < (1ull << 24);
  h.d.direct_values = dv;
  const u32 lb = std::min<u32>(kLutBits, maxlen);
  h.d.lb = lb;
  h.lut.assign(1u << lb, 0u);
  h.values.resize(entries);
  auto new_node = [&]() {
    h.tree.push_back(kTreeNull);
    h.tree.push_back(kTreeNull);
    return (u32)(h.tree.size() / 2 - 1);
  };
  for (u32 i = 0; i < entries; ++i) {
    const E& e = es[i];
    h.values[i] = e.sym;
    if (e.len <= lb) {
      const u32 lo = e.bits << (lb - e.len), hi = (e.bits + 1) << (lb - e.len);
      for (u32 x = lo; x < hi; ++x) {
        if (h.lut[x]) throw Error(XNYZ_ECORRUPT, "Huffman codes are not prefix-free");
        h.lut[x] = (e.len << 24) | (dv ? (u32)e.sym : i);
      }
    } else {
      const u32 pre = e.bits >> (e.len - lb);
      u32 node;
      if (!h.lut[pre]) {
        node = new_node();
        h.lut[pre] = node;  // length 0: continue in the tree
        if (node == 0) {    // node 0 would read as "no entry": keep it unused
          node = new_node();
          h.lut[pre] = node;
        }
      } else if ((h.lut[pre] >> 24) == 0) {
        node = h.lut[pre] & 0xFFFFFFu;
      } else {
        throw Error(XNYZ_ECORRUPT, "Huffman codes are not prefix-free");
      }
      for (u32 b = e.len - lb; b-- > 0;) {
        const u32 bit = (e.bits >> b) & 1;
        u32& ch = h.tree[2 * node + bit];
        if (b == 0) {
          if (ch != kTreeNull) throw Error(XNYZ_ECORRUPT, "Huffman codes are not prefix-free");
          ch = kTreeLeaf | (dv ? (u32)e.sym : i);
        } else {
          if (ch == kTreeNull) {
            const u32 nn = new_node();
            h.tree[2 * node + bit] = nn;
            node = nn;
          } else if (ch & kTreeLeaf) {
            throw Error(XNYZ_ECORRUPT, "Huffman codes are not prefix-free");
          } else {
            node = ch;
          }
        }
      }
    }
  }
  if (h.tree.empty()) h.tree.assign(2, kTreeNull);
  return h;
}

int decompress_device_impl(xnyz_ctx* ctx, const uint8_t* in, u64 len, const xnyz_opts* o,
                           float* d_out, u64 cap_points, uint64_t* n_points,
                           xnyz_buf* host_out, double* ms_device) {
  if (!in || !o) throw Error(XNYZ_EINVAL, "null argument");
  if (o->quantizer != XNYZ_CUBE && o->quantizer != XNYZ_TRUNCATED_OCTAHEDRON)
    throw Error(XNYZ_EINVAL, "bad quantizer");
  std::vector<uint8_t> raw;
  const uint8_t* buf = in;
  u64 blen = len;
  if (o->frame_mode != XNYZ_FRAME_NONE) {  // source/composer/utils.hpp:13-29
    const unsigned long long fs = ZSTD_getFrameContentSize(in, len);
    if (fs == ZSTD_CONTENTSIZE_ERROR) throw Error(XNYZ_ECORRUPT, "Not a zstd frame");
    if (fs == ZSTD_CONTENTSIZE_UNKNOWN) throw Error(XNYZ_ECORRUPT, "Original size unknown");
    raw.resize(fs);
    const size_t d = ZSTD_decompress(raw.data(), fs, in, len);
    if (ZSTD_isError(d)) throw Error(XNYZ_ECORRUPT, ZSTD_getErrorName(d));
    if (d != fs) throw Error(XNYZ_ECORRUPT, "Decompressed size mismatch");
    buf = raw.data();
    blen = raw.size();
  }
  // ---- header and section records (block_decompressor.hpp:37-60) ------------
  u64 at = 0;
  auto get = [&](void* dst, u64 n) {
    if (blen - at < n || at > blen) throw Error(XNYZ_ECORRUPT, "truncated stream");
    std::memcpy(dst, buf + at, n);
    at += n;
  };
  u64 num = 0, params = 0;
  get(&num, 8);
  get(&params, 8);
  if (params > (blen - at) / 4) throw Error(XNYZ_ECORRUPT, "truncated stream");
  at += params * 4;
  int32_t rng[3];
  get(rng, 12);
  uint8_t hil = 0, direct = 0;
  get(&hil, 1);
  get(&direct, 1);
  auto section = [&]() {
    DevSec r;
    get(&r.data_len, 8);
    get(&r.meta_len, 8);
    r.meta_at = at;
    if (r.meta_len > blen - at) throw Error(XNYZ_ECORRUPT, "truncated section");
    at += r.meta_len;
    r.data_at = at;
    if (r.data_len > blen - at) throw Error(XNYZ_ECORRUPT, "truncated section");
    at += r.data_len;
    return r;
  };
  const bool rle = o->rle != 0;
  DevSec s_blk{}, s_cnt{}, s_gr{}, s_huf{}, s_qgr{}, s_quad{}, s_rle{};
  if (!direct) {
    s_blk = section();
    s_cnt = section();
  }
  s_gr = section();
  s_huf = section();
  if (!direct) {
    s_qgr = section();
    s_quad = section();
  }
  if (rle) s_rle = section();
  float off[3];
  get(off, 12);
  if (num == 0) throw Error(XNYZ_EUNSUPPORTED, "stream of zero points");
  if (num >= 0xFFFFFFFFull) throw Error(XNYZ_EUNSUPPORTED, "more than 2^32-2 coded points");
  if (!direct && s_qgr.meta_len + s_qgr.data_len)
    throw Error(XNYZ_EUNSUPPORTED, "quads patches (a quad delta above 254)");
  auto le32_at = [&](u64 p) {
    uint32_t v;
    std::memcpy(&v, buf + p, 4);
    return v;
  };
  // BitPacker metas (bitpacking_encoder.hpp:27-32): u32 LE n, u8 width
  u64 B = 0;
  u32 w1 = 0, w2 = 0;
  if (!direct) {
    if (s_blk.meta_len != 5 || s_cnt.meta_len != 5) throw Error(XNYZ_ECORRUPT, "BitPacker meta");
    B = le32_at(s_blk.meta_at);
    w1 = buf[s_blk.meta_at + 4];
    w2 = buf[s_cnt.meta_at + 4];
    if (le32_at(s_cnt.meta_at) != B || w1 == 0 || w1 > 64 || w2 == 0 || w2 > 64 ||
        B * w1 > s_blk.data_len * 8 || B * w2 > s_cnt.data_len * 8 || B == 0 || B > num)
      throw Error(XNYZ_ECORRUPT, "BitPacker sections");
  }
  // Golomb-Rice meta (golomb_rice_encoder.hpp:126-133): u32 LE n, u8 k
  u64 P = 0;
  HostCode gr;
  if (s_gr.meta_len || s_gr.data_len) {
    if (s_gr.meta_len != 5) throw Error(XNYZ_ECORRUPT, "Golomb-Rice meta");
    P = le32_at(s_gr.meta_at);
    gr.d.kind = kCodeGR;
    gr.d.k = buf[s_gr.meta_at + 4];
    if (gr.d.k > 62) throw Error(XNYZ_ECORRUPT, "Golomb-Rice k too large");
  }
  HostCode hs = huffman_code(buf + s_huf.meta_at, s_huf.meta_len, 8, false, 255);
  HostCode hq;
  if (!direct) hq = huffman_code(buf + s_quad.meta_at, s_quad.meta_len, 1, false, 255);
  HostCode hc;
  if (rle) hc = huffman_code(buf + s_rle.meta_at, s_rle.meta_len, 4, true, 0);

  // ---- device buffers -------------------------------------------------------
  hipStream_t st = ctx->stream;
  const double t0 = now_ms();
  ctx->h_dflag.ensure(64);
  ctx->dstream.ensure(blen + 64);
  // The stream crosses PCIe in pieces on the copy stream: the Golomb-Rice
  // patches first (their decoding takes longest), the sections after them
  // (Huffman symbols, quads, RLE counts), the block counts, and last the
  // header and block ids, which only the final pass needs.  Each section's
  // decoding waits only for the piece that holds it, so decoding overlaps the
  // upload of what follows.
  if (!ctx->copy_stream)
    check_hip(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking), "hipStreamCreate");
  for (hipEvent_t& e : ctx->dec_ev)
    if (!e) check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  check_hip(hipEventRecord(ctx->dec_ev[4], st), "hipEventRecord");  // the stream buffer is free
  check_hip(hipStreamWaitEvent(ctx->copy_stream, ctx->dec_ev[4], 0), "hipStreamWaitEvent");
  const u64 gr0 = s_gr.data_at, gr1 = s_gr.data_at + s_gr.data_len;
  const u64 cn0 = direct ? gr0 : s_cnt.data_at;
  const u64 pieces[4][2] = {{gr0, gr1}, {gr1, blen}, {cn0, gr0}, {0, cn0}};
  for (int i = 0; i < 4; ++i) {
    const u64 a0 = pieces[i][0], a1 = pieces[i][1];
    if (a1 > a0)
      check_hip(hipMemcpyAsync(ctx->dstream.as<uint8_t>() + a0, buf + a0, a1 - a0,
                               hipMemcpyHostToDevice, ctx->copy_stream),
                "stream H2D");
    check_hip(hipEventRecord(ctx->dec_ev[i], ctx->copy_stream), "hipEventRecord");
  }
  auto arrived = [&](int piece) {
    check_hip(hipStreamWaitEvent(st, ctx->dec_ev[piece], 0), "hipStreamWaitEvent");
    ctx->copied();  // per-kernel timing: the next interval starts after the wait
  };
  // The host stream (the caller's, or the decompressed frame) must outlive
  // the copies, also when decoding stops early with an error.
  struct UploadDone {
    hipStream_t s;
    ~UploadDone() { (void)hipStreamSynchronize(s); }
  } upload_done{ctx->copy_stream};
  ctx->copied();
  launch_zero_u8(ctx->dstream.as<uint8_t>() + blen, 64, st);
  // code tables of the sections, one buffer
  std::vector<uint8_t> tab;
  auto put_tab = [&](HostCode& h) {
    auto add = [&](const void* p, size_t n) {
      const size_t a = (tab.size() + 255) & ~size_t(255);
      tab.resize(a + n);
      std::memcpy(tab.data() + a, p, n);
      return a;
    };
    size_t lut = 0, tree = 0, vals = 0;
    if (h.d.kind == kCodeHuf) {
      lut = add(h.lut.data(), h.lut.size() * 4);
      tree = add(h.tree.data(), h.tree.size() * 4);
      vals = add(h.values.data(), h.values.size() * 8);
    }
    return std::array<size_t, 3>{lut, tree, vals};
  };
  const auto ts = put_tab(hs);
  const auto tq = put_tab(hq);
  const auto tc = put_tab(hc);
  // through pinned memory and the copy kernel: the copy engine is busy with
  // the stream's upload
  ctx->dtab.ensure(tab.size() + 256);
  if (!tab.empty()) {
    ctx->h_big.ensure(tab.size() + 16);
    std::memcpy(ctx->h_big.p, tab.data(), tab.size());
    launch_copy_words(ctx->h_big.p, ctx->dtab.p, tab.size(), st);
  }
  auto bind = [&](HostCode& h, const std::array<size_t, 3>& t, const DevSec& sec) {
    h.d.base_bit = sec.data_at * 8;
    h.d.bits = sec.data_len * 8;
    if (h.d.kind == kCodeHuf) {
      const uint8_t* b0 = ctx->dtab.as<uint8_t>();
      h.d.lut = reinterpret_cast<const u32*>(b0 + t[0]);
      h.d.tree = reinterpret_cast<const u32*>(b0 + t[1]);
      h.d.values = reinterpret_cast<const u64*>(b0 + t[2]);
    }
  };
  bind(hs, ts, s_huf);
  if (!direct) bind(hq, tq, s_quad);
  if (rle) bind(hc, tc, s_rle);
  bind(gr, {0, 0, 0}, s_gr);
  // scratch: the largest section's sub-sequence tables, scans, tile states
  u64 maxbits = std::max({s_huf.data_len, s_quad.data_len, s_rle.data_len, s_gr.data_len}) * 8;
  const u64 nsub = maxbits / kSub + 2;
  const size_t scratch = (size_t)nsub * 48 + lcp_scratch_bytes(num) + (num / 4096 + 64) * 16 +
                         (8u << 20);
  ctx->darena.ensure(scratch);
  Arena ar{ctx->darena.as<char>(), ctx->darena.cap, 0};
  const uint8_t* dS = ctx->dstream.as<uint8_t>();
  DecScratch ws{ctx->h_dflag.as<u32>()};

  // ---- sections, in upload order ---------------------------------------------------
  arrived(0);
  ctx->dps.ensure((P + 1) * 8 + 16);
  u64* dps = ctx->dps.as<u64>();
  if (P) {
    ctx->dpatch.ensure(P * 8 + 16);
    decode_section<u64>(dS, gr.d, P, ctx->dpatch.as<u64>(), ws, ar, st);
    scan_excl_u64(ctx->dpatch.as<u64>(), dps, P, dps + P, ar, st);
  } else {
    launch_zero_u8(reinterpret_cast<uint8_t*>(dps), 8, st);
  }
  arrived(1);
  ctx->dsym.ensure(num + 16);
  decode_section<uint8_t>(dS, hs.d, num, ctx->dsym.as<uint8_t>(), ws, ar, st);
  if (!direct) {
    ctx->dquad.ensure(num + 16);
    decode_section<uint8_t>(dS, hq.d, num, ctx->dquad.as<uint8_t>(), ws, ar, st);
  }
  u64 total_pts = num;
  const u64* out_at = nullptr;
  const u64* counts = nullptr;
  if (rle) {  // counts, then each cell's first output point
    ctx->dcnt.ensure(num * 16 + 64);
    u64* c = ctx->dcnt.as<u64>();
    u64* at2 = c + num;
    decode_section<u64>(dS, hc.d, num, c, ws, ar, st);
    u64* total = ar.take<u64>(1);
    scan_excl_u64(c, at2, num, total, ar, st);
    launch_copy_words(total, ws.host_flag, 8, st);
    ctx->sync();
    std::memcpy(&total_pts, ws.host_flag, 8);
    if (total_pts >= (1ull << 40)) throw Error(XNYZ_EUNSUPPORTED, "RLE counts beyond 2^40 points");
    out_at = at2;
    counts = c;
  }
  // block counts (block_decompressor.hpp:163-165: prefix sums of the
  // BitPacker deltas), then the first element of each block
  u64* bid = nullptr;
  if (!direct) {
    arrived(2);
    ctx->dblk.ensure(B * 8 * 4 + 64);
    bid = ctx->dblk.as<u64>();
    u64* bcnt = bid + B;
    u64* tmp = bcnt + B;
    u64* bstart = tmp + B;
    launch_bitunpack(dS, s_cnt.data_at * 8, w2, B, bcnt, st);
    scan_excl_u64(bcnt, tmp, B, nullptr, ar, st);
    launch_add_u64(bcnt, tmp, B, st);
    u64* total = ar.take<u64>(1);
    scan_excl_u64(bcnt, bstart, B, total, ar, st);
    launch_copy_words(total, ws.host_flag, 8, st);
    ctx->sync();
    u64 tot;
    std::memcpy(&tot, ws.host_flag, 8);
    if (tot != num) throw Error(XNYZ_ECORRUPT, "block counts do not add up to the points");
    ctx->dhead.ensure(num + 16);
    launch_zero_u8(ctx->dhead.as<uint8_t>(), num, st);
    launch_mark_heads(bstart, B, ctx->dhead.as<uint8_t>(), st);
  }
  if (n_points) *n_points = total_pts;
  float* dst = d_out;
  if (d_out && total_pts > cap_points)
    throw Error(XNYZ_ENOSPC, "output holds " + std::to_string(cap_points) + " points, " +
                                 std::to_string(total_pts) + " needed");
  if (!dst) {
    ctx->dpts.ensure(total_pts * 12 + 64);
    dst = ctx->dpts.as<float>();
  }
  LcpArgs a{};
  a.n = num;
  a.direct = direct;
  a.hilbert = hil;
  a.to = o->quantizer == XNYZ_TRUNCATED_OCTAHEDRON;
  a.rle = rle;
  const int mr = std::max({rng[0], rng[1], rng[2]});
  u32 bits = 0;
  for (int t = mr; t > 0; t >>= 1) ++bits;
  a.bits = bits ? bits : 1;
  if (direct && hil && a.bits > 21)
    throw Error(XNYZ_EUNSUPPORTED, "direct-mode Hilbert code of more than 21 digits");
  a.nx = (u64)(uint32_t)rng[0] / 128 + 1;
  a.ny = (u64)(uint32_t)rng[1] / 128 + 1;
  a.b = o->l2_bound;
  a.rec_div = a.to ? float(std::sqrt(5.0)) : float(std::sqrt(3.0));
  a.rec_inv = 1.0f / a.rec_div;
  for (int k = 0; k < 3; ++k) a.off[k] = off[k];
  a.sym = ctx->dsym.as<uint8_t>();
  a.quad = direct ? nullptr : ctx->dquad.as<uint8_t>();
  a.head = direct ? nullptr : ctx->dhead.as<uint8_t>();
  a.ps = dps;
  a.np = P;
  a.total = total_pts;
  a.out_at = out_at;
  a.counts = counts;
  a.out = dst;
  char* lcp_scratch = ar.take<char>(lcp_scratch_bytes(num));
  launch_lcp_scan(a, lcp_scratch, st);
  // block ids: prefix sums of their deltas (block_decompressor.hpp:163-165)
  arrived(3);
  if (!direct) {
    u64* tmp = bid + 2 * B;
    launch_bitunpack(dS, s_blk.data_at * 8, w1, B, bid, st);
    scan_excl_u64(bid, tmp, B, nullptr, ar, st);
    launch_add_u64(bid, tmp, B, st);
    a.blkst = bid;
  }
  launch_lcp_apply(a, ctx->hinv, lcp_scratch, st);
  if (host_out) {
    uint8_t* p = xmalloc(total_pts * 12 + 16);
    host_out->data = p;
    host_out->len = total_pts * 12;
    check_hip(hipMemcpyAsync(p, dst, total_pts * 12, hipMemcpyDeviceToHost, st), "points D2H");
  }
  ctx->sync();
  if (ms_device) *ms_device = now_ms() - t0;
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

// Batch offsets relative to the first cloud (advancing d_xyz to it).
std::vector<u64> rebase(const uint64_t* offsets, uint32_t count, const float*& d_xyz) {
  std::vector<u64> off(offsets, offsets + count + 1);
  for (u32 s = 0; s < count; ++s)
    if (off[s + 1] < off[s]) throw Error(XNYZ_EINVAL, "offsets must not decrease");
  if (off[0] != 0) {
    for (auto& v : off) v -= offsets[0];
    d_xyz += 3 * offsets[0];
  }
  return off;
}

// Exhaustive mode search of every segment: one statistics pass per curve x
// RLE variant, then the (k, width) grid from the size model.
void search_impl(xnyz_ctx* ctx, const float* d_xyz, const std::vector<u64>& off,
                 const xnyz_opts* opts, xnyz_search_result* results) {
  const u32 count = (u32)(off.size() - 1);
  for (u32 s = 0; s < count; ++s) {
    std::memset(&results[s], 0, sizeof(results[s]));
    results[s].best_bytes = ~0ull;
  }
  std::vector<xnyz_opts> vo(opts, opts + count);
  for (int var = 0; var < 4; ++var) {  // 2 * is_hilbert + rle
    for (u32 s = 0; s < count; ++s) {
      vo[s] = opts[s];
      vo[s].is_hilbert = var >> 1;
      vo[s].rle = var & 1;
      vo[s].mode_flags = 0;
    }
    std::vector<ModeStats> ms;
    CallOut co{nullptr, nullptr, nullptr, nullptr};
    co.search = &ms;
    compress_impl(ctx, d_xyz, off, vo.data(), co, now_ms());
    parallel_for(count, host_threads(0), [&](size_t s) {
      ModeChoice ref{}, best{};
      search_modes(ms[s], ref, best);
      xnyz_search_result& r = results[s];
      r.variant[var] = xnyz_mode{var >> 1, var & 1, best.k, best.w};
      r.variant_bytes[var] = best.bytes;
      if ((opts[s].is_hilbert != 0) == (var >> 1) && (opts[s].rle != 0) == (var & 1))
        r.reference_bytes = ref.bytes;
    });
  }
  for (u32 s = 0; s < count; ++s) {
    xnyz_search_result& r = results[s];
    for (int var = 0; var < 4; ++var)
      if (r.variant_bytes[var] < r.best_bytes) {
        r.best_bytes = r.variant_bytes[var];
        r.best = r.variant[var];
      }
  }
}

xnyz_opts with_mode(xnyz_opts o, const xnyz_mode& m) {
  o.is_hilbert = m.is_hilbert;
  o.rle = m.rle;
  o.gr_k = m.gr_k;
  o.pfd_width = m.pfd_width;
  o.mode_flags = XNYZ_MODE_GR_K | XNYZ_MODE_PFD_WIDTH;
  return o;
}

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
    c->name_buffers();
    for (xnyz_tuning& t : c->tuning) t = default_tuning();
    try {
      use_device(c);
      check_hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
      c->hs = build_hilbert_sm();
      c->hinv = build_hilbert_inv();
      c->na_dev = noalloc_scratch(device);
      c->na_host.ensure(kNoAllocHostBytes);
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
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);  // pending copies
    for (DBuf* b : ctx->dbufs()) b->release();
    for (HBuf* b : ctx->hbufs()) b->release();
    ctx->na_host.release();
    ctx->prof.destroy();
    ctx->stage2.clear();
    for (hipEvent_t e : ctx->slice_ev) (void)hipEventDestroy(e);
    if (ctx->ev_pipe) (void)hipEventDestroy(ctx->ev_pipe);
    if (ctx->ev_fixed) (void)hipEventDestroy(ctx->ev_fixed);
    if (ctx->ev_sizes) (void)hipEventDestroy(ctx->ev_sizes);
    for (hipEvent_t e : ctx->aev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->dec_ev)
      if (e) (void)hipEventDestroy(e);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    delete ctx->fs;
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
    validate(opts, n);
    use_device(ctx);
    return compress_impl(ctx, d_xyz, {0, n}, opts, CallOut{out_frame, out_prezstd, stats, perm},
                         now_ms());
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
    return compress_impl(ctx, ctx->upload.as<float>(), {0, n}, opts,
                         CallOut{out_frame, out_prezstd, stats, perm}, t0);
  });
}

int xnyz_compress_batch_device(xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets,
                               uint32_t count, const xnyz_opts* opts, xnyz_buf* frames,
                               xnyz_buf* prezstd, xnyz_stats* stats) {
  return guarded([&] {
    if (!ctx || !d_xyz || !offsets || !opts || !count)
      throw Error(XNYZ_EINVAL, "null argument or empty batch");
    std::vector<u64> off(offsets, offsets + count + 1);
    for (u32 s = 0; s < count; ++s)
      if (off[s + 1] < off[s]) throw Error(XNYZ_EINVAL, "offsets must not decrease");
    if (off[0] != 0) {  // segments are addressed from the first cloud
      for (auto& v : off) v -= offsets[0];
      d_xyz += 3 * offsets[0];
    }
    use_device(ctx);
    return compress_impl(ctx, d_xyz, off, opts, CallOut{frames, prezstd, stats, nullptr},
                         now_ms());
  });
}

int xnyz_compress_batch_device_into(xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets,
                                    uint32_t count, const xnyz_opts* opts, uint8_t* out,
                                    uint64_t cap, uint64_t* out_offsets, xnyz_stats* stats) {
  return guarded([&] {
    if (!ctx || !d_xyz || !offsets || !opts || !count || !out_offsets || (!out && cap))
      throw Error(XNYZ_EINVAL, "null argument or empty batch");
    std::vector<u64> off = rebase(offsets, count, d_xyz);
    use_device(ctx);
    CallOut co{nullptr, nullptr, stats, nullptr};
    co.dst = out;
    co.cap = cap;
    co.dst_off = out_offsets;
    return compress_impl(ctx, d_xyz, off, opts, co, now_ms());
  });
}

int xnyz_compress_batch_device_into_async(xnyz_ctx* ctx, const float* d_xyz,
                                          const uint64_t* offsets, uint32_t count,
                                          const xnyz_opts* opts, uint8_t* out, uint64_t cap,
                                          uint64_t* out_offsets, xnyz_stats* stats) {
  return guarded([&] {
    if (!ctx || !d_xyz || !offsets || !opts || !count || !out_offsets || !out || !cap)
      throw Error(XNYZ_EINVAL, "null argument or empty batch");
    std::vector<u64> off = rebase(offsets, count, d_xyz);
    use_device(ctx);
    CallOut co{nullptr, nullptr, stats, nullptr};
    co.dst = out;
    co.cap = cap;
    co.dst_off = out_offsets;
    co.async_out = true;
    return compress_impl(ctx, d_xyz, off, opts, co, now_ms());
  });
}

int xnyz_wait(xnyz_ctx* ctx) {
  return guarded([&] {
    if (!ctx) throw Error(XNYZ_EINVAL, "null context");
    use_device(ctx);
    if (ctx->copy_stream)
      check_hip(hipStreamSynchronize(ctx->copy_stream), "hipStreamSynchronize");
    ctx->abusy[0] = ctx->abusy[1] = false;
    return XNYZ_OK;
  });
}

int xnyz_compress_batch_device_noalloc(xnyz_ctx* ctx, const float* d_xyz,
                                       const uint64_t* offsets, uint32_t count,
                                       const xnyz_opts* opts, uint8_t* out, uint64_t cap,
                                       uint64_t* out_offsets, xnyz_stats* stats) {
  return guarded([&] {
    if (!ctx || !d_xyz || !offsets || !opts || !count || !out_offsets || (!out && cap))
      throw Error(XNYZ_EINVAL, "null argument or empty batch");
    if (count > XNYZ_NOALLOC_MAX_CLOUDS)
      throw Error(XNYZ_EUNSUPPORTED, "no-allocation calls take at most " +
                                         std::to_string(XNYZ_NOALLOC_MAX_CLOUDS) + " clouds");
    std::vector<u64> off = rebase(offsets, count, d_xyz);
    if (off[count] > XNYZ_NOALLOC_MAX_POINTS)
      throw Error(XNYZ_EUNSUPPORTED, "no-allocation calls take at most " +
                                         std::to_string(XNYZ_NOALLOC_MAX_POINTS) + " points");
    for (u32 s = 0; s < count; ++s)
      if (opts[s].frame_mode != XNYZ_FRAME_NONE)
        throw Error(XNYZ_EUNSUPPORTED, "no-allocation calls return assembled streams only "
                                       "(frame_mode XNYZ_FRAME_NONE)");
    use_device(ctx);
    // Every pooled buffer of the context is set aside for the call, so the
    // pipeline's requests all carve from the fixed regions; the context's own
    // buffers come back untouched afterwards, error or not.
    struct Swap {
      xnyz_ctx* c;
      std::vector<DBuf> d;
      std::vector<HBuf> h;
      Bump bump;
      explicit Swap(xnyz_ctx* ctx) : c(ctx) {
        for (DBuf* b : c->dbufs()) {
          d.push_back(*b);
          b->p = nullptr;
          b->cap = 0;
        }
        for (HBuf* b : c->hbufs()) {
          h.push_back(*b);
          b->p = nullptr;
          b->cap = 0;
        }
        // the array's last 64 KiB stay unused: loads that run a few words
        // past a buffer's end (as they may into an allocation's growth
        // margin) stay inside the array
        bump = Bump{static_cast<char*>(c->na_dev), kNoAllocScratchBytes - (64u << 10), 0,
                    c->na_host.as<char>(), kNoAllocHostBytes, 0};
        g_bump = &bump;
      }
      ~Swap() {
        g_bump = nullptr;
        size_t i = 0;
        for (DBuf* b : c->dbufs()) *b = d[i++];
        i = 0;
        for (HBuf* b : c->hbufs()) *b = h[i++];
      }
    } swap(ctx);
    CallOut co{nullptr, nullptr, stats, nullptr};
    co.dst = out;
    co.cap = cap;
    co.dst_off = out_offsets;
    co.noalloc = true;
    return compress_impl(ctx, d_xyz, off, opts, co, now_ms());
  });
}

int xnyz_fit_profile_device(xnyz_ctx* ctx, const float* d_xyz, uint64_t n, const xnyz_opts* opts,
                            xnyz_buf* out_frame, xnyz_buf* out_prezstd, xnyz_buf* profile) {
  return guarded([&] {
    if (!ctx || !d_xyz || !profile) throw Error(XNYZ_EINVAL, "null argument");
    validate(opts, n);
    if (opts->rle) throw Error(XNYZ_EUNSUPPORTED, "shared codebooks are not defined for RLE");
    use_device(ctx);
    xnyz_buf scratch{};
    CallOut co{out_frame ? out_frame : &scratch, out_prezstd, nullptr, nullptr};
    std::vector<uint8_t> prof;
    co.shared = 1;
    co.profile_out = &prof;
    compress_impl(ctx, d_xyz, {0, n}, opts, co, now_ms());
    if (scratch.data) xnyz_free(&scratch);
    profile->data = xmalloc(prof.size());
    std::memcpy(profile->data, prof.data(), prof.size());
    profile->len = prof.size();
    return XNYZ_OK;
  });
}

namespace {

// deserialize_profile (block_encoding_profile.hpp:48-70): the repos and the
// quads metas.
void read_profile(const uint8_t* profile, uint64_t profile_len, std::vector<uint8_t>* metas) {
  u64 pos = 0;
  for (int k = 0; k < 2; ++k) {
    if (pos + 4 > profile_len) throw Error(XNYZ_EINVAL, "truncated profile");
    const uint32_t n = uint32_t(profile[pos]) | uint32_t(profile[pos + 1]) << 8 |
                       uint32_t(profile[pos + 2]) << 16 | uint32_t(profile[pos + 3]) << 24;
    pos += 4;
    if (pos + n > profile_len) throw Error(XNYZ_EINVAL, "truncated profile");
    metas[k].assign(profile + pos, profile + pos + n);
    pos += n;
  }
  if (metas[0].empty() || metas[1].empty()) throw Error(XNYZ_EINVAL, "empty profile");
}

int with_profile(xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets, uint32_t count,
                 const xnyz_opts* opts, const uint8_t* profile, uint64_t profile_len,
                 CallOut& co) {
  if (!ctx || !d_xyz || !offsets || !opts || !count || !profile)
    throw Error(XNYZ_EINVAL, "null argument or empty batch");
  std::vector<u64> off = rebase(offsets, count, d_xyz);
  for (u32 s = 0; s < count; ++s)
    if (opts[s].rle) throw Error(XNYZ_EUNSUPPORTED, "shared codebooks are not defined for RLE");
  std::vector<uint8_t> metas[2];
  read_profile(profile, profile_len, metas);
  use_device(ctx);
  co.shared = 2;
  co.profile_metas = metas;
  return compress_impl(ctx, d_xyz, off, opts, co, now_ms());
}

int global_rle(xnyz_ctx* ctx, const float* d_xyz, uint64_t n, const xnyz_opts* opts,
               CallOut& co) {
  if (!ctx || !d_xyz) throw Error(XNYZ_EINVAL, "null argument");
  validate(opts, n);
  xnyz_opts o = *opts;
  o.rle = 1;
  o.mode_flags = 0;
  use_device(ctx);
  co.global_rle = true;
  return compress_impl(ctx, d_xyz, {0, n}, &o, co, now_ms());
}

}  // namespace

int xnyz_compress_with_profile_batch_device(xnyz_ctx* ctx, const float* d_xyz,
                                            const uint64_t* offsets, uint32_t count,
                                            const xnyz_opts* opts, const uint8_t* profile,
                                            uint64_t profile_len, xnyz_buf* frames,
                                            xnyz_buf* prezstd, xnyz_stats* stats) {
  return guarded([&] {
    if (!frames) throw Error(XNYZ_EINVAL, "null frames");
    CallOut co{frames, prezstd, stats, nullptr};
    return with_profile(ctx, d_xyz, offsets, count, opts, profile, profile_len, co);
  });
}

int xnyz_compress_with_profile_batch_device_into(xnyz_ctx* ctx, const float* d_xyz,
                                                 const uint64_t* offsets, uint32_t count,
                                                 const xnyz_opts* opts, const uint8_t* profile,
                                                 uint64_t profile_len, uint8_t* out,
                                                 uint64_t cap, uint64_t* out_offsets,
                                                 xnyz_stats* stats) {
  return guarded([&] {
    if (!out_offsets || (!out && cap)) throw Error(XNYZ_EINVAL, "null output");
    CallOut co{nullptr, nullptr, stats, nullptr};
    co.dst = out;
    co.cap = cap;
    co.dst_off = out_offsets;
    return with_profile(ctx, d_xyz, offsets, count, opts, profile, profile_len, co);
  });
}

int xnyz_compress_global_rle_device(xnyz_ctx* ctx, const float* d_xyz, uint64_t n,
                                    const xnyz_opts* opts, xnyz_buf* out_frame,
                                    xnyz_buf* out_prezstd, xnyz_stats* stats) {
  return guarded([&] {
    if (!out_frame) throw Error(XNYZ_EINVAL, "null argument");
    CallOut co{out_frame, out_prezstd, stats, nullptr};
    return global_rle(ctx, d_xyz, n, opts, co);
  });
}

int xnyz_compress_global_rle_device_into(xnyz_ctx* ctx, const float* d_xyz, uint64_t n,
                                         const xnyz_opts* opts, uint8_t* out, uint64_t cap,
                                         uint64_t* out_offsets, xnyz_stats* stats) {
  return guarded([&] {
    if (!out_offsets || (!out && cap)) throw Error(XNYZ_EINVAL, "null output");
    CallOut co{nullptr, nullptr, stats, nullptr};
    co.dst = out;
    co.cap = cap;
    co.dst_off = out_offsets;
    return global_rle(ctx, d_xyz, n, opts, co);
  });
}

int xnyz_search_batch_device(xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets,
                             uint32_t count, const xnyz_opts* opts,
                             xnyz_search_result* results) {
  return guarded([&] {
    if (!ctx || !d_xyz || !offsets || !opts || !count || !results)
      throw Error(XNYZ_EINVAL, "null argument or empty batch");
    std::vector<u64> off = rebase(offsets, count, d_xyz);
    use_device(ctx);
    search_impl(ctx, d_xyz, off, opts, results);
    return XNYZ_OK;
  });
}

int xnyz_features_batch_device(xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets,
                               uint32_t count, const xnyz_opts* opts, float* features) {
  return guarded([&] {
    if (!ctx || !d_xyz || !offsets || !opts || !count || !features)
      throw Error(XNYZ_EINVAL, "null argument or empty batch");
    std::vector<u64> off = rebase(offsets, count, d_xyz);
    use_device(ctx);
    std::vector<float> f;
    CallOut co{nullptr, nullptr, nullptr, nullptr};
    co.features = &f;
    compress_impl(ctx, d_xyz, off, opts, co, now_ms());
    std::memcpy(features, f.data(), f.size() * sizeof(float));
    return XNYZ_OK;
  });
}

namespace {

static_assert(kNoAllocScratchBytes == XNYZ_NOALLOC_SCRATCH_BYTES,
              "no-allocation scratch size differs from include/xnyz.h");
static_assert(kPTotal == XNYZ_MLP_PARAMS && kFeat == XNYZ_MLP_FEATURES,
              "selector layout differs from include/xnyz.h");

MlpLabel label_of(const xnyz_mode& m) {
  return MlpLabel{std::min(std::max(m.gr_k, 0), 31), std::min(std::max(m.pfd_width, 1), 8) - 1,
                  m.rle ? 1 : 0, m.is_hilbert ? 1 : 0};
}

void upload(DBuf& d, const void* src, size_t bytes, hipStream_t st) {
  d.ensure(bytes + 16);
  check_hip(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, st), "H2D");
}

void predict_impl(xnyz_ctx* ctx, const float* params, const float* features, u64 n,
                  std::vector<MlpLabel>& out, std::vector<float>& conf) {
  hipStream_t st = ctx->stream;
  constexpr u64 kChunk = 1 << 16;
  const u64 m0 = std::min(n, kChunk);
  // pinned staging (h_mlp): parameters, then one chunk's features, labels
  // and confidences
  const size_t pb = kPTotal * sizeof(float), xb = (size_t)m0 * kFeat * sizeof(float);
  const size_t ob = (size_t)m0 * sizeof(MlpLabel), cb = (size_t)m0 * sizeof(float);
  const size_t xo = (pb + 255) & ~size_t(255), oo = (xo + xb + 255) & ~size_t(255);
  const size_t co = (oo + ob + 255) & ~size_t(255);
  ctx->h_mlp.ensure(co + cb);
  char* h = ctx->h_mlp.as<char>();
  std::memcpy(h, params, pb);
  ctx->mP.ensure(pb + 16);
  launch_copy_words(h, ctx->mP.p, pb, st);
  ctx->mWh.ensure(kHalfWeights * sizeof(__half) + 16);
  mlp_refresh_half(ctx->mP.as<float>(), ctx->mWh.as<__half>(), st);
  out.resize(n);
  conf.resize(n);
  ctx->mScratch.ensure(mlp_scratch_floats((u32)m0) * sizeof(float));
  ctx->mOut.ensure(ob);
  ctx->mConf.ensure(cb);
  ctx->mX.ensure(xb + 16);
  for (u64 a = 0; a < n; a += kChunk) {
    const u32 m = (u32)std::min(kChunk, n - a);
    std::memcpy(h + xo, features + a * kFeat, (size_t)m * kFeat * sizeof(float));
    launch_copy_words(h + xo, ctx->mX.p, (size_t)m * kFeat * sizeof(float), st);
    mlp_predict(ctx->mX.as<float>(), m, ctx->mP.as<float>(), ctx->mWh.as<__half>(),
                ctx->mOut.as<MlpLabel>(), ctx->mConf.as<float>(), ctx->mScratch.as<float>(), st);
    launch_copy_words(ctx->mOut.p, h + oo, m * sizeof(MlpLabel), st);
    launch_copy_words(ctx->mConf.p, h + co, m * sizeof(float), st);
    check_hip(hipStreamSynchronize(st), "sync");
    std::memcpy(out.data() + a, h + oo, m * sizeof(MlpLabel));
    std::memcpy(conf.data() + a, h + co, m * sizeof(float));
  }
}

// Batched compression with a mode per cloud: the selector's curve x RLE
// where it is confident, the exhaustive search's optimum elsewhere.  Every
// cloud's Golomb-Rice k and PFD width are then settled exactly from the
// statistics of the compressing pass itself (CallOut::refine): the
// selector's k / width stay where the size model finds them optimal.
void selected_impl(xnyz_ctx* ctx, const float* d_xyz, const std::vector<u64>& off,
                   const xnyz_opts* opts, const float* params, float min_confidence,
                   CallOut co, xnyz_mode* chosen, int32_t* searched) {
  const u32 count = (u32)(off.size() - 1);
  for (u32 s = 0; s < count; ++s) validate(&opts[s], off[s + 1] - off[s]);
  std::vector<xnyz_mode> mode(count);
  std::vector<char> fall(count, 1);
  if (params) {  // features -> MLP
    std::vector<float> f;
    CallOut fo{nullptr, nullptr, nullptr, nullptr};
    fo.features = &f;
    compress_impl(ctx, d_xyz, off, opts, fo, now_ms());
    std::vector<MlpLabel> out;
    std::vector<float> conf;
    predict_impl(ctx, params, f.data(), count, out, conf);
    for (u32 s = 0; s < count; ++s) {
      mode[s] = xnyz_mode{out[s].curve, out[s].rle, out[s].k, out[s].w + 1};
      fall[s] = conf[s] < min_confidence;
    }
  }
  std::vector<u32> fb;
  for (u32 s = 0; s < count; ++s)
    if (fall[s]) fb.push_back(s);
  if (!fb.empty()) {  // exhaustive search of the fallback clouds, gathered
    const bool all = fb.size() == count;
    std::vector<u64> so{0};
    std::vector<xnyz_opts> sopt;
    const float* src = d_xyz;
    if (!all) {
      u64 tot = 0;
      for (u32 s : fb) tot += off[s + 1] - off[s];
      ctx->tmp.ensure((size_t)tot * 12);
      for (u32 s : fb) {
        const u64 n = off[s + 1] - off[s];
        check_hip(hipMemcpyAsync(ctx->tmp.as<float>() + 3 * so.back(), d_xyz + 3 * off[s],
                                 n * 12, hipMemcpyDeviceToDevice, ctx->stream), "gather");
        so.push_back(so.back() + n);
      }
      src = ctx->tmp.as<float>();
    } else {
      so = off;
    }
    for (u32 s : fb) sopt.push_back(opts[s]);
    std::vector<xnyz_search_result> r(fb.size());
    search_impl(ctx, src, so, sopt.data(), r.data());
    for (size_t i = 0; i < fb.size(); ++i) mode[fb[i]] = r[i].best;
  }
  std::vector<xnyz_opts> vo(count);
  for (u32 s = 0; s < count; ++s) vo[s] = with_mode(opts[s], mode[s]);
  std::vector<char> refine(count, 1), corrected(count, 0);
  std::vector<xnyz_mode> used(count);
  co.refine = refine.data();
  co.modes_out = used.data();
  co.corrected = corrected.data();
  compress_impl(ctx, d_xyz, off, vo.data(), co, now_ms());
  for (u32 s = 0; s < count; ++s) {
    if (chosen) chosen[s] = used[s];
    if (searched) searched[s] = (fall[s] ? 1 : 0) | (corrected[s] ? 2 : 0);
  }
}

}  // namespace

int xnyz_mlp_train(xnyz_ctx* ctx, const float* features, const xnyz_mode* labels, uint64_t n,
                   const xnyz_train_opts* to, float* params, float* final_loss) {
  return guarded([&] {
    if (!ctx || !features || !labels || !n || !to || !params)
      throw Error(XNYZ_EINVAL, "null argument or no samples");
    if (to->epochs <= 0 || to->batch <= 0 || to->batch > 65536 || !(to->lr > 0.f))
      throw Error(XNYZ_EINVAL, "bad training options");
    if (n > 0xFFFFFFFFull) throw Error(XNYZ_EUNSUPPORTED, "more than 2^32-1 samples");
    use_device(ctx);
    hipStream_t st = ctx->stream;
    if (to->init) {  // normaliser from the data, He-initialised layers, zero biases
      std::vector<double> mean(kFeat, 0.0), var(kFeat, 0.0);
      for (u64 i = 0; i < n; ++i)
        for (int f = 0; f < kFeat; ++f) mean[f] += features[i * kFeat + f];
      for (int f = 0; f < kFeat; ++f) mean[f] /= (double)n;
      for (u64 i = 0; i < n; ++i)
        for (int f = 0; f < kFeat; ++f) {
          const double d = features[i * kFeat + f] - mean[f];
          var[f] += d * d;
        }
      std::mt19937_64 rng(to->seed);
      std::normal_distribution<float> nd(0.f, 1.f);
      std::fill(params, params + kPTotal, 0.f);
      for (int f = 0; f < kFeat; ++f) {
        params[kPMean + f] = (float)mean[f];
        const double sd = std::sqrt(var[f] / (double)n);
        params[kPIstd + f] = sd > 1e-6 ? (float)(1.0 / sd) : 0.f;
      }
      auto he = [&](int off, int fan_in, int count) {
        const float sc = std::sqrt(2.f / (float)fan_in);
        for (int i = 0; i < count; ++i) params[off + i] = nd(rng) * sc;
      };
      he(kPW1, kFeat, kFeat * kHid);
      he(kPW2, kHid, kHid * kHid);
      he(kPW3, kHid, kHid * kOut);
    }
    std::vector<MlpLabel> y(n);
    for (u64 i = 0; i < n; ++i) y[i] = label_of(labels[i]);
    upload(ctx->mX, features, n * kFeat * sizeof(float), st);
    upload(ctx->mY, y.data(), n * sizeof(MlpLabel), st);
    upload(ctx->mP, params, kPTotal * sizeof(float), st);
    ctx->mWh.ensure(kHalfWeights * sizeof(__half) + 16);
    for (DBuf* b : {&ctx->mG, &ctx->mM, &ctx->mV}) {
      b->ensure(kTrainable * sizeof(float) + 16);
      check_hip(hipMemsetAsync(b->p, 0, kTrainable * sizeof(float), st), "memset");
    }
    const u32 nb = (u32)std::min<u64>(to->batch, n);
    ctx->mScratch.ensure(mlp_scratch_floats(nb) * sizeof(float));
    ctx->mLoss.ensure(4 * sizeof(float));
    ctx->mIdx.ensure(n * 4 + 16);
    mlp_refresh_half(ctx->mP.as<float>(), ctx->mWh.as<__half>(), st);
    std::vector<u32> perm(n);
    std::mt19937_64 rng(to->seed ^ 0x5DEECE66Dull);
    int t = 0;
    float loss = 0.f;
    for (int e = 0; e < to->epochs; ++e) {
      for (u64 i = 0; i < n; ++i) perm[i] = (u32)i;
      std::shuffle(perm.begin(), perm.end(), rng);
      check_hip(hipMemcpyAsync(ctx->mIdx.p, perm.data(), n * 4, hipMemcpyHostToDevice, st), "H2D");
      check_hip(hipMemsetAsync(ctx->mLoss.p, 0, sizeof(float), st), "memset");
      u32 steps = 0;
      for (u64 a = 0; a + nb <= n; a += nb) {  // whole batches only
        mlp_train_step(ctx->mX.as<float>(), ctx->mY.as<MlpLabel>(), ctx->mIdx.as<u32>() + a, nb,
                       ctx->mP.as<float>(), ctx->mWh.as<__half>(), ctx->mScratch.as<float>(),
                       ctx->mG.as<float>(), ctx->mM.as<float>(), ctx->mV.as<float>(), ++t,
                       to->lr, ctx->mLoss.as<float>(), st);
        ++steps;
      }
      float acc = 0.f;
      check_hip(hipMemcpyAsync(&acc, ctx->mLoss.p, sizeof(float), hipMemcpyDeviceToHost, st),
                "D2H");
      check_hip(hipStreamSynchronize(st), "sync");  // also orders the next perm upload
      loss = steps ? acc / (float)steps : 0.f;
    }
    check_hip(hipMemcpyAsync(params, ctx->mP.p, kPTotal * sizeof(float), hipMemcpyDeviceToHost,
                             st), "D2H");
    check_hip(hipStreamSynchronize(st), "sync");
    if (final_loss) *final_loss = loss;
    return XNYZ_OK;
  });
}

int xnyz_mlp_predict(xnyz_ctx* ctx, const float* params, const float* features, uint64_t n,
                     xnyz_mode* modes, float* confidence) {
  return guarded([&] {
    if (!ctx || !params || !features || !n || !modes)
      throw Error(XNYZ_EINVAL, "null argument or no samples");
    use_device(ctx);
    std::vector<MlpLabel> out;
    std::vector<float> conf;
    predict_impl(ctx, params, features, n, out, conf);
    for (u64 i = 0; i < n; ++i) {
      modes[i] = xnyz_mode{out[i].curve, out[i].rle, out[i].k, out[i].w + 1};
      if (confidence) confidence[i] = conf[i];
    }
    return XNYZ_OK;
  });
}

int xnyz_compress_selected_batch_device(xnyz_ctx* ctx, const float* d_xyz,
                                        const uint64_t* offsets, uint32_t count,
                                        const xnyz_opts* opts, const float* params,
                                        float min_confidence, xnyz_buf* frames,
                                        xnyz_buf* prezstd, xnyz_stats* stats,
                                        xnyz_mode* chosen, int32_t* searched) {
  return guarded([&] {
    if (!ctx || !d_xyz || !offsets || !opts || !count || !frames)
      throw Error(XNYZ_EINVAL, "null argument or empty batch");
    std::vector<u64> off = rebase(offsets, count, d_xyz);
    use_device(ctx);
    selected_impl(ctx, d_xyz, off, opts, params, min_confidence,
                  CallOut{frames, prezstd, stats, nullptr}, chosen, searched);
    return XNYZ_OK;
  });
}

int xnyz_compress_selected_batch_device_into(xnyz_ctx* ctx, const float* d_xyz,
                                             const uint64_t* offsets, uint32_t count,
                                             const xnyz_opts* opts, const float* params,
                                             float min_confidence, uint8_t* out, uint64_t cap,
                                             uint64_t* out_offsets, xnyz_stats* stats,
                                             xnyz_mode* chosen, int32_t* searched) {
  return guarded([&] {
    if (!ctx || !d_xyz || !offsets || !opts || !count || !out_offsets || (!out && cap))
      throw Error(XNYZ_EINVAL, "null argument or empty batch");
    std::vector<u64> off = rebase(offsets, count, d_xyz);
    use_device(ctx);
    CallOut co{nullptr, nullptr, stats, nullptr};
    co.dst = out;
    co.cap = cap;
    co.dst_off = out_offsets;
    selected_impl(ctx, d_xyz, off, opts, params, min_confidence, co, chosen, searched);
    return XNYZ_OK;
  });
}

int xnyz_compress_selected_batch_device_into_async(
    xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets, uint32_t count,
    const xnyz_opts* opts, const float* params, float min_confidence, uint8_t* out, uint64_t cap,
    uint64_t* out_offsets, xnyz_stats* stats, xnyz_mode* chosen, int32_t* searched) {
  return guarded([&] {
    if (!ctx || !d_xyz || !offsets || !opts || !count || !out_offsets || !out || !cap)
      throw Error(XNYZ_EINVAL, "null argument or empty batch");
    std::vector<u64> off = rebase(offsets, count, d_xyz);
    use_device(ctx);
    CallOut co{nullptr, nullptr, stats, nullptr};
    co.dst = out;
    co.cap = cap;
    co.dst_off = out_offsets;
    co.async_out = true;
    selected_impl(ctx, d_xyz, off, opts, params, min_confidence, co, chosen, searched);
    return XNYZ_OK;
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
    const u32 N = (u32)n, ntiles = (N + kTile - 1) / kTile;
    ctx->h_tiles.ensure(sizeof(Tile) * ntiles);
    ctx->h_sp.ensure(sizeof(SegParam));
    ctx->h_sc.ensure(sizeof(SegCtl));
    for (u32 l = 0; l < ntiles; ++l) ctx->h_tiles.as<Tile>()[l] = Tile{0, l};
    SegParam& p = *ctx->h_sp.as<SegParam>();
    std::memset(&p, 0, sizeof(p));
    p.n = N;
    p.ntiles = ntiles;
    p.qp = quant_params(&o);
    ctx->tiles.ensure(sizeof(Tile) * ntiles);
    ctx->sp.ensure(sizeof(SegParam));
    ctx->sc.ensure(sizeof(SegCtl));
    ctx->sh.ensure(sizeof(SegHist));
    ctx->nlarge.ensure(8);
    ctx->tmp.ensure((size_t)n * 12);
    ctx->h2d(ctx->tiles, ctx->h_tiles, sizeof(Tile) * ntiles);
    ctx->h2d(ctx->sp, ctx->h_sp, sizeof(SegParam));
    Batch bt{};
    bt.S = 1;
    bt.ntiles = ntiles;
    bt.N = N;
    bt.tiles = ctx->tiles.as<Tile>();
    bt.sp = ctx->sp.as<SegParam>();
    bt.sc = ctx->sc.as<SegCtl>();
    bt.sh = ctx->sh.as<SegHist>();
    bt.n_large = ctx->nlarge.as<u32>();
    launch_init(bt, s);
    launch_bbox(bt, d_xyz, s);
    launch_quantize_nodes(bt, d_xyz, ctx->tmp.as<int32_t>(), s);
    check_hip(hipMemcpyAsync(nodes, ctx->tmp.p, (size_t)n * 12, hipMemcpyDeviceToHost, s),
              "nodes D2H");
    ctx->d2h(ctx->h_sc, ctx->sc, sizeof(SegCtl));
    ctx->sync();
    if (offset)
      for (int a = 0; a < 3; ++a) offset[a] = ctx->h_sc.as<SegCtl>()->offset[a];
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

int xnyz_decompress(const uint8_t* in, uint64_t len, const xnyz_opts* opts,
                    xnyz_buf* out_points) {
  return guarded([&] { return decompress_impl(in, len, opts, out_points); });
}

int xnyz_decompress_device(xnyz_ctx* ctx, const uint8_t* in, uint64_t len,
                           const xnyz_opts* opts, xnyz_buf* out_points) {
  return guarded([&] {
    if (!ctx || !out_points) throw Error(XNYZ_EINVAL, "null argument");
    use_device(ctx);
    ProfGuard pg(ctx);
    return decompress_device_impl(ctx, in, len, opts, nullptr, 0, nullptr, out_points, nullptr);
  });
}

int xnyz_decompress_to_device(xnyz_ctx* ctx, const uint8_t* in, uint64_t len,
                              const xnyz_opts* opts, float* d_points, uint64_t cap_points,
                              uint64_t* n_points) {
  return guarded([&] {
    if (!ctx || !d_points || !n_points) throw Error(XNYZ_EINVAL, "null argument");
    use_device(ctx);
    ProfGuard pg(ctx);
    return decompress_device_impl(ctx, in, len, opts, d_points, cap_points, n_points, nullptr,
                                  nullptr);
  });
}

int xnyz_container_write(const char* part_path, const char* manifest_path, const uint8_t* out,
                         const uint64_t* out_offsets, uint32_t count, xnyz_chunk* chunks) {
  return guarded([&] {
    if (!part_path || !manifest_path || !out_offsets || !chunks || (!out && count))
      throw Error(XNYZ_EINVAL, "null argument");
    write_container(part_path, manifest_path, out, out_offsets, count, chunks);
    return XNYZ_OK;
  });
}

int xnyz_container_read(const char* manifest_path, xnyz_buf* chunks) {
  return guarded([&] {
    if (!manifest_path || !chunks) throw Error(XNYZ_EINVAL, "null argument");
    const std::vector<xnyz_chunk> v = read_manifest(manifest_path);
    const size_t bytes = v.size() * sizeof(xnyz_chunk);
    chunks->data = xmalloc(bytes);
    if (bytes) std::memcpy(chunks->data, v.data(), bytes);
    chunks->len = bytes;
    return XNYZ_OK;
  });
}

int xnyz_container_decompress(xnyz_ctx* ctx, const char* part_path, const char* manifest_path,
                              xnyz_buf* out_points) {
  return guarded([&] {
    if (!part_path || !manifest_path || !out_points) throw Error(XNYZ_EINVAL, "null argument");
    const std::vector<xnyz_chunk> v = read_manifest(manifest_path);
    std::vector<uint8_t> all;
    if (ctx) use_device(ctx);
    for (const xnyz_chunk& c : v) {
      const std::vector<uint8_t> b = read_part(part_path, c);
      xnyz_opts o{};
      o.quantizer = c.quantizer;
      o.l2_bound = c.l2_bound;
      o.rle = c.rle;
      o.frame_mode = c.frame_mode == XNYZ_FRAME_NONE ? XNYZ_FRAME_NONE : XNYZ_FRAME_ZSTD3;
      if (!(o.l2_bound > 0.f)) throw Error(XNYZ_EINVAL, "chunk without an L2 bound");
      xnyz_buf pts{};
      if (ctx) {
        ProfGuard pg(ctx);
        decompress_device_impl(ctx, b.data(), b.size(), &o, nullptr, 0, nullptr, &pts, nullptr);
      } else {
        decompress_impl(b.data(), b.size(), &o, &pts);
      }
      all.insert(all.end(), pts.data, pts.data + pts.len);
      xnyz_free(&pts);
    }
    out_points->data = xmalloc(all.size());
    if (!all.empty()) std::memcpy(out_points->data, all.data(), all.size());
    out_points->len = all.size();
    return XNYZ_OK;
  });
}

int xnyz_hilbert_to_morton(xnyz_ctx* ctx, const uint64_t* in, uint64_t n, uint32_t iteration,
                           uint64_t* out) {
  return guarded([&] {
    if (!ctx || !in || !out) throw Error(XNYZ_EINVAL, "null argument");
    if (iteration > 21) throw Error(XNYZ_EUNSUPPORTED, "iteration above 21");
    if (n == 0) return XNYZ_OK;
    use_device(ctx);
    hipStream_t s = ctx->stream;
    ctx->tmp.ensure((size_t)n * 16);
    u64* d = ctx->tmp.as<u64>();
    check_hip(hipMemcpyAsync(d, in, n * 8, hipMemcpyHostToDevice, s), "codes H2D");
    launch_h2m(d, n, iteration, ctx->hinv, d + n, s);
    check_hip(hipMemcpyAsync(out, d + n, n * 8, hipMemcpyDeviceToHost, s), "codes D2H");
    ctx->sync();
    return XNYZ_OK;
  });
}

int xnyz_graphs_enable(xnyz_ctx* ctx, int on) {
  return guarded([&] {
    if (!ctx) throw Error(XNYZ_EINVAL, "null context");
    ctx->graphs = on != 0;
    if (!ctx->graphs) {
      use_device(ctx);
      ctx->stage2.clear();
    }
    return XNYZ_OK;
  });
}

int xnyz_tuning_set(xnyz_ctx* ctx, int size_class, const xnyz_tuning* t) {
  return guarded([&] {
    if (!ctx || !t) throw Error(XNYZ_EINVAL, "null argument");
    if (size_class < -1 || size_class >= XNYZ_SIZE_CLASSES)
      throw Error(XNYZ_EINVAL, "size class out of range");
    check_tuning(*t);
    if (size_class < 0) {
      for (xnyz_tuning& x : ctx->tuning) x = *t;
    } else {
      ctx->tuning[size_class] = *t;
    }
    return XNYZ_OK;
  });
}

int xnyz_tuning_get(xnyz_ctx* ctx, int size_class, xnyz_tuning* t) {
  return guarded([&] {
    if (!ctx || !t) throw Error(XNYZ_EINVAL, "null argument");
    if (size_class < 0 || size_class >= XNYZ_SIZE_CLASSES)
      throw Error(XNYZ_EINVAL, "size class out of range");
    *t = ctx->tuning[size_class];
    return XNYZ_OK;
  });
}

int xnyz_tuning_save(xnyz_ctx* ctx, const char* path) {
  return guarded([&] {
    if (!ctx || !path) throw Error(XNYZ_EINVAL, "null argument");
    FILE* f = std::fopen(path, "w");
    if (!f) throw Error(XNYZ_EIO, std::string("cannot write ") + path);
    std::fprintf(f,
                 "# class zstd_chunk_log copy_slice_log gpu_frame_share persist_blocks copy_first "
                 "radix_bits radix_tile\n");
    for (int c = 0; c < XNYZ_SIZE_CLASSES; ++c) {
      const xnyz_tuning& t = ctx->tuning[c];
      std::fprintf(f, "%d %u %u %.4f %u %u %u %u\n", c, t.zstd_chunk_log, t.copy_slice_log,
                   (double)t.gpu_frame_share, t.persist_blocks, t.copy_first, t.radix_bits,
                   t.radix_tile);
    }
    const bool ok = std::fclose(f) == 0;
    if (!ok) throw Error(XNYZ_EIO, std::string("cannot write ") + path);
    return XNYZ_OK;
  });
}

int xnyz_tuning_load(xnyz_ctx* ctx, const char* path) {
  return guarded([&] {
    if (!ctx || !path) throw Error(XNYZ_EINVAL, "null argument");
    FILE* f = std::fopen(path, "r");
    if (!f) throw Error(XNYZ_EIO, std::string("cannot read ") + path);
    std::vector<std::pair<int, xnyz_tuning>> rows;
    char line[256];
    int lineno = 0;
    bool bad = false;
    while (std::fgets(line, sizeof(line), f)) {
      ++lineno;
      const char* p = line;
      while (*p == ' ' || *p == '\t') ++p;
      if (*p == '#' || *p == '\n' || *p == 0) continue;
      int c;
      xnyz_tuning t;
      if (std::sscanf(p, "%d %u %u %f %u %u %u %u", &c, &t.zstd_chunk_log, &t.copy_slice_log,
                      &t.gpu_frame_share, &t.persist_blocks, &t.copy_first, &t.radix_bits,
                      &t.radix_tile) != 8 ||
          c < 0 || c >= XNYZ_SIZE_CLASSES) {
        bad = true;
        break;
      }
      rows.emplace_back(c, t);
    }
    std::fclose(f);
    if (bad) throw Error(XNYZ_EINVAL, std::string(path) + ":" + std::to_string(lineno) +
                                          ": expected 'class chunk_log slice_log share blocks copy_first radix_bits "
                                          "radix_tile'");
    for (const auto& r : rows) check_tuning(r.second);  // all or nothing
    for (const auto& r : rows) ctx->tuning[r.first] = r.second;
    return XNYZ_OK;
  });
}

int xnyz_autotune(xnyz_ctx* ctx, const float* d_xyz, uint64_t n, const xnyz_opts* opts,
                  uint32_t reps, xnyz_tuning* best, double* best_ms) {
  return guarded([&] {
    if (!ctx || !d_xyz || !opts || !n) throw Error(XNYZ_EINVAL, "null argument or empty cloud");
    if (reps == 0 || reps > 1000) throw Error(XNYZ_EINVAL, "reps must be in [1, 1000]");
    validate(opts, n);
    use_device(ctx);
    const int c = size_class(n);
    const std::vector<u64> off{0, n};
    // A pinned output, as a serving loop reusing one would have (a split
    // frame's device blocks are copied straight into it).
    struct Out {
      uint8_t* p = nullptr;
      u64 n = 0;
      void alloc(u64 bytes) {
        release();
        p = static_cast<uint8_t*>(pinned_alloc(bytes));
        n = bytes;
      }
      void release() {
        if (p) pinned_free(p);
        p = nullptr;
      }
      ~Out() {
        try {
          release();
        } catch (...) {
        }
      }
    } out;
    out.alloc(13 * n + (1u << 20));
    uint64_t out_off[2];
    auto call = [&] {
      for (int attempt = 0;; ++attempt) {
        CallOut co{nullptr, nullptr, nullptr, nullptr};
        co.dst = out.p;
        co.cap = out.n;
        co.dst_off = out_off;
        try {
          compress_impl(ctx, d_xyz, off, opts, co, now_ms());
          return;
        } catch (const Error& e) {
          if (e.code != XNYZ_ENOSPC || attempt) throw;
          out.alloc(out_off[1]);
        }
      }
    };
    const xnyz_tuning saved = ctx->tuning[c];
    // The fastest of `reps` timed calls (after one untimed call): the
    // minimum is the estimate least disturbed by other activity on the host.
    auto timed = [&](const xnyz_tuning& t) {
      ctx->tuning[c] = t;
      call();  // untimed: buffers, captured graphs and host pages settle
      double best = 1e30;
      for (u32 r = 0; r < reps; ++r) {
        const double t0 = now_ms();
        call();
        best = std::min(best, now_ms() - t0);
      }
      return best;
    };
    try {
      xnyz_tuning cur = saved;
      double cur_ms = timed(cur);
      auto sweep = [&](auto field, std::initializer_list<double> values) {
        for (double v : values) {
          xnyz_tuning t = cur;
          field(t, v);
          if (std::memcmp(&t, &cur, sizeof(t)) == 0) continue;
          const double ms = timed(t);
          if (ms < cur_ms) {
            cur_ms = ms;
            cur = t;
          }
        }
      };
      // Kernel shapes are scored by the device time of the stream-building
      // kernels (HIP events around each launch): their differences are a
      // fraction of a millisecond, below the noise of wall-clock calls.
      auto dev_ms = [&](const xnyz_tuning& t) {
        ctx->tuning[c] = t;
        const bool was_on = ctx->prof_on;
        ctx->prof_on = true;
        double best = 1e30;
        try {
          call();
          for (u32 r = 0; r < reps; ++r) {
            ctx->prof.acc.clear();
            call();
            double ms = 0;
            for (const auto& a : ctx->prof.acc)
              if (a.first.rfind("k_z", 0) != 0) ms += a.second.second;
            best = std::min(best, ms);
          }
        } catch (...) {
          ctx->prof_on = was_on;
          throw;
        }
        ctx->prof_on = was_on;
        ctx->prof.acc.clear();
        return best;
      };
      auto shapes = [&] {
        double base = dev_ms(cur);
        auto sweep_dev = [&](auto field, std::initializer_list<double> values) {
          for (double v : values) {
            xnyz_tuning t = cur;
            field(t, v);
            if (std::memcmp(&t, &cur, sizeof(t)) == 0) continue;
            const double ms = dev_ms(t);
            if (ms < base) {
              base = ms;
              cur = t;
            }
          }
        };
        sweep_dev([](xnyz_tuning& t, double v) { t.persist_blocks = (u32)v; },
                  {512, 1024, 2048, 4096, 8192});
        sweep_dev([](xnyz_tuning& t, double v) { t.radix_bits = (u32)v; }, {6, 7, 8});
        sweep_dev([](xnyz_tuning& t, double v) { t.radix_tile = (u32)v; }, {2048, 4096});
        cur_ms = timed(cur);
      };
      if (opts->frame_mode == XNYZ_FRAME_ZSTD3_MT) {
        sweep([](xnyz_tuning& t, double v) { t.zstd_chunk_log = (u32)v; }, {18, 19, 20, 21, 22});
        // device share and copy order together (they interact)
        for (u32 first = 0; first < 2; ++first)
          sweep([first](xnyz_tuning& t, double v) {
                  t.copy_first = first;
                  t.gpu_frame_share = (float)v;
                },
                {0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7});
        shapes();
        const double s0 = cur.gpu_frame_share;
        if (s0 > 0)
          sweep([](xnyz_tuning& t, double v) { t.gpu_frame_share = (float)v; },
                {std::max(0.0, s0 - 0.05), std::max(0.0, s0 - 0.025), std::min(0.9, s0 + 0.025),
                 std::min(0.9, s0 + 0.05)});
        sweep([](xnyz_tuning& t, double v) { t.copy_slice_log = (u32)v; }, {22, 24, 26});
        sweep([](xnyz_tuning& t, double v) { t.zstd_chunk_log = (u32)v; }, {18, 19, 20, 21});
      } else {
        shapes();
      }
      ctx->tuning[c] = cur;
      if (best) *best = cur;
      if (best_ms) *best_ms = cur_ms;
    } catch (...) {
      ctx->tuning[c] = saved;
      throw;
    }
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

int xnyz_host_alloc(uint64_t bytes, void** out) {
  return guarded([&] {
    if (!out || !bytes) throw Error(XNYZ_EINVAL, "null out or zero size");
    *out = nullptr;
    *out = pinned_alloc(bytes);
    return XNYZ_OK;
  });
}

int xnyz_host_free(void* p) {
  if (!p) return XNYZ_OK;
  return guarded([&] {
    pinned_free(p);
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
