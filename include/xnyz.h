/*
 * xnyz.h -- C ABI of libxnyz.so, the MI355X (gfx950) implementation of
 * XnYZip's compression path.
 *
 * The reference exposes no FFI: its interface is the header-only C++ API in
 * namespace XnYZip plus the CLI (source/main.cpp).  Each entry point below
 * names the reference interface it replaces (paths relative to the reference
 * repository root).  Plain pointers and sizes only; no exceptions cross the
 * boundary: every call returns XNYZ_OK (0) or a negative XNYZ_E* code, and
 * xnyz_last_error() returns the message of the calling thread's last error.
 *
 * Threading: one xnyz_ctx per (device, HIP stream); calls on distinct
 * contexts may run concurrently from different threads.  A context owns its
 * device and pinned host buffers and reuses them across calls.
 */
#ifndef XNYZ_H
#define XNYZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XNYZ_ABI_VERSION 7

enum xnyz_status {
  XNYZ_OK = 0,
  XNYZ_EINVAL = -1,       /* bad argument (N = 0, null pointer, bad enum)  */
  XNYZ_EHIP = -2,         /* HIP runtime error                              */
  XNYZ_ENOMEM = -3,       /* device or host allocation failed               */
  XNYZ_EZSTD = -4,        /* libzstd error                                  */
  XNYZ_EUNSUPPORTED = -5, /* input outside the implemented range (see msg) */
  XNYZ_ECORRUPT = -6,     /* malformed compressed stream                    */
  XNYZ_EINTERNAL = -7,    /* device-side invariant violated                 */
  XNYZ_ENOSPC = -8,       /* caller's output buffer too small               */
  XNYZ_EIO = -9           /* a file could not be read or written            */
};

/* QUANTIZER_TYPE (source/composer/utils.hpp:55-59); ADAPTIVE is not
 * reachable from the reference CLI and is not offered. */
enum xnyz_quantizer { XNYZ_CUBE = 0, XNYZ_TRUNCATED_OCTAHEDRON = 1 };

/* How the assembled stream is wrapped.  XNYZ_FRAME_ZSTD3 is the reference's
 * compress_u8_vector (source/composer/utils.hpp:35-53): one single-shot
 * ZSTD_compress at level 3 -- byte-identical frames given the same libzstd.
 * XNYZ_FRAME_ZSTD3_MT is the same level split over libzstd worker threads: a
 * valid single frame with the content size, decodable by the reference's
 * decompress_buffer, but not byte-identical to the single-shot frame.
 * XNYZ_FRAME_NONE returns only the assembled (pre-zstd) stream.
 * XNYZ_FRAME_ZSTD_GPU builds the frame on the device with this library's
 * own zstd-format encoder (16 KiB blocks: LZ matches, Huffman literals, FSE
 * sequences): a valid single frame with the content size for any zstd
 * decoder, about 2% larger than level 3 on point-cloud streams, and the
 * assembled stream never leaves device memory. */
enum xnyz_frame_mode {
  XNYZ_FRAME_ZSTD3 = 0,
  XNYZ_FRAME_ZSTD3_MT = 1,
  XNYZ_FRAME_NONE = 2,
  XNYZ_FRAME_ZSTD_GPU = 3
};

/* Constructor arguments of BlockCompressor[RLE]<float>
 * (source/composer/block_compressor.hpp:35, block_compressor_rle.hpp:34)
 * plus the compress(points, is_hilbert) curve flag and the class choice. */
typedef struct xnyz_opts {
  int32_t quantizer;        /* enum xnyz_quantizer                          */
  float l2_bound;           /* L2 error bound                               */
  int32_t is_hilbert;       /* 0: Z-order (-z), 1: Hilbert (-h)             */
  int32_t rle;              /* 0: BlockCompressor, 1: BlockCompressorRLE    */
  int32_t direct_threshold; /* default 1024 (source/main.cpp:76)            */
  int32_t frame_mode;       /* enum xnyz_frame_mode                         */
  int32_t zstd_workers;     /* XNYZ_FRAME_ZSTD3_MT only; 0 = all host cores */
  /* Mode overrides (no reference counterpart; zero = the reference's own
   * choices).  Streams stay decodable by the reference's decoder: the
   * Golomb-Rice k is stored in its meta (golomb_rice_encoder.hpp:135-139)
   * and the decoder ignores the PFD limit (p_for_delta_encoder.hpp:45-57). */
  int32_t mode_flags;       /* XNYZ_MODE_* bits                              */
  int32_t gr_k;             /* XNYZ_MODE_GR_K: Golomb-Rice k, 0..62          */
  int32_t pfd_width;        /* XNYZ_MODE_PFD_WIDTH: 1..8; repos / sfc-delta
                               values above 2^w - 2 become patches (w = 8 is
                               the reference's limit 254)                    */
} xnyz_opts;

enum xnyz_mode_flags { XNYZ_MODE_GR_K = 1, XNYZ_MODE_PFD_WIDTH = 2 };

/* One coding mode of a chunk: curve, RLE, Golomb-Rice k, PFD width. */
typedef struct xnyz_mode {
  int32_t is_hilbert;
  int32_t rle;
  int32_t gr_k;       /* 0..62 */
  int32_t pfd_width;  /* 1..8 */
} xnyz_mode;

/* Result of the exhaustive per-chunk search.  Sizes are exact assembled
 * (pre-zstd) stream bytes. */
typedef struct xnyz_search_result {
  xnyz_mode best;              /* argmin over the whole space               */
  uint64_t best_bytes;
  uint64_t reference_bytes;    /* the opts' curve and RLE flag with the
                                  reference's k rule and width 8            */
  xnyz_mode variant[4];        /* [2*is_hilbert + rle]: best (k, width)     */
  uint64_t variant_bytes[4];
} xnyz_search_result;

/* Library-owned byte buffer; release with xnyz_free(). */
typedef struct xnyz_buf {
  uint8_t* data;
  uint64_t len;
} xnyz_buf;

/* What the reference prints from inside compress() (Encoding mode, Unique
 * points, patches sizes, Max L2 error, ...), plus per-stage times. */
typedef struct xnyz_stats {
  uint64_t input_points;     /* N                                          */
  uint64_t num_points;       /* N, or unique cells U in RLE mode           */
  int32_t range[3];          /* max quantized x, y, z                      */
  int32_t is_direct;         /* "Encoding mode: direct|block"             */
  uint64_t blocks;           /* occupied 128^3 lattice blocks              */
  uint64_t patches;          /* repos (or sfc-delta) PFD patches          */
  uint64_t quad_patches;     /* quads PFD patches                          */
  uint64_t wrapped_points;   /* points whose node z is -1                  */
  int32_t counts_exp_golomb; /* RLE counts coded with Exp-Golomb           */
  int32_t counts_eg_k;
  double entropy_repos;      /* "entropy:" lines of HuffmanEncoder        */
  double entropy_quads;
  double entropy_counts;
  double max_l2;             /* RLE: "Max L2 error:"                      */
  double mse;                /* RLE: "MSE:"                                */
  uint64_t prezstd_bytes;    /* assembled stream                           */
  uint64_t frame_bytes;      /* returned frame                             */
  double ms_device;          /* upload (host input) + kernels + readback   */
  double ms_zstd;            /* frame stage on the host                    */
  double ms_total;
} xnyz_stats;

typedef struct xnyz_ctx xnyz_ctx;

/* Library version / ABI; no reference counterpart. */
int xnyz_abi_version(void);
const char* xnyz_last_error(void);

/* Context on HIP device `device` with its own non-blocking stream.  No
 * reference counterpart (the reference is single-threaded host code). */
int xnyz_create(int device, xnyz_ctx** out);
int xnyz_destroy(xnyz_ctx* ctx);

/* Replaces BlockCompressor<float>::compress and BlockCompressorRLE<float>::
 * compress (source/composer/block_compressor.hpp:39-222,
 * block_compressor_rle.hpp:38-321) on points in HOST memory: n points of
 * interleaved float x,y,z.  The input is const: the reference's in-place
 * reordering of `points` is returned through `perm` instead (nullable;
 * n entries; non-RLE only, else untouched).  `out_frame` receives the
 * frame per opts->frame_mode; `out_prezstd` (nullable) the assembled
 * stream.  `stats` is nullable.  N = 0 is XNYZ_EINVAL (the reference
 * indexes an empty vector there). */
int xnyz_compress(xnyz_ctx* ctx, const float* xyz, uint64_t n, const xnyz_opts* opts,
                  xnyz_buf* out_frame, xnyz_buf* out_prezstd, uint32_t* perm,
                  xnyz_stats* stats);

/* Same, with the points already resident in device memory of ctx's device
 * (d_xyz: n*3 floats, 4-byte aligned).  No reference counterpart beyond
 * compress(); this is the entry point the throughput figures use. */
int xnyz_compress_device(xnyz_ctx* ctx, const float* d_xyz, uint64_t n,
                         const xnyz_opts* opts, xnyz_buf* out_frame,
                         xnyz_buf* out_prezstd, uint32_t* perm, xnyz_stats* stats);

/* Many clouds in one launch set.  Replaces the MPI driver's per-block loop
 * (source/mpi_driver.cpp:819-943: one compress() per block, blocks spread
 * over ranks) on one GPU.  Cloud s is points [offsets[s], offsets[s+1]) of
 * d_xyz (device memory, interleaved x,y,z floats), compressed with opts[s]
 * -- each cloud its own quantizer, bound, curve, RLE flag and frame mode --
 * into frames[s]; prezstd and stats are nullable arrays of `count` entries.
 * Every cloud's stream is the one xnyz_compress_device would produce for it
 * alone.  No permutation output. */
int xnyz_compress_batch_device(xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets,
                               uint32_t count, const xnyz_opts* opts, xnyz_buf* frames,
                               xnyz_buf* prezstd, xnyz_stats* stats);

/* xnyz_compress_batch_device writing into caller memory instead of
 * library-allocated buffers (no allocation or page faulting per call):
 * cloud s's output, as opts[s].frame_mode selects, is out[out_offsets[s],
 * out_offsets[s + 1]) (count + 1 entries, outputs back to back).
 * out_offsets[0] is 0, except for a frame split between the device and the
 * host (a single XNYZ_FRAME_ZSTD3_MT cloud with a device share) when cap
 * also holds the bound of both parts: then the host's blocks are placed at
 * a fixed offset while they are compressed, the device's part just before
 * them, and the frame starts at out_offsets[0] > 0.  If the output does not
 * fit, nothing is written, out_offsets[count] holds the size needed and the
 * call returns XNYZ_ENOSPC.  No reference counterpart beyond compress(). */
int xnyz_compress_batch_device_into(xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets,
                                    uint32_t count, const xnyz_opts* opts, uint8_t* out,
                                    uint64_t cap, uint64_t* out_offsets, xnyz_stats* stats);

/* xnyz_compress_batch_device_into returning as soon as the assembled
 * streams are in device memory and their copy into `out` is queued on the
 * context's copy stream: the next call's kernels run while this call's
 * output crosses PCIe.  No reference counterpart beyond compress(); it is
 * for ingest and serving loops that compress one cloud or batch after
 * another.  opts[s].frame_mode must be XNYZ_FRAME_NONE and `out` pinned
 * (xnyz_host_alloc), else XNYZ_EINVAL.  out_offsets and stats are final
 * when the call returns; the bytes in `out` only after xnyz_wait(ctx) (or
 * xnyz_destroy).  The context keeps two stream buffers: a call whose
 * buffer's previous copy is still running waits for it first, so at most
 * two copies are in flight.  Copies run in call order; the caller must not
 * read, free or reuse `out` of a call before xnyz_wait. */
int xnyz_compress_batch_device_into_async(xnyz_ctx* ctx, const float* d_xyz,
                                          const uint64_t* offsets, uint32_t count,
                                          const xnyz_opts* opts, uint8_t* out, uint64_t cap,
                                          uint64_t* out_offsets, xnyz_stats* stats);

/* Waits until every copy queued by xnyz_compress_batch_device_into_async on
 * this context has landed in host memory. */
int xnyz_wait(xnyz_ctx* ctx);

/* xnyz_compress_batch_device_into without allocating: no hipMalloc,
 * hipFree, hipHostMalloc or graph instantiation during the call.  No
 * reference counterpart; it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * Every device buffer of the pipeline is carved from one file-scope device
 * array of XNYZ_NOALLOC_SCRATCH_BYTES per GPU (held from the library's load
 * until the process ends; its address is looked up by xnyz_create), and the
 * host-side tables from a pinned block xnyz_create sets up.  Limits: at most
 * XNYZ_NOALLOC_MAX_CLOUDS clouds and XNYZ_NOALLOC_MAX_POINTS points per call,
 * opts[s].frame_mode XNYZ_FRAME_NONE (the assembled streams, back to back in
 * `out` as out_offsets says); beyond them, or when a call's buffers (keys,
 * cells, stream) outgrow the array, it returns XNYZ_EUNSUPPORTED and
 * launches nothing.  XNYZ_ENOSPC (size needed in out_offsets[count]) as for
 * the _into form.  All calls on one GPU share its array: calls that overlap
 * on one GPU (from any context) race and must be serialised by the caller.
 * Streams are the same bytes xnyz_compress_batch_device_into returns. */
#define XNYZ_NOALLOC_SCRATCH_BYTES (256ull << 20)
#define XNYZ_NOALLOC_MAX_POINTS (1ull << 22)
#define XNYZ_NOALLOC_MAX_CLOUDS 256
int xnyz_compress_batch_device_noalloc(xnyz_ctx* ctx, const float* d_xyz,
                                       const uint64_t* offsets, uint32_t count,
                                       const xnyz_opts* opts, uint8_t* out, uint64_t cap,
                                       uint64_t* out_offsets, xnyz_stats* stats);

/* BlockCompressor<float>::fit_profile (source/composer/block_compressor.hpp:
 * 241-246, compress_internal :278-431) on device input: one cloud coded in
 * block mode (whatever its range) with repos and quads Huffman codebooks
 * that cover every symbol 0..255 (HuffmanEncoder::build_complete,
 * huffman_encoder.hpp:69-84).  `profile` receives the serialized profile
 * (serialize_profile, block_encoding_profile.hpp:34-44: u32 LE size and
 * bytes of the repos meta, then of the quads meta); out_frame and
 * out_prezstd (nullable) the cloud's own output.  RLE is
 * XNYZ_EUNSUPPORTED, as in the reference's driver. */
int xnyz_fit_profile_device(xnyz_ctx* ctx, const float* d_xyz, uint64_t n, const xnyz_opts* opts,
                            xnyz_buf* out_frame, xnyz_buf* out_prezstd, xnyz_buf* profile);

/* BlockCompressor<float>::compress_with_profile (block_compressor.hpp:
 * 248-256) over a batch (layout as xnyz_compress_batch_device): every cloud
 * in block mode with the profile's codebooks (their metas embedded as the
 * reference's load_meta + get_meta writes them), so all chunks share their
 * codebook bytes -- the MPI driver's --shared-codebook mode
 * (source/mpi_driver.cpp:722-779) on one GPU.  A cloud with a symbol the
 * profile lacks is XNYZ_EINVAL. */
int xnyz_compress_with_profile_batch_device(xnyz_ctx* ctx, const float* d_xyz,
                                            const uint64_t* offsets, uint32_t count,
                                            const xnyz_opts* opts, const uint8_t* profile,
                                            uint64_t profile_len, xnyz_buf* frames,
                                            xnyz_buf* prezstd, xnyz_stats* stats);

/* xnyz_compress_with_profile_batch_device writing into caller memory, laid
 * out as by xnyz_compress_batch_device_into (out_offsets: count + 1 entries;
 * XNYZ_ENOSPC with the size needed in out_offsets[count]).  Same
 * counterpart: compress_with_profile (block_compressor.hpp:248-256). */
int xnyz_compress_with_profile_batch_device_into(xnyz_ctx* ctx, const float* d_xyz,
                                                 const uint64_t* offsets, uint32_t count,
                                                 const xnyz_opts* opts, const uint8_t* profile,
                                                 uint64_t profile_len, uint8_t* out,
                                                 uint64_t cap, uint64_t* out_offsets,
                                                 xnyz_stats* stats);

/* The MPI driver's mapreduce mode (source/mpi_driver.cpp:365-648: global
 * bounding box, quantize every chunk against it, shuffle and reduce the cell
 * counts across ranks, then BlockCompressorRLE::encode_cells_with_counts,
 * block_compressor_rle.hpp:341-470) on one GPU: the cross-chunk dedup is
 * the device's one sort of the whole input, so the shuffle and reduce
 * vanish.  The stream is the one a single rank's driver writes: RLE cells
 * and counts, block mode always, the largest quantized coordinate as the
 * range of all three axes, the global minimum as the trailer.  opts supply
 * quantizer, l2_bound, curve and frame_mode (rle is implied).  The minimum
 * follows shift_points (shifting.hpp:11-36): it differs from the driver's
 * FLT_MAX-seeded one only when the first point holds a NaN. */
int xnyz_compress_global_rle_device(xnyz_ctx* ctx, const float* d_xyz, uint64_t n,
                                    const xnyz_opts* opts, xnyz_buf* out_frame,
                                    xnyz_buf* out_prezstd, xnyz_stats* stats);

/* xnyz_compress_global_rle_device writing into caller memory (out_offsets:
 * 2 entries, layout and XNYZ_ENOSPC as for xnyz_compress_batch_device_into).
 * Same counterpart: the mapreduce mode (mpi_driver.cpp:365-648). */
int xnyz_compress_global_rle_device_into(xnyz_ctx* ctx, const float* d_xyz, uint64_t n,
                                         const xnyz_opts* opts, uint8_t* out, uint64_t cap,
                                         uint64_t* out_offsets, xnyz_stats* stats);

/* Exhaustive per-chunk mode search; no reference counterpart (the
 * reference codes every chunk with the options it is given).  For each cloud
 * of the batch (layout as xnyz_compress_batch_device; opts supply quantizer,
 * l2_bound, direct_threshold and the reference curve / RLE flag), the device
 * pipeline runs up to the symbol statistics once per curve x RLE variant
 * (four sorts), and the Golomb-Rice k (0..31 and the reference's rule) and
 * PFD width (1..8) are costed exactly from the device histograms and the
 * patch bit populations.  results: `count` entries. */
int xnyz_search_batch_device(xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets,
                             uint32_t count, const xnyz_opts* opts,
                             xnyz_search_result* results);

/* ---- Learned mode selector (no reference counterpart) -------------------
 * An MLP 32 -> 64 -> 64 (ReLU) -> heads {Golomb-Rice k 0..31, PFD width
 * 1..8, RLE off/on, curve Z/Hilbert}, FP16 weights with FP32 accumulation,
 * trained and run by HIP kernels.  Parameter blob: XNYZ_MLP_PARAMS floats,
 * layout feature mean[32], 1/std[32], W1[32][64], b1[64], W2[64][64], b2[64],
 * W3[64][44], b3[44] (W[i][o]: input i -> output o; logits 0..31 k,
 * 32..39 width 1..8, 40..41 RLE, 42..43 curve). */
#define XNYZ_MLP_FEATURES 32
#define XNYZ_MLP_PARAMS 9196

/* Per cloud, XNYZ_MLP_FEATURES floats from one device pass after the range
 * scan (no sort): log sizes and extents, HyperLogLog estimates of the
 * distinct cells / 64^3 sub-blocks / 128^3 blocks, quantizer and bound, and
 * input-order locality (consecutive points sharing a cell / sub-block /
 * block, histogram of their Morton-code xor magnitudes).  features: host,
 * count * XNYZ_MLP_FEATURES floats. */
int xnyz_features_batch_device(xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets,
                               uint32_t count, const xnyz_opts* opts, float* features);

typedef struct xnyz_train_opts {
  int32_t epochs;
  int32_t batch;   /* samples per Adam step, <= 65536 */
  float lr;        /* Adam learning rate */
  int32_t init;    /* 1: fit the normaliser, He-initialise; 0: continue from params */
  uint64_t seed;
} xnyz_train_opts;

/* Softmax cross-entropy on the four heads, Adam; labels from the exhaustive
 * search (xnyz_search_result.best).  params: in/out, XNYZ_MLP_PARAMS floats
 * (host).  final_loss (nullable): mean loss of the last epoch. */
int xnyz_mlp_train(xnyz_ctx* ctx, const float* features, const xnyz_mode* labels, uint64_t n,
                   const xnyz_train_opts* opts, float* params, float* final_loss);

/* Predicted mode per sample and its confidence (the smallest, over the
 * heads, of the top softmax probability); confidence nullable. */
int xnyz_mlp_predict(xnyz_ctx* ctx, const float* params, const float* features, uint64_t n,
                     xnyz_mode* modes, float* confidence);

/* Batched compression with a per-cloud mode: the selector's prediction
 * where its confidence is >= min_confidence, the exhaustive search's optimum
 * otherwise (params NULL: search every cloud).  opts supply quantizer,
 * l2_bound, direct_threshold and frame_mode; the chosen curve / RLE / k /
 * width replace the rest.  The compressing pass then settles each cloud's
 * Golomb-Rice k and PFD width exactly from its own statistics (the size
 * model of the search, at no extra device pass): a predicted k / width
 * stays where it is optimal and is replaced by the optimum where it costs
 * more.  chosen (nullable, `count` entries) gets the modes used; searched
 * (nullable) gets bit 0 = the cloud fell back to the search, bit 1 = the
 * predicted k / width were replaced.  The RLE flag is not recorded in the
 * stream: decode each cloud with chosen[s].rle (or write the outputs into a
 * container, xnyz_container_*). */
int xnyz_compress_selected_batch_device(xnyz_ctx* ctx, const float* d_xyz,
                                        const uint64_t* offsets, uint32_t count,
                                        const xnyz_opts* opts, const float* params,
                                        float min_confidence, xnyz_buf* frames,
                                        xnyz_buf* prezstd, xnyz_stats* stats,
                                        xnyz_mode* chosen, int32_t* searched);

/* xnyz_compress_selected_batch_device writing into caller memory, laid out
 * as by xnyz_compress_batch_device_into (out_offsets: count + 1 entries;
 * XNYZ_ENOSPC with the size needed in out_offsets[count] when cap is too
 * small).  No reference counterpart beyond compress(). */
int xnyz_compress_selected_batch_device_into(xnyz_ctx* ctx, const float* d_xyz,
                                             const uint64_t* offsets, uint32_t count,
                                             const xnyz_opts* opts, const float* params,
                                             float min_confidence, uint8_t* out, uint64_t cap,
                                             uint64_t* out_offsets, xnyz_stats* stats,
                                             xnyz_mode* chosen, int32_t* searched);

/* xnyz_compress_selected_batch_device_into returning once the assembled
 * streams are in device memory and their copy into `out` is queued on the
 * context's copy stream, as xnyz_compress_batch_device_into_async does (same
 * rules: XNYZ_FRAME_NONE, pinned `out`, two stream buffers, bytes in `out`
 * after xnyz_wait).  The modes are chosen (features, predictions, fallback
 * search) and out_offsets, stats, chosen and searched are final when the
 * call returns; the next call's selection and kernels run while this call's
 * output crosses PCIe.  No reference counterpart beyond compress(). */
int xnyz_compress_selected_batch_device_into_async(
    xnyz_ctx* ctx, const float* d_xyz, const uint64_t* offsets, uint32_t count,
    const xnyz_opts* opts, const float* params, float min_confidence, uint8_t* out, uint64_t cap,
    uint64_t* out_offsets, xnyz_stats* stats, xnyz_mode* chosen, int32_t* searched);

/* Replaces TruncatedOctahedronQuantizer/CubeQuantizer::quantize after
 * shift_points (source/quantizer/truncated_octahedron_quantizer.hpp:21-55,
 * cube_quantizer.hpp:15-27, source/preprocessor/shifting.hpp:11-36) on
 * device input: writes n int32 x,y,z lattice nodes to host `nodes` and the
 * shift offset to host `offset[3]`. */
int xnyz_quantize_device(xnyz_ctx* ctx, const float* d_xyz, uint64_t n,
                         int32_t quantizer, float l2_bound, int32_t* nodes,
                         float* offset);

/* Replaces mve::mortonToHilbert3 (source/quantizer/hilbert-curves-cpp/
 * hilbert.cpp:423-448) on device: out[i] = mortonToHilbert3(in[i], it),
 * host arrays of n codes. */
int xnyz_morton_to_hilbert(xnyz_ctx* ctx, const uint64_t* in, uint64_t n,
                           uint32_t iteration, uint64_t* out);

/* Replaces BlockDecompressor<float>::decompress and BlockDecompressorRLE<
 * float>::decompress (source/composer/block_decompressor.hpp:31-192,
 * block_decompressor_rle.hpp:32-207).  As there, the caller supplies what
 * the stream does not record: opts->quantizer, l2_bound and rle;
 * opts->frame_mode XNYZ_FRAME_NONE means `in` is an assembled (pre-zstd)
 * stream, anything else a zstd frame.  `out_points` receives n*3 floats
 * (len = 12*n bytes) in stream order.  Runs on the host (no device or
 * context needed); xnyz_decompress_device below is the device form. */
int xnyz_decompress(const uint8_t* in, uint64_t len, const xnyz_opts* opts,
                    xnyz_buf* out_points);

/* xnyz_decompress on the device: the host reads the header, section records
 * and metas; the device decodes the Huffman / Exp-Golomb / Golomb-Rice
 * sections in parallel (sub-sequences that synchronise on common code-word
 * starts), rebuilds the block and LCP metadata with one segmented scan
 * (recover_from_lcp_meta, source/quantizer/block_quantizer.hpp:153-207),
 * recovers and unshifts the points.  Same inputs and the same points, in
 * the same order, as xnyz_decompress / BlockDecompressor[RLE]<float>::
 * decompress (source/composer/block_decompressor.hpp:31-192,
 * block_decompressor_rle.hpp:32-207); a zstd frame is opened on the host
 * first.  out_points: host floats, as xnyz_decompress. */
int xnyz_decompress_device(xnyz_ctx* ctx, const uint8_t* in, uint64_t len,
                           const xnyz_opts* opts, xnyz_buf* out_points);

/* The same, leaving the points in device memory: d_points holds room for
 * cap_points points (x, y, z floats); *n_points gets the count (also when
 * the call returns XNYZ_ENOSPC because the room is too small). */
int xnyz_decompress_to_device(xnyz_ctx* ctx, const uint8_t* in, uint64_t len,
                              const xnyz_opts* opts, float* d_points, uint64_t cap_points,
                              uint64_t* n_points);

/* Replaces mve::hilbertToMorton3 (source/quantizer/hilbert-curves-cpp/
 * hilbert.cpp:396-421) on device: out[i] = hilbertToMorton3(in[i], it),
 * host arrays of n codes. */
int xnyz_hilbert_to_morton(xnyz_ctx* ctx, const uint64_t* in, uint64_t n,
                           uint32_t iteration, uint64_t* out);

/* ---- Multi-chunk container ------------------------------------------------
 * A part file holding the chunks' outputs back to back and a manifest of
 * one JSON line per chunk, in the shape of the MPI driver's per-rank part
 * file and manifest (source/mpi_driver.cpp:709-714, :902-914: block_id,
 * input_offset, input_size, rank, part_offset, compressed_size), with each
 * chunk's coding mode added: quantizer, l2_bound, curve, rle, gr_k,
 * pfd_width, direct_threshold, frame ("zstd" or "stream").  The stream does
 * not record the RLE flag or the quantizer, and a learned batch codes every
 * chunk its own way, so the container keeps them.  Manifests written by the
 * reference's driver read too (mode fields absent: TO, zstd frames; the
 * caller supplies bound, curve and RLE). */
typedef struct xnyz_chunk {
  uint64_t input_offset;     /* byte offset of the chunk in the input (12 per point) */
  uint64_t input_points;
  uint64_t part_offset;      /* set by xnyz_container_write */
  uint64_t compressed_size;  /* set by xnyz_container_write */
  int32_t quantizer;         /* enum xnyz_quantizer */
  float l2_bound;
  int32_t is_hilbert;
  int32_t rle;
  int32_t gr_k;              /* -1: not recorded */
  int32_t pfd_width;
  int32_t direct_threshold;
  int32_t frame_mode;        /* XNYZ_FRAME_NONE: assembled stream; otherwise a zstd frame */
} xnyz_chunk;

/* Writes chunk s's output out[out_offsets[s], out_offsets[s + 1]) (the layout
 * of the _into entry points) to the part file and its line to the manifest;
 * fills each chunk's part_offset and compressed_size. */
int xnyz_container_write(const char* part_path, const char* manifest_path, const uint8_t* out,
                         const uint64_t* out_offsets, uint32_t count, xnyz_chunk* chunks);

/* The manifest's chunks: chunks->data holds len / sizeof(xnyz_chunk)
 * entries (release with xnyz_free). */
int xnyz_container_read(const char* manifest_path, xnyz_buf* chunks);

/* Decodes every chunk with its own mode, in manifest order, on the device of
 * ctx (or on the host when ctx is NULL): out_points gets all points, chunk
 * after chunk, each in its decoder's order (x, y, z floats). */
int xnyz_container_decompress(xnyz_ctx* ctx, const char* part_path, const char* manifest_path,
                              xnyz_buf* out_points);

/* Stage-2 launch sequences (key generation, radix passes, RLE cells, block
 * runs, statistics: some 40 kernels) are captured into a hipGraph per launch
 * shape and replayed while the shape repeats (same cloud sizes and options
 * class); on by default, off while profiling.  No reference counterpart. */
int xnyz_graphs_enable(xnyz_ctx* ctx, int on);

/* Run-time parameters of a context, kept per size class (class c holds the
 * calls with 2^c <= total points < 2^(c+1); c in [0, 32]).  No reference
 * counterpart: the reference's comparable knobs are compile-time constants
 * or absent (one ZSTD_compress call, source/composer/utils.hpp:35-53).
 *   zstd_chunk_log   host-threaded frames: log2 bytes per chunk, 16..24 (20)
 *   copy_slice_log   copy-back of the stream in slices of 2^x bytes, 20..30 (24)
 *   gpu_frame_share  share of a single cloud's host-threaded frame whose
 *                    blocks the device encodes (the frame's tail, while the
 *                    host threads compress the head), 0..0.9 (0)
 *   persist_blocks   workgroups of the persistent reduction kernels,
 *                    64..16384 (2048)
 *   copy_first       split frames: 1 = the head's copy-back runs before the
 *                    device encoder on the context's stream, 0 = on a second
 *                    stream beside it (1)
 *   radix_bits       digit width of the key sort, 6..8 bits (8)
 *   radix_tile       keys per radix-sort workgroup, 2048 or 4096 (4096)
 * The defaults are given in parentheses.  Every setting yields a valid frame
 * of the same stream; only the frame bytes and the speed change. */
typedef struct xnyz_tuning {
  uint32_t zstd_chunk_log;
  uint32_t copy_slice_log;
  float gpu_frame_share;
  uint32_t persist_blocks;
  uint32_t copy_first;
  uint32_t radix_bits;
  uint32_t radix_tile;
} xnyz_tuning;

#define XNYZ_SIZE_CLASSES 33

/* size_class -1 sets every class.  No reference counterpart. */
int xnyz_tuning_set(xnyz_ctx* ctx, int size_class, const xnyz_tuning* t);
int xnyz_tuning_get(xnyz_ctx* ctx, int size_class, xnyz_tuning* t);

/* The table as text, one line per class ("class chunk_log slice_log share
 * blocks copy_first radix_bits radix_tile"); load accepts any subset of the classes.  No reference
 * counterpart. */
int xnyz_tuning_save(xnyz_ctx* ctx, const char* path);
int xnyz_tuning_load(xnyz_ctx* ctx, const char* path);

/* Autotuner: times compressions of the device-resident cloud (n points,
 * options `opts`, frames into scratch owned by the context) over candidate
 * settings -- one parameter at a time with the others held at their best so
 * far, the device share and copy order of a split frame as a grid, the
 * kernel shapes once the share is set -- `reps` timed calls per candidate
 * after one untimed call (a candidate scores its fastest call: wall time,
 * or for the kernel shapes the device time of the stream-building kernels),
 * and stores the best setting as the tuning of n's size class.  *best gets
 * it, *best_ms its wall time per call.  No reference counterpart. */
int xnyz_autotune(xnyz_ctx* ctx, const float* d_xyz, uint64_t n, const xnyz_opts* opts,
                  uint32_t reps, xnyz_tuning* best, double* best_ms);

/* Per-kernel device time (HIP events on the context's stream) accumulated
 * over the calls made while enabled; enabling clears the totals.  The report
 * is a JSON object {"kernel": {"calls": c, "ms": t}, ...} written to buf
 * (NUL-terminated, truncated to cap); *needed gets the full size.  No
 * reference counterpart (the reference's only timer is source/statistics/
 * timer.hpp around the whole compress call). */
int xnyz_profile_enable(xnyz_ctx* ctx, int on);
int xnyz_profile_report(xnyz_ctx* ctx, char* buf, uint64_t cap, uint64_t* needed);

/* Releases a buffer returned in an xnyz_buf (also for the reference's
 * returned std::vector, which owned its bytes). */
void xnyz_free(xnyz_buf* buf);

/* Pinned (page-locked) host memory for the `out` of
 * xnyz_compress_batch_device_into: when the output lies inside one such
 * buffer, the device-built blocks of a split frame are copied there by DMA
 * directly instead of through a staging buffer.  No reference counterpart
 * (the reference returns a std::vector).  xnyz_host_free takes only
 * pointers from xnyz_host_alloc (XNYZ_EINVAL otherwise); free a buffer only
 * when no call writing into it is in flight. */
int xnyz_host_alloc(uint64_t bytes, void** out);
int xnyz_host_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
