/*
 * xnyz_oracle.h -- CPU restatement of XnYZip's compression path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path in xnyzip_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the timed CPU
 * baseline -- never as the thing measured or shipped.  The product library
 * (libxnyz.so) does not link or call it.
 *
 * Every function follows the reference file:line cited beside it in
 * xnyz_oracle.cpp.  Paths are relative to the reference repository root.
 */
#ifndef XNYZ_ORACLE_H
#define XNYZ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_CUBE = 0, ORACLE_TO = 1 };

typedef struct {
  uint64_t num_points;      /* N, or U (unique cells) in RLE mode          */
  uint64_t input_points;    /* N                                           */
  int32_t range[3];         /* max quantized x, y, z (clamped at 0)        */
  int32_t is_direct;
  uint64_t blocks;          /* occupied 128^3 blocks (block mode)          */
  uint64_t patches;         /* repos / sfc-delta PFD patches               */
  uint64_t quad_patches;    /* quads PFD patches (always 0 in practice)    */
  int32_t counts_exp_golomb;/* RLE counts stream coded with Exp-Golomb     */
  int32_t counts_eg_k;
  double max_l2;            /* RLE diagnostic: max |p - recover(q(p))|     */
  double mse;               /* RLE diagnostic                              */
  double seconds_total;     /* construct + compress(), like main.cpp:85-91 */
  double seconds_stream;    /* same, without the RLE diagnostic pass        */
  uint64_t wrapped_points;  /* points with quantized z == -1               */
} oracle_info;

/* BlockCompressor[RLE]<float>::compress (block_compressor.hpp:39-222,
 * block_compressor_rle.hpp:38-321).  Returns 0 on success.  `pre` receives
 * the assembled buffer handed to zstd, `frame` the ZSTD_compress(level 3)
 * output.  `perm` (nullable, n entries) receives the reordering the
 * reference applies to the caller's points (ords / sfc order). */
int oracle_compress(const float* xyz, uint64_t n, int quantizer, float l2_bound,
                    int is_hilbert, int rle, int direct_threshold,
                    uint8_t** pre, uint64_t* pre_len,
                    uint8_t** frame, uint64_t* frame_len,
                    uint64_t* perm, oracle_info* info);

/* Same pipeline, timed; `with_zstd` 0 stops at the assembled buffer. */
/* oracle_compress with the mode overrides of the repos / sfc-delta stream
 * (test infrastructure for the mode search): gr_k < 0 keeps the reference's
 * Golomb-Rice k rule; pfd_width 0 keeps the PFD limit 254, 1..8 sets it to
 * 2^w - 2.  Assembled stream only. */
int oracle_compress_mode(const float* xyz, uint64_t n, int quantizer, float l2_bound,
                         int is_hilbert, int rle, int direct_threshold, int gr_k, int pfd_width,
                         uint8_t** pre, uint64_t* pre_len, oracle_info* info);

int oracle_compress_timed(const float* xyz, uint64_t n, int quantizer,
                          float l2_bound, int is_hilbert, int rle,
                          int direct_threshold, int with_zstd,
                          uint64_t* out_bytes, oracle_info* info);

/* BlockDecompressor[RLE]<float>::decompress on the assembled (pre-zstd)
 * buffer or on a zstd frame (block_decompressor.hpp:31-192,
 * block_decompressor_rle.hpp:32-207). */
int oracle_decompress_buffer(const uint8_t* buf, uint64_t len, int quantizer,
                             float l2_bound, int rle, float** xyz,
                             uint64_t* n);
int oracle_decompress_frame(const uint8_t* frame, uint64_t len, int quantizer,
                            float l2_bound, int rle, float** xyz, uint64_t* n);

/* Building blocks, exported so tests can pin each one separately. */
void oracle_shift(const float* xyz, uint64_t n, float* shifted, float* offset);
void oracle_quantize(const float* xyz, uint64_t n, int quantizer,
                     float l2_bound, int32_t* nodes, float* offset);
void oracle_recover(const int32_t* nodes, uint64_t n, int quantizer,
                    float l2_bound, float* out);
uint64_t oracle_morton_code(int32_t x, int32_t y, int32_t z);
void oracle_decode_morton(uint64_t m, int32_t* xyz);
uint64_t oracle_morton_to_hilbert3(uint64_t morton, unsigned iteration);
uint64_t oracle_hilbert_to_morton3(uint64_t hilbert, unsigned iteration);
/* m2h_perm[8], h2m_perm[8], m2h_perms[64], h2m_perms[64] */
void oracle_hilbert_tables(uint8_t* m2h_perm, uint8_t* h2m_perm,
                           uint8_t* m2h_perms, uint8_t* h2m_perms);

int oracle_bitpack_u64(const uint64_t* v, uint64_t n, uint8_t** meta,
                       uint64_t* meta_len, uint8_t** data, uint64_t* data_len);
int oracle_gr_pack_u64(const uint64_t* v, uint64_t n, uint8_t** meta,
                       uint64_t* meta_len, uint8_t** data, uint64_t* data_len);
int oracle_gr_pack_u8(const uint8_t* v, uint64_t n, uint8_t** meta,
                      uint64_t* meta_len, uint8_t** data, uint64_t* data_len);
/* Huffman/Exp-Golomb: sym_type 0 = uint64_t, 1 = uint8_t, 2 = int.
 * mode 0: one build() (repos / quads streams);
 * mode 1: the RLE counts double build (block_compressor_rle.hpp:290-303);
 * mode 2: build() with use_exp_golomb preset. */
/* BlockCompressor<float>::fit_profile (profile_in NULL; profile_out gets the
 * serialize_profile blob) and compress_with_profile (profile_in given)
 * (block_compressor.hpp:241-431).  Assembled stream only. */
int oracle_compress_shared(const float* xyz, uint64_t n, int quantizer, float l2_bound,
                           int is_hilbert, const uint8_t* profile_in, uint64_t profile_len,
                           uint8_t** pre, uint64_t* pre_len, uint8_t** profile_out,
                           uint64_t* profile_out_len);

/* The MPI driver's mapreduce mode on one rank (mpi_driver.cpp:365-648) and
 * BlockCompressorRLE::encode_cells_with_counts (block_compressor_rle.hpp:
 * 341-470): the assembled stream of the whole input's global RLE. */
int oracle_compress_global_rle(const float* xyz, uint64_t n, int quantizer, float l2_bound,
                               int is_hilbert, uint8_t** pre, uint64_t* pre_len);

int oracle_huffman(const void* v, uint64_t n, int sym_type, int mode,
                   uint8_t** meta, uint64_t* meta_len, uint8_t** data,
                   uint64_t* data_len);
/* Decoders for the same streams (round-trip pins). */
int oracle_huffman_decode(const uint8_t* meta, uint64_t meta_len,
                          const uint8_t* data, uint64_t data_len,
                          uint64_t count, int sym_type, void* out);
int oracle_gr_unpack_u64(const uint8_t* meta, uint64_t meta_len,
                         const uint8_t* data, uint64_t data_len,
                         uint64_t* out, uint64_t* n);
int oracle_bitunpack_u64(const uint8_t* meta, uint64_t meta_len,
                         const uint8_t* data, uint64_t data_len,
                         uint64_t* out, uint64_t* n);

uint64_t oracle_zstd_version(void);
const char* oracle_last_error(void);
void oracle_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
