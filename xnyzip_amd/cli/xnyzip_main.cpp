// XnYZip CLI drop-in (source/main.cpp:48-203) on libxnyz.so.
//
//   XnYZip <input_file> <cube|to> <L2 bound> <-z|-h> <-rle|-normal>
//          <decompression_file_name> [direct_threshold] [--offset=B] [--count=B]
//
// and, for many chunks of one file on one GPU (the MPI driver's per-block
// loop, source/mpi_driver.cpp:819-943, as one batched call), a container of
// part_rank0000.bin + manifest_rank0000.jsonl in the driver's shape:
//
//   XnYZip --batch <input_file> <cube|to> <L2 bound> <-z|-h|-auto> <-rle|-normal|-auto>
//          <output_dir> [--chunk-points=N] [--stream]
//   XnYZip --unbatch <output_dir> <decompression_file_name> [--cpu]
//   XnYZip --mapreduce <input_file> <cube|to> <L2 bound> <-z|-h> <output_dir>
//
// --mapreduce is the MPI driver's global RLE (--mode mapreduce,
// source/mpi_driver.cpp:365-648) on one GPU: one stream of the whole file's
// cells and counts, its part file and one manifest line; --unbatch reads it.
//
// -auto lets the learned selector (falling back to the exhaustive search)
// pick each chunk's curve and RLE flag; the manifest records the choice.
//
// Same arguments, chunk flags (main.cpp:15-45, source/io/fileUtils.hpp:44-97)
// and stdout lines as the reference: the lines its harness scripts parse
// ("compression ratio:", "chunk input bytes:", "chunk compressed bytes:",
// "XnYZip compression time = ...s", "Encoding mode:", "Unique points:",
// "Duplicate ratio:", "Max L2 error:", "p2p psnr:") keep their formats.
// Compression runs on GPU 0 through the C ABI; the reordering the reference
// applies to its input vector comes back as a permutation.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>
#include <unistd.h>

#include "xnyz.h"

namespace {

struct ChunkSpec {
  bool active = false;
  uint64_t offset = 0, count = 0;
};

ChunkSpec parse_chunk_flags(int argc, char** argv) {
  ChunkSpec s;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("--offset=", 0) == 0) {
      s.offset = std::stoull(a.substr(9));
      s.active = true;
    } else if (a.rfind("--count=", 0) == 0) {
      s.count = std::stoull(a.substr(8));
      s.active = true;
    }
  }
  return s;
}

uint64_t file_size(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f.is_open()) throw std::runtime_error("Failed to open file: " + path);
  auto pos = f.tellg();
  if (pos < 0) throw std::runtime_error("Failed to query file size: " + path);
  return uint64_t(pos);
}

// fileUtils.hpp:13-29 (whole file, trailing partial point dropped) and
// :44-97 (byte range: offset and count multiples of 12; count 0 or past the
// end means the rest of the file).
std::vector<float> read_points(const std::string& path, const ChunkSpec& c) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) throw std::runtime_error("Failed to open file: " + path);
  const uint64_t size = file_size(path);
  uint64_t off = 0, count = size - size % 12;
  if (c.active) {
    if (c.offset % 12) throw std::runtime_error("offset_bytes must be a multiple of 12");
    if (c.count % 12) throw std::runtime_error("count_bytes must be a multiple of 12");
    if (c.offset >= size) return {};
    const uint64_t rem = size - c.offset;
    off = c.offset;
    count = (c.count == 0 || c.count > rem) ? rem - rem % 12 : c.count;
  }
  std::vector<float> v(count / 4);
  f.seekg(std::streamoff(off), std::ios::beg);
  if (!f.read(reinterpret_cast<char*>(v.data()), std::streamsize(count)))
    throw std::runtime_error("read_file_chunk: short read");
  return v;
}

void write_floats(const std::string& path, const float* p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  if (!f.is_open()) throw std::runtime_error("Failed to open file: " + path);
  f.write(reinterpret_cast<const char*>(p), std::streamsize(n * sizeof(float)));
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int die(const char* what) {
  std::fprintf(stderr, "XnYZip: %s: %s\n", what, xnyz_last_error());
  return 1;
}

std::string flag_value(int argc, char** argv, const char* name) {
  const std::string k = std::string(name) + "=";
  for (int i = 1; i < argc; ++i)
    if (std::string(argv[i]).rfind(k, 0) == 0) return std::string(argv[i]).substr(k.size());
  return "";
}

bool has_flag(int argc, char** argv, const char* name) {
  for (int i = 1; i < argc; ++i)
    if (std::string(argv[i]) == name) return true;
  return false;
}

// The packaged selector parameters (xnyzip_amd/data/selector_params.npy,
// next to the binary's ../data): a version-1 .npy of float32, read as raw
// little-endian floats after its header.
std::vector<float> load_selector() {
  char self[4096];
  const ssize_t n = readlink("/proc/self/exe", self, sizeof(self) - 1);
  if (n <= 0) throw std::runtime_error("cannot locate the executable");
  self[n] = 0;
  std::string dir(self);
  dir = dir.substr(0, dir.rfind('/'));
  const std::string path = dir + "/../data/selector_params.npy";
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) throw std::runtime_error("cannot open " + path);
  std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (b.size() < 10 || std::memcmp(b.data(), "\x93NUMPY", 6) != 0)
    throw std::runtime_error(path + ": not a .npy file");
  const size_t hl = (unsigned char)b[8] | (size_t)(unsigned char)b[9] << 8;
  const size_t at = 10 + hl;
  if (b.size() != at + XNYZ_MLP_PARAMS * 4) throw std::runtime_error(path + ": wrong size");
  std::vector<float> p(XNYZ_MLP_PARAMS);
  std::memcpy(p.data(), b.data() + at, p.size() * 4);
  return p;
}

int run_batch(int argc, char** argv, const std::vector<char*>& pos) {
  if (pos.size() != 8) {
    std::printf("Usage: %s --batch <input_file> <cube|to> <L2 bound> <-z|-h|-auto> "
                "<-rle|-normal|-auto> <output_dir> [--chunk-points=N] [--stream]\n", argv[0]);
    return 1;
  }
  const std::vector<float> pts = read_points(pos[2], ChunkSpec{});
  const uint64_t n = pts.size() / 3;
  if (!n) {
    std::fprintf(stderr, "empty input\n");
    return 1;
  }
  const std::string cs = flag_value(argc, argv, "--chunk-points");
  const uint64_t chunk = cs.empty() ? (1u << 20) : std::stoull(cs);
  if (!chunk) {
    std::fprintf(stderr, "--chunk-points must be positive\n");
    return 1;
  }
  const std::string curve = pos[5], rmode = pos[6];
  const bool learned = curve == "-auto" || rmode == "-auto";
  std::vector<uint64_t> offs;
  for (uint64_t a = 0; a < n; a += chunk) offs.push_back(a);
  offs.push_back(n);
  const uint32_t count = uint32_t(offs.size() - 1);
  xnyz_opts base{};
  base.quantizer = std::string(pos[3]) == "to" ? XNYZ_TRUNCATED_OCTAHEDRON : XNYZ_CUBE;
  base.l2_bound = std::stof(pos[4]);
  base.is_hilbert = curve == "-h";
  base.rle = rmode == "-rle";
  base.direct_threshold = 1024;
  base.frame_mode = has_flag(argc, argv, "--stream") ? XNYZ_FRAME_NONE : XNYZ_FRAME_ZSTD3;
  std::vector<xnyz_opts> opts(count, base);
  auto t0 = std::chrono::steady_clock::now();
  xnyz_ctx* ctx = nullptr;
  if (xnyz_create(0, &ctx) != XNYZ_OK) return die("xnyz_create");
  float* d = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&d), n * 12) != hipSuccess ||
      hipMemcpy(d, pts.data(), n * 12, hipMemcpyHostToDevice) != hipSuccess) {
    std::fprintf(stderr, "XnYZip: device upload failed\n");
    return 1;
  }
  std::vector<uint64_t> oo(count + 1);
  std::vector<xnyz_stats> st(count);
  std::vector<xnyz_mode> modes(count);
  std::vector<uint8_t> out(n * 13 + (1u << 20));
  int rc;
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (learned) {
      const std::vector<float> params = load_selector();
      rc = xnyz_compress_selected_batch_device_into(ctx, d, offs.data(), count, opts.data(),
                                                    params.data(), 0.5f, out.data(), out.size(),
                                                    oo.data(), st.data(), modes.data(), nullptr);
    } else {
      rc = xnyz_compress_batch_device_into(ctx, d, offs.data(), count, opts.data(), out.data(),
                                           out.size(), oo.data(), st.data());
      for (uint32_t s = 0; s < count; ++s)
        modes[s] = xnyz_mode{base.is_hilbert, base.rle, -1, 8};
    }
    if (rc != XNYZ_ENOSPC) break;
    out.resize(oo[count] + (1u << 20));
  }
  if (rc != XNYZ_OK) return die("batched compression");
  const double tc = seconds_since(t0);
  std::vector<xnyz_chunk> ch(count);
  for (uint32_t s = 0; s < count; ++s) {
    xnyz_chunk& c = ch[s];
    c.input_offset = offs[s] * 12;
    c.input_points = offs[s + 1] - offs[s];
    c.quantizer = base.quantizer;
    c.l2_bound = base.l2_bound;
    c.is_hilbert = modes[s].is_hilbert;
    c.rle = modes[s].rle;
    c.gr_k = modes[s].gr_k;
    c.pfd_width = modes[s].pfd_width;
    c.direct_threshold = base.direct_threshold;
    c.frame_mode = base.frame_mode;
  }
  const std::string dir = pos[7];
  const std::string part = dir + "/part_rank0000.bin", man = dir + "/manifest_rank0000.jsonl";
  if (xnyz_container_write(part.c_str(), man.c_str(), out.data(), oo.data(), count, ch.data()) !=
      XNYZ_OK)
    return die("xnyz_container_write");
  const uint64_t bytes = oo[count] - oo[0];
  std::printf("chunks: %u\n", count);
  std::printf("chunk input bytes: %lu\n", (unsigned long)(n * 12));
  std::printf("chunk compressed bytes: %lu\n", (unsigned long)bytes);
  std::printf("compression ratio: %f\n", (double)n * 12 / (double)bytes);
  std::cout << "XnYZip compression time = " << tc << "s" << std::endl;
  (void)hipFree(d);
  xnyz_destroy(ctx);
  return 0;
}

int run_mapreduce(int argc, char** argv, const std::vector<char*>& pos) {
  (void)argc;
  if (pos.size() != 7) {
    std::printf("Usage: %s --mapreduce <input_file> <cube|to> <L2 bound> <-z|-h> <output_dir>\n",
                argv[0]);
    return 1;
  }
  const std::vector<float> pts = read_points(pos[2], ChunkSpec{});
  const uint64_t n = pts.size() / 3;
  if (!n) {
    std::fprintf(stderr, "empty input\n");
    return 1;
  }
  xnyz_opts o{};
  o.quantizer = std::string(pos[3]) == "to" ? XNYZ_TRUNCATED_OCTAHEDRON : XNYZ_CUBE;
  o.l2_bound = std::stof(pos[4]);
  o.is_hilbert = std::string(pos[5]) == "-h";
  o.rle = 1;
  o.direct_threshold = 1024;
  o.frame_mode = XNYZ_FRAME_ZSTD3;
  auto t0 = std::chrono::steady_clock::now();
  xnyz_ctx* ctx = nullptr;
  if (xnyz_create(0, &ctx) != XNYZ_OK) return die("xnyz_create");
  float* d = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&d), n * 12) != hipSuccess ||
      hipMemcpy(d, pts.data(), n * 12, hipMemcpyHostToDevice) != hipSuccess) {
    std::fprintf(stderr, "XnYZip: device upload failed\n");
    return 1;
  }
  xnyz_buf frame{};
  xnyz_stats st{};
  if (xnyz_compress_global_rle_device(ctx, d, n, &o, &frame, nullptr, &st) != XNYZ_OK)
    return die("xnyz_compress_global_rle_device");
  const double tc = seconds_since(t0);
  const std::string dir = pos[6];
  std::ofstream part(dir + "/part_rank0000.bin", std::ios::binary);
  std::ofstream meta(dir + "/manifest_rank0000.jsonl");
  if (!part.is_open() || !meta.is_open()) {
    std::fprintf(stderr, "cannot open output files in %s\n", dir.c_str());
    return 1;
  }
  part.write(reinterpret_cast<const char*>(frame.data), std::streamsize(frame.len));
  // mpi_driver.cpp:626-635, with the mode fields the decoder needs
  meta << "{\"rank\": 0, \"mode\": \"mapreduce\", \"input_bytes\": " << n * 12
       << ", \"input_blocks_owned\": 1, \"owned_cells\": " << st.num_points
       << ", \"part_offset\": 0, \"compressed_size\": " << frame.len
       << ", \"quantizer\": \"" << pos[3] << "\", \"l2_bound\": " << pos[4]
       << ", \"curve\": \"" << (o.is_hilbert ? "hilbert" : "z") << "\", \"rle\": 1}\n";
  std::printf("[mapreduce] rank 0: input=%llu B, encoded %llu cells -> %llu B\n",
              (unsigned long long)(n * 12), (unsigned long long)st.num_points,
              (unsigned long long)frame.len);
  std::printf("compression ratio: %f\n", (double)n * 12 / (double)frame.len);
  std::cout << "XnYZip compression time = " << tc << "s" << std::endl;
  xnyz_free(&frame);
  (void)hipFree(d);
  xnyz_destroy(ctx);
  return 0;
}

int run_unbatch(int argc, char** argv, const std::vector<char*>& pos) {
  if (pos.size() != 4) {
    std::printf("Usage: %s --unbatch <output_dir> <decompression_file_name> [--cpu]\n", argv[0]);
    return 1;
  }
  const std::string dir = pos[2];
  const std::string part = dir + "/part_rank0000.bin", man = dir + "/manifest_rank0000.jsonl";
  xnyz_ctx* ctx = nullptr;
  if (!has_flag(argc, argv, "--cpu") && xnyz_create(0, &ctx) != XNYZ_OK) return die("xnyz_create");
  auto t0 = std::chrono::steady_clock::now();
  xnyz_buf b{};
  if (xnyz_container_decompress(ctx, part.c_str(), man.c_str(), &b) != XNYZ_OK)
    return die("xnyz_container_decompress");
  std::cout << "XnYZip decompression time = " << seconds_since(t0) << "s" << std::endl;
  std::printf("decompressed finished, size: %lu\n", (unsigned long)(b.len / 12));
  write_floats(pos[3], reinterpret_cast<const float*>(b.data), b.len / 4);
  xnyz_free(&b);
  if (ctx) xnyz_destroy(ctx);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  {
    std::vector<char*> p;
    for (int i = 0; i < argc; ++i)
      if (std::string(argv[i]).rfind("--", 0) != 0 || i == 1) p.push_back(argv[i]);
    try {
      if (argc > 1 && std::string(argv[1]) == "--batch") return run_batch(argc, argv, p);
      if (argc > 1 && std::string(argv[1]) == "--unbatch") return run_unbatch(argc, argv, p);
      if (argc > 1 && std::string(argv[1]) == "--mapreduce") return run_mapreduce(argc, argv, p);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "%s\n", e.what());
      return 1;
    }
  }
  ChunkSpec chunk = parse_chunk_flags(argc, argv);
  std::vector<char*> pos;
  for (int i = 0; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("--offset=", 0) == 0 || a.rfind("--count=", 0) == 0) continue;
    pos.push_back(argv[i]);
  }
  const int pargc = int(pos.size());
  if (pargc < 7 || pargc > 8) {
    std::printf("Usage: %s <input_file> <quantizer_type (cube/to)> <L2 bound> <curve_type (-z/-h)> "
                "<-rle/-normal> <decompression_file_name> [direct_threshold (default 1024, "
                "0=force block)] [--offset=BYTES] [--count=BYTES]\n", argv[0]);
    return 1;
  }
  std::vector<float> pts;
  try {
    pts = read_points(pos[1], chunk);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  const uint64_t n = pts.size() / 3;
  if (chunk.active)
    std::cout << "Chunk mode: file_size=" << file_size(pos[1]) << " offset=" << chunk.offset
              << " count=" << chunk.count << " points=" << n << std::endl;
  const bool hilbert = std::string(pos[4]) == "-h";
  const bool rle = std::string(pos[5]) == "-rle";
  const int thr = pargc >= 8 ? std::stoi(pos[7]) : 1024;
  std::cout << "SFC type: " << (hilbert ? "Hilbert" : "Z-order") << std::endl;
  std::cout << "Use RLE: " << (rle ? "Yes" : "No") << std::endl;
  std::cout << "Direct threshold: " << thr << std::endl;
  const uint64_t input_bytes = n * 3 * sizeof(float);
  uint64_t compressed_bytes = 0;
  const std::string qname = pos[2];
  const float bound = std::stof(pos[3]);
  std::vector<float> rec;
  std::vector<uint32_t> perm;

  if (qname == "cube" || qname == "to") {
    xnyz_opts o{};
    o.quantizer = qname == "to" ? XNYZ_TRUNCATED_OCTAHEDRON : XNYZ_CUBE;
    o.l2_bound = bound;
    o.is_hilbert = hilbert;
    o.rle = rle;
    o.direct_threshold = thr;
    o.frame_mode = XNYZ_FRAME_ZSTD3;
    if (!rle) perm.resize(n);
    auto t0 = std::chrono::steady_clock::now();
    xnyz_ctx* ctx = nullptr;
    if (xnyz_create(0, &ctx) != XNYZ_OK) return die("xnyz_create");
    xnyz_buf frame{};
    xnyz_stats st{};
    if (xnyz_compress(ctx, pts.data(), n, &o, &frame, nullptr, rle ? nullptr : perm.data(),
                      &st) != XNYZ_OK)
      return die("xnyz_compress");
    const double tc = seconds_since(t0);
    if (rle) {  // lines printed inside BlockCompressorRLE::compress
      float mx[3] = {-1e9f, -1e9f, -1e9f}, mn[3] = {1e9f, 1e9f, 1e9f};
      for (uint64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
          mx[a] = std::max(mx[a], pts[3 * i + a]);
          mn[a] = std::min(mn[a], pts[3 * i + a]);
        }
      const double range = (mx[0] - mn[0]) * (mx[0] - mn[0]) + (mx[1] - mn[1]) * (mx[1] - mn[1]) +
                           (mx[2] - mn[2]) * (mx[2] - mn[2]);
      std::printf("Max L2 error: %lf\n", st.max_l2);
      std::printf("p2p psnr: %lf\n", 10.0 * std::log10(range / st.mse));
      std::printf("MSE: %f, range: %f\n", st.mse, range);
      std::printf("Max diff: %f, %f, %f\n", mx[0], mx[1], mx[2]);
      std::printf("Unique points: %lu\n", (unsigned long)st.num_points);
      std::printf("Duplicate ratio: %.3f\n", 1.0 - (double)st.num_points / n);
    }
    std::cout << "Encoding mode: " << (st.is_direct ? "direct" : "block") << std::endl;
    std::cout << "compressed size of coords (after zstd): " << frame.len << std::endl;
    std::cout << "compression ratio of coords (after zstd): "
              << (double)n * 3 * sizeof(float) / frame.len << std::endl;
    std::cout << "XnYZip compression time = " << tc << "s" << std::endl;
    compressed_bytes = frame.len;
    std::printf("compression ratio: %f\n",
                (float)n * 3 * sizeof(float) / frame.len * sizeof(uint8_t));
    auto t1 = std::chrono::steady_clock::now();
    xnyz_buf out{};
    if (xnyz_decompress(frame.data, frame.len, &o, &out) != XNYZ_OK) return die("xnyz_decompress");
    std::cout << "XnYZip decompression time = " << seconds_since(t1) << "s" << std::endl;
    rec.assign(reinterpret_cast<float*>(out.data), reinterpret_cast<float*>(out.data) + out.len / 4);
    xnyz_free(&out);
    xnyz_free(&frame);
    xnyz_destroy(ctx);
    if (qname == "to") std::printf("decompressed finished, size: %lu\n", (unsigned long)(rec.size() / 3));
  }

  std::printf("chunk input bytes: %lu\n", (unsigned long)input_bytes);
  std::printf("chunk compressed bytes: %lu\n", (unsigned long)compressed_bytes);
  try {
    write_floats(pos[6], rec.data(), rec.size());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  if (!rle) {  // main.cpp:161-201, against the reordered input
    double acc = 0;
    float max_l2 = 0;
    int overflow = 0;
    std::vector<float> errors;
    const uint64_t m = std::min<uint64_t>(n, rec.size() / 3);
    for (uint64_t i = 0; i < m; ++i) {
      const float* p = &pts[3 * uint64_t(perm.empty() ? i : perm[i])];
      const float* r = &rec[3 * i];
      float d2 = 0;
      for (int a = 0; a < 3; ++a) {
        acc += (p[a] - r[a]) * (p[a] - r[a]);
      }
      const float dx = p[0] - r[0], dy = p[1] - r[1], dz = p[2] - r[2];
      d2 = dx * dx + (dy * dy + dz * dz);
      const float e = std::sqrt(d2);
      errors.push_back(e);
      if (e > bound) ++overflow;
      max_l2 = std::max(max_l2, e);
    }
    write_floats("errors-" + qname + ".bin", errors.data(), errors.size());
    std::printf("overflow count: %d\n", overflow);
    std::printf("Max L2 error: %f\n", max_l2);
    acc /= n;
    float mx[3] = {-1e9f, -1e9f, -1e9f}, mn[3] = {1e9f, 1e9f, 1e9f};
    for (uint64_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        mx[a] = std::max(mx[a], pts[3 * i + a]);
        mn[a] = std::min(mn[a], pts[3 * i + a]);
      }
    const double range = (mx[0] - mn[0]) * (mx[0] - mn[0]) + (mx[1] - mn[1]) * (mx[1] - mn[1]) +
                         (mx[2] - mn[2]) * (mx[2] - mn[2]);
    std::printf("p2p psnr: %lf\n", 10.0 * std::log10(range / acc));
    std::printf("MSE: %f, range: %f\n", acc, range);
    std::printf("Max diff: %f, %f, %f\n", mx[0], mx[1], mx[2]);
  }
  return 0;
}
