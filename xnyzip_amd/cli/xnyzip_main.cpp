// XnYZip CLI drop-in (source/main.cpp:48-203) on libxnyz.so.
//
//   XnYZip <input_file> <cube|to> <L2 bound> <-z|-h> <-rle|-normal>
//          <decompression_file_name> [direct_threshold] [--offset=B] [--count=B]
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

}  // namespace

int main(int argc, char** argv) {
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
