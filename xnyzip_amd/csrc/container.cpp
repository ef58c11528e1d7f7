// container.cpp -- multi-chunk container: a part file holding the chunks'
// outputs back to back, and a manifest with one JSON line per chunk.
//
// The shape is the MPI driver's (source/mpi_driver.cpp:709-714 opens
// part_rankNNNN.bin and manifest_rankNNNN.jsonl; :902-914 writes, per block,
// block_id, input_offset, input_size, rank, part_offset, compressed_size).
// Each line here also carries the chunk's coding mode -- quantizer, bound,
// curve, RLE flag, Golomb-Rice k, PFD width, direct threshold -- and whether
// the chunk is a zstd frame or an assembled stream: the stream does not
// record the RLE flag or the quantizer, and chunks of one container may be
// coded differently (the learned selector picks per chunk).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/xnyz.h"
#include "kernels.hpp"

namespace xnyz {

namespace {

const char* quantizer_name(int q) { return q == XNYZ_TRUNCATED_OCTAHEDRON ? "to" : "cube"; }

std::string line_of(uint32_t id, const xnyz_chunk& c) {
  char b[512];
  std::snprintf(b, sizeof(b),
                "{\"block_id\":%u,\"input_offset\":%llu,\"input_size\":%llu,\"rank\":0,"
                "\"part_offset\":%llu,\"compressed_size\":%llu,\"points\":%llu,"
                "\"quantizer\":\"%s\",\"l2_bound\":%.9g,\"curve\":\"%s\",\"rle\":%d,"
                "\"gr_k\":%d,\"pfd_width\":%d,\"direct_threshold\":%d,\"frame\":\"%s\"}\n",
                id, (unsigned long long)c.input_offset, (unsigned long long)(c.input_points * 12),
                (unsigned long long)c.part_offset, (unsigned long long)c.compressed_size,
                (unsigned long long)c.input_points, quantizer_name(c.quantizer),
                (double)c.l2_bound, c.is_hilbert ? "hilbert" : "z", c.rle ? 1 : 0, c.gr_k,
                c.pfd_width, c.direct_threshold,
                c.frame_mode == XNYZ_FRAME_NONE ? "stream" : "zstd");
  return b;
}

// The value of "key" in one flat JSON line: the raw token (number) or the
// string without quotes.
bool field(const std::string& l, const char* key, std::string& out) {
  const std::string k = std::string("\"") + key + "\":";
  const size_t p = l.find(k);
  if (p == std::string::npos) return false;
  size_t a = p + k.size();
  while (a < l.size() && (l[a] == ' ' || l[a] == '\t')) ++a;
  if (a < l.size() && l[a] == '"') {
    const size_t e = l.find('"', a + 1);
    if (e == std::string::npos) return false;
    out = l.substr(a + 1, e - a - 1);
    return true;
  }
  size_t e = a;
  while (e < l.size() && l[e] != ',' && l[e] != '}' && l[e] != ' ') ++e;
  out = l.substr(a, e - a);
  return !out.empty();
}

xnyz_chunk parse_line(const std::string& l, size_t lineno) {
  auto need = [&](const char* key) {
    std::string v;
    if (!field(l, key, v))
      throw Error(XNYZ_ECORRUPT, "manifest line " + std::to_string(lineno) + ": no \"" + key + "\"");
    return v;
  };
  xnyz_chunk c{};
  std::string v;
  // a mapreduce line (mpi_driver.cpp:626-635) has no input_offset: one part
  // per rank holding the rank's owned cells
  c.input_offset = field(l, "input_offset", v) ? std::stoull(v) : 0;
  c.part_offset = std::stoull(need("part_offset"));
  c.compressed_size = std::stoull(need("compressed_size"));
  if (field(l, "points", v)) c.input_points = std::stoull(v);
  else if (field(l, "input_size", v)) c.input_points = std::stoull(v) / 12;
  else if (field(l, "input_bytes", v)) c.input_points = std::stoull(v) / 12;
  // The reference's own manifests carry no mode: the defaults are the MPI
  // driver's (TO quantizer, zstd frames); the bound, curve and RLE flag must
  // then be given by the caller.
  c.quantizer = XNYZ_TRUNCATED_OCTAHEDRON;
  if (field(l, "quantizer", v)) c.quantizer = v == "to" ? XNYZ_TRUNCATED_OCTAHEDRON : XNYZ_CUBE;
  c.l2_bound = field(l, "l2_bound", v) ? std::stof(v) : 0.f;
  c.is_hilbert = field(l, "curve", v) ? v == "hilbert" : 0;
  c.rle = field(l, "rle", v) ? std::stoi(v) != 0 : 0;
  if (field(l, "mode", v) && v == "mapreduce") c.rle = 1;
  c.gr_k = field(l, "gr_k", v) ? std::stoi(v) : -1;
  c.pfd_width = field(l, "pfd_width", v) ? std::stoi(v) : 8;
  c.direct_threshold = field(l, "direct_threshold", v) ? std::stoi(v) : 1024;
  c.frame_mode = field(l, "frame", v) && v == "stream" ? XNYZ_FRAME_NONE : XNYZ_FRAME_ZSTD3;
  return c;
}

}  // namespace

std::vector<xnyz_chunk> read_manifest(const char* path) {
  std::ifstream f(path);
  if (!f.is_open()) throw Error(XNYZ_EIO, std::string("cannot open ") + path);
  std::vector<xnyz_chunk> v;
  std::string l;
  size_t n = 0;
  while (std::getline(f, l)) {
    ++n;
    if (l.find_first_not_of(" \t\r") == std::string::npos) continue;
    v.push_back(parse_line(l, n));
  }
  return v;
}

void write_container(const char* part_path, const char* manifest_path, const uint8_t* out,
                     const uint64_t* out_offsets, uint32_t count, xnyz_chunk* chunks) {
  std::ofstream part(part_path, std::ios::binary);
  std::ofstream meta(manifest_path);
  if (!part.is_open() || !meta.is_open())
    throw Error(XNYZ_EIO, std::string("cannot open ") + part_path + " or " + manifest_path);
  uint64_t at = 0;
  for (uint32_t s = 0; s < count; ++s) {
    if (out_offsets[s + 1] < out_offsets[s]) throw Error(XNYZ_EINVAL, "offsets must not decrease");
    const uint64_t n = out_offsets[s + 1] - out_offsets[s];
    part.write(reinterpret_cast<const char*>(out + out_offsets[s]), std::streamsize(n));
    chunks[s].part_offset = at;
    chunks[s].compressed_size = n;
    at += n;
    meta << line_of(s, chunks[s]);
  }
  part.flush();
  meta.flush();
  if (!part.good() || !meta.good()) throw Error(XNYZ_EIO, "container write failed");
}

std::vector<uint8_t> read_part(const char* part_path, const xnyz_chunk& c) {
  std::ifstream f(part_path, std::ios::binary);
  if (!f.is_open()) throw Error(XNYZ_EIO, std::string("cannot open ") + part_path);
  std::vector<uint8_t> b(c.compressed_size);
  f.seekg(std::streamoff(c.part_offset));
  if (!f.read(reinterpret_cast<char*>(b.data()), std::streamsize(b.size())))
    throw Error(XNYZ_ECORRUPT, "part file shorter than the manifest says");
  return b;
}

}  // namespace xnyz
