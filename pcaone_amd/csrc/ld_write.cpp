// The .ld.gz writer of the LD r^2 path: the text of ld_r2_big / ld_r2_small
// (src/LD.cpp:419-442, 388-416), one line per in-window pair in window order,
//   CHR_A \t BP_A \t SNP_A \t CHR_B \t BP_B \t SNP_B \t R2
// where each SNP's three fields come from its .bim line (tokens 0, 3 and 1 of
// the line split at every space or tab, src/LD.hpp:18-28) and R2 is printed as
// std::to_string(double), i.e. "%f" (a NaN as "-nan", the reference's
// monomorphic-SNP NaN on x86).  The reference formats and gzwrite()s
// line by line on one thread (its own FIXME at src/LD.cpp:435).  Here the
// pairs are cut into chunks that worker threads format and deflate into
// independent gzip members; the members are written in order, so the file
// gunzips (gzip, zcat, zlib's gzread) to the same text.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace pcaone {
namespace {

// The reference's BIM record: getline, then token0 \t token3 \t token1 of the
// line split at every ' ' or '\t' (consecutive separators give empty tokens).
std::vector<std::string> read_bim_fields(const std::string& path) {
  std::ifstream in(path);
  PCAONE_REQUIRE(in.is_open(), -1, "write_ld_gz: can not open " + path);
  std::vector<std::string> out;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty()) {
      out.emplace_back();
      continue;
    }
    std::vector<std::string> tok;
    size_t b = 0;
    for (size_t i = 0; i <= line.size(); ++i) {
      if (i == line.size() || line[i] == ' ' || line[i] == '\t') {
        tok.emplace_back(line, b, i - b);
        b = i + 1;
      }
    }
    PCAONE_REQUIRE(tok.size() >= 4, -1, "write_ld_gz: .bim line " + std::to_string(out.size() + 1) +
                                            " has fewer than 4 fields");
    out.push_back(tok[0] + "\t" + tok[3] + "\t" + tok[1]);
  }
  return out;
}

// One gzip member holding `text`.
std::string gzip_member(const std::string& text, int level) {
  z_stream zs;
  std::memset(&zs, 0, sizeof(zs));
  PCAONE_REQUIRE(deflateInit2(&zs, level, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) == Z_OK, -1,
                 "write_ld_gz: deflateInit2 failed");
  std::string out;
  out.resize(deflateBound(&zs, text.size()) + 64);
  zs.next_in = reinterpret_cast<Bytef*>(const_cast<char*>(text.data()));
  zs.avail_in = (uInt)text.size();
  zs.next_out = reinterpret_cast<Bytef*>(&out[0]);
  zs.avail_out = (uInt)out.size();
  const int rc = deflate(&zs, Z_FINISH);
  const size_t produced = out.size() - zs.avail_out;
  deflateEnd(&zs);
  PCAONE_REQUIRE(rc == Z_STREAM_END, -1, "write_ld_gz: deflate failed");
  out.resize(produced);
  return out;
}

}  // namespace

void write_ld_gz(const char* fileout, const char* filebim, const int32_t* ws, const int32_t* we, uint32_t nwin,
                 const double* r2, uint64_t npairs, int level, int nthreads) {
  PCAONE_REQUIRE(fileout && filebim, -1, "write_ld_gz: no file name");
  PCAONE_REQUIRE(level >= -1 && level <= 9, -1, "write_ld_gz: compression level must be in [-1, 9]");
  const std::vector<std::string> bims = read_bim_fields(filebim);
  PCAONE_REQUIRE(nwin == 0 || (ws && we), -1, "write_ld_gz: no windows");
  std::vector<int64_t> off;
  std::vector<int32_t> win_of_row;
  const uint64_t total = window_offsets(ws, we, nwin, (uint32_t)bims.size(), off, win_of_row);
  PCAONE_REQUIRE(npairs == total, -1,
                 "write_ld_gz: the windows hold " + std::to_string(total) + " pairs, r2 has " + std::to_string(npairs));
  PCAONE_REQUIRE(total == 0 || r2, -1, "write_ld_gz: no r2");
  // chunks of whole windows, about kChunkPairs pairs each
  constexpr uint64_t kChunkPairs = 1u << 20;
  std::vector<uint32_t> cut{0};
  uint64_t acc = 0;
  for (uint32_t w = 0; w < nwin; ++w) {
    acc += (uint64_t)(we[w] - 1);
    if (acc >= kChunkPairs) {
      cut.push_back(w + 1);
      acc = 0;
    }
  }
  if (cut.back() != nwin) cut.push_back(nwin);
  const size_t nchunk = cut.size() - 1;
  FILE* fp = std::fopen(fileout, "wb");
  PCAONE_REQUIRE(fp != nullptr, -1, std::string("write_ld_gz: can not open ") + fileout);
  auto put = [&](const std::string& bytes) {
    PCAONE_REQUIRE(std::fwrite(bytes.data(), 1, bytes.size(), fp) == bytes.size(), -1,
                   "write_ld_gz: failed to write data to ld.gz file");
  };
  const int nt = std::max(1, std::min<int>(nthreads, 64));
  std::vector<std::string> mem(nchunk);
  auto format = [&](size_t c) {
    std::string text;
    const uint64_t p0 = (uint64_t)off[cut[c]];
    const uint64_t p1 = cut[c + 1] < nwin ? (uint64_t)off[cut[c + 1]] : total;
    text.reserve((p1 - p0) * 48);
    char num[64];
    uint64_t p = p0;
    for (uint32_t w = cut[c]; w < cut[c + 1]; ++w) {
      const int i = ws[w];
      for (int j = 1; j < we[w]; ++j, ++p) {
        // std::to_string(double) is "%f"; the reference's NaN (0 * inf for a
        // monomorphic SNP) is the x86 default NaN, sign bit set: "-nan"
        const double x = r2[p];
        const int n = x == x ? std::snprintf(num, sizeof(num), "%f", x) : std::snprintf(num, sizeof(num), "-nan");
        text += bims[i];
        text += '\t';
        text += bims[i + j];
        text += '\t';
        text.append(num, (size_t)n);
        text += '\n';
      }
    }
    mem[c] = gzip_member(text, level);
  };
  // rounds of nt chunks: formatted and deflated in parallel, written in order
  try {
    put(gzip_member("CHR_A\tBP_A\tSNP_A\tCHR_B\tBP_B\tSNP_B\tR2\n", level));
    for (size_t c0 = 0; c0 < nchunk; c0 += (size_t)nt) {
      const size_t c1 = std::min(nchunk, c0 + (size_t)nt);
      std::vector<std::thread> th;
      std::vector<std::string> errs(c1 - c0);
      for (size_t c = c0; c < c1; ++c)
        th.emplace_back([&, c] {
          try {
            format(c);
          } catch (const std::exception& e) {
            errs[c - c0] = e.what();
          }
        });
      for (auto& t : th) t.join();
      for (auto& e : errs) PCAONE_REQUIRE(e.empty(), -1, e);
      for (size_t c = c0; c < c1; ++c) {
        put(mem[c]);
        std::string().swap(mem[c]);
      }
    }
  } catch (...) {
    std::fclose(fp);
    throw;
  }
  PCAONE_REQUIRE(std::fclose(fp) == 0, -1, "write_ld_gz: failed to close the ld.gz file");
}

}  // namespace pcaone
