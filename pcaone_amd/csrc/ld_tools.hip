// The consumers and producers around the LD r^2 path:
//   write_residuals  Data::write_residuals (src/Data.cpp:196-245) for PLINK
//                    input: the centred genotypes G = BED2GENO - F (missing 0),
//                    minus U S V^T for --ld-stats 0, re-centred per SNP, as
//                    float32 rows per SNP after an (nsnps, nsamples) uint32
//                    header, in the .bed's SNP order.  Rows are made on the
//                    device in chunks and streamed to the file.
//   write_mbim       Data::save_snps_in_bim (src/Data.cpp:96-135): every .bim
//                    line followed by a tab and its allele frequency.
//   ld_prune         ld_prune_big (src/LD.cpp:194-242): windows of --ld-bp,
//                    their r^2 on the device in one call, then the greedy
//                    pass over the windows on the host; .ld.prune.in / .out.
//   ld_clump         ld_clump_single_pheno and its readers (src/LD.cpp:244-370,
//                    56-122): the pairs the greedy clumping can visit are
//                    collected first, their r^2 computed on the device in one
//                    call, then the greedy pass runs on the host.
// Both r^2 consumers use whatever LD matrix the handle holds: a dense matrix
// (.residuals) through ld_r2, PLINK genotypes through ld_r2_geno (whose
// centred genotypes have zero column means, as the reference's in-core G).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace pcaone {
namespace {

// split_string (src/Utils.cpp:207-221): a token at every separator (empty ones kept)
std::vector<std::string> split_keep(const std::string& s, const char* seps) {
  std::vector<std::string> out;
  size_t b = 0;
  for (size_t i = 0; i <= s.size(); ++i) {
    if (i == s.size() || std::strchr(seps, s[i]) != nullptr) {
      out.emplace_back(s, b, i - b);
      b = i + 1;
    }
  }
  return out;
}

int to_int(const std::string& s, const std::string& what) {
  try {
    return std::stoi(s);
  } catch (...) {
    throw Error(-1, what + ": not an integer: '" + s + "'");
  }
}
double to_double(const std::string& s, const std::string& what) {
  try {
    return std::stod(s);
  } catch (...) {
    throw Error(-1, what + ": not a number: '" + s + "'");
  }
}

// SNPld as get_snp_pos_bim (src/LD.cpp:56-78) fills it
struct SnpTable {
  std::vector<std::string> lines;
  std::vector<int> pos;
  std::vector<int> end_pos;           // each chromosome's last index
  std::vector<std::string> chr;       // chromosome of each run
  std::vector<double> af;             // the 7th column of a .mbim
  std::string header;
};

SnpTable read_snps(const std::string& path, bool header, int ichr, int ipos) {
  std::ifstream in(path);
  PCAONE_REQUIRE(in.is_open(), -1, "can not open " + path);
  SnpTable t;
  std::string line, prev;
  if (header) std::getline(in, t.header);
  int i = 0;
  while (std::getline(in, line)) {
    const auto tok = split_keep(line, " \t");
    PCAONE_REQUIRE((int)tok.size() > std::max(ichr, ipos), -1,
                   path + ": line " + std::to_string(i + 1 + (header ? 1 : 0)) + " has too few fields");
    if (tok.size() == 7 && !header) t.af.push_back(to_double(tok[6], path));
    const std::string& cur = tok[ichr];
    if (prev.empty()) t.chr.push_back(cur);
    if (!prev.empty() && prev != cur) {
      t.end_pos.push_back(i - 1);
      t.chr.push_back(cur);
    }
    prev = cur;
    t.pos.push_back(to_int(tok[ipos], path));
    t.lines.push_back(line);
    ++i;
  }
  PCAONE_REQUIRE(i > 0, -1, path + ": no variants");
  t.end_pos.push_back(i - 1);
  return t;
}

void write_lines(const std::string& path, const std::vector<std::string>& lines, const std::vector<uint8_t>& keep,
                 uint8_t want) {
  std::ofstream o(path);
  PCAONE_REQUIRE(o.is_open(), -1, "can not open " + path);
  for (size_t i = 0; i < lines.size(); ++i)
    if (keep[i] == want) o << lines[i] << '\n';
  PCAONE_REQUIRE((bool)o, -1, "failed to write " + path);
}

// rows [r0, r0 + nr) of the residual file: thread block = 256 threads, 8 SNPs;
// pass 1 sums each SNP's values over the samples, pass 2 writes them centred
constexpr int kResSnps = 8;
constexpr int kResMaxK = 64;

__device__ inline double geno_value(const uint8_t* row, uint32_t j, double F) {
  const uint32_t c = (row[j >> 2] >> (2 * (j & 3))) & 3u;
  // BED2GENO = {1, missing, 0.5, 0}; missing imputed to the mean (0 after centring)
  return c == 1u ? 0.0 : (c == 0u ? 1.0 : (c == 2u ? 0.5 : 0.0)) - F;
}

__global__ __launch_bounds__(256) void k_residual_rows(const uint8_t* __restrict__ bed, uint64_t row_bytes,
                                                       uint32_t nsamples, uint32_t r0, uint32_t nr,
                                                       const double* __restrict__ F, const double* __restrict__ U,
                                                       const double* __restrict__ SV, uint32_t k,
                                                       float* __restrict__ out) {
  __shared__ double sv[kResSnps][kResMaxK];
  __shared__ double red[kResSnps][4];
  __shared__ double mean[kResSnps];
  const uint32_t s0 = blockIdx.x * kResSnps, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (uint32_t e = t; e < kResSnps * k; e += blockDim.x) {
    const uint32_t q = e / k, c = e % k;
    sv[q][c] = (s0 + q < nr) ? SV[(uint64_t)(r0 + s0 + q) * k + c] : 0.0;
  }
  __syncthreads();
  auto value = [&](uint32_t q, uint32_t j) {
    const uint32_t i = r0 + s0 + q;
    double v = geno_value(bed + (uint64_t)i * row_bytes, j, F[i]);
    for (uint32_t c = 0; c < k; ++c) v -= U[(uint64_t)c * nsamples + j] * sv[q][c];
    return v;
  };
  double acc[kResSnps];
#pragma unroll
  for (int q = 0; q < kResSnps; ++q) acc[q] = 0.0;
  for (uint32_t j = t; j < nsamples; j += blockDim.x)
#pragma unroll
    for (int q = 0; q < kResSnps; ++q)
      if (s0 + q < nr) acc[q] += value(q, j);
#pragma unroll
  for (int q = 0; q < kResSnps; ++q) {
    double x = acc[q];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o);
    if (lane == 0) red[q][wv] = x;
  }
  __syncthreads();
  if (t < kResSnps) mean[t] = (red[t][0] + red[t][1] + red[t][2] + red[t][3]) / (double)nsamples;
  __syncthreads();
  for (uint32_t j = t; j < nsamples; j += blockDim.x)
#pragma unroll
    for (int q = 0; q < kResSnps; ++q)
      if (s0 + q < nr) out[(uint64_t)(s0 + q) * nsamples + j] = (float)(value(q, j) - mean[q]);
}

}  // namespace

// ---------------------------------------------------------------------------
// .residuals / .mbim
// ---------------------------------------------------------------------------
void Engine::write_residuals(const char* fileout, const double* U, const double* S, const double* V, uint32_t k,
                             int ld_stats) {
  PCAONE_REQUIRE(have_bed_ && src_ == SRC_BED && !gl_mode_, -3, "write_residuals: load PLINK genotypes first");
  PCAONE_REQUIRE(fileout, -1, "write_residuals: no file name");
  PCAONE_REQUIRE(ld_stats == 0 || ld_stats == 1, -1, "write_residuals: ld_stats must be 0 or 1");
  if (ld_stats == 0) {
    PCAONE_REQUIRE(U && S && V && k >= 1 && k <= (uint32_t)kResMaxK, -1,
                   "write_residuals: --ld-stats 0 needs U (N x k), S (k) and V (M x k), 1 <= k <= 64");
  } else {
    k = 0;
  }
  PCAONE_HIP(hipSetDevice(device_));
  const uint64_t nb = (N_ + 3) >> 2;
  DevBuf<double> dF, dU, dSV;
  dF.reserve(M_);
  PCAONE_HIP(hipMemcpyAsync(dF.p, F_orig_.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
  std::vector<double> sv;
  if (k) {
    dU.reserve((uint64_t)N_ * k);
    PCAONE_HIP(hipMemcpyAsync(dU.p, U, (size_t)N_ * k * 8, hipMemcpyHostToDevice, stream_));
    sv.resize((size_t)M_ * k);  // row i: S_c V[i][c]
    for (uint32_t i = 0; i < M_; ++i)
      for (uint32_t c = 0; c < k; ++c) sv[(size_t)i * k + c] = S[c] * V[(size_t)c * M_ + i];
    dSV.reserve((uint64_t)M_ * k);
    PCAONE_HIP(hipMemcpyAsync(dSV.p, sv.data(), sv.size() * 8, hipMemcpyHostToDevice, stream_));
  }
  FILE* fp = std::fopen(fileout, "wb");
  PCAONE_REQUIRE(fp != nullptr, -1, std::string("write_residuals: can not open ") + fileout);
  try {
    const uint32_t hdr[2] = {M_, N_};
    PCAONE_REQUIRE(std::fwrite(hdr, 4, 2, fp) == 2, -1, "write_residuals: write failed");
    // chunks of about 256 MiB of float rows
    const uint32_t chunk = (uint32_t)std::max<uint64_t>(kResSnps, ((256ull << 20) / ((uint64_t)N_ * 4)) / kResSnps *
                                                                      kResSnps);
    DevBuf<float> dout;
    dout.reserve((uint64_t)std::min(chunk, M_) * N_);
    std::vector<float> host((size_t)std::min(chunk, M_) * N_);
    for (uint32_t r0 = 0; r0 < M_; r0 += chunk) {
      const uint32_t nr = std::min(chunk, M_ - r0);
      timed("residuals", [&] {
        hipLaunchKernelGGL(k_residual_rows, dim3((uint32_t)ceil_div(nr, kResSnps)), dim3(256), 0, stream_, resident_raw("write_residuals"),
                           nb, N_, r0, nr, dF.p, dU.p, dSV.p, k, dout.p);
      });
      PCAONE_HIP(hipMemcpyAsync(host.data(), dout.p, (size_t)nr * N_ * 4, hipMemcpyDeviceToHost, stream_));
      sync();
      PCAONE_REQUIRE(std::fwrite(host.data(), 4, (size_t)nr * N_, fp) == (size_t)nr * N_, -1,
                     "write_residuals: write failed");
    }
  } catch (...) {
    std::fclose(fp);
    throw;
  }
  PCAONE_REQUIRE(std::fclose(fp) == 0, -1, "write_residuals: close failed");
}

void write_mbim(const char* fileout, const char* filebim, const double* F, uint32_t nsnps) {
  PCAONE_REQUIRE(fileout && filebim && F, -1, "write_mbim: null argument");
  std::ifstream in(filebim);
  PCAONE_REQUIRE(in.is_open(), -1, std::string("write_mbim: can not open ") + filebim);
  std::ofstream o(fileout);
  PCAONE_REQUIRE(o.is_open(), -1, std::string("write_mbim: can not open ") + fileout);
  std::string line;
  uint32_t j = 0;
  while (std::getline(in, line)) {
    PCAONE_REQUIRE(j < nsnps, -1, "write_mbim: the .bim has more lines than SNPs");
    o << line << '\t' << F[j] << '\n';  // default stream format: 6 significant digits
    ++j;
  }
  PCAONE_REQUIRE(j == nsnps, -1, "write_mbim: the .bim has fewer lines than SNPs");
  PCAONE_REQUIRE((bool)o, -1, "write_mbim: write failed");
}

// ---------------------------------------------------------------------------
// r^2 of window pairs on whatever LD matrix the handle holds
// ---------------------------------------------------------------------------
std::vector<double> Engine::ld_pairs(const std::vector<int32_t>& ws, const std::vector<int32_t>& we) {
  uint64_t total = 0;
  for (size_t w = 0; w < ws.size(); ++w) total += (uint64_t)(we[w] - 1);
  std::vector<double> r2(total);
  if (ws.empty()) return r2;
  if (have_bed_ && src_ == SRC_BED && !gl_mode_ && !have_resid_) {
    int cfg = -1;  // the tuned tile shape
    ld_r2_geno(ws.data(), we.data(), (uint32_t)ws.size(), r2.data(), total, nullptr, &cfg);
  }
  else
    ld_r2(ws.data(), we.data(), (uint32_t)ws.size(), r2.data(), total);
  return r2;
}

// ---------------------------------------------------------------------------
// pruning
// ---------------------------------------------------------------------------
void Engine::ld_prune(const char* filebim, int ld_bp, double r2_tol, const char* prefix) {
  PCAONE_REQUIRE(filebim && prefix, -1, "ld_prune: null argument");
  PCAONE_REQUIRE(ld_bp > 0, -1, "ld_prune: --ld-bp must be > 0");
  const SnpTable snp = read_snps(filebim, false, 0, 3);
  PCAONE_REQUIRE(snp.pos.size() == M_, -1, "The number of variants is not matching the LD matrix");
  const bool random_one = snp.af.empty();
  PCAONE_REQUIRE(random_one || snp.af.size() == M_, -1, "ld_prune: allele frequencies missing on some lines");
  std::vector<int32_t> pos(snp.pos.begin(), snp.pos.end()), ce(snp.end_pos.begin(), snp.end_pos.end());
  std::vector<int32_t> ws(M_), we(M_);
  const uint32_t nw = ld_windows(pos.data(), M_, ce.data(), (uint32_t)ce.size(), ld_bp, ws.data(), we.data());
  ws.resize(nw);
  we.resize(nw);
  const std::vector<double> r2 = ld_pairs(ws, we);
  std::vector<uint8_t> keep(M_, 1);
  auto maf = [](double a) { return a > 0.5 ? 1.0 - a : a; };
  uint64_t p = 0;
  for (uint32_t w = 0; w < nw; ++w) {
    const int i = ws[w];
    if (!keep[i]) {
      p += (uint64_t)(we[w] - 1);
      continue;
    }
    for (int j = 1; j < we[w]; ++j, ++p) {
      const int k = i + j;
      if (!keep[k]) continue;
      if (r2[p] > r2_tol) {
        int o = k;
        if (!random_one) o = maf(snp.af[k]) > maf(snp.af[i]) ? i : k;
        keep[o] = 0;
      }
    }
  }
  write_lines(std::string(prefix) + ".ld.prune.in", snp.lines, keep, 1);
  write_lines(std::string(prefix) + ".ld.prune.out", snp.lines, keep, 0);
}

// ---------------------------------------------------------------------------
// clumping
// ---------------------------------------------------------------------------
void Engine::ld_clump(const char* fileassoc, const char* colnames, const char* filebim, int clump_bp, double clump_r2,
                      double clump_p1, double clump_p2, const char* fileout) {
  PCAONE_REQUIRE(fileassoc && filebim && fileout, -1, "ld_clump: null argument");
  const SnpTable snp = read_snps(filebim, false, 0, 3);
  PCAONE_REQUIRE(snp.pos.size() == M_, -1, "The number of variants is not matching the LD matrix");
  // valid_assoc_file (src/LD.cpp:244-270): header fields split at tabs
  std::vector<std::string> names{"CHR", "BP", "P"};
  if (colnames && *colnames) names = split_keep(colnames, ",");
  PCAONE_REQUIRE(names.size() >= 3, -1, "ld_clump: --clump-names needs three names");
  std::string head;
  {
    std::ifstream in(fileassoc);
    PCAONE_REQUIRE(in.is_open(), -1, std::string("can not open ") + fileassoc);
    std::getline(in, head);
  }
  int col[3] = {-1, -1, -1};
  {
    const auto f = split_keep(head, "\t");
    for (int j = 0; j < (int)f.size(); ++j)
      for (int q = 0; q < 3; ++q)
        if (f[j] == names[q]) {
          col[q] = j;
          break;
        }
    for (int q = 0; q < 3; ++q)
      PCAONE_REQUIRE(col[q] >= 0, -1, "the assoc-like file has no " + names[q] + " column");
  }
  // the target SNPs (get_snp_pos_bim with the header and the CHR, BP columns)
  const SnpTable tgt = read_snps(fileassoc, true, col[0], col[1]);
  // map_index_snps (src/LD.cpp:272-296): per chromosome run, bp -> (p, line) for
  // p <= clump_p2, first occurrence kept; as there, a line is filed before the
  // chromosome-change test, so a run's first line lands in the previous run's map
  std::vector<std::map<int, std::pair<double, std::string>>> pv(1);
  {
    std::ifstream in(fileassoc);
    std::string line, prev;
    std::getline(in, line);
    while (std::getline(in, line)) {
      const auto tok = split_keep(line, "\t");
      PCAONE_REQUIRE((int)tok.size() > std::max(col[0], std::max(col[1], col[2])), -1,
                     std::string(fileassoc) + ": a line has too few fields");
      const std::string& cur = tok[col[0]];
      const int bp = to_int(tok[col[1]], fileassoc);
      const double pval = to_double(tok[col[2]], fileassoc);
      if (pval <= clump_p2) pv.back().insert({bp, {pval, line}});
      if (!prev.empty() && prev != cur) pv.emplace_back();
      prev = cur;
    }
  }
  // get_target_snp_idx (src/LD.cpp:83-122), its index ranges included: each
  // chromosome's lists start at the previous chromosome's last SNP
  {
    std::vector<size_t> ord;
    for (const auto& c : snp.chr) {
      size_t i = 0;
      while (i < tgt.chr.size() && tgt.chr[i] != c) ++i;
      ord.push_back(i);
    }
    PCAONE_REQUIRE(std::is_sorted(ord.begin(), ord.end()), -1,
                   "the association file may be not sorted, hence not matching the bimfile ");
  }
  const size_t ntc = tgt.chr.size();
  std::vector<std::vector<int>> idx_per(ntc), bp_per(ntc);
  for (size_t tc = 0; tc < ntc; ++tc) {
    size_t c = 0;
    while (c < snp.chr.size() && snp.chr[c] != tgt.chr[tc]) ++c;
    PCAONE_REQUIRE(c < snp.chr.size(), -1, "ld_clump: chromosome " + tgt.chr[tc] + " is not in the .bim");
    std::unordered_map<int, int> mpos;
    int e = snp.end_pos[c], s = c > 0 ? snp.end_pos[c - 1] : 0;
    for (int i = s; i <= e; ++i) mpos[snp.pos[i]] = i;
    e = tgt.end_pos[tc];
    s = tc > 0 ? tgt.end_pos[tc - 1] : 0;
    for (int i = s; i <= e; ++i) {
      auto it = mpos.find(tgt.pos[i]);
      if (it != mpos.end()) {
        bp_per[tc].push_back(tgt.pos[i]);
        idx_per[tc].push_back(it->second);
      }
    }
  }
  PCAONE_REQUIRE(pv.size() == ntc, -1, "ld_clump: inconsistent chromosome runs in the assoc-like file");
  // index SNPs per chromosome in p order (ties: by position) and the pairs
  // their scans can reach (the scan bounds depend on positions only)
  struct Cand {
    double p;
    int bp;
  };
  std::vector<std::vector<Cand>> cands(ntc);
  std::vector<std::unordered_map<int, int>> mbp(ntc);
  std::map<int32_t, int32_t> reach;  // lead LD index -> farthest partner
  auto scan = [&](size_t c, int j, auto&& visit) {
    const auto& bp = bp_per[c];
    const int p = bp[j];
    for (int k = j - 1; k >= 0 && !(bp[k] < p - clump_bp); --k) visit(k);
    for (int k = j + 1; k < (int)bp.size() && !(bp[k] > p + clump_bp); ++k) visit(k);
  };
  for (size_t c = 0; c < ntc; ++c) {
    for (size_t j = 0; j < bp_per[c].size(); ++j) mbp[c].insert({bp_per[c][j], (int)j});
    for (const auto& kv : pv[c])
      if (kv.second.first <= clump_p1) cands[c].push_back({kv.second.first, kv.first});
    std::stable_sort(cands[c].begin(), cands[c].end(), [](const Cand& a, const Cand& b) { return a.p < b.p; });
    for (const auto& cd : cands[c]) {
      auto it = mbp[c].find(cd.bp);
      if (it == mbp[c].end()) continue;
      const int j = it->second;
      scan(c, j, [&](int k) {
        int a = idx_per[c][j], b = idx_per[c][k];
        if (a == b) return;
        if (a > b) std::swap(a, b);
        auto r = reach.find(a);
        if (r == reach.end()) reach[a] = b;
        else r->second = std::max(r->second, b);
      });
    }
  }
  std::vector<int32_t> ws, we;
  std::unordered_map<int32_t, uint64_t> off;
  uint64_t tot = 0;
  for (const auto& kv : reach) {
    ws.push_back(kv.first);
    we.push_back(kv.second - kv.first + 1);
    off[kv.first] = tot;
    tot += (uint64_t)(kv.second - kv.first);
  }
  const std::vector<double> r2 = ld_pairs(ws, we);
  auto pair_r2 = [&](int a, int b) {
    if (a == b) return 1.0;  // a variant with itself
    if (a > b) std::swap(a, b);
    return r2[off.at(a) + (uint64_t)(b - a - 1)];
  };
  std::ofstream o(fileout);
  PCAONE_REQUIRE(o.is_open(), -1, std::string("can not open ") + fileout);
  o << head << "\tSP2\n";
  for (size_t c = 0; c < ntc; ++c) {
    auto mpp = pv[c];
    const auto& bp = bp_per[c];
    for (const auto& cd : cands[c]) {
      const int p = cd.bp;
      if (!mpp.count(p)) continue;  // already clumped
      auto it = mbp[c].find(p);
      if (it == mbp[c].end()) continue;
      const int j = it->second;
      std::vector<int> clumped;
      scan(c, j, [&](int k) {
        const int p2 = bp[k];
        if (!mpp.count(p2)) return;
        if (pair_r2(idx_per[c][j], idx_per[c][k]) >= clump_r2) {
          clumped.push_back(p2);
          mpp.erase(p2);
        }
      });
      o << pv[c].at(p).second << '\t';
      if (clumped.empty()) {
        o << "NONE";
      } else {
        std::vector<size_t> ordc(clumped.size());
        for (size_t q = 0; q < ordc.size(); ++q) ordc[q] = q;
        std::stable_sort(ordc.begin(), ordc.end(), [&](size_t x, size_t y) {
          return pv[c].at(clumped[x]).first < pv[c].at(clumped[y]).first;
        });
        for (size_t q = 0; q < ordc.size(); ++q) o << (q ? "," : "") << clumped[ordc[q]];
      }
      o << '\n';
    }
  }
  PCAONE_REQUIRE((bool)o, -1, std::string("failed to write ") + fileout);
}

}  // namespace pcaone
