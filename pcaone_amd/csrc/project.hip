// Projection of new samples onto a reference PCA (run_projection,
// src/Projection.cpp:147-208, with the SNP matching of FileBed's constructor,
// src/FilePlink.hpp:31-93) for PLINK targets:
//   matching  target .bim lines against the reference .mbim by ID, then
//             chromosome and position; alleles equal (F = the .mbim
//             frequency) or swapped (F = 1 - it, the loading row negated)
//   values    g = BED2GENO - F (missing 0), divided by sqrt(2F(1-F)) when
//             standardising and that exceeds 1e-9
//   method 1  U = G V S^-1                          (no missingness model)
//   method 2  u_i = argmin |V_o S u - g_i,o|, o the sample's observed SNPs,
//             solved from the normal equations (V_o S)^T (V_o S) u = (V_o S)^T g_o
// One wave per sample walks the matched SNPs: lanes hold the right-hand side
// and, for method 2, the missing SNPs' rank-1 corrections of the Gram of V S
// (so the observed Gram is the full Gram minus them); the k x k system is then
// solved by Cholesky in the wave.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace pcaone {
namespace {

constexpr int kProjMaxK = 32;
constexpr int kProjPairs = kProjMaxK * (kProjMaxK + 1) / 2;  // 528
constexpr int kProjPer = (kProjPairs + 63) / 64;             // 9 pair slots per lane

std::vector<std::string> split_any(const std::string& s, const char* seps) {
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

struct ProjArgs {
  const uint8_t* bed;
  uint64_t row_bytes;
  uint32_t nsamples, nmatch, k;
  const int32_t* snp;   // target SNP of each matched row
  const double* F;      // per matched row
  const double* d;      // per matched row: 1/sd or 1
  const double* W;      // matched rows x k: V S^-1 (method 1) or V S (method 2)
  const double* gram;   // k x k: W^T W (method 2)
  double* U;            // nsamples x k column-major
  int method;
};

__global__ __launch_bounds__(256) void k_project(ProjArgs a) {
  __shared__ double A[4][kProjMaxK][kProjMaxK + 1];
  __shared__ double rhs[4][kProjMaxK];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * 4 + wv;
  if (i >= a.nsamples) return;  // whole waves; no block-wide barrier follows
  const int k = (int)a.k;
  double b = 0.0;  // lane c < k: sum over observed SNPs of g W_c
  double corr[kProjPer];
#pragma unroll
  for (int q = 0; q < kProjPer; ++q) corr[q] = 0.0;
  // this lane's Gram pairs (c <= d), packed row by row
  int pc[kProjPer], pd[kProjPer];
#pragma unroll
  for (int q = 0; q < kProjPer; ++q) {
    int e = lane + 64 * q, c = 0;
    pc[q] = pd[q] = -1;
    if (e < k * (k + 1) / 2) {
      while (e >= k - c) {
        e -= k - c;
        ++c;
      }
      pc[q] = c;
      pd[q] = c + e;
    }
  }
  // 64 matched SNPs at a time: lane q fetches SNP r0 + q's code, F and scale
  // for this sample, then the wave walks them with the values broadcast
  for (uint32_t r0 = 0; r0 < a.nmatch; r0 += 64) {
    const uint32_t rl = r0 + (uint32_t)lane;
    uint32_t mycode = 1u;
    double myg = 0.0;
    if (rl < a.nmatch) {
      const uint8_t byte = a.bed[(uint64_t)a.snp[rl] * a.row_bytes + (i >> 2)];
      mycode = (byte >> (2 * (i & 3))) & 3u;
      myg = (((mycode == 0u) ? 1.0 : (mycode == 2u) ? 0.5 : 0.0) - a.F[rl]) * a.d[rl];
    }
    const int nq = (int)min(64u, a.nmatch - r0);

    for (int q = 0; q < nq; ++q) {
      const uint32_t code = (uint32_t)__shfl((int)mycode, q);
      const double g = __shfl(myg, q);
      const double* w = a.W + (uint64_t)(r0 + q) * k;
      if (code == 1u) {  // missing: 0 in g; for method 2 its rank-1 term leaves the Gram
        if (a.method == 2) {
#pragma unroll
          for (int e = 0; e < kProjPer; ++e)
            if (pc[e] >= 0) corr[e] += w[pc[e]] * w[pd[e]];
        }
      } else if (lane < k) {
        b += g * w[lane];
      }
    }
  }
  double* u = a.U;
  if (a.method == 1) {
    if (lane < k) u[(uint64_t)lane * a.nsamples + i] = b;
    return;
  }
  // the observed Gram, then Cholesky and the two triangular solves (one wave)
#pragma unroll
  for (int q = 0; q < kProjPer; ++q)
    if (pc[q] >= 0) {
      const double v = a.gram[pc[q] * k + pd[q]] - corr[q];
      A[wv][pc[q]][pd[q]] = v;
      A[wv][pd[q]][pc[q]] = v;
    }
  if (lane < k) rhs[wv][lane] = b;
  __builtin_amdgcn_wave_barrier();
  for (int c = 0; c < k; ++c) {
    // A[c][c..] is final; column c below the diagonal is scaled, the rest updated
    const double p = A[wv][c][c];
    const double rp = p > 0.0 ? 1.0 / sqrt(p) : 0.0;
    __builtin_amdgcn_wave_barrier();
    if (lane >= c && lane < k) A[wv][lane][c] *= rp;  // L[lane][c] (lane == c: sqrt p)
    __builtin_amdgcn_wave_barrier();
    for (int rr = c + 1 + lane; rr < k; rr += 64)
      for (int cc = c + 1; cc <= rr; ++cc) A[wv][rr][cc] -= A[wv][rr][c] * A[wv][cc][c];
    __builtin_amdgcn_wave_barrier();
  }
  // L y = rhs, then L^T x = y, lane 0 (k <= 32)
  if (lane == 0) {
    double y[kProjMaxK];
    for (int c = 0; c < k; ++c) {
      double s = rhs[wv][c];
      for (int cc = 0; cc < c; ++cc) s -= A[wv][c][cc] * y[cc];
      y[c] = s / A[wv][c][c];
    }
    for (int c = k - 1; c >= 0; --c) {
      double s = y[c];
      for (int cc = c + 1; cc < k; ++cc) s -= A[wv][cc][c] * y[cc];
      y[c] = s / A[wv][c][c];
    }
    for (int c = 0; c < k; ++c) u[(uint64_t)c * a.nsamples + i] = y[c];
  }
}

}  // namespace

uint32_t Engine::project(const char* filebim, const char* filembim, const double* S, const double* Vref, uint32_t k,
                         uint32_t ref_nsnps, int method, int standardize, double* U, double* Wout) {
  PCAONE_REQUIRE(have_bed_ && src_ == SRC_BED && !gl_mode_, -3, "project: load the target PLINK genotypes first");
  PCAONE_REQUIRE(filebim && filembim && S && Vref && U, -1, "project: null argument");
  PCAONE_REQUIRE(method == 1 || method == 2, -1, "project: method must be 1 or 2");
  PCAONE_REQUIRE(k >= 1 && k <= (uint32_t)kProjMaxK, -1, "project: k must be in [1, 32]");
  // the reference sites: .mbim lines of exactly 7 tab-separated fields
  struct Ref {
    int index;
    std::string chrom, pos, a1, a2;
    double freq;
  };
  std::unordered_map<std::string, Ref> ref;
  {
    std::ifstream in(filembim);
    PCAONE_REQUIRE(in.is_open(), -1, "Cannot open the extended bim (.mbim) file.");
    std::string line;
    int idx = 0;
    while (std::getline(in, line)) {
      const auto t = split_any(line, "\t");
      PCAONE_REQUIRE(t.size() == 7, -1, std::string("the input file is not valid!\n => ") + filembim);
      double f = 0.0;
      try {
        f = std::stod(t[6]);
      } catch (...) {
        throw Error(-1, std::string("the input file is not valid!\n => ") + filembim);
      }
      ref.emplace(t[1], Ref{idx++, t[0], t[3], t[4], t[5], f});
    }
    PCAONE_REQUIRE((uint32_t)idx == ref_nsnps, -1, "project: the .mbim and the loadings differ in SNP count");
  }
  std::vector<int32_t> snp;
  std::vector<double> F, d, W;
  {
    std::ifstream in(filebim);
    PCAONE_REQUIRE(in.is_open(), -1, "Cannot open the target bim file.");
    std::string line;
    int32_t j = 0;
    while (std::getline(in, line)) {
      const auto t = split_any(line, "\t ");
      PCAONE_REQUIRE(t.size() >= 6, -1, std::string("the input file is not valid!\n => ") + filebim);
      PCAONE_REQUIRE((uint32_t)j < M_, -1, "project: the target .bim has more lines than the loaded SNPs");
      auto it = ref.find(t[1]);
      if (it != ref.end() && t[0] == it->second.chrom && t[3] == it->second.pos) {
        const Ref& r = it->second;
        int flip = -1;
        if (t[4] == r.a1 && t[5] == r.a2) flip = 0;
        else if (t[4] == r.a2 && t[5] == r.a1) flip = 1;
        if (flip >= 0) {
          const double f = flip ? 1.0 - r.freq : r.freq;
          snp.push_back(j);
          F.push_back(f);
          const double sd = std::sqrt(2.0 * f * (1.0 - f));
          d.push_back(standardize && sd > 1e-9 ? 1.0 / sd : 1.0);
          for (uint32_t c = 0; c < k; ++c) {
            double v = Vref[(size_t)c * ref_nsnps + r.index] * (flip ? -1.0 : 1.0);
            W.push_back(method == 1 ? v / S[c] : v * S[c]);
          }
        }
      }
      ++j;
    }
    PCAONE_REQUIRE((uint32_t)j == M_, -1, "project: the target .bim has fewer lines than the loaded SNPs");
  }
  const uint32_t nm = (uint32_t)snp.size();
  PCAONE_REQUIRE(nm > 0, -1, "no overlapping SNPs found between target .bim and reference .mbim");
  std::vector<double> gram((size_t)k * k, 0.0);
  if (method == 2)
    for (uint32_t r = 0; r < nm; ++r)
      for (uint32_t c = 0; c < k; ++c)
        for (uint32_t e = 0; e < k; ++e) gram[c * k + e] += W[(size_t)r * k + c] * W[(size_t)r * k + e];
  PCAONE_HIP(hipSetDevice(device_));
  DevBuf<int32_t> dsnp;
  DevBuf<double> dF, dd, dW, dG, dU;
  dsnp.reserve(nm);
  dF.reserve(nm);
  dd.reserve(nm);
  dW.reserve((uint64_t)nm * k);
  dG.reserve((uint64_t)k * k);
  dU.reserve((uint64_t)N_ * k);
  PCAONE_HIP(hipMemcpyAsync(dsnp.p, snp.data(), (size_t)nm * 4, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dF.p, F.data(), (size_t)nm * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dd.p, d.data(), (size_t)nm * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dW.p, W.data(), W.size() * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dG.p, gram.data(), gram.size() * 8, hipMemcpyHostToDevice, stream_));
  ProjArgs a{};
  a.bed = resident_raw("project");
  a.row_bytes = (N_ + 3) >> 2;
  a.nsamples = N_;
  a.nmatch = nm;
  a.k = k;
  a.snp = dsnp.p;
  a.F = dF.p;
  a.d = dd.p;
  a.W = dW.p;
  a.gram = dG.p;
  a.U = dU.p;
  a.method = method;
  timed("project", [&] {
    hipLaunchKernelGGL(k_project, dim3((uint32_t)ceil_div(N_, 4)), dim3(256), 0, stream_, a);
  });
  PCAONE_HIP(hipMemcpyAsync(U, dU.p, (size_t)N_ * k * 8, hipMemcpyDeviceToHost, stream_));
  sync();
  if (Wout)  // the loadings the reference writes: rows of the matched SNPs, column-major nm x k
    for (uint32_t r = 0; r < nm; ++r)
      for (uint32_t c = 0; c < k; ++c) Wout[(size_t)c * nm + r] = W[(size_t)r * k + c];
  return nm;
}

}  // namespace pcaone
