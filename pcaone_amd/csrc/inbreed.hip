// Per-site inbreeding coefficients and the HWE likelihood-ratio test under
// population structure (run_inbreeding_em, calc_inbreed_coef,
// calc_inbreed_site_lrt, src/Inbreeding.cpp:16-190) on PLINK hard calls:
// the individual allele frequencies are Pi = clamp((U S V^T + 2F) / 2,
// 1e-4, 1 - 1e-4) (FileUSV::read_all, src/FileUSV.cpp:11-22) with U, S, V a
// PCA's printed results and F the .mbim frequencies.  Pi is never stored:
// one block per SNP recomputes it from U (N x K, L2-resident) and S V_j in
// registers while it reads the SNP's packed codes, and reduces the site's sums
// over samples.  The SQUAREM-accelerated EM (three coefficient passes per
// iteration) and the convergence test run on the host over the per-site
// results, as the reference's loop does.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "engine.hpp"

namespace pcaone {
namespace {

constexpr int kInbMaxK = 64;

struct InbPass {
  const uint8_t* bed;  // packed .bed rows, the loaded order
  uint64_t row_bytes;
  uint32_t nsamples, nsnps, k;
  const double* U;   // N x K column-major
  const double* SV;  // M x K row-major: S_c V_jc
  const double* F;   // .mbim frequencies
  const double* coef_in;
  double* coef_out;  // mode 0: the updated coefficient
  double* delta;     // mode 0: new - old
  double* lrt;       // mode 1: 2 (log alt - log null)
  int mode;
};

__global__ __launch_bounds__(256) void k_inbreed(InbPass g) {
  __shared__ double sv[kInbMaxK];
  __shared__ double red[2][4];
  const uint32_t j = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (uint32_t c = t; c < g.k; c += blockDim.x) sv[c] = g.SV[(uint64_t)j * g.k + c];
  __syncthreads();
  const uint8_t* row = g.bed + (uint64_t)j * g.row_bytes;
  const double fj = g.F[j], fc = g.coef_in[j];
  double s0 = 0.0, s1 = 0.0;
  for (uint32_t i = t; i < g.nsamples; i += blockDim.x) {
    double z = 0.0;
    for (uint32_t c = 0; c < g.k; ++c) z += g.U[(uint64_t)c * g.nsamples + i] * sv[c];
    const double pi = fmin(fmax((z + 2.0 * fj) * 0.5, 1e-4), 1.0 - 1e-4);
    const uint32_t code = (row[i >> 2] >> (2 * (i & 3))) & 3u;  // 00 hom A1, 01 missing, 10 het, 11 hom A2
    const double fadj = (1.0 - pi) * pi * fc;
    double p0 = fmax(1e-4, (1.0 - pi) * (1.0 - pi) + fadj);
    double p1 = g.mode == 0 ? fmax(1e-4, 2.0 * pi * (1.0 - pi) * (1.0 - fc))
                            : fmax(1e-4, 2.0 * pi * (1.0 - pi) - 2.0 * fadj);
    double p2 = fmax(1e-4, pi * pi + fadj);
    const double ps = 1.0 / (p0 + p1 + p2);
    p0 *= ps;
    p1 *= ps;
    p2 *= ps;
    if (g.mode == 0) {
      double q0, q1, q2;
      if (code == 3u) {
        q0 = p0; q1 = 0.0; q2 = 0.0;
      } else if (code == 2u) {
        q0 = 0.0; q1 = p1; q2 = 0.0;
      } else if (code == 0u) {
        q0 = 0.0; q1 = 0.0; q2 = p2;
      } else {
        q0 = 0.333333 * p0; q1 = 0.333333 * p1; q2 = 0.333333 * p2;
      }
      s0 += q1 / (q0 + q1 + q2);           // observed heterozygosity
      s1 += 2.0 * pi * (1.0 - pi);         // expected
    } else {
      if (code == 3u) {
        s0 += log(p0);
        s1 += log((1.0 - pi) * (1.0 - pi));
      } else if (code == 2u) {
        s0 += log(p1);
        s1 += log(2.0 * pi * (1.0 - pi));
      } else if (code == 0u) {
        s0 += log(p2);
        s1 += log(pi * pi);
      } else {
        s0 += log(0.333333 * (p0 + p1 + p2));
        s1 += log(0.333333 * ((1.0 - pi) * (1.0 - pi) + 2.0 * pi * (1.0 - pi) + pi * pi));
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_down(s0, o);
    s1 += __shfl_down(s1, o);
  }
  if (lane == 0) {
    red[0][wv] = s0;
    red[1][wv] = s1;
  }
  __syncthreads();
  if (t == 0) {
    const double a = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    const double b = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    if (g.mode == 0) {
      const double f = fmin(fmax(-1.0, 1.0 - a / b), 1.0);
      g.delta[j] = f - fc;
      g.coef_out[j] = f;
    } else {
      g.lrt[j] = 2.0 * (a - b);
    }
  }
}

// AreClose() (src/Common.hpp:83-105) with its defaults: abs 0, rel 1e-9
bool are_close(double a, double b) {
  if (a == b) return true;
  if (std::isinf(a) || std::isinf(b)) return false;
  const double d = std::fabs(a - b);
  return d <= 0.0 || d <= std::fabs(1e-9 * a) || d <= std::fabs(1e-9 * b);
}

}  // namespace

int Engine::inbreeding(const double* U, const double* S, const double* V, uint32_t k, const double* F, int maxiter,
                       double tolem, double* coef, double* lrt, double* hwe_p) {
  PCAONE_REQUIRE(have_bed_ && src_ == SRC_BED && !gl_mode_, -3, "inbreeding: load PLINK genotypes first");
  PCAONE_REQUIRE(U && S && V && F && k >= 1 && k <= (uint32_t)kInbMaxK, -1,
                 "inbreeding: needs U (N x k), S (k), V (M x k), F (M), 1 <= k <= 64");
  PCAONE_REQUIRE(maxiter >= 0, -1, "inbreeding: maxiter must be >= 0");
  PCAONE_HIP(hipSetDevice(device_));
  const uint32_t M = M_;
  std::vector<double> sv((size_t)M * k);
  for (uint32_t j = 0; j < M; ++j)
    for (uint32_t c = 0; c < k; ++c) sv[(size_t)j * k + c] = S[c] * V[(size_t)c * M + j];
  DevBuf<double> dU, dSV, dF, dc0, dc1, dd;
  dU.reserve((uint64_t)N_ * k);
  dSV.reserve((uint64_t)M * k);
  dF.reserve(M);
  dc0.reserve(M);
  dc1.reserve(M);
  dd.reserve(M);
  PCAONE_HIP(hipMemcpyAsync(dU.p, U, (size_t)N_ * k * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dSV.p, sv.data(), sv.size() * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dF.p, F, (size_t)M * 8, hipMemcpyHostToDevice, stream_));
  InbPass g{};
  g.bed = resident_raw("inbreeding");
  g.row_bytes = (N_ + 3) >> 2;
  g.nsamples = N_;
  g.nsnps = M;
  g.k = k;
  g.U = dU.p;
  g.SV = dSV.p;
  g.F = dF.p;
  // one coefficient pass: F (host) in, F out and the change
  std::vector<double> Fc(M, 0.0), D(M);
  auto coef_pass = [&](std::vector<double>& f, std::vector<double>& d) {
    PCAONE_HIP(hipMemcpyAsync(dc0.p, f.data(), (size_t)M * 8, hipMemcpyHostToDevice, stream_));
    g.coef_in = dc0.p;
    g.coef_out = dc1.p;
    g.delta = dd.p;
    g.mode = 0;
    timed("inbreed", [&] { hipLaunchKernelGGL(k_inbreed, dim3(M), dim3(256), 0, stream_, g); });
    PCAONE_HIP(hipMemcpyAsync(f.data(), dc1.p, (size_t)M * 8, hipMemcpyDeviceToHost, stream_));
    PCAONE_HIP(hipMemcpyAsync(d.data(), dd.p, (size_t)M * 8, hipMemcpyDeviceToHost, stream_));
    sync();
  };
  std::vector<double> F0(M), D1(M), D2(M);
  int it = 0;
  for (; it < maxiter; ++it) {
    F0 = Fc;
    coef_pass(Fc, D1);  // F1
    double sr2 = 0.0;
    for (uint32_t j = 0; j < M; ++j) sr2 += D1[j] * D1[j];
    coef_pass(Fc, D2);  // F2
    double sv2 = 0.0;
    for (uint32_t j = 0; j < M; ++j) sv2 += (D2[j] - D1[j]) * (D2[j] - D1[j]);
    if (are_close(sv2, 0.0)) break;
    const double alpha = std::fmin(std::fmax(1.0, std::sqrt(sr2 / sv2)), 256.0);
    for (uint32_t j = 0; j < M; ++j) {
      const double f = F0[j] + 2.0 * alpha * D1[j] + alpha * alpha * (D2[j] - D1[j]);
      Fc[j] = f < -1.0 ? -1.0 : (f > 1.0 ? 1.0 : f);
    }
    coef_pass(Fc, D1);  // the stabilising step
    double e = 0.0;
    for (uint32_t j = 0; j < M; ++j) e += (F0[j] - Fc[j]) * (F0[j] - Fc[j]);
    if (std::sqrt(e / M) < tolem) break;
  }
  // the LRT at the final coefficients
  PCAONE_HIP(hipMemcpyAsync(dc0.p, Fc.data(), (size_t)M * 8, hipMemcpyHostToDevice, stream_));
  g.coef_in = dc0.p;
  g.lrt = dd.p;
  g.mode = 1;
  timed("inbreed", [&] { hipLaunchKernelGGL(k_inbreed, dim3(M), dim3(256), 0, stream_, g); });
  std::vector<double> T(M);
  PCAONE_HIP(hipMemcpyAsync(T.data(), dd.p, (size_t)M * 8, hipMemcpyDeviceToHost, stream_));
  sync();
  for (uint32_t j = 0; j < M; ++j) {
    if (coef) coef[j] = Fc[j];
    if (lrt) lrt[j] = T[j];
    // chisq1d (src/Utils.cpp:440-443): the chi-square(1) upper tail
    // Q(1/2, x/2) = erfc(sqrt(x/2)); NaN (x < 0) is 1
    const double p = std::erfc(std::sqrt(T[j] / 2.0));
    if (hwe_p) hwe_p[j] = std::isnan(p) ? 1.0 : p;
  }
  return std::min(it + 1, std::max(maxiter, 1));
}

void write_hwe(const char* fileout, const char* filebim, const double* hwe_p, const double* lrt, const double* coef,
               uint32_t nsnps) {
  PCAONE_REQUIRE(fileout && filebim && hwe_p && lrt && coef, -1, "write_hwe: null argument");
  std::ifstream in(filebim);
  PCAONE_REQUIRE(in.is_open(), -1, std::string("can not open ") + filebim);
  std::ofstream o(fileout);
  PCAONE_REQUIRE(o.is_open(), -1, std::string("can not open ") + fileout);
  o << "#ID\tHWE_P\tLRT\tInbreeding_coefficient\n";
  std::string line;
  uint32_t j = 0;
  while (std::getline(in, line)) {
    PCAONE_REQUIRE(j < nsnps, -1, "write_hwe: the .bim has more lines than sites");
    const size_t a = line.find('\t');
    PCAONE_REQUIRE(a != std::string::npos, -1, "write_hwe: a .bim line has no second field");
    const size_t b = line.find('\t', a + 1);
    o << line.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1) << '\t' << hwe_p[j] << '\t'
      << lrt[j] << '\t' << coef[j] << '\n';
    ++j;
  }
  PCAONE_REQUIRE(j == nsnps, -1, "write_hwe: the .bim has fewer lines than sites");
  PCAONE_REQUIRE((bool)o, -1, "write_hwe: write failed");
}

}  // namespace pcaone
