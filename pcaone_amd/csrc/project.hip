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
#include <zlib.h>

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

// One wave: A (k x k, symmetric, in LDS) = L L^T by Cholesky in place, then
// L y = rhs and L^T x = y by lane 0; x_c goes to out[c * stride].  A pivot
// that is not positive gives a zero reciprocal (its unknown left at 0).
__device__ void wave_cholesky_solve(double (*A)[kProjMaxK + 1], double* rhs, int k, int lane, double* out,
                                    uint64_t stride) {
  __builtin_amdgcn_wave_barrier();
  for (int c = 0; c < k; ++c) {
    // A[c][c..] is final; column c below the diagonal is scaled, the rest updated
    const double p = A[c][c];
    const double rp = p > 0.0 ? 1.0 / sqrt(p) : 0.0;
    __builtin_amdgcn_wave_barrier();
    if (lane >= c && lane < k) A[lane][c] *= rp;  // L[lane][c] (lane == c: sqrt p)
    __builtin_amdgcn_wave_barrier();
    for (int rr = c + 1 + lane; rr < k; rr += 64)
      for (int cc = c + 1; cc <= rr; ++cc) A[rr][cc] -= A[rr][c] * A[cc][c];
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0) {  // k <= 32
    double y[kProjMaxK];
    for (int c = 0; c < k; ++c) {
      double s = rhs[c];
      for (int cc = 0; cc < c; ++cc) s -= A[c][cc] * y[cc];
      y[c] = s / A[c][c];
    }
    for (int c = k - 1; c >= 0; --c) {
      double s = y[c];
      for (int cc = c + 1; cc < k; ++cc) s -= A[cc][c] * y[cc];
      y[c] = s / A[c][c];
    }
    for (int c = 0; c < k; ++c) out[(uint64_t)c * stride] = y[c];
  }
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
  wave_cholesky_solve(A[wv], rhs[wv], k, lane, u + i, a.nsamples);
}

// ---- BEAGLE likelihood projection (run_projection_beagle_gl) --------------

// E step and right-hand sides: thread = sample, blockIdx.y = a chunk of
// matched rows.  Initial step: pt = F; iterations: pt = (U_i S) L_j, times
// sqrt(2F(1-F)) when standardising, plus F; both clamped to [1e-4, 1 - 1e-4].
// L and VS = L S are zero-padded to KM columns (row stride KM) and, with F and
// norm = sqrt(2F(1-F)) per row, read at wave-uniform addresses.
template <int KM>
__global__ __launch_bounds__(256) void k_projgl_estep(
    const float2* __restrict__ P, const int32_t* __restrict__ site, const double* __restrict__ F,
    const double* __restrict__ nrm, const double* __restrict__ L, const double* __restrict__ VS,
    const double* __restrict__ S, const double* __restrict__ U, double* __restrict__ part, double* __restrict__ corr,
    uint32_t* __restrict__ nmiss, uint32_t nsamples, uint32_t nmatch, uint32_t k, uint32_t chunk, int init,
    int standardize) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nsamples) return;  // no barrier in this kernel
  double us[KM], b[KM];
#pragma unroll
  for (int c = 0; c < KM; ++c) {
    us[c] = (!init && c < (int)k) ? U[(uint64_t)i * k + c] * S[c] : 0.0;
    b[c] = 0.0;
  }
  const uint32_t r0 = blockIdx.y * chunk, r1 = min(nmatch, r0 + chunk);
  uint32_t miss = 0;
  // eight sites' likelihoods in flight before any is used
  constexpr int kAhead = 8;
  for (uint32_t rb = r0; rb < r1; rb += kAhead) {
    float2 glv[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
      glv[u] = rb + u < r1 ? P[(uint64_t)site[rb + u] * nsamples + i] : make_float2(0.f, 0.f);
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      const uint32_t r = rb + u;
      if (r >= r1) break;
      const double f = F[r], norm = nrm[r];
      const double* Lr = L + (uint64_t)r * KM;
      const double* Vr = VS + (uint64_t)r * KM;
      double pt = f;
      if (!init) {
        double z = 0.0;
#pragma unroll
        for (int c = 0; c < KM; ++c) z += us[c] * Lr[c];
        pt = (standardize ? z * norm : z) + f;
      }
      pt = fmin(fmax(pt, 1e-4), 1.0 - 1e-4);
      const double g11 = glv[u].x, g12 = glv[u].y, g22 = 1.0 - g11 - g12;
      const double p11 = g11 * pt * pt;
      const double p12 = g12 * 2.0 * pt * (1.0 - pt);
      const double p22 = g22 * (1.0 - pt) * (1.0 - pt);
      const double ps = p11 + p12 + p22;
      if (!isfinite(ps) || ps <= 0.0) {  // missing: out of this sample's fit
        ++miss;
        double* cr = corr + (uint64_t)i * k * k;
        for (int c = 0; c < (int)k; ++c)
          for (int d = 0; d < (int)k; ++d) atomicAdd(cr + c * k + d, Vr[c] * Vr[d]);
        continue;
      }
      double g = (2.0 * p11 + p12) / (2.0 * ps) - f;
      if (standardize) {
        if (norm > 1e-9) g /= norm;
        else if (!init) g = 0.0;
      }
#pragma unroll
      for (int c = 0; c < KM; ++c) b[c] += g * Vr[c];
    }
  }
  if (miss) atomicAdd(nmiss + i, miss);
  double* out = part + ((uint64_t)blockIdx.y * nsamples + i) * k;
#pragma unroll
  for (int c = 0; c < KM; ++c)
    if (c < (int)k) out[c] = b[c];
}

// per sample (one wave): the right-hand side summed over the chunks in order,
// the Gram of V S less the sample's missing rows, Cholesky and solves;
// U (nsamples x k row-major)
__global__ __launch_bounds__(256) void k_projgl_solve(const double* __restrict__ part, uint32_t nchunk,
                                                      const double* __restrict__ gram, const double* __restrict__ corr,
                                                      const uint32_t* __restrict__ nmiss, uint32_t nsamples,
                                                      uint32_t k, double* __restrict__ U) {
  __shared__ double A[4][kProjMaxK][kProjMaxK + 1];
  __shared__ double rhs[4][kProjMaxK];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * 4 + wv;
  if (i >= nsamples) return;  // whole waves; no block-wide barrier follows
  const int kk = (int)k;
  const bool has_miss = nmiss[i] != 0;
  for (int e = lane; e < kk * kk; e += 64) {
    const int c = e / kk, d = e % kk;
    A[wv][c][d] = gram[e] - (has_miss ? corr[(uint64_t)i * k * k + e] : 0.0);
  }
  if (lane < kk) {
    double s = 0.0;
    for (uint32_t ch = 0; ch < nchunk; ++ch) s += part[((uint64_t)ch * nsamples + i) * k + lane];
    rhs[wv][lane] = s;
  }
  wave_cholesky_solve(A[wv], rhs[wv], kk, lane, U + (uint64_t)i * k, 1);
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

namespace pcaone {

// run_projection_beagle_gl (src/Projection.cpp:32-137) on the likelihoods of
// load_gl, with FileBeagle::read_all's matching of the BEAGLE sites against
// the reference .mbim (src/FileBeagle.cpp:12-118)
uint32_t Engine::project_gl(const char* filebeagle, const char* filembim, const double* S, const double* Vref,
                            uint32_t k, uint32_t ref_nsnps, int standardize, int maxiter, double tolem, double* U,
                            double* VSout, int* iters, double* diffs) {
  PCAONE_REQUIRE(have_bed_ && gl_mode_ && !gl_bed_, -3, "project_gl: load the target likelihoods (load_gl) first");
  PCAONE_REQUIRE(filebeagle && filembim && S && Vref && U, -1, "project_gl: null argument");
  PCAONE_REQUIRE(k >= 1 && k <= (uint32_t)kProjMaxK, -1, "project_gl: k must be in [1, 32]");
  PCAONE_REQUIRE(maxiter >= 0, -1, "project_gl: maxiter must be >= 0");
  struct Ref {
    int index;
    std::string a1, a2;
    double freq;
  };
  std::unordered_map<std::string, Ref> by_id, by_pos;
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
      const Ref r{idx++, t[4], t[5], f};
      by_id.emplace(t[1], r);  // the first line of a key wins
      by_pos.emplace(t[0] + "_" + t[3], r);
    }
    PCAONE_REQUIRE((uint32_t)idx == ref_nsnps, -1, "project_gl: the .mbim and the loadings differ in SNP count");
  }
  // the BEAGLE sites: marker and the two alleles (0-3 read as A, C, G, T)
  auto allele = [](const std::string& a) -> std::string {
    if (a == "0") return "A";
    if (a == "1") return "C";
    if (a == "2") return "G";
    if (a == "3") return "T";
    return a;
  };
  std::vector<int32_t> ref_idx(M_, -1);
  std::vector<uint8_t> flips(M_, 0);
  std::vector<double> Fsite(M_, 0.0);
  {
    gzFile fp = gzopen(filebeagle, "rb");
    PCAONE_REQUIRE(fp != nullptr, -1, std::string("project_gl: cannot open ") + filebeagle);
    std::vector<char> buf(1 << 20);
    auto next_line = [&](std::string& head) -> bool {  // the line's first 1 MB; the rest skipped
      if (!gzgets(fp, buf.data(), (int)buf.size())) return false;
      head.assign(buf.data());
      bool nl = !head.empty() && head.back() == '\n';
      while (!nl) {
        if (!gzgets(fp, buf.data(), (int)buf.size())) break;
        const size_t len = std::strlen(buf.data());
        nl = len > 0 && buf[len - 1] == '\n';
      }
      return true;
    };
    std::string line;
    bool ok = next_line(line);  // header
    uint32_t j = 0;
    while (ok && next_line(line)) {
      if (j >= M_) {
        gzclose(fp);
        throw Error(-1, "project_gl: the BEAGLE file has more sites than the loaded likelihoods");
      }
      std::vector<std::string> tok;
      size_t p = 0;
      while (tok.size() < 3) {
        const size_t b = line.find_first_not_of("\t \n", p);
        if (b == std::string::npos) break;
        const size_t e = line.find_first_of("\t \n", b);
        tok.emplace_back(line, b, (e == std::string::npos ? line.size() : e) - b);
        if (e == std::string::npos) break;
        p = e;
      }
      while (tok.size() < 3) tok.emplace_back();
      const Ref* found = nullptr;
      const auto iid = by_id.find(tok[0]);
      if (iid != by_id.end()) {
        found = &iid->second;
      } else {
        const auto ipos = by_pos.find(tok[0]);
        if (ipos != by_pos.end()) found = &ipos->second;
      }
      if (found) {
        const Ref& r = *found;
        const std::string a1 = allele(tok[1]), a2 = allele(tok[2]);
        int flip = -1;
        if (a1 == r.a1 && a2 == r.a2) flip = 0;
        else if (a1 == r.a2 && a2 == r.a1) flip = 1;
        if (flip >= 0) {
          ref_idx[j] = r.index;
          flips[j] = (uint8_t)flip;
          Fsite[j] = flip ? 1.0 - r.freq : r.freq;
        }
      }
      ++j;
    }
    gzclose(fp);
    PCAONE_REQUIRE(j == M_, -1, "project_gl: the BEAGLE file and the loaded likelihoods differ in site count");
  }
  return project_gl_matched(ref_idx.data(), flips.data(), Fsite.data(), S, Vref, k, ref_nsnps, standardize, maxiter,
                            tolem, U, VSout, iters, diffs);
}

// the EM of run_projection_beagle_gl once the sites are matched: ref_idx (-1
// unmatched), flip and F per loaded site, as FileBeagle::read_all leaves them
uint32_t Engine::project_gl_matched(const int32_t* ref_idx, const uint8_t* flip, const double* Fsite,
                                    const double* S, const double* Vref, uint32_t k, uint32_t ref_nsnps,
                                    int standardize, int maxiter, double tolem, double* U, double* VSout, int* iters,
                                    double* diffs) {
  PCAONE_REQUIRE(have_bed_ && gl_mode_ && !gl_bed_, -3, "project_gl: load the target likelihoods (load_gl) first");
  PCAONE_REQUIRE(ref_idx && flip && Fsite && S && Vref && U, -1, "project_gl: null argument");
  PCAONE_REQUIRE(k >= 1 && k <= (uint32_t)kProjMaxK, -1, "project_gl: k must be in [1, 32]");
  PCAONE_REQUIRE(maxiter >= 0, -1, "project_gl: maxiter must be >= 0");
  std::vector<int32_t> site;
  std::vector<double> F, L;
  for (uint32_t j = 0; j < M_; ++j) {
    if (ref_idx[j] < 0) continue;
    PCAONE_REQUIRE((uint32_t)ref_idx[j] < ref_nsnps, -1, "project_gl: a reference index beyond the loadings");
    site.push_back((int32_t)j);
    F.push_back(Fsite[j]);
    for (uint32_t c = 0; c < k; ++c) L.push_back(Vref[(size_t)c * ref_nsnps + ref_idx[j]] * (flip[j] ? -1.0 : 1.0));
  }
  const uint32_t nm = (uint32_t)site.size();
  PCAONE_REQUIRE(nm > 0, -1, "no overlapping SNPs found between target BEAGLE and reference .mbim");
  const uint32_t km = k <= 4 ? 4 : k <= 8 ? 8 : k <= 16 ? 16 : 32;  // padded row length
  std::vector<double> gram((size_t)k * k, 0.0), Sv(S, S + k), nrm(nm), Lp((size_t)nm * km, 0.0),
      VSp((size_t)nm * km, 0.0);
  for (uint32_t r = 0; r < nm; ++r) {
    nrm[r] = std::sqrt(2.0 * F[r] * (1.0 - F[r]));
    for (uint32_t c = 0; c < k; ++c) {
      Lp[(size_t)r * km + c] = L[(size_t)r * k + c];
      VSp[(size_t)r * km + c] = L[(size_t)r * k + c] * S[c];
    }
    for (uint32_t c = 0; c < k; ++c)
      for (uint32_t e = 0; e < k; ++e) gram[c * k + e] += VSp[(size_t)r * km + c] * VSp[(size_t)r * km + e];
  }
  PCAONE_HIP(hipSetDevice(device_));
  const uint32_t bx = (uint32_t)ceil_div(N_, 256);
  const uint32_t nchunk = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(nm, 16), ceil_div(2048, bx)));
  const uint32_t chunk = (uint32_t)ceil_div(nm, nchunk);
  DevBuf<int32_t> dsite;
  DevBuf<double> dF, dn, dL, dV, dS, dG, dU, dUn, dpart, dcorr;
  DevBuf<uint32_t> dmiss;
  dsite.reserve(nm);
  dF.reserve(nm);
  dn.reserve(nm);
  dL.reserve((uint64_t)nm * km);
  dV.reserve((uint64_t)nm * km);
  dS.reserve(k);
  dG.reserve((uint64_t)k * k);
  dU.reserve((uint64_t)N_ * k);
  dUn.reserve((uint64_t)N_ * k);
  dpart.reserve((uint64_t)nchunk * N_ * k);
  dcorr.reserve((uint64_t)N_ * k * k);
  dmiss.reserve(N_);
  PCAONE_HIP(hipMemcpyAsync(dsite.p, site.data(), (size_t)nm * 4, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dF.p, F.data(), (size_t)nm * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dn.p, nrm.data(), (size_t)nm * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dL.p, Lp.data(), Lp.size() * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dV.p, VSp.data(), VSp.size() * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dS.p, Sv.data(), (size_t)k * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dG.p, gram.data(), gram.size() * 8, hipMemcpyHostToDevice, stream_));
  const float2* P2 = reinterpret_cast<const float2*>(glP_.p);
  const int std1 = standardize ? 1 : 0;
  // one E step and solve: scores into dUn
  auto step = [&](bool init) {
    const int in = init ? 1 : 0;
    PCAONE_HIP(hipMemsetAsync(dcorr.p, 0, (size_t)N_ * k * k * 8, stream_));
    PCAONE_HIP(hipMemsetAsync(dmiss.p, 0, (size_t)N_ * 4, stream_));
    timed("projgl_estep", [&] {
      const dim3 grid(bx, nchunk);
#define PCAONE_PROJGL(KM)                                                                                          \
  hipLaunchKernelGGL(k_projgl_estep<KM>, grid, dim3(256), 0, stream_, P2, dsite.p, dF.p, dn.p, dL.p, dV.p, dS.p, \
                     dU.p, dpart.p, dcorr.p, dmiss.p, N_, nm, k, chunk, in, std1)
      if (km == 4) PCAONE_PROJGL(4);
      else if (km == 8) PCAONE_PROJGL(8);
      else if (km == 16) PCAONE_PROJGL(16);
      else PCAONE_PROJGL(32);
#undef PCAONE_PROJGL
    });
    timed("projgl_solve", [&] {
      hipLaunchKernelGGL(k_projgl_solve, dim3((uint32_t)ceil_div(N_, 4)), dim3(256), 0, stream_, dpart.p, nchunk,
                         dG.p, dcorr.p, dmiss.p, N_, k, dUn.p);
    });
  };
  const size_t ub = (size_t)N_ * k * 8;
  std::vector<double> Uc((size_t)N_ * k), Up((size_t)N_ * k);  // row-major
  step(true);
  PCAONE_HIP(hipMemcpyAsync(Uc.data(), dUn.p, ub, hipMemcpyDeviceToHost, stream_));
  sync();
  int it = 0;
  for (int iter = 0; iter < maxiter; ++iter) {
    Up = Uc;
    PCAONE_HIP(hipMemcpyAsync(dU.p, Up.data(), ub, hipMemcpyHostToDevice, stream_));
    step(false);
    PCAONE_HIP(hipMemcpyAsync(Uc.data(), dUn.p, ub, hipMemcpyDeviceToHost, stream_));
    sync();
    // align each score column with the previous one, then the relative change
    for (uint32_t c = 0; c < k; ++c) {
      double dot = 0.0;
      for (uint32_t i = 0; i < N_; ++i) dot += Uc[(size_t)i * k + c] * Up[(size_t)i * k + c];
      if (dot < 0.0)
        for (uint32_t i = 0; i < N_; ++i) Uc[(size_t)i * k + c] = -Uc[(size_t)i * k + c];
    }
    double num = 0.0, den = 0.0;
    for (size_t e = 0; e < Uc.size(); ++e) {
      num += (Uc[e] - Up[e]) * (Uc[e] - Up[e]);
      den += Up[e] * Up[e];
    }
    den = std::sqrt(den);
    if (den < 1e-12) den = 1.0;
    const double diff = std::sqrt(num) / den;
    if (diffs) diffs[iter] = diff;
    it = iter + 1;
    if (diff < tolem) break;
  }
  if (iters) *iters = it;
  for (uint32_t i = 0; i < N_; ++i)
    for (uint32_t c = 0; c < k; ++c) U[(size_t)c * N_ + i] = Uc[(size_t)i * k + c];
  if (VSout)  // the loadings the reference writes: V S on the matched sites, column-major nm x k
    for (uint32_t r = 0; r < nm; ++r)
      for (uint32_t c = 0; c < k; ++c) VSout[(size_t)c * nm + r] = L[(size_t)r * k + c] * S[c];
  return nm;
}

}  // namespace pcaone
