// Genotype-likelihood kernels for PCAngsd on BEAGLE input.  The likelihoods
// stay FP32 in HBM as [site][sample][2] (the reference keeps two GLs per
// sample and uses 1 - P0 - P1 for the third, src/Utils.cpp:391-397); every
// sum over samples or sites is FP64.
//
//   posterior(pt) = (p1 + 2 p2) / (p0 + p1 + p2),
//   p0 = P0 (1-pt)^2, p1 = P1 2 pt (1-pt), p2 = (1 - P0 - P1) pt^2
//   emMAF_with_GL:  F_j <- 1/(2N) sum_i posterior(F_j)          (src/Utils.cpp:501-531)
//   initial E:      E_ij = posterior(F_j) - 2 F_j                (src/FileBeagle.cpp:133-143)
//   PCAngsd update: pt = clamp((U_i S V_j + 2 F_j) / 2, 1e-4, 1 - 1e-4),
//                   E_ij = posterior(pt) - 2 F_j                 (src/Data.cpp:249-270)
//   standardise:    E_ij / sqrt(2 F (1 - F)) and the GRM diagonal
//                   Dc_i = sum_j sum_g (g - 2F)^2 post_g / (2F(1-F))  (src/Data.cpp:301-343)
//
// PLINK input with --pcangsd: the likelihoods are the hard calls of the
// packed codes (src/FilePlink.cpp:89-104), F the genotype-count frequency and
// the initial E the centred genotype BED2GENO - F, missing 0
// (src/FilePlink.cpp:106-114), in half-dosage units as the reference has it.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {

__device__ inline double gl_post(double P0, double P1, double pt, double& p0, double& p1, double& p2) {
  p0 = P0 * (1.0 - pt) * (1.0 - pt);
  p1 = P1 * 2.0 * pt * (1.0 - pt);
  p2 = (1.0 - P0 - P1) * pt * pt;
  return (p1 + 2.0 * p2) / (p0 + p1 + p2);
}

// one wave per site; F_out = 1/(2N) sum_i posterior(F_in); dpart[block] = sum (F_out - F_in)^2
__global__ __launch_bounds__(256) void k_gl_em_maf(const float* __restrict__ P, uint32_t nsamples, uint32_t nsites,
                                                   const double* __restrict__ Fin, double* __restrict__ Fout,
                                                   double* __restrict__ dpart) {
  __shared__ double sh[4];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t j = (uint64_t)blockIdx.x * 4 + wv;
  double d2 = 0.0;
  if (j < nsites) {
    const float* row = P + j * (uint64_t)nsamples * 2;
    const double f = Fin[j];
    double s = 0.0, p0, p1, p2;
    for (uint32_t i = lane; i < nsamples; i += 64) s += gl_post(row[2 * i], row[2 * i + 1], f, p0, p1, p2);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane == 0) {
      const double fn = s * (1.0 / (2.0 * nsamples));
      Fout[j] = fn;
      d2 = (fn - f) * (fn - f);
    }
  }
  if (lane == 0) sh[wv] = d2;
  __syncthreads();
  if (threadIdx.x == 0) dpart[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

void launch_gl_em_maf(const float* P, uint32_t nsamples, uint32_t nsites, const double* Fin, double* Fout,
                      double* dpart, hipStream_t s) {
  hipLaunchKernelGGL(k_gl_em_maf, dim3((uint32_t)ceil_div(nsites, 4)), dim3(256), 0, s, P, nsamples, nsites, Fin,
                     Fout, dpart);
}

uint32_t gl_em_maf_blocks(uint32_t nsites) { return (uint32_t)ceil_div(nsites, 4); }

// FileBed::read_all's two likelihoods per sample for --pcangsd
// (src/FilePlink.cpp:89-104): code 3 (RR) (1, 0), code 0 (AA) (0, 0), code 2
// (RA) (0, 1), missing (0.333333, 0.333333).  raw: nsites rows of
// ceil(nsamples / 4) bytes, sample j in bits 2 (j % 4) of byte j / 4.
__global__ __launch_bounds__(256) void k_bed_to_gl(const uint8_t* __restrict__ raw, uint32_t nsamples,
                                                   uint32_t nsites, float2* __restrict__ P) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (uint64_t)nsamples * nsites) return;
  const uint64_t site = gid / nsamples;
  const uint32_t j = (uint32_t)(gid - site * nsamples);
  const uint32_t nb = (nsamples + 3) >> 2;
  const uint32_t code = (raw[site * nb + (j >> 2)] >> (2 * (j & 3))) & 3u;
  float2 p;
  p.x = code == 3 ? 1.0f : (code == 1 ? 0.333333f : 0.0f);
  p.y = code == 2 ? 1.0f : (code == 1 ? 0.333333f : 0.0f);
  P[gid] = p;
}

void launch_bed_to_gl(const uint8_t* raw, uint32_t nsamples, uint32_t nsites, float* P, hipStream_t s) {
  const uint64_t total = (uint64_t)nsamples * nsites;
  if (!total) return;
  hipLaunchKernelGGL(k_bed_to_gl, dim3((uint32_t)ceil_div(total, 256)), dim3(256), 0, s, raw, nsamples, nsites,
                     reinterpret_cast<float2*>(P));
}

// E rows (laid-out row r holds site perm[r]) into out[r * ld + i]; with U == null
// the prior is pt = F (initial E), else the PCAngsd update from U (N x k),
// S (k), V (M x k, laid-out rows).  Optional standardisation and Dc partials
// (dpart[blockIdx.y][i], rows of block y).  U_i S stays in registers (KMAX
// bounds k at compile time).
template <int KMAX>
__global__ __launch_bounds__(256) void k_gl_E(GlE e) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t r0 = blockIdx.y * e.rows_per_block;
  const uint32_t r1 = min(r0 + e.rows_per_block, e.nsites);
  double dc = 0.0;
  if (i < e.nsamples) {
    double us[KMAX];
#pragma unroll
    for (int c = 0; c < KMAX; ++c) us[c] = (e.U && c < (int)e.k) ? e.U[(uint64_t)i * e.k + c] * e.S[c] : 0.0;
    for (uint32_t r = r0; r < r1; ++r) {
      const uint32_t site = e.perm ? (uint32_t)e.perm[r] : r;
      const float2 pp = reinterpret_cast<const float2*>(e.P + (uint64_t)site * e.nsamples * 2)[i];
      const double F = e.F[site];
      double pt = F;
      if (e.U) {
        double dot = 0.0;
#pragma unroll
        for (int c = 0; c < KMAX; ++c)
          if (c < (int)e.k) dot += us[c] * e.V[(uint64_t)r * e.k + c];
        pt = fmin(fmax((dot + 2.0 * F) / 2.0, 1e-4), 1.0 - 1e-4);
      }
      double p0, p1, p2;
      const double post = gl_post(pp.x, pp.y, pt, p0, p1, p2);
      double v = post - 2.0 * F;
      if (!e.U && e.bed_init) {
        // hard calls: dosage P1 + 2 (1 - P0 - P1); missing (both 0.333333) is 0
        const bool miss = pp.x > 0.3f && pp.y > 0.3f;
        v = miss ? 0.0 : 0.5 * ((double)pp.y + 2.0 * (1.0 - (double)pp.x - (double)pp.y)) - F;
      }
      if (e.standardize) {
        const double norm = sqrt(2.0 * F * (1.0 - F));
        if (norm > 1e-9) v /= norm;
        const double ps = p0 + p1 + p2, f2 = 2.0 * F;
        const double t = (0.0 - f2) * (0.0 - f2) * (p0 / ps) + (1.0 - f2) * (1.0 - f2) * (p1 / ps) +
                         (2.0 - f2) * (2.0 - f2) * (p2 / ps);
        dc += t / (2.0 * F * (1.0 - F));
      }
      e.out[(uint64_t)r * e.ld + i] = (float)v;
    }
    if (e.dpart) e.dpart[(uint64_t)blockIdx.y * e.nsamples + i] = dc;
  }
}

void launch_gl_E(const GlE& e, hipStream_t s) {
  PCAONE_REQUIRE(e.k <= 64, -1, "PCAngsd update supports k <= 64");
  dim3 grid((uint32_t)ceil_div(e.nsamples, 256), (uint32_t)ceil_div(e.nsites, e.rows_per_block));
  const uint32_t k = e.U ? e.k : 0;
  if (k <= 8)
    hipLaunchKernelGGL(k_gl_E<8>, grid, dim3(256), 0, s, e);
  else if (k <= 16)
    hipLaunchKernelGGL(k_gl_E<16>, grid, dim3(256), 0, s, e);
  else if (k <= 32)
    hipLaunchKernelGGL(k_gl_E<32>, grid, dim3(256), 0, s, e);
  else
    hipLaunchKernelGGL(k_gl_E<64>, grid, dim3(256), 0, s, e);
}

// out[i] = sum_y part[y][i] over ny rows, fixed order
__global__ void k_sum_rows(const double* __restrict__ part, uint32_t ny, uint32_t n, double* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s0 = 0.0, s1 = 0.0;
  uint32_t y = 0;
  for (; y + 1 < ny; y += 2) {
    s0 += part[(uint64_t)y * n + i];
    s1 += part[(uint64_t)(y + 1) * n + i];
  }
  if (y < ny) s0 += part[(uint64_t)y * n + i];
  out[i] = s0 + s1;
}

void launch_sum_rows(const double* part, uint32_t ny, uint32_t n, double* out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_sum_rows, dim3((uint32_t)ceil_div(n, 256)), dim3(256), 0, s, part, ny, n, out);
}

uint32_t gl_rows_per_block(uint32_t nsites) {
  return (uint32_t)std::max<uint64_t>(64, ceil_div(nsites, 256));
}

// C[a][b] = sum_r E[r][a] E[r][b] / M (FP64 accumulation of FP32 rows), with
// the diagonal replaced by Dc / M (src/Halko.cpp:317-331).  64 x 64 tiles.
__global__ __launch_bounds__(256) void k_grm(const float* __restrict__ E, uint64_t ld, uint32_t nsamples,
                                             uint32_t nsites, const double* __restrict__ Dc, double* __restrict__ C) {
  __shared__ float ta[32][64], tb[32][64];
  const uint32_t a0 = blockIdx.y * 64, b0 = blockIdx.x * 64;
  if (b0 < a0) return;  // upper triangle of tiles; mirrored on write
  const uint32_t tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4] = {};
  for (uint32_t r0 = 0; r0 < nsites; r0 += 32) {
    for (uint32_t e = threadIdx.x; e < 32 * 64; e += 256) {
      const uint32_t rr = e / 64, cc = e % 64, r = r0 + rr;
      ta[rr][cc] = (r < nsites && a0 + cc < nsamples) ? E[(uint64_t)r * ld + a0 + cc] : 0.0f;
      tb[rr][cc] = (r < nsites && b0 + cc < nsamples) ? E[(uint64_t)r * ld + b0 + cc] : 0.0f;
    }
    __syncthreads();
    for (uint32_t rr = 0; rr < 32; ++rr) {
      double va[4], vb[4];
      for (int q = 0; q < 4; ++q) {
        va[q] = ta[rr][ty + 16 * q];
        vb[q] = tb[rr][tx + 16 * q];
      }
      for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) acc[p][q] = fma(va[p], vb[q], acc[p][q]);
    }
    __syncthreads();
  }
  for (int p = 0; p < 4; ++p)
    for (int q = 0; q < 4; ++q) {
      const uint32_t a = a0 + ty + 16 * p, b = b0 + tx + 16 * q;
      if (a < nsamples && b < nsamples) {
        const double v = (a == b) ? Dc[a] / nsites : acc[p][q] / nsites;
        C[(uint64_t)a * nsamples + b] = v;
        C[(uint64_t)b * nsamples + a] = v;
      }
    }
}

void launch_grm(const float* E, uint64_t ld, uint32_t nsamples, uint32_t nsites, const double* Dc, double* C,
                hipStream_t s) {
  const uint32_t t = (uint32_t)ceil_div(nsamples, 64);
  hipLaunchKernelGGL(k_grm, dim3(t, t), dim3(256), 0, s, E, ld, nsamples, nsites, Dc, C);
}

}  // namespace pcaone
