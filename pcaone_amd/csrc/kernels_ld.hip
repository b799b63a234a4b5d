// LD r^2 over the windows of divide_pos_by_window on a dense feature-major
// matrix: the FP32 .residuals input (src/LD.cpp:419-442 on
// src/FileBinary.cpp:21-30's centred FP64 matrix), or the FP64 ancestry-
// adjusted genotypes of the --USV path on PLINK input (src/LD.cpp:460-468).
// The band of the Gram X X^T that the windows touch is computed in 64 x 64
// tiles on the FP64 MFMA (products of centred FP32 inputs are exact in FP64),
// then r = <x_i, x_k> (sds_i sds_k df) and r^2 land in the window-ordered
// output.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// sds[i] = 1 / sqrt(sum_k (x_ik - mu_i)^2 / (N - 1))   (calc_sds, src/LD.cpp:20-23)
template <typename T>
__global__ __launch_bounds__(256) void k_ld_sds(const T* __restrict__ X, uint32_t n, uint32_t m,
                                                const double* __restrict__ mu, double* __restrict__ sds) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (r >= m) return;
  const T* p = X + r * n;
  const double u = mu[r];
  double s = 0.0;
  for (uint32_t j = lane; j < n; j += 64) {
    const double v = (double)p[j] - u;
    s += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) sds[r] = 1.0 / sqrt(s * (1.0 / (double)(n - 1)));
}

void launch_ld_sds(const void* X, int x_f64, uint32_t n, uint32_t m, const double* mu, double* sds,
                   hipStream_t s) {
  const dim3 grid((uint32_t)ceil_div((uint64_t)m * 64, 256));
  if (x_f64)
    hipLaunchKernelGGL(k_ld_sds<double>, grid, dim3(256), 0, s, static_cast<const double*>(X), n, m, mu, sds);
  else
    hipLaunchKernelGGL(k_ld_sds<float>, grid, dim3(256), 0, s, static_cast<const float*>(X), n, m, mu, sds);
}

constexpr int kLdK = 32;         // samples per LDS stage
constexpr int kLdLd = kLdK + 1;  // padded row stride (doubles)

// One 64 x 64 tile (row block I, column block J) per workgroup; wave (wr, wc)
// owns the 32 x 32 quadrant, four 16 x 16 MFMA tiles.  Lane (k = lane / 16,
// i = lane % 16) feeds A[i][k] = x_row(i)[k] and B[k][j] = x_col(j)[k].
template <typename T>
__global__ __launch_bounds__(256) void k_ld_tile(LdTiles t) {
  const T* __restrict__ X = static_cast<const T*>(t.X);
  __shared__ double xa[64 * kLdLd];
  __shared__ double xb[64 * kLdLd];
  const int2 tile = t.tiles[blockIdx.x];
  const uint32_t r0 = (uint32_t)tile.x * 64, c0 = (uint32_t)tile.y * 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wr = wv >> 1, wc = wv & 1;
  const int kr = lane >> 4, ci = lane & 15;
  f64x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
  for (uint32_t k0 = 0; k0 < t.n; k0 += kLdK) {
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * kLdK; e += 256) {
      const int rr = e / kLdK, kk = e % kLdK;
      const uint32_t k = k0 + kk, ra = r0 + rr, rb = c0 + rr;
      xa[rr * kLdLd + kk] = (k < t.n && ra < t.m) ? (double)X[(uint64_t)ra * t.n + k] - t.mu[ra] : 0.0;
      xb[rr * kLdLd + kk] = (k < t.n && rb < t.m) ? (double)X[(uint64_t)rb * t.n + k] - t.mu[rb] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kLdK / 4; ++ks) {
      const int k = ks * 4 + kr;
      double a[2], b[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        a[q] = xa[(wr * 32 + q * 16 + ci) * kLdLd + k];
        b[q] = xb[(wc * 32 + q * 16 + ci) * kLdLd + k];
      }
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) acc[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[p], b[q], acc[p][q], 0, 0, 0);
    }
  }
  const double df = 1.0 / (double)(t.n - 1);
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // accumulator element v of this lane: row (lane / 16) + 4 v, column lane % 16
        const uint32_t row = r0 + wr * 32 + p * 16 + (lane >> 4) + 4 * v;
        const uint32_t col = c0 + wc * 32 + q * 16 + ci;
        if (row >= t.m || col >= t.m || col <= row) continue;
        const int32_t w = t.win_of_row[row];
        if (w < 0) continue;
        const uint32_t d = col - row;
        if (d >= (uint32_t)t.we[w]) continue;
        const double r = acc[p][q][v] * (t.sds[row] * t.sds[col] * df);
        t.r2[t.off[w] + d - 1] = r * r;
      }
}

void launch_ld_tiles(const LdTiles& t, uint32_t ntiles, hipStream_t s) {
  if (!ntiles) return;
  if (t.x_f64)
    hipLaunchKernelGGL(k_ld_tile<double>, dim3(ntiles), dim3(256), 0, s, t);
  else
    hipLaunchKernelGGL(k_ld_tile<float>, dim3(ntiles), dim3(256), 0, s, t);
}

// ---------------------------------------------------------------------------
// --USV on PLINK input (run_ld_stuff, src/LD.cpp:460-466): G = BED2GENO - F
// with missing 0 (FileBed::read_all, src/FilePlink.cpp:106-114), adjusted as
// G <- (I - U U^T) G with U the N x k .eigvecs as read (not re-orthonormalised,
// so the product is formed as G - U (U^T G), not as a projection identity).
// ---------------------------------------------------------------------------
__device__ inline double centred_geno(const uint8_t* __restrict__ raw, uint32_t nb, uint32_t r, uint32_t j, double F) {
  const uint32_t code = (raw[(uint64_t)r * nb + (j >> 2)] >> (2 * (j & 3))) & 3u;
  // BED2GENO (src/Common.hpp:58): code 0 -> 1, 2 -> 0.5, 3 -> 0; missing -> 0 after centring
  return code == 1 ? 0.0 : (code == 0 ? 1.0 : (code == 2 ? 0.5 : 0.0)) - F;
}

// W = U^T G as [snp][c]: a block takes 64 SNPs and all k (<= 64) columns over
// 64-sample stages; thread (tr, tc) accumulates rows 4 tr .. 4 tr + 3 and
// columns tc + 16 q, FP64.
__global__ __launch_bounds__(256) void k_usv_w(const uint8_t* __restrict__ raw, uint32_t n, uint32_t m,
                                               const double* __restrict__ F, const double* __restrict__ U,
                                               uint32_t k, double* __restrict__ W) {
  __shared__ double g[64][65];  // [snp][sample]
  __shared__ double u[64][65];  // [sample][column]
  const uint32_t r0 = blockIdx.x * 64, nb = (n + 3) >> 2;
  const int t = threadIdx.x, tr = t >> 4, tc = t & 15;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[i][q] = 0.0;
  for (uint32_t s0 = 0; s0 < n; s0 += 64) {
    __syncthreads();
    for (int e = t; e < 64 * 64; e += 256) {
      const int a = e >> 6, b = e & 63;
      const uint32_t r = r0 + a, j = s0 + b, js = s0 + a;
      g[a][b] = (r < m && j < n) ? centred_geno(raw, nb, r, j, F[r]) : 0.0;
      u[a][b] = (js < n && (uint32_t)b < k) ? U[(uint64_t)js * k + b] : 0.0;
    }
    __syncthreads();
    for (int sj = 0; sj < 64; ++sj) {
      double x[4], y[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = g[4 * tr + i][sj];
#pragma unroll
      for (int q = 0; q < 4; ++q) y[q] = u[sj][tc + 16 * q];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[i][q] = fma(x[i], y[q], acc[i][q]);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t r = r0 + 4 * tr + i, c = tc + 16 * q;
      if (r < m && c < k) W[(uint64_t)r * k + c] = acc[i][q];
    }
}

// R[r][j] = G[r][j] - sum_c U[j][c] W[r][c]: thread j keeps U's row in
// registers (KMAX bounds k) and walks the block's SNP rows.
template <int KMAX>
__global__ __launch_bounds__(256) void k_usv_resid(const uint8_t* __restrict__ raw, uint32_t n, uint32_t m,
                                                   const double* __restrict__ F, const double* __restrict__ U,
                                                   uint32_t k, const double* __restrict__ W, uint32_t rows_per_block,
                                                   double* __restrict__ R) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t r0 = blockIdx.y * rows_per_block, r1 = min(r0 + rows_per_block, m), nb = (n + 3) >> 2;
  double uj[KMAX];
#pragma unroll
  for (int c = 0; c < KMAX; ++c) uj[c] = c < (int)k ? U[(uint64_t)j * k + c] : 0.0;
  for (uint32_t r = r0; r < r1; ++r) {
    double dot = 0.0;
#pragma unroll
    for (int c = 0; c < KMAX; ++c)
      if (c < (int)k) dot = fma(uj[c], W[(uint64_t)r * k + c], dot);
    R[(uint64_t)r * n + j] = centred_geno(raw, nb, r, j, F[r]) - dot;
  }
}

void launch_usv_residual(const uint8_t* raw, uint32_t n, uint32_t m, const double* F, const double* U, uint32_t k,
                         double* W, double* R, hipStream_t s) {
  PCAONE_REQUIRE(k >= 1 && k <= 64, -1, "ld_usv: k must be in [1, 64]");
  hipLaunchKernelGGL(k_usv_w, dim3((uint32_t)ceil_div(m, 64)), dim3(256), 0, s, raw, n, m, F, U, k, W);
  const uint32_t rpb = 64;
  const dim3 grid((uint32_t)ceil_div(n, 256), (uint32_t)ceil_div(m, rpb));
  if (k <= 8)
    hipLaunchKernelGGL(k_usv_resid<8>, grid, dim3(256), 0, s, raw, n, m, F, U, k, W, rpb, R);
  else if (k <= 16)
    hipLaunchKernelGGL(k_usv_resid<16>, grid, dim3(256), 0, s, raw, n, m, F, U, k, W, rpb, R);
  else if (k <= 32)
    hipLaunchKernelGGL(k_usv_resid<32>, grid, dim3(256), 0, s, raw, n, m, F, U, k, W, rpb, R);
  else
    hipLaunchKernelGGL(k_usv_resid<64>, grid, dim3(256), 0, s, raw, n, m, F, U, k, W, rpb, R);
}

}  // namespace pcaone
