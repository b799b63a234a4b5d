// LD r^2 over the windows of divide_pos_by_window on a dense feature-major
// FP32 matrix (the .residuals input, src/LD.cpp:419-442 on
// src/FileBinary.cpp:21-30's centred FP64 matrix).  The band of the Gram
// X X^T that the windows touch is computed in 64 x 64 tiles on the FP64
// MFMA (products of centred FP32 inputs are exact in FP64), then
// r = <x_i, x_k> (sds_i sds_k df) and r^2 land in the window-ordered output.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// sds[i] = 1 / sqrt(sum_k (x_ik - mu_i)^2 / (N - 1))   (calc_sds, src/LD.cpp:20-23)
__global__ __launch_bounds__(256) void k_ld_sds(const float* __restrict__ X, uint32_t n, uint32_t m,
                                                const double* __restrict__ mu, double* __restrict__ sds) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (r >= m) return;
  const float* p = X + r * n;
  const double u = mu[r];
  double s = 0.0;
  for (uint32_t j = lane; j < n; j += 64) {
    const double v = (double)p[j] - u;
    s += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) sds[r] = 1.0 / sqrt(s * (1.0 / (double)(n - 1)));
}

void launch_ld_sds(const float* X, uint32_t n, uint32_t m, const double* mu, double* sds, hipStream_t s) {
  hipLaunchKernelGGL(k_ld_sds, dim3((uint32_t)ceil_div((uint64_t)m * 64, 256)), dim3(256), 0, s, X, n, m, mu, sds);
}

constexpr int kLdK = 32;         // samples per LDS stage
constexpr int kLdLd = kLdK + 1;  // padded row stride (doubles)

// One 64 x 64 tile (row block I, column block J) per workgroup; wave (wr, wc)
// owns the 32 x 32 quadrant, four 16 x 16 MFMA tiles.  Lane (k = lane / 16,
// i = lane % 16) feeds A[i][k] = x_row(i)[k] and B[k][j] = x_col(j)[k].
__global__ __launch_bounds__(256) void k_ld_tile(LdTiles t) {
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
      xa[rr * kLdLd + kk] = (k < t.n && ra < t.m) ? (double)t.X[(uint64_t)ra * t.n + k] - t.mu[ra] : 0.0;
      xb[rr * kLdLd + kk] = (k < t.n && rb < t.m) ? (double)t.X[(uint64_t)rb * t.n + k] - t.mu[rb] : 0.0;
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
  hipLaunchKernelGGL(k_ld_tile, dim3(ntiles), dim3(256), 0, s, t);
}

}  // namespace pcaone
