// Small jobs that need nothing from the launch they ride on: they run as the
// first blocks of the next missing-entry gather (k_csr_gather), or on their own
// (k_side_jobs) when no gather follows.  256-thread blocks, in a fixed
// summation order either way.
//
//   signs   flipOmg's decision from the flip pass's block statistics
//           (src/Halko.cpp:213-216, src/RSVD.hpp:81-90): each column's sign
//           is the flip on the signed Omega (the Householder sign of a column
//           swaps its two distances) times that sign; colsum is the unsigned
//           column sum.  One block.
//   colsum  the window sums S2[w][c] of the finalize chunks' partials
//           s2p[w][chunk][c].  One block per window and 64 columns.
#pragma once

#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {

constexpr int kSideThreads = 256;
constexpr int kSideWaves = kSideThreads / 64;
constexpr int kSideStat = 3 * 64;  // statistics per partial at most (3 W, W <= 64)
// LDS doubles a side block needs
constexpr int kSideLds = kSideWaves * kSideStat;

// wave w sums partials w, w + 4, ... (four loads in flight per lane), the waves'
// sums add in wave order
__device__ inline void side_signs(const SideJobs& j, double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t t = threadIdx.x, W = j.W, stat = 3 * W;
  double acc[3] = {0.0, 0.0, 0.0};
  for (uint32_t b0 = wv; b0 < j.nparts; b0 += 4 * kSideWaves) {
    double v[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const uint32_t b = b0 + u * kSideWaves, e = lane + 64 * q;
        v[u][q] = (b < j.nparts && e < stat) ? j.stats[(uint64_t)b * stat + e] : 0.0;
      }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < 3; ++q) acc[q] += v[u][q];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (lane + 64 * q < (int)stat) red[wv * kSideStat + lane + 64 * q] = acc[q];
  __syncthreads();
  if (t < j.l) {
    double tot[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint32_t e = k * W + t;
      tot[k] = (red[e] + red[kSideStat + e]) + (red[2 * kSideStat + e] + red[3 * kSideStat + e]);
    }
    const double hs = j.sign[t];
    const double a = (hs > 0.0) ? tot[0] : tot[1];
    const double b = (hs > 0.0) ? tot[1] : tot[0];
    j.fsign[t] = ((a > 2.0 * b) ? -1.0 : 1.0) * hs;
    j.colsum[t] = tot[2];
  }
}

// window w's columns [64 cg, 64 cg + 64): wave w sums chunks w, w + 4, ...
// (eight loads in flight per lane), the waves' sums add in wave order
__device__ inline void side_colsum(const SideJobs& j, uint32_t w, uint32_t cg, double* red) {
  constexpr int kPer = 8;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t c = cg * 64 + lane;
  double tot = 0.0;
  if (c < j.l2)
    for (uint32_t q0 = wv; q0 < j.nchunk; q0 += kSideWaves * kPer) {
      double v[kPer];
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        const uint32_t q = q0 + kSideWaves * i;
        v[i] = q < j.nchunk ? j.s2p[((uint64_t)w * j.nchunk + q) * j.l2 + c] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < kPer; ++i) tot += v[i];
    }
  red[wv * 64 + lane] = tot;
  __syncthreads();
  if (threadIdx.x < 64 && c < j.l2)
    j.S2[(uint64_t)w * j.l2 + c] = (red[lane] + red[64 + lane]) + (red[128 + lane] + red[192 + lane]);
}

// blocks the jobs take
__host__ __device__ inline uint32_t side_blocks(const SideJobs& j) {
  return (j.stats ? 1u : 0u) + (j.s2p ? (j.w1 - j.w0) * ((j.l2 + 63) / 64) : 0u);
}

// side block b of the jobs (0 <= b < side_blocks(j)); red: kSideLds doubles of LDS
__device__ inline void side_block(const SideJobs& j, uint32_t b, double* red) {
  if (j.stats) {
    if (b == 0) {
      side_signs(j, red);
      return;
    }
    --b;
  }
  const uint32_t ncg = (j.l2 + 63) / 64;
  side_colsum(j, j.w0 + b / ncg, b % ncg, red);
}

}  // namespace pcaone
