// FP64 tall-skinny linear algebra for the RSVD steps whose stability needs it:
// CholeskyQR2 of the N x l / M x l blocks (Gram -> Cholesky -> triangular
// inverse -> apply), the Householder R-diagonal signs that make the CholeskyQR
// Q equal to the Householder Q the reference uses for Omega, the sign
// stabiliser flipOmg (src/RSVD.hpp:81-90), the l x l symmetric eigensolve for
// the SVD of B, and the mev convergence metric (src/Utils.cpp:192-198).
//
// All reductions are two-level (per-block partials, then one reduction in a
// fixed order), so every result is deterministic run to run.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {

constexpr int kTile = 32;  // rows staged per LDS tile
constexpr int kMaxL = 64;  // largest l the single-workgroup kernels hold in LDS

int gram_blocks(uint64_t nrows) {
  const uint64_t b = ceil_div(nrows, 64);
  return (int)std::min<uint64_t>(std::max<uint64_t>(b, 1), 1024);
}

// part[e] = sum_p part[p * n + e], in place, fixed order (deterministic)
__global__ void k_reduce_parts(double* part, int nparts, uint32_t n) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int p = 0; p < nparts; ++p) s += part[(uint64_t)p * n + e];
  part[e] = s;
}

static int reduce_parts(const double* part, int nparts, uint32_t n, hipStream_t s) {
  if (nparts <= 1) return nparts;
  hipLaunchKernelGGL(k_reduce_parts, dim3((uint32_t)ceil_div(n, 256)), dim3(256), 0, s, const_cast<double*>(part),
                     nparts, n);
  return 1;
}

// acc[e] += sum_r X[r][i(e)] * Y[r][j(e)] for the entries e = t + 256 q owned
// by this thread; X and Y are LDS tiles [rows][ldx].
template <int E>
__device__ inline void tile_xty(const double* X, const double* Y, int rows, int ldx, int ncx, int ncy,
                                double (&acc)[E]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = t + 256 * q;
    if (e < ncx * ncy) {
      const int i = e / ncy, j = e % ncy;
      double s = acc[q];
      for (int r = 0; r < rows; ++r) s = fma(X[r * ldx + i], Y[r * ldx + j], s);
      acc[q] = s;
    }
  }
}

template <int E>
__device__ inline void store_acc(double* part, int ncx, int ncy, const double (&acc)[E]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = t + 256 * q;
    if (e < ncx * ncy) part[e] = acc[q];
  }
}

static int entries_per_thread(int n) {
  const int e = (n + 255) / 256;
  int p = 1;
  while (p < e) p <<= 1;
  return p;
}

// ---------------------------------------------------------------------------
// partial Gram, optionally forming A + A2 on the fly
// ---------------------------------------------------------------------------
template <int E>
__global__ __launch_bounds__(256) void k_gram_partial(const double* __restrict__ A, const double* __restrict__ A2,
                                                      double* __restrict__ Aout, uint64_t nrows, uint32_t row0,
                                                      uint32_t l, double* __restrict__ part) {
  __shared__ double tile[kTile * kMaxL];
  const uint64_t live = nrows - row0;
  const uint64_t rpb = ceil_div(live, gridDim.x);
  const uint64_t rs = row0 + blockIdx.x * rpb;
  const uint64_t re = std::min<uint64_t>(rs + rpb, nrows);
  double acc[E];
#pragma unroll
  for (int q = 0; q < E; ++q) acc[q] = 0.0;
  for (uint64_t r0 = rs; r0 < re; r0 += kTile) {
    const int rows = (int)std::min<uint64_t>(kTile, re - r0);
    __syncthreads();
    for (int it = threadIdx.x; it < rows * (int)l; it += 256) {
      const uint64_t gi = (r0 + it / l) * l + it % l;
      double v = A[gi];
      if (A2) {
        v += A2[gi];
        Aout[gi] = v;
      }
      tile[it] = v;
    }
    __syncthreads();
    tile_xty<E>(tile, tile, rows, l, l, l, acc);
  }
  // rows above row0 still need A + A2 written
  if (A2 && blockIdx.x == 0)
    for (uint32_t it = threadIdx.x; it < row0 * l; it += 256) Aout[it] = A[it] + A2[it];
  store_acc<E>(part + (uint64_t)blockIdx.x * l * l, l, l, acc);
}

void launch_gram_partial(const double* A, const double* A2, double* Aout, uint64_t nrows, uint32_t row0,
                         uint32_t l, double* part, int nblocks, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kMaxL, -1, "l must be in [1, 64]");
  PCAONE_REQUIRE(nrows > row0, -1, "Gram over an empty row range");
  const int E = entries_per_thread((int)(l * l));
  switch (E) {
#define PCAONE_E(EPT) \
  case EPT: hipLaunchKernelGGL(k_gram_partial<EPT>, dim3(nblocks), dim3(256), 0, s, A, A2, Aout, nrows, row0, l, part); break;
    PCAONE_E(1) PCAONE_E(2) PCAONE_E(4) PCAONE_E(8) PCAONE_E(16)
#undef PCAONE_E
    default: throw Error(-1, "gram: unsupported l");
  }
}

// ---------------------------------------------------------------------------
// single-workgroup Cholesky (+ Householder signs, inverse, products)
// ---------------------------------------------------------------------------
// Upper Cholesky in place (a[l][l] in LDS, R^T R = a).  Returns via *bad the
// first failing pivot (1-based) or 0.
__device__ void chol_upper_lds(double* a, int l, int* bad) {
  const int t = threadIdx.x;
  for (int k = 0; k < l; ++k) {
    if (t == 0) {
      double d = a[k * l + k];
      if (!(d > 0.0) || !isfinite(d)) {
        if (*bad == 0) *bad = k + 1;
        d = 1e-300;
      }
      a[k * l + k] = sqrt(d);
    }
    __syncthreads();
    const double dk = a[k * l + k];
    for (int j = k + 1 + t; j < l; j += 256) a[k * l + j] /= dk;
    __syncthreads();
    const int m = l - k - 1;
    for (int e = t; e < m * m; e += 256) {
      const int i = k + 1 + e / m, j = k + 1 + e % m;
      if (j >= i) a[i * l + j] -= a[k * l + i] * a[k * l + j];
    }
    __syncthreads();
  }
  for (int e = t; e < l * l; e += 256)
    if (e / l > e % l) a[e] = 0.0;
  __syncthreads();
}

// Householder QR of S (rows x l, LDS, row-major) keeping only the sign of each
// R diagonal: beta_k = -sign(alpha_k) ||x_k|| as in LAPACK dlarfg and Eigen's
// makeHouseholder (zero tail: beta_k = alpha_k).  Wave 0 forms the tail norm,
// then thread j applies the reflection to column j (blockDim >= l).
__device__ void householder_signs_lds(double* S, int rows, int l, double* red, double* sign) {
  const int t = threadIdx.x;
  for (int k = 0; k < l; ++k) {
    if (t < 64) {
      double p = 0.0;
      for (int r = k + 1 + t; r < rows; r += 64) p += S[r * l + k] * S[r * l + k];
      for (int off = 32; off > 0; off >>= 1) p += __shfl_down(p, off);
      if (t == 0) {
        const double alpha = S[k * l + k];
        double beta, tau = 0.0, inv = 0.0;
        if (p <= 2.2250738585072014e-308) {
          beta = alpha;
        } else {
          beta = sqrt(alpha * alpha + p);
          if (alpha >= 0.0) beta = -beta;
          tau = (beta - alpha) / beta;
          inv = 1.0 / (alpha - beta);
        }
        sign[k] = (beta < 0.0) ? -1.0 : 1.0;
        red[0] = tau;
        red[1] = inv;
      }
    }
    __syncthreads();
    const double tau = red[0], inv = red[1];
    if (tau != 0.0 && t > k && t < l) {
      // u = [1; x_tail * inv]: w = u^T S[:, j]; S[:, j] -= tau u w
      const int j = t;
      double w = S[k * l + j];
      for (int r = k + 1; r < rows; ++r) w += S[r * l + k] * inv * S[r * l + j];
      S[k * l + j] -= tau * w;
      for (int r = k + 1; r < rows; ++r) S[r * l + j] -= tau * (S[r * l + k] * inv) * w;
    }
    __syncthreads();
  }
}

// X = R^-1 (upper) in LDS, rows solved bottom-up, one column per thread.
__device__ void upper_inverse_lds(const double* a, double* X, int l) {
  const int t = threadIdx.x;
  for (int e = t; e < l * l; e += 256) X[e] = (e / l == e % l) ? 1.0 / a[e] : 0.0;
  __syncthreads();
  for (int i = l - 2; i >= 0; --i) {
    if (t > i && t < l) {
      const int j = t;
      double s = 0.0;
      for (int p = i + 1; p <= j; ++p) s += a[i * l + p] * X[p * l + j];
      X[i * l + j] = -s / a[i * l + i];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_chol(const double* __restrict__ part, int nparts,
                                              const double* __restrict__ A, uint32_t row0, uint32_t l,
                                              double* R, double* Rinv, const double* Rprev,
                                              const double* Rprev_inv, double* Rtot, double* Rtot_inv,
                                              double* sign, int* status) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  double* a = reinterpret_cast<double*>(smem_raw);  // l*l
  double* S = a + l * l;                             // 2l*l
  __shared__ double red[4];
  __shared__ int bad;
  const int t = threadIdx.x;
  const int L = (int)l;
  if (t == 0) bad = 0;
  for (int e = t; e < L * L; e += 256) {
    double v = 0.0;
    for (int p = 0; p < nparts; ++p) v += part[(uint64_t)p * L * L + e];
    a[e] = v;
  }
  __syncthreads();
  if (row0 > 0) {
    if (sign) {
      // a = Gram of rows >= row0; R_b = chol(a); S = [A_top; R_b]
      chol_upper_lds(a, L, &bad);
      for (int e = t; e < (int)row0 * L; e += 256) S[e] = A[e];
      for (int e = t; e < L * L; e += 256) S[row0 * L + e] = a[e];
      __syncthreads();
      householder_signs_lds(S, (int)row0 + L, L, red, sign);
      // full Gram = R_b^T R_b + A_top^T A_top
      for (int e = t; e < L * L; e += 256) {
        const int i = e / L, j = e % L;
        double v = 0.0;
        for (int p = 0; p <= min(i, j); ++p) v += a[p * L + i] * a[p * L + j];
        S[e] = v;
      }
      __syncthreads();
      for (int e = t; e < L * L; e += 256) a[e] = S[e];
      __syncthreads();
    }
    for (int e = t; e < L * L; e += 256) {
      const int i = e / L, j = e % L;
      double v = a[e];
      for (uint32_t r = 0; r < row0; ++r) v += A[(uint64_t)r * L + i] * A[(uint64_t)r * L + j];
      a[e] = v;
    }
    __syncthreads();
  }
  chol_upper_lds(a, L, &bad);
  double* X = S;  // R^-1
  upper_inverse_lds(a, X, L);
  for (int e = t; e < L * L; e += 256) {
    R[e] = a[e];
    Rinv[e] = X[e];
  }
  if (Rprev) {
    for (int e = t; e < L * L; e += 256) {
      const int i = e / L, j = e % L;
      double v = 0.0, w = 0.0;
      for (int p = 0; p < L; ++p) {
        v += a[i * L + p] * Rprev[p * L + j];
        w += Rprev_inv[i * L + p] * X[p * L + j];
      }
      Rtot[e] = v;
      Rtot_inv[e] = w;
    }
  }
  if (t == 0 && status) *status = bad;
}

void launch_chol(const double* part, int nparts, const double* A, uint32_t row0, uint32_t l, double* R,
                 double* Rinv, const double* Rprev, const double* Rprev_inv, double* Rtot, double* Rtot_inv,
                 double* sign, int* status, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kMaxL, -1, "l must be in [1, 64]");
  nparts = reduce_parts(part, nparts, l * l, s);
  const size_t shm = (size_t)3 * l * l * sizeof(double);
  hipLaunchKernelGGL(k_chol, dim3(1), dim3(256), shm, s, part, nparts, A, row0, l, R, Rinv, Rprev, Rprev_inv,
                     Rtot, Rtot_inv, sign, status);
}

// ---------------------------------------------------------------------------
// Q = A * Rinv * diag(sign), plus the partial Gram of Q
// ---------------------------------------------------------------------------
template <int E>
__global__ __launch_bounds__(256) void k_apply_rinv(const double* __restrict__ A, const double* __restrict__ Rinv,
                                                    const double* __restrict__ sign, double* __restrict__ Q,
                                                    uint64_t nrows, uint32_t l, double* __restrict__ part) {
  __shared__ double ri[kMaxL * kMaxL];
  __shared__ double at[kTile * kMaxL];
  __shared__ double qt[kTile * kMaxL];
  const int L = (int)l;
  for (int e = threadIdx.x; e < L * L; e += 256) ri[e] = Rinv[e] * (sign ? sign[e % L] : 1.0);
  const uint64_t rpb = ceil_div(nrows, gridDim.x);
  const uint64_t rs = blockIdx.x * rpb;
  const uint64_t re = std::min<uint64_t>(rs + rpb, nrows);
  double acc[E];
#pragma unroll
  for (int q = 0; q < E; ++q) acc[q] = 0.0;
  for (uint64_t r0 = rs; r0 < re; r0 += kTile) {
    const int rows = (int)std::min<uint64_t>(kTile, re - r0);
    __syncthreads();
    for (int it = threadIdx.x; it < rows * L; it += 256) at[it] = A[r0 * L + it];
    __syncthreads();
    for (int it = threadIdx.x; it < rows * L; it += 256) {
      const int r = it / L, c = it % L;
      double v = 0.0;
      for (int p = 0; p <= c; ++p) v = fma(at[r * L + p], ri[p * L + c], v);
      qt[it] = v;
      Q[r0 * L + it] = v;
    }
    __syncthreads();
    if (part) tile_xty<E>(qt, qt, rows, L, L, L, acc);
  }
  if (part) store_acc<E>(part + (uint64_t)blockIdx.x * L * L, L, L, acc);
}

void launch_apply_rinv(const double* A, const double* Rinv, const double* sign, double* Q, uint64_t nrows,
                       uint32_t l, double* gram_part, int nblocks, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kMaxL, -1, "l must be in [1, 64]");
  const int E = entries_per_thread((int)(l * l));
  switch (E) {
#define PCAONE_E(EPT)                                                                                        \
  case EPT:                                                                                                \
    hipLaunchKernelGGL(k_apply_rinv<EPT>, dim3(nblocks), dim3(256), 0, s, A, Rinv, sign, Q, nrows, l, gram_part); \
    break;
    PCAONE_E(1) PCAONE_E(2) PCAONE_E(4) PCAONE_E(8) PCAONE_E(16)
#undef PCAONE_E
    default: throw Error(-1, "apply: unsupported l");
  }
}

// ---------------------------------------------------------------------------
// flipOmg
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_flip_stats(const double* __restrict__ Q, const double* __restrict__ omg2,
                                                    uint64_t nrows, uint32_t l, double* __restrict__ part) {
  __shared__ double sh[3][256];
  const uint64_t rpb = ceil_div(nrows, gridDim.x);
  const uint64_t rs = blockIdx.x * rpb, re = std::min<uint64_t>(rs + rpb, nrows);
  const uint32_t t = threadIdx.x, R = 256 / l, c = t % l, rr = t / l;
  double dm = 0.0, dp = 0.0, cs = 0.0;
  if (rr < R)
    for (uint64_t r = rs + rr; r < re; r += R) {
      const double q = Q[r * l + c], o = omg2[r * l + c];
      dm += fabs(o - q);
      dp += fabs(o + q);
      cs += q;
    }
  sh[0][t] = (rr < R) ? dm : 0.0;
  sh[1][t] = (rr < R) ? dp : 0.0;
  sh[2][t] = (rr < R) ? cs : 0.0;
  __syncthreads();
  if (t < l) {
    double a = 0, b = 0, d = 0;
    for (uint32_t q = 0; q < R; ++q) {
      a += sh[0][q * l + t];
      b += sh[1][q * l + t];
      d += sh[2][q * l + t];
    }
    double* p = part + (uint64_t)blockIdx.x * 3 * l;
    p[t] = a;
    p[l + t] = b;
    p[2 * l + t] = d;
  }
}

__global__ __launch_bounds__(256) void k_flip_apply(double* __restrict__ Q, double* __restrict__ omg2,
                                                    uint64_t nrows, uint32_t l, const double* __restrict__ part,
                                                    int nparts, double* __restrict__ colsum,
                                                    double* __restrict__ zero_acc) {
  __shared__ double flip[kMaxL];
  const uint32_t t = threadIdx.x;
  if (t < l) {
    double a = 0, b = 0, d = 0;
    for (int p = 0; p < nparts; ++p) {
      a += part[(uint64_t)p * 3 * l + t];
      b += part[(uint64_t)p * 3 * l + l + t];
      d += part[(uint64_t)p * 3 * l + 2 * l + t];
    }
    const double f = (a > 2.0 * b) ? -1.0 : 1.0;
    flip[t] = f;
    if (blockIdx.x == 0) colsum[t] = f * d;
  }
  __syncthreads();
  const uint64_t n = nrows * l;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + t; i < n; i += (uint64_t)gridDim.x * 256) {
    const double q = Q[i] * flip[i % l];
    Q[i] = q;
    omg2[i] = q;
    if (zero_acc) zero_acc[i] = 0.0;
  }
}

void launch_flip_stats(const double* Q, const double* omg2, uint64_t nrows, uint32_t l, double* part,
                       int nblocks, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kMaxL, -1, "l must be in [1, 64]");
  hipLaunchKernelGGL(k_flip_stats, dim3(nblocks), dim3(256), 0, s, Q, omg2, nrows, l, part);
}

void launch_flip_apply(double* Q, double* omg2, uint64_t nrows, uint32_t l, const double* part, int nblocks,
                       double* colsum, double* zero_acc, hipStream_t s) {
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(ceil_div(nrows * l, 256), 1024);
  hipLaunchKernelGGL(k_flip_apply, dim3(blocks), dim3(256), 0, s, Q, omg2, nrows, l, part, nblocks, colsum,
                     zero_acc);
}

// ---------------------------------------------------------------------------
// small dense helpers
// ---------------------------------------------------------------------------
__global__ void k_rows_times_small(const double* __restrict__ A, uint64_t n, uint32_t l,
                                   const double* __restrict__ W, uint32_t ldw, uint32_t kc,
                                   const double* __restrict__ scale, double* __restrict__ C) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * kc) return;
  const uint64_t r = gid / kc;
  const uint32_t j = (uint32_t)(gid % kc);
  double v = 0.0;
  for (uint32_t a = 0; a < l; ++a) v = fma(A[r * l + a], W[(uint64_t)a * ldw + j], v);
  C[gid] = scale ? v * scale[j] : v;
}

void launch_rows_times_small(const double* A, uint64_t n, uint32_t l, const double* W, uint32_t ldw,
                             uint32_t kc, const double* scale, double* C, hipStream_t s) {
  const uint64_t total = n * kc;
  if (!total) return;
  hipLaunchKernelGGL(k_rows_times_small, dim3((uint32_t)ceil_div(total, 256)), dim3(256), 0, s, A, n, l, W, ldw,
                     kc, scale, C);
}

// Cyclic Jacobi with round-robin pair ordering: every round applies n/2
// disjoint rotations at once (rows, then columns), Numerical-Recipes angle.
__global__ __launch_bounds__(256) void k_sym_eig(const double* __restrict__ part, int nparts, uint32_t l,
                                                 double* __restrict__ evals, double* __restrict__ W,
                                                 double* __restrict__ sinv, uint32_t kc) {
  __shared__ double a[kMaxL * kMaxL];
  __shared__ double v[kMaxL * kMaxL];
  __shared__ double cs[kMaxL / 2 + 1][2];
  __shared__ int pr[kMaxL / 2 + 1][2];
  __shared__ int pos[kMaxL + 1];
  __shared__ double red[256];
  __shared__ int order[kMaxL];
  const int t = threadIdx.x;
  const int L = (int)l;
  const int n = (L % 2) ? L + 1 : L;  // dummy index L when odd
  for (int e = t; e < L * L; e += 256) {
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += part[(uint64_t)p * L * L + e];
    a[e] = s;
    v[e] = (e / L == e % L) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 40; ++sweep) {
    // convergence check
    double off = 0.0, dg = 0.0;
    for (int e = t; e < L * L; e += 256) {
      const double x = a[e] * a[e];
      if (e / L == e % L) dg += x; else off += x;
    }
    red[t] = off;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (t < o) red[t] += red[t + o];
      __syncthreads();
    }
    const double offs = red[0];
    __syncthreads();
    red[t] = dg;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (t < o) red[t] += red[t + o];
      __syncthreads();
    }
    const double dgs = red[0];
    __syncthreads();
    // off-diagonal Frobenius norm below 1e-13 of the diagonal's: FP64 rounding level
    if (offs <= 1e-26 * dgs || offs == 0.0) break;
    for (int round = 0; round < n - 1; ++round) {
      if (t == 0) {
        pos[0] = 0;
        for (int k = 1; k < n; ++k) pos[k] = 1 + (k - 1 + round) % (n - 1);
      }
      __syncthreads();
      if (t < n / 2) {
        int p = pos[t], q = pos[n - 1 - t];
        if (p > q) { const int tmp = p; p = q; q = tmp; }
        double c = 1.0, s = 0.0;
        if (q < L) {
          const double apq = a[p * L + q];
          if (fabs(apq) > 1e-300) {
            const double theta = (a[q * L + q] - a[p * L + p]) / (2.0 * apq);
            double tt = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
            if (theta < 0.0) tt = -tt;
            c = 1.0 / sqrt(tt * tt + 1.0);
            s = tt * c;
          }
        }
        pr[t][0] = p;
        pr[t][1] = q;
        cs[t][0] = c;
        cs[t][1] = s;
      }
      __syncthreads();
      for (int e = t; e < (n / 2) * L; e += 256) {  // rows
        const int k = e / L, j = e % L;
        const int p = pr[k][0], q = pr[k][1];
        if (q < L) {
          const double c = cs[k][0], s = cs[k][1];
          const double x = a[p * L + j], y = a[q * L + j];
          a[p * L + j] = c * x - s * y;
          a[q * L + j] = s * x + c * y;
        }
      }
      __syncthreads();
      for (int e = t; e < (n / 2) * L; e += 256) {  // columns of a and v
        const int k = e / L, i = e % L;
        const int p = pr[k][0], q = pr[k][1];
        if (q < L) {
          const double c = cs[k][0], s = cs[k][1];
          double x = a[i * L + p], y = a[i * L + q];
          a[i * L + p] = c * x - s * y;
          a[i * L + q] = s * x + c * y;
          x = v[i * L + p];
          y = v[i * L + q];
          v[i * L + p] = c * x - s * y;
          v[i * L + q] = s * x + c * y;
        }
      }
      __syncthreads();
    }
  }
  if (t == 0) {  // descending order of the eigenvalues
    uint64_t used = 0;
    for (int i = 0; i < L; ++i) {
      int best = -1;
      double bv = 0.0;
      for (int j = 0; j < L; ++j)
        if (!((used >> j) & 1ull) && (best < 0 || a[j * L + j] > bv)) {
          best = j;
          bv = a[j * L + j];
        }
      used |= 1ull << best;
      order[i] = best;
    }
  }
  __syncthreads();
  for (int e = t; e < L * L; e += 256) {
    const int i = e / L, j = e % L;
    W[e] = v[i * L + order[j]];
  }
  if (t < L) {
    const double ev = a[order[t] * L + order[t]];
    evals[t] = ev;
    if (t < (int)kc) sinv[t] = 1.0 / sqrt(fmax(ev, 1e-300));
  }
}

void launch_sym_eig(const double* part, int nparts, uint32_t l, double* evals, double* W, double* sinv,
                    uint32_t kc, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kMaxL, -1, "l must be in [1, 64]");
  nparts = reduce_parts(part, nparts, l * l, s);
  hipLaunchKernelGGL(k_sym_eig, dim3(1), dim3(256), 0, s, part, nparts, l, evals, W, sinv, kc);
}

// ---------------------------------------------------------------------------
// mev: M = X^T Y (k x k), value = mean_i ||M[:, i]||
// ---------------------------------------------------------------------------
template <int E>
__global__ __launch_bounds__(256) void k_xty_partial(const double* __restrict__ X, const double* __restrict__ Y,
                                                     uint64_t n, uint32_t k, double* __restrict__ part) {
  __shared__ double xt[kTile * kMaxL];
  __shared__ double yt[kTile * kMaxL];
  const int K = (int)k;
  const uint64_t rpb = ceil_div(n, gridDim.x);
  const uint64_t rs = blockIdx.x * rpb, re = std::min<uint64_t>(rs + rpb, n);
  double acc[E];
#pragma unroll
  for (int q = 0; q < E; ++q) acc[q] = 0.0;
  for (uint64_t r0 = rs; r0 < re; r0 += kTile) {
    const int rows = (int)std::min<uint64_t>(kTile, re - r0);
    __syncthreads();
    for (int it = threadIdx.x; it < rows * K; it += 256) {
      xt[it] = X[r0 * K + it];
      yt[it] = Y[r0 * K + it];
    }
    __syncthreads();
    tile_xty<E>(xt, yt, rows, K, K, K, acc);
  }
  store_acc<E>(part + (uint64_t)blockIdx.x * K * K, K, K, acc);
}

void launch_xty_partial(const double* X, const double* Y, uint64_t n, uint32_t k, double* part, int nblocks,
                        hipStream_t s) {
  PCAONE_REQUIRE(k >= 1 && k <= kMaxL, -1, "k must be in [1, 64]");
  const int E = entries_per_thread((int)(k * k));
  switch (E) {
#define PCAONE_E(EPT) \
  case EPT: hipLaunchKernelGGL(k_xty_partial<EPT>, dim3(nblocks), dim3(256), 0, s, X, Y, n, k, part); break;
    PCAONE_E(1) PCAONE_E(2) PCAONE_E(4) PCAONE_E(8) PCAONE_E(16)
#undef PCAONE_E
    default: throw Error(-1, "xty: unsupported k");
  }
}

__global__ __launch_bounds__(256) void k_mev_finish(const double* __restrict__ part, int nparts, uint32_t k,
                                                    double* __restrict__ out) {
  __shared__ double m[kMaxL * kMaxL];
  __shared__ double nrm[kMaxL];
  const int K = (int)k, t = threadIdx.x;
  for (int e = t; e < K * K; e += 256) {
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += part[(uint64_t)p * K * K + e];
    m[e] = s;
  }
  __syncthreads();
  if (t < K) {
    double s = 0.0;
    for (int a = 0; a < K; ++a) s += m[a * K + t] * m[a * K + t];
    nrm[t] = sqrt(s);
  }
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int i = 0; i < K; ++i) s += nrm[i];
    out[0] = s / K;
  }
}

void launch_mev_finish(const double* part, int nparts, uint32_t k, double* out, hipStream_t s) {
  nparts = reduce_parts(part, nparts, k * k, s);
  hipLaunchKernelGGL(k_mev_finish, dim3(1), dim3(256), 0, s, part, nparts, k, out);
}

}  // namespace pcaone
