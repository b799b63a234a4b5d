// FP64 tall-skinny linear algebra of the RSVD operator for l in (64, 128]
// (k + oversamples above 64: --oversamples is raised to k, src/Cmd.cpp:190,
// so any k >= 33 lands here).  The same steps as kernels_qr.hip /
// kernels_linalg.hip -- CholeskyQR2 with the Householder R-diagonal signs,
// flipOmg, computeUSV without forming Q, the Jacobi eigensolve of B B^T --
// in general form for a 128-column padded width:
//
//   k_wide_gram    Gram partials of the rows (optionally of A1 + A2, written),
//                  64-row LDS stages, the 36 upper 16 x 16 tile pairs split over
//                  the block's eight waves (FP64 MFMA)
//   k_wide_sum     fixed-order sum of the partials (the full symmetric l x l)
//   k_wide_apply   Out = A S for an l x l S, 128-row blocks, S staged in 16-row
//                  K chunks; optionally flipOmg's column statistics against a
//                  previous Omega, per block
//   k_wide_chol    R^-1 of the Cholesky factor of a Gram (one block, in LDS)
//   k_wide_r2      the second factor, Rt^-1 = R1^-1 R2^-1 and the Householder
//                  signs of the top l x l block of Q1 R2^-1 (one block)
//   k_wide_write   the signed, flipped Omega into every consumer
//   k_sym_eig_wide cyclic Jacobi of the l x l B B^T (one block; the matrix in
//                  LDS, the eigenvectors in global memory)
//
// Every reduction has a fixed order, so the results are deterministic.  The
// tall data is read once per pass; the l x l steps run in one workgroup each
// (the one-wave register-grid factorisations of the l <= 64 path do not fit
// 128 columns).
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4_f64: lane (k = lane / 16, i = lane % 16) supplies A[i][k]
// and B[k][i]; accumulator element v holds D[lane / 16 + 4 v][lane % 16].
__device__ inline f64x4 mfma64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int WW = (int)kWideMaxL;  // padded width
constexpr int WNB = WW / 16;        // 16-column tiles
constexpr int WNP = WNB * (WNB + 1) / 2;
constexpr int kGramRows = 64;
constexpr int kGramLd = WW + 1;
constexpr int kGramWaves = 8;
constexpr int kPairsPerWave = (WNP + kGramWaves - 1) / kGramWaves;
constexpr int kGramPart = WNP * 256;  // doubles per Gram partial (tile layout)
constexpr int kMaxGramBlocks = 512;  // two 66 KB-LDS blocks per CU

// pair p (row by row over ca <= cb) -> tile indices
__device__ inline void pair_of(int p, int& ca, int& cb) {
  ca = 0;
  int rem = p;
  while (rem >= WNB - ca) {
    rem -= WNB - ca;
    ++ca;
  }
  cb = ca + rem;
}

__global__ __launch_bounds__(kGramWaves * 64) void k_wide_gram(const double* __restrict__ A1,
                                                              const double* __restrict__ A2, double* __restrict__ Aout,
                                                              uint64_t n, uint32_t l, double* __restrict__ part) {
  __shared__ double st[kGramRows * kGramLd];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, kr = lane >> 4, ci = lane & 15;
  int pa[kPairsPerWave], pb[kPairsPerWave];
  f64x4 acc[kPairsPerWave];
#pragma unroll
  for (int u = 0; u < kPairsPerWave; ++u) {
    const int p = wv + kGramWaves * u;
    pa[u] = pb[u] = -1;
    if (p < WNP) pair_of(p, pa[u], pb[u]);
    acc[u] = f64x4{0.0, 0.0, 0.0, 0.0};
  }
  const uint64_t stages = ceil_div(n, kGramRows);
  for (uint64_t sg = blockIdx.x; sg < stages; sg += gridDim.x) {
    __syncthreads();
    const uint64_t r0 = sg * kGramRows;
    for (int e = threadIdx.x; e < kGramRows * WW; e += blockDim.x) {
      const int rr = e / WW, c = e % WW;
      const uint64_t r = r0 + rr;
      double v = 0.0;
      if (r < n && (uint32_t)c < l) {
        v = A1[r * l + c];
        if (A2) {
          v += A2[r * l + c];
          Aout[r * l + c] = v;
        }
      }
      st[rr * kGramLd + c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPairsPerWave; ++u) {
      if (pa[u] < 0) continue;
#pragma unroll 4
      for (int ks = 0; ks < kGramRows / 4; ++ks) {
        const double* row = st + (ks * 4 + kr) * kGramLd;
        acc[u] = mfma64(row[pa[u] * 16 + ci], row[pb[u] * 16 + ci], acc[u]);
      }
    }
  }
  double* dst = part + (uint64_t)blockIdx.x * kGramPart;
#pragma unroll
  for (int u = 0; u < kPairsPerWave; ++u) {
    const int p = wv + kGramWaves * u;
    if (p >= WNP) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) dst[p * 256 + v * 64 + lane] = acc[u][v];
  }
}

// out[e] = sum over parts in part order; with `square` the tile-layout Gram
// becomes the full symmetric l x l matrix (row-major)
__global__ __launch_bounds__(256) void k_wide_sum(const double* __restrict__ part, int nparts, uint64_t stride,
                                                  int count, uint32_t l, int square, double* __restrict__ out) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (e < count)
    for (int b0 = wv * 16; b0 < nparts; b0 += 64) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = (b0 + u < nparts) ? part[(uint64_t)(b0 + u) * stride + e] : 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u) s += v[u];
    }
  red[wv][lane] = s;
  __syncthreads();
  if (wv != 0 || e >= count) return;
  const double tot = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
  if (!square) {
    out[e] = tot;
    return;
  }
  const int p = e >> 8, v = (e >> 6) & 3, ln = e & 63;
  int ca, cb;
  pair_of(p, ca, cb);
  const uint32_t r = ca * 16 + (ln >> 4) + 4 * v, c = cb * 16 + (ln & 15);
  if (r < l && c < l && r <= c) {
    out[r * l + c] = tot;
    out[c * l + r] = tot;
  }
}

// ---------------------------------------------------------------------------
constexpr int kApRows = 128;  // eight waves x one 16-row tile
constexpr int kApK = 16;
constexpr int kApBlocks = 512;  // two blocks per CU (58 KB of LDS, 128 VGPRs)

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_wide_apply(WideApply w) {
  __shared__ double sa[kApRows * (kApK + 1)];
  __shared__ double ss[kApK * (WW + 1)];
  __shared__ double red[8][3 * WW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, kr = lane >> 4, ci = lane & 15;
  const uint32_t l = w.l;
  double dm[WNB], dp[WNB], cs[WNB];
#pragma unroll
  for (int c = 0; c < WNB; ++c) dm[c] = dp[c] = cs[c] = 0.0;
  const uint64_t groups = ceil_div(w.n, kApRows);
  for (uint64_t gi = blockIdx.x; gi < groups; gi += gridDim.x) {
    const uint64_t r0 = gi * kApRows;
    f64x4 acc[WNB];
#pragma unroll
    for (int c = 0; c < WNB; ++c) acc[c] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (uint32_t k0 = 0; k0 < l; k0 += kApK) {
      __syncthreads();
      for (int e = threadIdx.x; e < kApRows * kApK; e += blockDim.x) {
        const int rr = e / kApK, kk = e % kApK;
        const uint64_t r = r0 + rr;
        const uint32_t k = k0 + kk;
        sa[rr * (kApK + 1) + kk] = (r < w.n && k < l) ? w.A[r * l + k] : 0.0;
      }
      for (int e = threadIdx.x; e < kApK * WW; e += blockDim.x) {
        const int kk = e / WW, c = e % WW;
        const uint32_t k = k0 + kk;
        ss[kk * (WW + 1) + c] = (k < l && (uint32_t)c < l) ? w.S[(uint64_t)k * l + c] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < kApK / 4; ++ks) {
        const double a = sa[(wv * 16 + ci) * (kApK + 1) + ks * 4 + kr];
#pragma unroll
        for (int ct = 0; ct < WNB; ++ct) acc[ct] = mfma64(a, ss[(ks * 4 + kr) * (WW + 1) + ct * 16 + ci], acc[ct]);
      }
    }
#pragma unroll
    for (int ct = 0; ct < WNB; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint64_t row = r0 + wv * 16 + kr + 4 * v;
        const uint32_t col = ct * 16 + ci;
        if (row >= w.n || col >= l) continue;
        const double y = acc[ct][v];
        w.out[row * l + col] = y;
        if (w.prev) {
          const double p = w.prev[row * l + col];
          dm[ct] += fabs(p - y);
          dp[ct] += fabs(p + y);
          cs[ct] += y;
        }
      }
  }
  if (!w.stat_part) return;
  // a column's rows are spread over lane / 16: fold them, then the waves
#pragma unroll
  for (int ct = 0; ct < WNB; ++ct)
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
      dm[ct] += __shfl_xor(dm[ct], off);
      dp[ct] += __shfl_xor(dp[ct], off);
      cs[ct] += __shfl_xor(cs[ct], off);
    }
  if (kr == 0)
#pragma unroll
    for (int ct = 0; ct < WNB; ++ct) {
      red[wv][ct * 16 + ci] = dm[ct];
      red[wv][WW + ct * 16 + ci] = dp[ct];
      red[wv][2 * WW + ct * 16 + ci] = cs[ct];
    }
  __syncthreads();
  for (int e = threadIdx.x; e < 3 * WW; e += blockDim.x) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += red[q][e];
    w.stat_part[(uint64_t)blockIdx.x * 3 * WW + e] = s;
  }
}

// ---------------------------------------------------------------------------
// one-block l x l steps; the matrix lives in LDS with row stride kSm
constexpr int kSm = WW + 1;
constexpr int kSmallThreads = 1024;
// the W x W array, isd (W), and two row and two column buffers of the
// register-resident eliminations (4 W)
constexpr size_t kSmallLds = (size_t)WW * kSm * sizeof(double) + 5 * WW * sizeof(double);
// The elimination steps map threads as rows x kStepCols columns (l <= WW), so
// no step divides a thread index by a run-time width.
constexpr uint32_t kStepCols = WW;
static_assert(kSmallThreads % kStepCols == 0, "whole rows of step threads");

// Upper Cholesky A = R^T R and R^-1 by elimination in one W x W array: row
// operations reduce the upper triangle to U = D L^T while the strictly lower
// triangle collects L^-1 (unit diagonal implied), so R^-1[i][j] =
// L^-1[j][i] / sqrt(d_j).  On entry the upper triangle of a holds the Gram;
// isd (LDS, l) receives 1 / sqrt(d_j).  Returns the first failing pivot
// (1-based), 0 when none.
//
// The array lives in registers during the elimination: thread (ti, tj) =
// (t / kStepCols, t % kStepCols) holds entries (ti + 8 r, tj), r < kRegRows.
// Step k publishes row k through one of two LDS row buffers (rows, 2 WW
// doubles), so a step costs one barrier, a few LDS reads and at most kRegRows
// FMAs per thread; the operations on each entry are those of the plain
// elimination.  The result goes back to a.
constexpr int kRowGroups = kSmallThreads / kStepCols;  // 8
constexpr int kRegRows = WW / kRowGroups;               // 16
__device__ int chol_inverse_lds(double* a, uint32_t l, double* isd, double* rows) {
  const uint32_t t = threadIdx.x, nt = blockDim.x, tj = t % kStepCols, ti = t / kStepCols;
  double v[kRegRows];
#pragma unroll
  for (int r = 0; r < kRegRows; ++r) {
    const uint32_t i = ti + kRowGroups * r;
    v[r] = (i < l && tj < l && tj >= i) ? a[i * kSm + tj] : 0.0;
  }
  int bad = 0;
#pragma unroll
  for (int r = 0; r < kRegRows; ++r) {
    for (uint32_t p = 0; p < (uint32_t)kRowGroups; ++p) {
      const uint32_t k = kRowGroups * r + p;
      if (k >= l) break;
      double* rowk = rows + (k & 1) * WW;
      if (ti == p) rowk[tj] = v[r];  // row k (its L^-1 part left of the diagonal)
      __syncthreads();
      double d = rowk[k];
      const bool ok = d > 0.0 && isfinite(d);
      if (!ok && bad == 0) bad = (int)k + 1;
      d = ok ? d : 1e-300;
      const double inv = 1.0 / d;
      const double akj = (tj == k) ? 1.0 : rowk[tj];
#pragma unroll
      for (int rr = 0; rr < kRegRows; ++rr) {
        const uint32_t i = ti + kRowGroups * rr;
        if (i > k && i < l && tj < l && (tj >= i || tj <= k)) v[rr] = fma(-(rowk[i] * inv), akj, v[rr]);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kRegRows; ++r) {
    const uint32_t i = ti + kRowGroups * r;
    if (i < l && tj < l) a[i * kSm + tj] = v[r];
  }
  __syncthreads();
  for (uint32_t j = t; j < l; j += nt) {
    const double d = a[j * kSm + j];
    isd[j] = 1.0 / sqrt((d > 0.0 && isfinite(d)) ? d : 1e-300);
  }
  __syncthreads();
  return bad;
}

// R^-1[i][j] from the eliminated array (i <= j, else 0)
__device__ inline double rinv_at(const double* a, const double* isd, uint32_t i, uint32_t j) {
  if (i > j) return 0.0;
  return (i == j ? 1.0 : a[j * kSm + i]) * isd[j];
}

__device__ inline void load_gram_lds(const double* __restrict__ gram, uint32_t l, double* a) {
  for (uint32_t e = threadIdx.x; e < l * l; e += blockDim.x) a[(e / l) * kSm + e % l] = gram[e];
}

__global__ __launch_bounds__(kSmallThreads) void k_wide_chol(const double* __restrict__ gram, uint32_t l,
                                                            double* __restrict__ rinv, int* status) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  double* a = reinterpret_cast<double*>(smem_raw);
  double* isd = a + WW * kSm;
  load_gram_lds(gram, l, a);
  const int bad = chol_inverse_lds(a, l, isd, isd + WW);
  for (uint32_t e = threadIdx.x; e < l * l; e += blockDim.x) rinv[e] = rinv_at(a, isd, e / l, e % l);
  if (threadIdx.x == 0 && bad && *status == 0) *status = bad;
}

// R2^-1 from the Gram of Q1, Rt^-1 = R1^-1 R2^-1; with q1 (the first l rows
// of Q1, row stride l) also the Householder signs of the top block of
// Q1 R2^-1: LU of B - S with S_kk = -sign(pivot) (Ballard et al. 2014).
__global__ __launch_bounds__(kSmallThreads) void k_wide_r2(const double* __restrict__ gram, uint32_t l,
                                                          const double* __restrict__ r1i, double* __restrict__ r2i,
                                                          double* __restrict__ rti, const double* __restrict__ q1,
                                                          double* __restrict__ sign, int* status) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  double* a = reinterpret_cast<double*>(smem_raw);
  double* isd = a + WW * kSm;
  const uint32_t t = threadIdx.x, nt = blockDim.x;
  double* rows = isd + WW;  // 2 W row buffers, then 2 W column buffers
  load_gram_lds(gram, l, a);
  const int bad = chol_inverse_lds(a, l, isd, rows);
  if (t == 0 && bad && *status == 0) *status = bad;
  for (uint32_t e = t; e < l * l; e += nt) {
    const uint32_t i = e / l, j = e % l;
    r2i[e] = rinv_at(a, isd, i, j);
    double s = 0.0;  // (R1^-1 R2^-1)[i][j], both upper triangular
    for (uint32_t k = i; k <= j; ++k) s = fma(r1i[(uint64_t)i * l + k], rinv_at(a, isd, k, j), s);
    rti[e] = s;
  }
  if (!q1) return;
  // B = Q1[0:l] R2^-1 into registers (16 entries per thread at most), then LDS
  constexpr int kPer = WW * WW / kSmallThreads;
  double b[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const uint32_t e = t + nt * u;
    b[u] = 0.0;
    if (e < l * l) {
      const uint32_t i = e / l, j = e % l;
      double s = 0.0;
      for (uint32_t k = 0; k <= j; ++k) s = fma(q1[(uint64_t)i * l + k], rinv_at(a, isd, k, j), s);
      b[u] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const uint32_t e = t + nt * u;
    if (e < l * l) a[(e / l) * kSm + e % l] = b[u];
  }
  // the LU of B - S in registers, as the Cholesky: step k publishes row k and
  // column k
  __syncthreads();
  const uint32_t tj = t % kStepCols, ti = t / kStepCols;
  double v[kRegRows];
#pragma unroll
  for (int r = 0; r < kRegRows; ++r) {
    const uint32_t i = ti + kRowGroups * r;
    v[r] = (i < l && tj < l) ? a[i * kSm + tj] : 0.0;
  }
  double* cols = rows + 2 * WW;
#pragma unroll
  for (int r = 0; r < kRegRows; ++r) {
    for (uint32_t p = 0; p < (uint32_t)kRowGroups; ++p) {
      const uint32_t k = kRowGroups * r + p;
      if (k >= l) break;
      double* rowk = rows + (k & 1) * WW;
      double* colk = cols + (k & 1) * WW;
      if (ti == p) rowk[tj] = v[r];
      if (tj == k)
#pragma unroll
        for (int rr = 0; rr < kRegRows; ++rr) colk[ti + kRowGroups * rr] = v[rr];
      __syncthreads();
      const double bkk = rowk[k];
      const double sk = (bkk >= 0.0) ? -1.0 : 1.0;
      const double ip = 1.0 / (bkk - sk);
      if (t == 0) sign[k] = sk;
      const double akj = rowk[tj];
#pragma unroll
      for (int rr = 0; rr < kRegRows; ++rr) {
        const uint32_t i = ti + kRowGroups * rr;
        if (i > k && i < l && tj > k && tj < l) v[rr] = fma(-colk[i] * ip, akj, v[rr]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Rank repair of a Gram (robust refresh): column k is (numerically) in the
// span of the columns before it when its Cholesky pivot falls below tau times
// its own squared norm (the pivot is the squared distance of the column from
// that span).  Such columns are marked and dropped from the elimination, so
// later columns are judged against the kept ones only.  scale receives the
// RMS entry of the n x l matrix, sqrt(trace / (n l)).
__global__ __launch_bounds__(kSmallThreads) void k_wide_deficient(const double* __restrict__ gram, uint32_t l,
                                                                 uint64_t n, double tau, int* __restrict__ mask,
                                                                 double* __restrict__ scale) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  double* a = reinterpret_cast<double*>(smem_raw);
  const uint32_t t = threadIdx.x, tj = t % kStepCols, ti = t / kStepCols, nr = blockDim.x / kStepCols;
  load_gram_lds(gram, l, a);
  for (uint32_t k = 0; k < l; ++k) {
    __syncthreads();
    const double d = a[k * kSm + k];
    const bool keep = d > tau * gram[(uint64_t)k * l + k] && d > 0.0 && isfinite(d);
    if (t == 0) mask[k] = keep ? 0 : 1;
    if (!keep) continue;  // uniform: every thread read the same d
    const double inv = 1.0 / d;
    for (uint32_t i = k + 1 + ti; i < l; i += nr)
      for (uint32_t j = tj; j < l; j += kStepCols)
        if (j >= i) a[i * kSm + j] = fma(-a[k * kSm + i] * inv, a[k * kSm + j], a[i * kSm + j]);
  }
  if (t == 0) {
    double tr = 0.0;
    for (uint32_t c = 0; c < l; ++c) tr += gram[(uint64_t)c * l + c];
    *scale = sqrt(fmax(tr, 1e-300) / ((double)n * (double)l));
  }
}

// out = A with every marked column replaced by a deterministic pseudo-random
// normal column of RMS entry *scale (SplitMix64 of (row, column), Box-Muller)
__device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ __launch_bounds__(256) void k_fill_deficient(const double* __restrict__ A, uint64_t n, uint32_t l,
                                                       const int* __restrict__ mask, const double* __restrict__ scale,
                                                       double* __restrict__ out) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * l) return;
  const uint32_t c = (uint32_t)(gid % l);
  if (!mask[c]) {
    out[gid] = A[gid];
    return;
  }
  const uint64_t h = mix64(gid * 0x2545F4914F6CDD1Dull + 0x5851F42D4C957F2Dull);
  const double u1 = ((double)(h >> 11) + 0.5) * 0x1.0p-53, u2 = (double)(mix64(h) >> 11) * 0x1.0p-53;
  out[gid] = *scale * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// ---------------------------------------------------------------------------
constexpr int kWrRows = 64;

__global__ __launch_bounds__(256) void k_wide_write(WideWrite q) {
  __shared__ double fs[WW];
  __shared__ double tile[kWrRows * (WW + 1)];
  const uint32_t t = threadIdx.x, l = q.l;
  // flipOmg on the signed Omega (the sign of a column swaps its two distances)
  if (t < l) {
    const double s = q.sign[t];
    const double a = (s > 0.0) ? q.stat[t] : q.stat[WW + t];
    const double b = (s > 0.0) ? q.stat[WW + t] : q.stat[t];
    const double f = (a > 2.0 * b) ? -1.0 : 1.0;
    fs[t] = f * s;
    if (blockIdx.x == 0) q.colsum[t] = f * s * q.stat[2 * WW + t];
  }
  const uint64_t groups = ceil_div(q.n, kWrRows);
  for (uint64_t gi = blockIdx.x; gi < groups; gi += gridDim.x) {
    __syncthreads();
    const uint64_t r0 = gi * kWrRows;
    for (uint32_t e = t; e < kWrRows * l; e += blockDim.x) {
      const uint32_t rr = e / l, c = e % l;
      const uint64_t row = r0 + rr;
      double x = 0.0;
      if (row < q.n) {
        const uint64_t gidx = row * l + c;
        x = q.omu[gidx] * fs[c];
        q.omega[gidx] = x;
        if (q.zero_acc) q.zero_acc[gidx] = 0.0;
        if (q.omega32) q.omega32[gidx] = (float)x;
      }
      tile[rr * (WW + 1) + c] = x;
    }
    __syncthreads();
    // the contraction operand, column-major: consecutive lanes on consecutive rows
    for (uint32_t e = t; e < kWrRows * l; e += blockDim.x) {
      const uint32_t c = e / kWrRows, rr = e % kWrRows;
      const uint64_t row = r0 + rr;
      if (row >= q.n) continue;
      const double x = tile[rr * (WW + 1) + c];
      const uint64_t kpos = q.kperm ? kperm16(row) : row;
      if (q.bprec == PREC_SPLIT_BF16) {
        const __bf16 hi = (__bf16)(float)x;
        reinterpret_cast<__bf16*>(q.bg)[(uint64_t)c * q.kpad + kpos] = hi;
        reinterpret_cast<__bf16*>(q.bg)[(uint64_t)(q.lo_off + c) * q.kpad + kpos] =
            (__bf16)(float)(x - (double)(float)hi);
      } else if (q.bprec == PREC_FP32) {
        reinterpret_cast<float*>(q.bg)[(uint64_t)c * q.kpad + kpos] = (float)x;
      } else {
        reinterpret_cast<double*>(q.bg)[(uint64_t)c * q.kpad + kpos] = x;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// sort key of an eigenvalue: NaN below every number
__device__ inline double eig_key(double x) { return x == x ? x : -__builtin_huge_val(); }

// Cyclic Jacobi with round-robin pair ordering (n/2 disjoint rotations per
// round, Numerical-Recipes angle), the symmetric matrix in LDS and the
// eigenvectors in global memory (W).  A round: every wave computes the
// round's angles from the matrix as it stands (lanes 0 .. h-1), a barrier,
// then each thread rotates its own 2 x 2 blocks of the matrix and its own row
// pairs of W in place (the blocks and row pairs of a round are disjoint), and
// a barrier.
constexpr int kEigWaves = kSmallThreads / 64;
constexpr size_t kEigLds = (size_t)WW * kSm * sizeof(double) + (size_t)kEigWaves * 64 * 2 * sizeof(double);

__device__ inline void rr_pair(int t, int round, int n, int& p, int& q) {
  auto pos = [&](int k) {
    const int x = k - 1 + round;
    return k == 0 ? 0 : 1 + (x >= n - 1 ? x - (n - 1) : x);
  };
  p = pos(t);
  q = pos(n - 1 - t);
  if (p > q) {
    const int x = p;
    p = q;
    q = x;
  }
}

__global__ __launch_bounds__(kSmallThreads) void k_sym_eig_wide(const double* __restrict__ A, uint32_t l,
                                                               double* evals, double* W, double* sinv, uint32_t kc,
                                                               const double* W0, double* tmp) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  double* a = reinterpret_cast<double*>(smem_raw);
  double* csw = a + WW * kSm;  // [wave][pair][2]
  __shared__ double red[kEigWaves][2];
  __shared__ int order[WW];
  const int L = (int)l, t = threadIdx.x, lane = t & 63, wv = t >> 6, nt = blockDim.x;
  const int n = (L % 2) ? L + 1 : L;  // dummy index L when odd: a zero row and column
  const int h = n / 2;
  for (int e = t; e < WW * kSm; e += nt) a[e] = 0.0;
  __syncthreads();
  for (int e = t; e < L * L; e += nt) a[(e / L) * kSm + e % L] = A[e];
  __syncthreads();
  if (W0) {
    // start from the previous eigenvectors: a <- W0^T a W0, W <- W0
    for (int e = t; e < L * L; e += nt) {
      const int i = e / L, j = e % L;
      double s = 0.0;
      for (int k = 0; k < L; ++k) s = fma(a[i * kSm + k], W0[k * L + j], s);
      tmp[e] = s;
    }
    __syncthreads();
    for (int e = t; e < L * L; e += nt) {
      const int i = e / L, j = e % L;
      if (i > j) continue;
      double s = 0.0;
      for (int k = 0; k < L; ++k) s = fma(W0[k * L + i], tmp[k * L + j], s);
      a[i * kSm + j] = s;
      a[j * kSm + i] = s;  // exact symmetry for the rotations
    }
    if (W0 != W)
      for (int e = t; e < L * L; e += nt) W[e] = W0[e];
  } else {
    for (int e = t; e < L * L; e += nt) W[e] = (e / L == e % L) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int e = t; e < L * L; e += nt) {
      const int i = e / L, j = e % L;
      const double x = a[i * kSm + j] * a[i * kSm + j];
      if (i == j) dg += x; else off += x;
    }
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_down(off, o);
      dg += __shfl_down(dg, o);
    }
    if (lane == 0) {
      red[wv][0] = off;
      red[wv][1] = dg;
    }
    __syncthreads();
    double offs = 0.0, dgs = 0.0;
    for (int q = 0; q < kEigWaves; ++q) {
      offs += red[q][0];
      dgs += red[q][1];
    }
    __syncthreads();  // red is rewritten by the next sweep's check
    if (offs <= 1e-26 * dgs || offs == 0.0) break;
    for (int round = 0; round < n - 1; ++round) {
      if (lane < h) {
        int p, q;
        rr_pair(lane, round, n, p, q);
        double c = 1.0, sn = 0.0;
        const double apq = a[p * kSm + q];
        if (q < L && fabs(apq) > 1e-300) {
          const double theta = (a[q * kSm + q] - a[p * kSm + p]) / (2.0 * apq);
          double tt = 1.0 / (fabs(theta) + sqrt(fma(theta, theta, 1.0)));
          if (theta < 0.0) tt = -tt;
          c = 1.0 / sqrt(fma(tt, tt, 1.0));
          sn = tt * c;
        }
        csw[(wv * 64 + lane) * 2] = c;
        csw[(wv * 64 + lane) * 2 + 1] = sn;
      }
      __syncthreads();
      const double* cw = csw + wv * 64 * 2;
      for (int e = t; e < h * h + h * L; e += nt) {
        if (e < h * h) {  // a block {p1, q1} x {p2, q2}: rows, then columns
          const int k1 = e / h, k2 = e % h;
          int p1, q1, p2, q2;
          rr_pair(k1, round, n, p1, q1);
          rr_pair(k2, round, n, p2, q2);
          const double c1 = cw[2 * k1], s1 = cw[2 * k1 + 1], c2 = cw[2 * k2], s2 = cw[2 * k2 + 1];
          const double x0 = a[p1 * kSm + p2], y0 = a[q1 * kSm + p2];
          const double x1 = a[p1 * kSm + q2], y1 = a[q1 * kSm + q2];
          const double rp0 = c1 * x0 - s1 * y0, rq0 = s1 * x0 + c1 * y0;
          const double rp1 = c1 * x1 - s1 * y1, rq1 = s1 * x1 + c1 * y1;
          a[p1 * kSm + p2] = c2 * rp0 - s2 * rp1;
          a[p1 * kSm + q2] = s2 * rp0 + c2 * rp1;
          a[q1 * kSm + p2] = c2 * rq0 - s2 * rq1;
          a[q1 * kSm + q2] = s2 * rq0 + c2 * rq1;
        } else {  // W entries (row, p), (row, q)
          const int f = e - h * h, k1 = f / L, row = f % L;
          int p1, q1;
          rr_pair(k1, round, n, p1, q1);
          if (q1 >= L) continue;  // the dummy index: identity rotation
          const double c1 = cw[2 * k1], s1 = cw[2 * k1 + 1];
          const double x = W[row * L + p1], y = W[row * L + q1];
          W[row * L + p1] = c1 * x - s1 * y;
          W[row * L + q1] = s1 * x + c1 * y;
        }
      }
      __syncthreads();
    }
  }
  // descending eigenvalues (ties by index)
  // (a total order even when a breakdown upstream left NaNs on the diagonal:
  // NaN ranks last, ties by index, so order is a permutation of 0 .. L-1;
  // every entry holds a valid index before the ranks land)
  if (t < L) order[t] = t;
  __syncthreads();
  if (t < L) {
    const double di = eig_key(a[t * kSm + t]);
    int rank = 0;
    for (int j = 0; j < L; ++j) {
      const double dj = eig_key(a[j * kSm + j]);
      rank += (dj > di) || (dj == di && j < t);
    }
    order[rank] = t;
  }
  for (int e = t; e < L * L; e += nt) tmp[e] = W[e];
  __syncthreads();
  for (int e = t; e < L * L; e += nt) W[e] = tmp[(e / L) * L + order[e % L]];
  if (t < L) {
    const double ev = a[order[t] * kSm + order[t]];
    evals[t] = ev;
    if (t < (int)kc) sinv[t] = 1.0 / sqrt(fmax(ev, 1e-300));
  }
}

// The same Jacobi with the eigenvectors in LDS as well, for l up to about
// 115: the symmetric matrix is kept as its packed upper triangle (row i holds
// columns i .. l-1) and W as l rows of stride l | 1.  Only the 2 x 2 blocks
// (k1, k2) with k1 <= k2 of a round are rotated (the others are their
// transposes); the diagonal block keeps its one off-diagonal entry.  A round:
// lanes 0 .. h-1 of wave 0 compute the angles and pairs, a barrier, every
// thread rotates its blocks and its W row pairs in place, a barrier.  The
// starting basis W0, when given, enters as W0^T A W0 by FP64 MFMA on the
// global copies of A and W0 (tmp holds A W0).
__device__ inline int pk_idx(int i, int j, int L) {  // packed upper index of (min, max)
  const int a = i < j ? i : j, b = i < j ? j : i;
  return a * L - ((a * (a - 1)) >> 1) + (b - a);
}

// C = X^T Y for l x l row-major X, Y in global memory, FP64 MFMA: every
// wave takes output tiles w, w + nwaves, ...; emit(row, col, value) per entry
template <class F>
__device__ inline void xty_global(const double* __restrict__ X, const double* __restrict__ Y, int L, F emit) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
  const int kr = lane >> 4, ci = lane & 15, nb = (L + 15) / 16;
  for (int tt = wv; tt < nb * nb; tt += nw) {
    const int ti = tt / nb, tj = tt % nb;
    const int cx = 16 * ti + ci, cy = 16 * tj + ci;
    f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < L; k0 += 4) {
      const int k = k0 + kr;
      const double x = (k < L && cx < L) ? X[k * L + cx] : 0.0;
      const double y = (k < L && cy < L) ? Y[k * L + cy] : 0.0;
      acc = mfma64(x, y, acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * ti + kr + 4 * v;
      if (r < L && cy < L) emit(r, cy, acc[v]);
    }
  }
}

constexpr int kEigLwItems = (64 * 65 / 2 + 64 * 128 + kSmallThreads - 1) / kSmallThreads;  // per thread and round

size_t eig_lw_bytes(uint32_t l) {
  const size_t L = l, lw = L | 1;
  return (L * (L + 1) / 2 + L * lw) * sizeof(double);
}

__global__ __launch_bounds__(kSmallThreads) void k_sym_eig_lw(const double* __restrict__ A, uint32_t l,
                                                             double* evals, double* W, double* sinv, uint32_t kc,
                                                             const double* W0, double* tmp) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int L = (int)l, LW = L | 1, t = threadIdx.x, lane = t & 63, wv = t >> 6, nt = blockDim.x;
  double* a = reinterpret_cast<double*>(smem_raw);  // packed upper triangle
  double* v = a + L * (L + 1) / 2;                  // W, L rows of stride LW
  __shared__ double csw[64][2];
  __shared__ int pq[64][2];
  __shared__ double red[kEigWaves][2];
  __shared__ int order[WW];
  const int n = (L % 2) ? L + 1 : L;  // dummy index L when odd: a zero row and column
  const int h = n / 2, nblk = h * (h + 1) / 2;
  // this thread's items of every round: triangle blocks (k1 <= k2), then (pair, row) of W
  int asg[kEigLwItems];
#pragma unroll
  for (int u = 0; u < kEigLwItems; ++u) {
    const int e = t + kSmallThreads * u;
    asg[u] = -1;
    if (e < nblk) {
      int k1 = 0, rem = e;
      while (rem >= h - k1) {
        rem -= h - k1;
        ++k1;
      }
      asg[u] = (1 << 16) | (k1 << 8) | (k1 + rem);
    } else if (e < nblk + h * L) {
      const int f = e - nblk;
      asg[u] = (2 << 16) | ((f / L) << 8) | (f % L);
    }
  }
  if (W0) {
    // a <- W0^T A W0 (upper triangle), v <- W0
    xty_global(A, W0, L, [&](int r, int c, double x) { tmp[r * L + c] = x; });  // A^T W0 = A W0
    __syncthreads();
    xty_global(W0, tmp, L, [&](int r, int c, double x) {
      if (r <= c) a[pk_idx(r, c, L)] = x;
    });
    for (int e = t; e < L * L; e += nt) v[(e / L) * LW + e % L] = W0[e];
  } else {
    for (int e = t; e < L * L; e += nt) {
      const int i = e / L, j = e % L;
      if (i <= j) a[pk_idx(i, j, L)] = A[e];
      v[i * LW + j] = (i == j) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int e = t; e < L * L; e += nt) {
      const int i = e / L, j = e % L;
      if (j < i) continue;
      const double y = a[pk_idx(i, j, L)], x = y * y;
      if (i == j) dg += x; else off += 2.0 * x;  // both triangles, as the full-matrix test
    }
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_down(off, o);
      dg += __shfl_down(dg, o);
    }
    if (lane == 0) {
      red[wv][0] = off;
      red[wv][1] = dg;
    }
    __syncthreads();
    double offs = 0.0, dgs = 0.0;
    for (int q = 0; q < kEigWaves; ++q) {
      offs += red[q][0];
      dgs += red[q][1];
    }
    __syncthreads();  // red is rewritten by the next sweep's check
    if (offs <= 1e-26 * dgs || offs == 0.0) break;
    for (int round = 0; round < n - 1; ++round) {
      if (t < h) {
        int p, q;
        rr_pair(t, round, n, p, q);
        double c = 1.0, sn = 0.0;
        if (q < L) {
          const double apq = a[pk_idx(p, q, L)];
          if (fabs(apq) > 1e-300) {
            const double theta = (a[pk_idx(q, q, L)] - a[pk_idx(p, p, L)]) / (2.0 * apq);
            double tt = 1.0 / (fabs(theta) + sqrt(fma(theta, theta, 1.0)));
            if (theta < 0.0) tt = -tt;
            c = 1.0 / sqrt(fma(tt, tt, 1.0));
            sn = tt * c;
          }
        }
        csw[t][0] = c;
        csw[t][1] = sn;
        pq[t][0] = p;
        pq[t][1] = q;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < kEigLwItems; ++u) {
        if (asg[u] < 0) continue;
        const int kind = asg[u] >> 16, k1 = (asg[u] >> 8) & 255, lo = asg[u] & 255;
        const int p1 = pq[k1][0], q1 = pq[k1][1];
        const double c1 = csw[k1][0], s1 = csw[k1][1];
        if (kind == 1) {
          const int k2 = lo, p2 = pq[k2][0], q2 = pq[k2][1];
          const double c2 = csw[k2][0], s2 = csw[k2][1];
          if (k1 == k2) {  // the diagonal block {p, q}: p < q
            if (q1 >= L) continue;
            const int ip = pk_idx(p1, p1, L), iq = pk_idx(q1, q1, L), ipq = pk_idx(p1, q1, L);
            const double x0 = a[ip], y0 = a[ipq], x1 = a[ipq], y1 = a[iq];
            const double rp0 = c1 * x0 - s1 * y0, rq0 = s1 * x0 + c1 * y0;
            const double rp1 = c1 * x1 - s1 * y1, rq1 = s1 * x1 + c1 * y1;
            a[ip] = c2 * rp0 - s2 * rp1;
            a[ipq] = s2 * rp0 + c2 * rp1;
            a[iq] = s2 * rq0 + c2 * rq1;
          } else {  // rows {p1, q1} x columns {p2, q2}; the dummy index holds zeros
            const bool vq1 = q1 < L, vq2 = q2 < L;
            const int i00 = pk_idx(p1, p2, L);
            const int i10 = vq1 ? pk_idx(q1, p2, L) : 0;
            const int i01 = vq2 ? pk_idx(p1, q2, L) : 0;
            const int i11 = (vq1 && vq2) ? pk_idx(q1, q2, L) : 0;
            const double x0 = a[i00], y0 = vq1 ? a[i10] : 0.0;
            const double x1 = vq2 ? a[i01] : 0.0, y1 = (vq1 && vq2) ? a[i11] : 0.0;
            const double rp0 = c1 * x0 - s1 * y0, rq0 = s1 * x0 + c1 * y0;
            const double rp1 = c1 * x1 - s1 * y1, rq1 = s1 * x1 + c1 * y1;
            a[i00] = c2 * rp0 - s2 * rp1;
            if (vq2) a[i01] = s2 * rp0 + c2 * rp1;
            if (vq1) a[i10] = c2 * rq0 - s2 * rq1;
            if (vq1 && vq2) a[i11] = s2 * rq0 + c2 * rq1;
          }
        } else {  // W entries (row, p1), (row, q1)
          if (q1 >= L) continue;
          const double x = v[lo * LW + p1], y = v[lo * LW + q1];
          v[lo * LW + p1] = c1 * x - s1 * y;
          v[lo * LW + q1] = s1 * x + c1 * y;
        }
      }
      __syncthreads();
    }
  }
  // descending eigenvalues (ties by index; NaN last): a permutation of 0 .. L-1
  if (t < L) order[t] = t;
  __syncthreads();
  if (t < L) {
    const double di = eig_key(a[pk_idx(t, t, L)]);
    int rank = 0;
    for (int j = 0; j < L; ++j) {
      const double dj = eig_key(a[pk_idx(j, j, L)]);
      rank += (dj > di) || (dj == di && j < t);
    }
    order[rank] = t;
  }
  __syncthreads();
  for (int e = t; e < L * L; e += nt) W[e] = v[(e / L) * LW + order[e % L]];
  if (t < L) {
    const double ev = a[pk_idx(order[t], order[t], L)];
    evals[t] = ev;
    if (t < (int)kc) sinv[t] = 1.0 / sqrt(fmax(ev, 1e-300));
  }
}

template <class K>
void big_lds(K kern, size_t bytes) {
  PCAONE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)bytes));
}

}  // namespace

int wide_gram_blocks(uint64_t n) {
  return (int)std::min<uint64_t>(std::max<uint64_t>(ceil_div(n, 4 * kGramRows), 1), kMaxGramBlocks);
}
uint64_t wide_part_doubles() {
  return std::max<uint64_t>((uint64_t)kMaxGramBlocks * kGramPart, (uint64_t)kApBlocks * 3 * WW);
}

void launch_wide_gram(const double* A1, const double* A2, double* Aout, uint64_t n, uint32_t l, double* part,
                      double* gram, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kWideMaxL, -1, "wide Gram: l must be in [1, 128]");
  const int nb = wide_gram_blocks(n);
  hipLaunchKernelGGL(k_wide_gram, dim3(nb), dim3(kGramWaves * 64), 0, s, A1, A2, Aout, n, l, part);
  hipLaunchKernelGGL(k_wide_sum, dim3((uint32_t)ceil_div(kGramPart, 64)), dim3(256), 0, s, part, nb,
                     (uint64_t)kGramPart, kGramPart, l, 1, gram);
}

void launch_wide_apply(const WideApply& w, double* stat, hipStream_t s) {
  PCAONE_REQUIRE(w.l >= 1 && w.l <= kWideMaxL, -1, "wide apply: l must be in [1, 128]");
  const int nb = (int)std::min<uint64_t>(std::max<uint64_t>(ceil_div(w.n, kApRows), 1), kApBlocks);
  hipLaunchKernelGGL(k_wide_apply, dim3(nb), dim3(512), 0, s, w);
  if (w.stat_part)
    hipLaunchKernelGGL(k_wide_sum, dim3((uint32_t)ceil_div(3 * WW, 64)), dim3(256), 0, s, w.stat_part, nb,
                       (uint64_t)3 * WW, 3 * WW, w.l, 0, stat);
}

void launch_wide_chol(const double* gram, uint32_t l, double* rinv, int* status, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kWideMaxL, -1, "wide Cholesky: l must be in [1, 128]");
  big_lds(k_wide_chol, kSmallLds);
  hipLaunchKernelGGL(k_wide_chol, dim3(1), dim3(kSmallThreads), kSmallLds, s, gram, l, rinv, status);
}

void launch_wide_r2(const double* gram, uint32_t l, const double* r1i, double* r2i, double* rti, const double* q1,
                    double* sign, int* status, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kWideMaxL, -1, "wide Cholesky: l must be in [1, 128]");
  big_lds(k_wide_r2, kSmallLds);
  hipLaunchKernelGGL(k_wide_r2, dim3(1), dim3(kSmallThreads), kSmallLds, s, gram, l, r1i, r2i, rti, q1, sign,
                     status);
}

void launch_wide_write(const WideWrite& q, hipStream_t s) {
  PCAONE_REQUIRE(q.l >= 1 && q.l <= kWideMaxL, -1, "wide write: l must be in [1, 128]");
  PCAONE_REQUIRE(!q.kperm || q.kpad % 16 == 0, -1, "omega operand: permuted K needs 16-code groups");
  const uint32_t nb = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(ceil_div(q.n, kWrRows), 1), 1024);
  hipLaunchKernelGGL(k_wide_write, dim3(nb), dim3(256), 0, s, q);
}

void launch_wide_repair(const double* A, const double* gram, uint64_t n, uint32_t l, int* mask, double* scale,
                        double* out, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kWideMaxL, -1, "rank repair: l must be in [1, 128]");
  big_lds(k_wide_deficient, kSmallLds);
  // a column whose part outside the earlier ones is below 1e-5 of its norm
  // (pivot below 1e-10 of its Gram diagonal) is replaced: at 1e-12 a column of
  // split-BF16 rounding noise (about 1e-6 relative) could be kept and the
  // Cholesky of the repaired matrix, eliminating in another order, still fail
  hipLaunchKernelGGL(k_wide_deficient, dim3(1), dim3(kSmallThreads), kSmallLds, s, gram, l, n, 1e-10, mask, scale);
  hipLaunchKernelGGL(k_fill_deficient, dim3((uint32_t)ceil_div(n * l, 256)), dim3(256), 0, s, A, n, l, mask, scale,
                     out);
}

__global__ void k_whiten(const double* __restrict__ W, const double* __restrict__ evals, uint32_t l, double tau,
                         double* __restrict__ T) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= l * l) return;
  const uint32_t j = e % l;
  const double ev = evals[j], top = evals[0];
  T[e] = (ev > tau * top && ev > 0.0) ? W[e] / sqrt(ev) : 0.0;
}

void launch_whiten(const double* W, const double* evals, uint32_t l, double tau, double* T, hipStream_t s) {
  hipLaunchKernelGGL(k_whiten, dim3((uint32_t)ceil_div((uint64_t)l * l, 256)), dim3(256), 0, s, W, evals, l, tau, T);
}

void launch_sym_eig_wide(const double* A, uint32_t l, double* evals, double* W, double* sinv, uint32_t kc,
                         const double* W0, double* tmp, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kWideMaxL, -1, "wide eigensolve: l must be in [1, 128]");
  // the matrix and the eigenvectors both in LDS while they fit beside the
  // kernel's static arrays (l up to about 115), else the eigenvectors in global memory
  if (eig_lw_bytes(l) + 4096 <= 160 * 1024) {
    big_lds(k_sym_eig_lw, eig_lw_bytes(l));
    hipLaunchKernelGGL(k_sym_eig_lw, dim3(1), dim3(kSmallThreads), eig_lw_bytes(l), s, A, l, evals, W, sinv, kc, W0,
                       tmp);
    return;
  }
  big_lds(k_sym_eig_wide, kEigLds);
  hipLaunchKernelGGL(k_sym_eig_wide, dim3(1), dim3(kSmallThreads), kEigLds, s, A, l, evals, W, sinv, kc, W0, tmp);
}

}  // namespace pcaone
