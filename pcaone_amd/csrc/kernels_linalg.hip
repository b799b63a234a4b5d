// FP64 small dense linear algebra of the RSVD steps: the l x l symmetric
// eigensolve for the SVD of B, the mev convergence metric
// (src/Utils.cpp:192-198) and row-block products.  The CholeskyQR2 passes
// (Omega update and computeUSV) are in kernels_qr.hip.
//
// All reductions are two-level (per-block partials, then one reduction in a
// fixed order), so every result is deterministic run to run.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {

constexpr int kMaxL = 64;  // largest l the single-workgroup kernels hold in LDS

typedef double f64x4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4_f64: lane (k = lane / 16, i = lane % 16) supplies A[i][k]
// and B[k][i]; accumulator element v of a lane holds D[mfma64_row(lane, v)][lane % 16].
__device__ inline int mfma64_row(int lane, int v) { return (lane >> 4) + 4 * v; }
__device__ inline f64x4 mfma64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

int gram_blocks(uint64_t nrows) {
  const uint64_t b = ceil_div(nrows, 64);
  return (int)std::min<uint64_t>(std::max<uint64_t>(b, 1), 256);
}

// part[e] = sum_p part[p * n + e] (e < n), in place.  Block b owns entries
// [64 b, 64 b + 64); wave w sums parts w, w + 4, ... in four interleaved
// chains, then the waves combine in a fixed order, so the result is
// deterministic.
__global__ __launch_bounds__(256) void k_reduce_parts(double* part, int nparts, uint32_t n) {
  __shared__ double sh[4][64];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t e = blockIdx.x * 64 + lane;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  if (e < n) {
    int p = (int)wv;
    for (; p + 12 < nparts; p += 16) {
      s0 += part[(uint64_t)p * n + e];
      s1 += part[(uint64_t)(p + 4) * n + e];
      s2 += part[(uint64_t)(p + 8) * n + e];
      s3 += part[(uint64_t)(p + 12) * n + e];
    }
    for (; p < nparts; p += 4) s0 += part[(uint64_t)p * n + e];
  }
  sh[wv][lane] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (wv == 0 && e < n) part[e] = (sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]);
}

static int reduce_parts(const double* part, int nparts, uint32_t n, hipStream_t s) {
  if (nparts <= 1) return nparts;
  hipLaunchKernelGGL(k_reduce_parts, dim3((uint32_t)ceil_div(n, 64)), dim3(256), 0, s, const_cast<double*>(part),
                     nparts, n);
  return 1;
}

// The NB x NB tile accumulators of a wave, summed over the block's four waves
// in a fixed order, written as the block's full l x l partial.  Tile (ca, cb)
// is skipped when `upper` and ca > cb (then mirrored from (cb, ca)).
template <int NB, int NP>
__device__ inline void block_tiles_out(const f64x4 (&acc)[NP], bool upper, uint32_t l, double* red, double* out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int w = 0; w < 4; ++w) {
    if (wv == w)
      for (int p = 0; p < NP; ++p)
        for (int v = 0; v < 4; ++v) {
          const int i = p * 256 + v * 64 + lane;
          red[i] = (w == 0) ? acc[p][v] : red[i] + acc[p][v];
        }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < NP * 256; e += 256) {
    const int p = e >> 8, v = (e >> 6) & 3, ln = e & 63;
    int ca, cb;
    if (upper) {  // p enumerates ca <= cb row by row
      ca = 0;
      int rem = p;
      while (rem >= NB - ca) {
        rem -= NB - ca;
        ++ca;
      }
      cb = ca + rem;
    } else {
      ca = p / NB;
      cb = p % NB;
    }
    const uint32_t r = ca * 16 + mfma64_row(ln, v), c = cb * 16 + (ln & 15);
    if (upper && ca == cb && r > c) continue;  // the diagonal tile's upper half writes both
    if (r < l && c < l) {
      out[(uint64_t)r * l + c] = red[e];
      if (upper) out[(uint64_t)c * l + r] = red[e];
    }
  }
}

// ---------------------------------------------------------------------------
// small dense helpers
// ---------------------------------------------------------------------------
// C (n x kc, row-major) = A (n x l, row-major) W[:, :kc] (W row stride ldw),
// times scale[j] when given.  A thread takes four consecutive outputs of one
// row: the row of A is read once for the four, and each output sums its l
// terms in order.
__global__ void k_rows_times_small(const double* __restrict__ A, uint64_t n, uint32_t l,
                                   const double* __restrict__ W, uint32_t ldw, uint32_t kc,
                                   const double* __restrict__ scale, double* __restrict__ C) {
  const uint32_t q = (kc + 3) / 4;  // column quads per row
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * q) return;
  const uint64_t r = gid / q;
  const uint32_t j0 = (uint32_t)(gid % q) * 4;
  const uint32_t nj = min(4u, kc - j0);
  const double* __restrict__ ar = A + r * l;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  for (uint32_t a = 0; a < l; ++a) {
    const double x = ar[a];
    const double* __restrict__ wr = W + (uint64_t)a * ldw + j0;
    v0 = fma(x, wr[0], v0);
    if (nj > 1) v1 = fma(x, wr[1], v1);
    if (nj > 2) v2 = fma(x, wr[2], v2);
    if (nj > 3) v3 = fma(x, wr[3], v3);
  }
  double* cr = C + r * kc + j0;
  cr[0] = scale ? v0 * scale[j0] : v0;
  if (nj > 1) cr[1] = scale ? v1 * scale[j0 + 1] : v1;
  if (nj > 2) cr[2] = scale ? v2 * scale[j0 + 2] : v2;
  if (nj > 3) cr[3] = scale ? v3 * scale[j0 + 3] : v3;
}

void launch_rows_times_small(const double* A, uint64_t n, uint32_t l, const double* W, uint32_t ldw,
                             uint32_t kc, const double* scale, double* C, hipStream_t s) {
  const uint64_t total = n * ((kc + 3) / 4);
  if (!total) return;
  hipLaunchKernelGGL(k_rows_times_small, dim3((uint32_t)ceil_div(total, 256)), dim3(256), 0, s, A, n, l, W, ldw,
                     kc, scale, C);
}

// out (column-major, n x kc) <- A (row-major, n x kc) with row r moved to
// perm[r] (identity when perm is null): one thread per element, reads coalesced.
__global__ void k_rows_to_cols(const double* __restrict__ A, uint64_t n, uint32_t kc,
                               const int32_t* __restrict__ perm, double* __restrict__ out) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * kc) return;
  const uint64_t r = gid / kc;
  const uint32_t c = (uint32_t)(gid % kc);
  out[(uint64_t)c * n + (perm ? (uint64_t)perm[r] : r)] = A[gid];
}

void launch_rows_to_cols(const double* A, uint64_t n, uint32_t kc, const int32_t* perm, double* out,
                         hipStream_t s) {
  const uint64_t total = n * kc;
  if (!total) return;
  hipLaunchKernelGGL(k_rows_to_cols, dim3((uint32_t)ceil_div(total, 256)), dim3(256), 0, s, A, n, kc, perm, out);
}

// sort key of an eigenvalue: NaN below every number
__device__ inline double eig_key(double x) { return x == x ? x : -__builtin_huge_val(); }

// Cyclic Jacobi with round-robin pair ordering: each round applies n/2
// disjoint rotations at once (Numerical-Recipes angle).  A thread owns 2 x 2
// blocks {p1, q1} x {p2, q2} of the matrix (rows, then columns rotated from
// one buffer into the other) or row pairs of the eigenvector matrix (in
// place).  Every wave computes the round's angles itself, so a round needs
// one barrier.  LDS: a, v and tmp, W x W each (W = 16 NB, zero padded past l, so an odd
// l's dummy index l is an empty row and column and its rotation the identity)
// with the odd row stride W + 1, so a wave's walk down a column (the
// eigenvector items) or along a row touches distinct banks.
__device__ inline void rr_pair(int t, int round, int n, int& p, int& q) {
  // position k of the round-robin table: 0 stays, the rest rotate; k - 1 and
  // round are both below n - 1, so the wrap is one conditional subtraction
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

// 1 / x and 1 / sqrt(x) for a normal x > 0: the hardware estimates (about 23
// bits) and two Newton steps each, to FP64 rounding.
__device__ inline double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ inline double rsqrt_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  r = r * fma(-hx * r, r, 1.5);
  return r * fma(-hx * r, r, 1.5);
}

// C = X^T Y for W x W LDS matrices (W = 16 NB, row stride LD) on FP64 MFMA:
// wave w takes the output tiles w, w + 8, ...
template <int NB, int LD>
__device__ inline void xty_lds(const double* X, const double* Y, double* C) {
  constexpr int W = 16 * NB;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kr = lane >> 4, ci = lane & 15;
  for (int tt = wv; tt < NB * NB; tt += (int)(blockDim.x >> 6)) {
    const int ti = tt / NB, tj = tt % NB;
    f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < W / 4; ++ks)
      acc = mfma64(X[(4 * ks + kr) * LD + 16 * ti + ci], Y[(4 * ks + kr) * LD + 16 * tj + ci], acc);
#pragma unroll
    for (int v = 0; v < 4; ++v) C[(16 * ti + kr + 4 * v) * LD + 16 * tj + ci] = acc[v];
  }
}

constexpr int kEigThreads = 512;

// W0 (optional): a starting basis (the previous eigenvectors); the sweeps then
// run on W0^T A W0 and the rotations accumulate onto W0.  W may alias W0.
template <int NB>
__global__ __launch_bounds__(kEigThreads) void k_sym_eig(const double* __restrict__ part, int nparts, uint32_t l,
                                                         double* evals, double* W, double* sinv, uint32_t kc,
                                                         const double* W0) {
  constexpr int WD = 16 * NB, LD = WD + 1;
  // items per thread and round: h^2 blocks of a plus h l row pairs of v, h <= 8 NB
  constexpr int U = (192 * NB * NB + kEigThreads - 1) / kEigThreads;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int L = (int)l, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double* a = reinterpret_cast<double*>(smem_raw);
  double* v = a + WD * LD;
  double* tmp = v + WD * LD;
  __shared__ __attribute__((aligned(16))) double csw[kEigThreads / 64][8 * NB][2];  // per wave: (c, s) of each pair
  __shared__ unsigned char rrp[16 * NB][8 * NB][2];  // [round][pair] -> (p, q), p < q
  __shared__ double red[kEigThreads / 64][2];
  __shared__ int order[WD];
  const int n = (L % 2) ? L + 1 : L;  // dummy index L when odd (a zero row of the padding)
  const int h = n / 2;
  // this thread's fixed items: rotation blocks (k1, k2) of a, then (pair,
  // row) items of v, packed as kind << 16 | hi << 8 | lo (kind 0: none)
  int asg[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int e = t + kEigThreads * u;
    if (e < h * h)
      asg[u] = (1 << 16) | ((e / h) << 8) | (e % h);
    else if (e < h * h + h * L)
      asg[u] = (2 << 16) | (((e - h * h) / L) << 8) | ((e - h * h) % L);
    else
      asg[u] = 0;
  }
  for (int e = t; e < WD * WD; e += kEigThreads) {
    const int i = e / WD, j = e % WD;
    const bool in = i < L && j < L;
    double s = 0.0;
    if (in)
      for (int p = 0; p < nparts; ++p) s += part[(uint64_t)p * L * L + i * L + j];
    a[i * LD + j] = s;
    v[i * LD + j] = in ? (W0 ? W0[i * L + j] : (i == j ? 1.0 : 0.0)) : 0.0;
  }
  // the round-robin pairs of every round, looked up instead of recomputed per item
  for (int e = t; e < (n - 1) * h; e += kEigThreads) {
    int p, q;
    rr_pair(e % h, e / h, n, p, q);
    rrp[e / h][e % h][0] = (unsigned char)p;
    rrp[e / h][e % h][1] = (unsigned char)q;
  }
  __syncthreads();
  if (W0) {
    xty_lds<NB, LD>(a, v, tmp);  // a W0 (a is symmetric)
    __syncthreads();
    xty_lds<NB, LD>(v, tmp, a);  // W0^T a W0
    __syncthreads();
    for (int e = t; e < WD * WD; e += kEigThreads) {  // exact symmetry for the rotations
      const int i = e / WD, j = e % WD;
      if (i > j) a[i * LD + j] = a[j * LD + i];
    }
    __syncthreads();
  }
  // the rounds write the next matrix into tmp (an) and swap; what they never
  // write there (past the n x n block) stays zero, as in a
  double* an = tmp;
  for (int e = t; e < WD * LD; e += kEigThreads) an[e] = 0.0;
  __syncthreads();
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int e = t; e < WD * WD; e += kEigThreads) {
      const int i = e / WD, j = e % WD;
      const double x = a[i * LD + j] * a[i * LD + j];
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
#pragma unroll
    for (int w = 0; w < kEigThreads / 64; ++w) {
      offs += red[w][0];
      dgs += red[w][1];
    }
    // off-diagonal Frobenius norm below 1e-13 of the diagonal's: FP64 rounding level
    if (offs <= 1e-26 * dgs || offs == 0.0) break;
    for (int round = 0; round < n - 1; ++round) {
      // Items read this round's a and write the next (an): every read of the
      // round, the angle inputs included, sees the matrix the round started
      // with, so one barrier per round suffices.
      int i0[U], i1[U], i2[U], i3[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kind = asg[u] >> 16, hi = (asg[u] >> 8) & 255, lo = asg[u] & 255;
        i0[u] = i1[u] = i2[u] = i3[u] = 0;
        if (kind == 0) continue;  // whole waves past the items skip the lookups
        const int p1 = rrp[round][hi][0], q1 = rrp[round][hi][1];
        const int k2 = kind == 1 ? lo : 0;
        const int p2 = rrp[round][k2][0], q2 = rrp[round][k2][1];
        if (kind == 1) {  // a entries (p1, p2), (q1, p2), (p1, q2), (q1, q2)
          i0[u] = p1 * LD + p2;
          i1[u] = q1 * LD + p2;
          i2[u] = p1 * LD + q2;
          i3[u] = q1 * LD + q2;
        } else {  // v entries (row, p1), (row, q1)
          i0[u] = lo * LD + p1;
          i1[u] = lo * LD + q1;
          i2[u] = i0[u];
          i3[u] = i1[u];
        }
      }
      double x0[U], y0[U], x1[U], y1[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        x0[u] = y0[u] = x1[u] = y1[u] = 0.0;
        if (asg[u] == 0) continue;  // whole waves past the items skip their loads
        const double* m = (asg[u] >> 16) == 2 ? v : a;
        x0[u] = m[i0[u]];
        y0[u] = m[i1[u]];
        x1[u] = m[i2[u]];
        y1[u] = m[i3[u]];
      }
      // every wave computes the round's h angles itself (lanes 0 .. h-1) and
      // hands them to its other lanes through its own LDS slots: LDS accesses
      // of one wave complete in order, so no barrier is needed
      if (lane < h) {
        const int p = rrp[round][lane][0], q = rrp[round][lane][1];
        double c = 1.0, sn = 0.0;
        const double apq = a[p * LD + q];
        if (q < L && fabs(apq) > 1e-300) {
          const double theta = (a[q * LD + q] - a[p * LD + p]) * rcp_nr(2.0 * apq);
          const double th2 = fma(theta, theta, 1.0);
          // sqrt(theta^2 + 1) as x rsqrt(x); past 1e150 the rotation is the identity
          const double root = th2 < 1e300 ? th2 * rsqrt_nr(th2) : 1e300;
          double tt = rcp_nr(fabs(theta) + root);
          if (theta < 0.0) tt = -tt;
          c = rsqrt_nr(fma(tt, tt, 1.0));
          sn = tt * c;
        }
        csw[wv][lane][0] = c;
        csw[wv][lane][1] = sn;
      }
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kind = asg[u] >> 16, hi = (asg[u] >> 8) & 255, lo = asg[u] & 255;
        if (kind == 0) continue;
        const double c1 = csw[wv][hi][0], s1 = csw[wv][hi][1];
        if (kind == 1) {
          const double c2 = csw[wv][lo][0], s2 = csw[wv][lo][1];
          // rows: (p1, q1) <- (c x - s y, s x + c y), then columns (p2, q2)
          const double rp0 = c1 * x0[u] - s1 * y0[u], rq0 = s1 * x0[u] + c1 * y0[u];
          const double rp1 = c1 * x1[u] - s1 * y1[u], rq1 = s1 * x1[u] + c1 * y1[u];
          an[i0[u]] = c2 * rp0 - s2 * rp1;
          an[i2[u]] = s2 * rp0 + c2 * rp1;
          an[i1[u]] = c2 * rq0 - s2 * rq1;
          an[i3[u]] = s2 * rq0 + c2 * rq1;
        } else {
          v[i0[u]] = c1 * x0[u] - s1 * y0[u];
          v[i1[u]] = s1 * x0[u] + c1 * y0[u];
        }
      }
      __syncthreads();  // also orders the convergence reads of red before its next write
      double* const sw = a;
      a = an;
      an = sw;
    }
  }
  // descending order of the eigenvalues (ties by index): rank of each diagonal entry
  // (a total order even when a breakdown upstream left NaNs on the diagonal:
  // NaN ranks last, ties by index, so order is a permutation of 0 .. L-1;
  // every entry holds a valid index before the ranks land)
  if (t < L) order[t] = t;
  __syncthreads();
  if (t < L) {
    const double di = eig_key(a[t * LD + t]);
    int rank = 0;
    for (int j = 0; j < L; ++j) {
      const double dj = eig_key(a[j * LD + j]);
      rank += (dj > di) || (dj == di && j < t);
    }
    order[rank] = t;
  }
  __syncthreads();
  for (int e = t; e < L * L; e += kEigThreads) {
    const int i = e / L, j = e % L;
    W[e] = v[i * LD + order[j]];
  }
  if (t < L) {
    const double ev = a[order[t] * LD + order[t]];
    evals[t] = ev;
    if (t < (int)kc) sinv[t] = 1.0 / sqrt(fmax(ev, 1e-300));
  }
}

void launch_sym_eig(const double* part, int nparts, uint32_t l, double* evals, double* W, double* sinv,
                    uint32_t kc, const double* W0, hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= kMaxL, -1, "l must be in [1, 64]");
  nparts = reduce_parts(part, nparts, l * l, s);
  const uint32_t nb = (l + 15) / 16;
  const size_t shm = (size_t)3 * (16 * nb) * (16 * nb + 1) * sizeof(double);
  switch (nb) {
#define PCAONE_EIG(NB)                                                                                         \
  case NB:                                                                                                     \
    if (shm > 65536)                                                                                           \
      PCAONE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_sym_eig<NB>),                             \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));                   \
    hipLaunchKernelGGL(k_sym_eig<NB>, dim3(1), dim3(kEigThreads), shm, s, part, nparts, l, evals, W, sinv, kc, \
                       W0);                                                                                    \
    break;
    PCAONE_EIG(1) PCAONE_EIG(2) PCAONE_EIG(3) PCAONE_EIG(4)
#undef PCAONE_EIG
  }
}

// ---------------------------------------------------------------------------
// mev: M = X^T Y (k x k), value = mean_i ||M[:, i]||
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(256) void k_xty_partial(const double* __restrict__ X, const double* __restrict__ Y,
                                                     uint64_t n, uint32_t k, double* __restrict__ part) {
  constexpr int NP = NB * NB;
  __shared__ double red[NP * 256];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kr = lane >> 4, ci = lane & 15;
  f64x4 acc[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) acc[p] = f64x4{0.0, 0.0, 0.0, 0.0};
  const uint64_t steps = ceil_div(n, 4);
  const uint64_t nw = (uint64_t)gridDim.x * 4, wid = (uint64_t)blockIdx.x * 4 + wv;
  const uint64_t sps = ceil_div(steps, nw);
  const uint64_t s0 = wid * sps, s1 = std::min<uint64_t>(s0 + sps, steps);
  for (uint64_t st = s0; st < s1; ++st) {
    const uint64_t r = st * 4 + kr;
    double x[NB], y[NB];
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
      const uint32_t c = cb * 16 + ci;
      const bool ok = r < n && c < k;
      x[cb] = ok ? X[r * k + c] : 0.0;
      y[cb] = ok ? Y[r * k + c] : 0.0;
    }
#pragma unroll
    for (int ca = 0; ca < NB; ++ca)
#pragma unroll
      for (int cb = 0; cb < NB; ++cb) acc[ca * NB + cb] = mfma64(x[ca], y[cb], acc[ca * NB + cb]);
  }
  block_tiles_out<NB, NP>(acc, false, k, red, part + (uint64_t)blockIdx.x * k * k);
}

void launch_xty_partial(const double* X, const double* Y, uint64_t n, uint32_t k, double* part, int nblocks,
                        hipStream_t s) {
  PCAONE_REQUIRE(k >= 1 && k <= kMaxL, -1, "k must be in [1, 64]");
  switch ((k + 15) / 16) {
#define PCAONE_NB(NBV) \
  case NBV: hipLaunchKernelGGL(k_xty_partial<NBV>, dim3(nblocks), dim3(256), 0, s, X, Y, n, k, part); break;
    PCAONE_NB(1) PCAONE_NB(2) PCAONE_NB(3) PCAONE_NB(4)
#undef PCAONE_NB
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
