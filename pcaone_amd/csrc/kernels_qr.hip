// The winSVD Omega update in three launches: Omega <- Householder Q of H
// (H = H1 + H2), then flipOmg (src/Halko.cpp:213-216, src/RSVD.hpp:81-90 on
// Eigen's HouseholderQR).  The Q is CholeskyQR2's, with the Householder
// R-diagonal signs reproduced:
//
//   k_qr_gram_h  H = H1 + H2 (written); block partials of A1 = H^T H
//   k_qr_gram_q  prologue: A1 = R1^T R1, R1^-1;  pass: Q1 = H R1^-1 on the fly,
//                block partials of A2 = Q1^T Q1
//   k_qr_flip    prologue: A2 = R2^T R2, R2^-1, the Householder signs of the top
//                l x l block of Q1 R2^-1;  pass: Omega_u = Q1 R2^-1 on the fly
//                into every consumer (the FP64 matrix, the contraction operand
//                in split BF16 / FP32 / FP64, K-permuted for the packed-genotype
//                kernels, the FP32 table of the missing-entry gathers; the
//                emptied H accumulator is zeroed) and block partials of
//                flipOmg's column statistics against the previous Omega
//   (signs)      one block sums the statistics and settles each column's sign:
//                a side job (side_jobs.hpp) of the next missing-entry gather,
//                or a launch of its own
//
// Omega is kept unsigned: a column's sign (Householder sign times flipOmg's
// flip) lives in `fsign`, and G = X Omega takes it in its epilogue, where every
// term (the product, the column sums, the missing-entry sums) is linear in the
// column.  The tall N x l data is read once per pass.  A small launch (or
// side job) after each pass sums its block partials in a fixed order
// (deterministic); each
// prologue reads that sum and runs the l x l step redundantly in every block
// (block 0 stores the results), so no block ever waits for another inside a
// launch.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 512;  // 8 waves per block for the reduction passes
constexpr int kWaves = kThreads / 64;
constexpr int kMaxBlocks = 64;

// v_mfma_f64_16x16x4_f64: lane (k = lane / 16, i = lane % 16) supplies A[i][k]
// and B[k][i]; accumulator element v holds D[lane / 16 + 4 v][lane % 16].
__device__ inline f64x4 mfma64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <int NB>
struct Shape {
  static constexpr int W = NB * 16;                 // padded l
  static constexpr int NP = NB * (NB + 1) / 2;      // upper Gram tiles
  static constexpr int GRAM = NP * 256;             // doubles per Gram partial
  static constexpr int STAT = 3 * W;                // doubles per flip partial
  static constexpr int SLOTS = (kWaves / 2) * GRAM; // block_sum_waves scratch
};

// The eight waves' accumulators summed into wave 0 (fixed tree order).
template <int NP>
__device__ inline void block_sum_waves(f64x4 (&g)[NP], double* slots) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int half = kWaves / 2; half >= 1; half >>= 1) {
    __syncthreads();
    if (wv >= half && wv < 2 * half)
#pragma unroll
      for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v) slots[((wv - half) * NP + p) * 256 + v * 64 + lane] = g[p][v];
    __syncthreads();
    if (wv < half)
#pragma unroll
      for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v) g[p][v] += slots[(wv * NP + p) * 256 + v * 64 + lane];
  }
}

// Wave 0's Gram tiles to the block's partial.
template <int NP>
__device__ inline void store_gram(const f64x4 (&g)[NP], double* dst) {
  const int lane = threadIdx.x & 63;
  if ((threadIdx.x >> 6) == 0)
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int v = 0; v < 4; ++v) dst[p * 256 + v * 64 + lane] = g[p][v];
}

// A reduced Gram (tile p enumerates ca <= cb row by row; element v * 64 +
// lane of a tile is D[lane / 16 + 4 v][lane % 16]) into the upper triangle of
// a (W x W, row stride W).
template <int NB>
__device__ void gram_to_upper(const double* __restrict__ gram, uint32_t l, double* a) {
  using S = Shape<NB>;
  for (int e = threadIdx.x; e < S::GRAM; e += blockDim.x) {
    const int p = e >> 8, v = (e >> 6) & 3, ln = e & 63;
    int ca = 0, rem = p;
    while (rem >= NB - ca) {
      rem -= NB - ca;
      ++ca;
    }
    const int cb = ca + rem;
    const uint32_t r = ca * 16 + (ln >> 4) + 4 * v, c = cb * 16 + (ln & 15);
    if (r >= l || c >= l || r > c) continue;
    a[r * S::W + c] = gram[e];
  }
}

// Upper Cholesky A = R^T R and R^-1 in l elimination steps: row operations
// reduce A to U = D L^T and turn the identity E into L^-1, so
// R^-1[i][j] = E[j][i] / sqrt(d_j).  a (upper triangle used), E and rinv are
// W x W in LDS; rinv comes out zero padded.  Returns the first failing pivot
// (1-based), 0 when none.
template <int W>
__device__ int chol_inverse(double* a, double* E, double* isd, uint32_t l, double* rinv) {
  const uint32_t t = threadIdx.x, nt = blockDim.x;
  for (uint32_t e = t; e < W * W; e += nt) E[e] = (e / W == e % W) ? 1.0 : 0.0;
  int bad = 0;
  for (uint32_t k = 0; k < l; ++k) {
    __syncthreads();
    double d = a[k * W + k];
    if (!(d > 0.0) || !isfinite(d)) {
      if (bad == 0) bad = (int)k + 1;
      d = 1e-300;
    }
    const double inv = 1.0 / d;
    const uint32_t items = (l - k - 1) * W;
    for (uint32_t e = t; e < items; e += nt) {
      const uint32_t i = k + 1 + e / W, j = e % W;
      const double m = a[k * W + i] * inv;
      if (j >= i && j < l)
        a[i * W + j] = fma(-m, a[k * W + j], a[i * W + j]);
      else if (j <= k)
        E[i * W + j] = fma(-m, E[k * W + j], E[i * W + j]);
    }
  }
  __syncthreads();
  for (uint32_t j = t; j < l; j += nt) {
    double d = a[j * W + j];
    if (!(d > 0.0) || !isfinite(d)) d = 1e-300;
    isd[j] = 1.0 / sqrt(d);
  }
  __syncthreads();
  for (uint32_t e = t; e < W * W; e += nt) {
    const uint32_t i = e / W, j = e % W;
    rinv[e] = (i <= j && j < l) ? E[j * W + i] * isd[j] : 0.0;
  }
  __syncthreads();
  return bad;
}

// 1 / d for a normal d: the hardware reciprocal estimate (about 23 bits) and
// two Newton steps r <- r + r (1 - d r), which leave the error at FP64
// rounding; a short dependent chain where the IEEE division sequence is long.
__device__ inline double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  return fma(r, fma(-d, r, 1.0), r);
}

__device__ inline double bcast(double v, int lane) {  // v of `lane` (uniform) in every lane
  const int2 w = __builtin_bit_cast(int2, v);
  return __builtin_bit_cast(double, int2{__builtin_amdgcn_readlane(w.x, lane), __builtin_amdgcn_readlane(w.y, lane)});
}

// The one-wave l x l factorisations use the wave as an 8 x 8 lane grid: lane
// (p, q) = (lane / 8, lane % 8) holds entry (8 u + p, 8 v + q) of the
// matrix as x[u][v], u, v < V (V = R / 8, l <= R).  Step k publishes row k
// (and for the LU column k) through LDS; every lane then updates its own
// entries, so a step costs a few LDS operations and at most V^2 FMAs per lane.
// The step index is a template parameter, so all register indices and the
// block ranges a step touches are compile-time constants.

// Step K of chol_inverse_2d.  Rows and columns in [l, 8 V) hold the identity,
// which factors into itself.  Only the upper triangle of A is kept exact
// (entries below it are updated too but never read); E is unit lower
// triangular, so row K of E is zero past column K.
template <int V, int K>
__device__ __attribute__((always_inline)) inline void chol2d_step(double (&a)[V][V], double (&e)[V][V], int p, int q,
                                                                  double* rowA, double* rowE, int& bad) {
  if constexpr (K < 8 * V) {
    constexpr int uk = K / 8, pk = K % 8;
    double d = bcast(a[uk][uk], 9 * pk);  // lane (pk, pk) holds A[K][K]
    const bool ok = d > 0.0 && isfinite(d);
    bad = (bad == 0 && !ok) ? K + 1 : bad;
    d = ok ? d : 1e-300;
    const double inv = rcp_nr(d);
    if (p == pk) {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        rowA[8 * v + q] = a[uk][v];
        rowE[8 * v + q] = e[uk][v];
      }
    }
    __builtin_amdgcn_wave_barrier();
    double ak[V], ek[V], m[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      ak[v] = (v >= uk) ? rowA[8 * v + q] : 0.0;  // A[K][j], j = 8 v + q (j >= i > K needs v >= uk)
      ek[v] = (v <= uk) ? rowE[8 * v + q] : 0.0;  // E[K][j], zero past j = K
      m[v] = (v >= uk) ? rowA[8 * v + p] * inv : 0.0;  // A[K][i] / A[K][K], i = 8 v + p
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = uk; u < V; ++u) {
      const bool live = u > uk || p > pk;  // row 8 u + p below the pivot
      const double mu = live ? m[u] : 0.0;
#pragma unroll
      for (int v = u; v < V; ++v) a[u][v] = fma(-mu, ak[v], a[u][v]);
#pragma unroll
      for (int v = 0; v <= uk; ++v) e[u][v] = fma(-mu, ek[v], e[u][v]);
    }
    // the step's updates complete here (no sinking into later steps: registers)
#pragma unroll
    for (int u = 0; u < V; ++u)
#pragma unroll
      for (int v = 0; v < V; ++v) asm volatile("" : "+v"(a[u][v]), "+v"(e[u][v]));
    chol2d_step<V, K + 1>(a, e, p, q, rowA, rowE, bad);
  }
}

// R^-1 of the Cholesky factor of A (upper triangle of a W x W LDS matrix,
// l <= 8 V) by elimination: row operations reduce A to U = D L^T and turn the
// identity E into L^-1, so R^-1[i][c] = E[c][i] / sqrt(d_c).  Called by one
// whole wave; rinv (W x W LDS, row stride W) comes out zero padded; scratch is
// 3 * 8 V doubles of LDS.  Returns the first failing pivot (1-based), 0 if none.
template <int W, int V>
__device__ __attribute__((always_inline)) inline int chol_inverse_2d(const double* A, uint32_t l, double* rinv,
                                                                     double* scratch) {
  const int lane = threadIdx.x & 63, p = lane >> 3, q = lane & 7;
  const int L = __builtin_amdgcn_readfirstlane((int)l);
  double a[V][V], e[V][V];
#pragma unroll
  for (int u = 0; u < V; ++u)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int r = 8 * u + p, c = 8 * v + q;
      a[u][v] = (r <= c && c < L) ? A[r * W + c] : (r == c ? 1.0 : 0.0);
      e[u][v] = (r == c) ? 1.0 : 0.0;
    }
  double* rowA = scratch;
  double* rowE = scratch + 8 * V;
  double* isd = scratch + 16 * V;
  int bad = 0;
  chol2d_step<V, 0>(a, e, p, q, rowA, rowE, bad);
  // Row K is final after step K - 1, so the diagonal now holds every pivot:
  // lane (p, p) takes 1 / sqrt of pivots 8 u + p.
  if (p == q)
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const double d = a[u][u];
      isd[8 * u + p] = 1.0 / sqrt((d > 0.0 && isfinite(d)) ? d : 1e-300);
    }
  __builtin_amdgcn_wave_barrier();
  for (int t = lane; t < W * W; t += 64) rinv[t] = 0.0;
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int u = 0; u < V; ++u)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int c = 8 * u + p, i = 8 * v + q;  // e[u][v] = E[c][i]
      if (i <= c && c < L) rinv[i * W + c] = e[u][v] * isd[c];
    }
  __builtin_amdgcn_wave_barrier();
  return bad > L ? 0 : bad;
}

// Step K of householder_signs_2d: LU of B - S with S_KK = -sign(pivot).
// Rows and columns at or left of K are dead after step K, so updates over
// whole blocks are harmless.
template <int V, int K>
__device__ __attribute__((always_inline)) inline void signs2d_step(double (&b)[V][V], int p, int q, int L,
                                                                   double* colK, double* rowK, double* sign) {
  if constexpr (K < 8 * V) {
    constexpr int kb = K / 8, kc = K % 8;
    const double bkk = bcast(b[kb][kb], 9 * kc);
    const double sk = (bkk >= 0.0) ? -1.0 : 1.0;
    const double ip = rcp_nr(bkk - sk);
    if (p == 0 && q == 0 && K < L) sign[K] = sk;
    if (q == kc)
#pragma unroll
      for (int u = 0; u < V; ++u) colK[8 * u + p] = b[u][kb];
    if (p == kc)
#pragma unroll
      for (int v = 0; v < V; ++v) rowK[8 * v + q] = b[kb][v];
    __builtin_amdgcn_wave_barrier();
    double lm[V], rk[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      lm[u] = (u >= kb) ? colK[8 * u + p] * ip : 0.0;  // B[i][K] / pivot, i = 8 u + p
      rk[u] = (u >= kb) ? rowK[8 * u + q] : 0.0;       // B[K][j], j = 8 u + q
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = kb; u < V; ++u)
#pragma unroll
      for (int v = kb; v < V; ++v) b[u][v] = fma(-lm[u], rk[v], b[u][v]);
#pragma unroll
    for (int u = 0; u < V; ++u)
#pragma unroll
      for (int v = 0; v < V; ++v) asm volatile("" : "+v"(b[u][v]));
    signs2d_step<V, K + 1>(b, p, q, L, colK, rowK, sign);
  }
}

// Householder R-diagonal signs of the top l x l block B (LDS, row stride W,
// l <= 8 V) of a CholeskyQR factor, in one wave: LU of B - S with
// S_kk = -sign(pivot) (Householder reconstruction, Ballard et al. 2014).
// Rows and columns past l only touch entries past l.  scratch: 16 V doubles.
template <int W, int V>
__device__ __attribute__((always_inline)) inline void householder_signs_2d(const double* B, uint32_t l, double* sign,
                                                                          double* scratch) {
  const int lane = threadIdx.x & 63, p = lane >> 3, q = lane & 7;
  double b[V][V];
#pragma unroll
  for (int u = 0; u < V; ++u)
#pragma unroll
    for (int v = 0; v < V; ++v) b[u][v] = B[(8 * u + p) * W + 8 * v + q];
  signs2d_step<V, 0>(b, p, q, __builtin_amdgcn_readfirstlane((int)l), scratch, scratch + 8 * V, sign);
}

// The one-wave routines hold 8 V = 16 NB - 8 rows and columns: l up to 8,
// 24, 40 or 56 for NB = 1 .. 4 (the register budget of two waves per SIMD);
// larger l of an NB falls back to the whole-block versions.
template <int NB>
__device__ __host__ constexpr int wave_v() { return NB <= 3 ? 2 * NB - 1 : 0; }

// R^-1 of the Gram in a (upper triangle, W x W LDS) into rinv (W x W LDS):
// one wave when l fits wave_v, the whole block above.  Ends with a block
// barrier.
template <int NB>
__device__ __attribute__((always_inline)) inline int factor(double* a, double* E, double* isd, uint32_t l,
                                                            double* rinv) {
  constexpr int W = NB * 16, V = wave_v<NB>();
  if constexpr (V > 0) {
    if (l <= 8 * V) {
      __shared__ int bad_sh;
      __syncthreads();
      if ((threadIdx.x >> 6) == 0) {
        const int bad = chol_inverse_2d<W, V>(a, l, rinv, E);
        if (threadIdx.x == 0) bad_sh = bad;
      }
      __syncthreads();
      return bad_sh;
    }
  }
  return chol_inverse<W>(a, E, isd, l, rinv);
}

__device__ inline void stage_padded(const double* __restrict__ src, uint32_t l, int W, double* dst) {
  for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
    const uint32_t k = e / W, j = e % W;
    dst[e] = (k < l && j < l) ? src[k * l + j] : 0.0;
  }
}

__device__ inline void store_compact(const double* src, uint32_t l, int W, double* dst) {
  for (uint32_t e = threadIdx.x; e < l * l; e += blockDim.x) dst[e] = src[(e / l) * W + e % l];
}

// The 16-row tile `tile` of an n x l row-major matrix in MFMA operand
// layout: lane (k, i) = (lane / 16, lane % 16) gets A[16 tile + i][4 ks + k] in
// x[ks], zero outside the matrix (predicated loads: none for the padding
// columns).
template <int NB>
__device__ inline void load_tile(const double* __restrict__ A, uint64_t n, uint32_t l, uint64_t tile,
                                 double (&x)[NB * 4]) {
  const int lane = threadIdx.x & 63, kr = lane >> 4, ci = lane & 15;
  const uint64_t row = tile * 16 + ci;
#pragma unroll
  for (int ks = 0; ks < NB * 4; ++ks) {
    const uint32_t k = ks * 4 + kr;
    x[ks] = (row < n && k < l) ? A[row * l + k] : 0.0;
  }
}

// The previous Omega at the positions tile_omega's om[c2][v] takes for tile
// `tile`: o[c2][v] = P[16 tile + lane % 16][16 c2 + lane / 16 + 4 v] (zero
// outside; clamped, unconditional loads).
template <int NB>
__device__ inline void load_prev(const double* __restrict__ P, uint64_t n, uint32_t l, uint64_t tile,
                                 double (&o)[NB][4]) {
  const int lane = threadIdx.x & 63, kr = lane >> 4, ci = lane & 15;
  const uint64_t row = tile * 16 + ci;
  const uint64_t rc = row < n ? row : n - 1;
#pragma unroll
  for (int c2 = 0; c2 < NB; ++c2)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint32_t c = c2 * 16 + kr + 4 * v;
      const double y = P[rc * l + (c < l ? c : l - 1)];
      o[c2][v] = (row < n && c < l) ? y : 0.0;
    }
}

// Omega^T of one 16-row tile in registers from the tile of H in b (load_tile
// layout, which is the B operand of step ks): Q1^T = R1^-T H^T, then
// Omega^T = Rb^T Q1^T, whose D layout (lane (k, i): Q1[r0 + i][16 cb + k + 4 v])
// is already the B operand of the second product.
// om[cb][v] = Omega[r0 + lane % 16][16 cb + lane / 16 + 4 v].
template <int NB>
__device__ inline void tile_omega(const double (&b)[NB * 4], uint32_t l, const double* r1, const double* rb,
                                  f64x4 (&om)[NB]) {
  constexpr int W = NB * 16;
  const int lane = threadIdx.x & 63, kr = lane >> 4, ci = lane & 15;
  f64x4 qt[NB];
#pragma unroll
  for (int cb = 0; cb < NB; ++cb) qt[cb] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < W / 4; ++ks) {
    if (ks * 4 >= (int)l) break;
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) qt[cb] = mfma64(r1[(ks * 4 + kr) * W + cb * 16 + ci], b[ks], qt[cb]);
  }
#pragma unroll
  for (int cb = 0; cb < NB; ++cb) om[cb] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int cb = 0; cb < NB; ++cb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ks = cb * 4 + v;
      if (ks * 4 >= (int)l) break;
#pragma unroll
      for (int c2 = 0; c2 < NB; ++c2) om[c2] = mfma64(rb[(ks * 4 + kr) * W + c2 * 16 + ci], qt[cb][v], om[c2]);
    }
}

// Tile range of this wave in a pass over ceil(n / 16) tiles; with skip0,
// wave 0 of block 0 takes none (it has other work).
__device__ inline void wave_tiles(uint64_t n, bool skip0, uint64_t& t0, uint64_t& t1) {
  const uint64_t tiles = ceil_div(n, 16);
  uint64_t nw = (uint64_t)gridDim.x * kWaves, wid = (uint64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (skip0) {
    if (wid == 0) {
      t0 = t1 = 0;
      return;
    }
    --nw;
    --wid;
  }
  const uint64_t tpw = ceil_div(tiles, nw);
  t0 = std::min<uint64_t>(wid * tpw, tiles);
  t1 = std::min<uint64_t>(t0 + tpw, tiles);
}

// ---------------------------------------------------------------------------
// Gram partial of the rows of A1 (+ A2, written to Aout) over this block's
// share of the 4-row K steps, summed over the block's waves into dst.
template <int NB>
__device__ __attribute__((always_inline)) inline void gram_rows_pass(const double* __restrict__ A1,
                                                                     const double* __restrict__ A2, double* Aout,
                                                                     uint64_t n, uint32_t l, double* slots,
                                                                     double* dst) {
  using S = Shape<NB>;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kr = lane >> 4, ci = lane & 15;
  f64x4 acc[S::NP];
#pragma unroll
  for (int p = 0; p < S::NP; ++p) acc[p] = f64x4{0.0, 0.0, 0.0, 0.0};
  // 4-row K steps: lane (k, i) loads A[row k][16 cb + i], at once the A and
  // the B operand of the tile products
  const uint64_t steps = ceil_div(n, 4);
  const uint64_t nw = (uint64_t)gridDim.x * kWaves, wid = (uint64_t)blockIdx.x * kWaves + wv;
  const uint64_t sps = ceil_div(steps, nw);
  const uint64_t s0 = std::min<uint64_t>(wid * sps, steps), s1 = std::min<uint64_t>(s0 + sps, steps);
  for (uint64_t st0 = s0; st0 < s1; st0 += 4) {  // four K steps' loads in flight
    double a[4][NB];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t r = (st0 + u) * 4 + kr;
      const bool live = st0 + u < s1 && r < n;
#pragma unroll
      for (int cb = 0; cb < NB; ++cb) {
        const uint32_t c = cb * 16 + ci;
        a[u][cb] = (live && c < l) ? A1[r * l + c] : 0.0;
      }
    }
    if (A2) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint64_t r = (st0 + u) * 4 + kr;
        const bool live = st0 + u < s1 && r < n;
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
          const uint32_t c = cb * 16 + ci;
          if (live && c < l) {
            a[u][cb] += A2[r * l + c];
            Aout[r * l + c] = a[u][cb];
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int p = 0;
#pragma unroll
      for (int ca = 0; ca < NB; ++ca)
#pragma unroll
        for (int cb = ca; cb < NB; ++cb, ++p) acc[p] = mfma64(a[u][ca], a[u][cb], acc[p]);
    }
  }
  block_sum_waves<S::NP>(acc, slots);
  store_gram<S::NP>(acc, dst);
}

// Q = A Rinv (rows of A in 16-row tiles, Rinv a zero-padded W x W LDS matrix)
// over the wave's tiles [t0, t1), written to Qout when given, and the Gram
// partial of Q summed over the block's waves into dst.  With AHEAD (for waves
// with a tile or two), x holds tile t0 (load_tile) on entry and each next
// tile is loaded while the current computes; without it (long ranges, where
// the registers buy occupancy instead) each tile is loaded in turn.
template <int NB, bool AHEAD>
__device__ __attribute__((always_inline)) inline void apply_gram_pass(const double* __restrict__ A, uint64_t n,
                                                                      uint32_t l, const double* rinv, double* Qout,
                                                                      double* slots, double* dst, uint64_t t0,
                                                                      uint64_t t1, double (&x)[NB * 4]) {
  using S = Shape<NB>;
  constexpr int W = S::W;
  const int lane = threadIdx.x & 63;
  const int kr = lane >> 4, ci = lane & 15;
  f64x4 g[S::NP];
#pragma unroll
  for (int p = 0; p < S::NP; ++p) g[p] = f64x4{0.0, 0.0, 0.0, 0.0};
  for (uint64_t tile = t0; tile < t1; ++tile) {
    double xn[AHEAD ? W / 4 : 1];
    if constexpr (AHEAD) {
      if (tile + 1 < t1) load_tile<NB>(A, n, l, tile + 1, xn);
    } else {
      load_tile<NB>(A, n, l, tile, x);
    }
    // Q tile = A tile Rinv: lane (k, i) supplies A[r0 + i][4 ks + k]
    f64x4 acc[NB];
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) acc[cb] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < W / 4; ++ks) {
      if (ks * 4 >= (int)l) break;
#pragma unroll
      for (int cb = 0; cb < NB; ++cb) acc[cb] = mfma64(x[ks], rinv[(ks * 4 + kr) * W + cb * 16 + ci], acc[cb]);
    }
    if (Qout)
#pragma unroll
      for (int cb = 0; cb < NB; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const uint64_t r = tile * 16 + kr + 4 * v;
          const uint32_t c = cb * 16 + ci;
          if (r < n && c < l) Qout[r * l + c] = acc[cb][v];
        }
    // accumulator element v is the 4-row slab (v, lane / 16) of the tile, in
    // the operand layout of the Gram's K step
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int p = 0;
#pragma unroll
      for (int ca = 0; ca < NB; ++ca)
#pragma unroll
        for (int cb = ca; cb < NB; ++cb, ++p) g[p] = mfma64(acc[ca][v], acc[cb][v], g[p]);
    }
    if constexpr (AHEAD) {
      if (tile + 1 < t1)
#pragma unroll
        for (int ks = 0; ks < W / 4; ++ks) x[ks] = xn[ks];
    }
  }
  block_sum_waves<S::NP>(g, slots);
  store_gram<S::NP>(g, dst);
}

template <int NB>
__global__ __launch_bounds__(kThreads) void k_qr_gram_h(QrPass q) {
  using S = Shape<NB>;
  __shared__ double slots[S::SLOTS];
  gram_rows_pass<NB>(q.H1, q.H2, q.H, q.n, q.l, slots, q.part + (uint64_t)blockIdx.x * S::GRAM);
}

template <int NB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_qr_gram_q(QrPass q) {
  using S = Shape<NB>;
  constexpr int W = S::W;
  constexpr int PRO = 2 * W * W + W;  // a, E, isd
  __shared__ double r1[W * W];
  __shared__ double sh[S::SLOTS > PRO ? S::SLOTS : PRO];
  const uint32_t l = q.l;
  // the wave's first tile of H is in flight during the prologue
  uint64_t t0, t1;
  wave_tiles(q.n, false, t0, t1);
  double x[NB * 4];
  if (t0 < t1) load_tile<NB>(q.H, q.n, l, t0, x);
  // prologue: R1^-1 from the Gram of H
  {
    double* a = sh;
    double* E = a + W * W;
    double* isd = E + W * W;
    gram_to_upper<NB>(q.gram, l, a);
    const int bad = factor<NB>(a, E, isd, l, r1);
    if (blockIdx.x == 0) {
      store_compact(r1, l, W, q.R1i);
      if (threadIdx.x == 0 && bad && q.status[0] == 0) q.status[0] = bad;
    }
  }
  apply_gram_pass<NB, true>(q.H, q.n, l, r1, nullptr, sh, q.part + (uint64_t)(kMaxBlocks + blockIdx.x) * S::GRAM, t0,
                            t1, x);
}

// The top W rows of the unsigned Omega (row stride W) from its first tiles,
// computed by `waves` waves of the block, the same way as every other row.
template <int NB>
__device__ void omega_top(const QrPass& q, const double* r1, const double* r2, int waves, double* bt) {
  constexpr int W = NB * 16;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kr = lane >> 4, ci = lane & 15;
  for (int tb = wv; tb < NB; tb += waves) {
    double b[NB * 4];
    load_tile<NB>(q.H, q.n, q.l, tb, b);
    f64x4 om[NB];
    tile_omega<NB>(b, q.l, r1, r2, om);
#pragma unroll
    for (int c2 = 0; c2 < NB; ++c2)
#pragma unroll
      for (int v = 0; v < 4; ++v) bt[(tb * 16 + ci) * W + c2 * 16 + kr + 4 * v] = om[c2][v];
  }
}

// LU of B - S for the signs with the whole block (l > 48), B in LDS (stride W).
template <int W>
__device__ void householder_signs_block(double* B, uint32_t l, double* sign) {
  const uint32_t t = threadIdx.x, nt = blockDim.x;
  for (uint32_t k = 0; k < l; ++k) {
    __syncthreads();
    const double bkk = B[k * W + k];
    const double sk = (bkk >= 0.0) ? -1.0 : 1.0;
    const double ip = 1.0 / (bkk - sk);
    if (t == 0) sign[k] = sk;
    const uint32_t items = (l - k - 1) * W;
    for (uint32_t e = t; e < items; e += nt) {
      const uint32_t i = k + 1 + e / W, j = e % W;
      if (j > k && j < l) B[i * W + j] = fma(-B[i * W + k] * ip, B[k * W + j], B[i * W + j]);
    }
  }
  __syncthreads();
}

// The unsigned Omega of one tile (tile_omega layout) into every consumer,
// through the wave's 16 x (W + 1) LDS rows `stg`: the tile's rows are
// contiguous in the row-major copies, and the split operand takes four
// positions per 8-byte store (zeros for rows past n: the operand's padding).
// Rolled loops keep the write phase's registers few.
// The same writes element by element from the registers (l > 48, where the
// staging rows do not fit the LDS beside the prologue's matrices).
template <int NB>
__device__ inline void write_omega_tile_direct(const QrPass& q, uint64_t tile, const f64x4 (&om)[NB]) {
  const int lane = threadIdx.x & 63, kr = lane >> 4, ci = lane & 15;
  const uint64_t row = tile * 16 + ci;
  if (row >= q.n) return;
  const uint32_t l = q.l;
  const uint64_t kpos = q.kperm ? kperm16(row) : row;
#pragma unroll
  for (int c2 = 0; c2 < NB; ++c2)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint32_t c = c2 * 16 + kr + 4 * v;
      if (c >= l) continue;
      const double x = om[c2][v];
      const uint64_t gi = row * l + c;
      q.omega[gi] = x;
      if (q.zero_acc) q.zero_acc[gi] = 0.0;
      if (q.omega32) q.omega32[gi] = (float)x;
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

template <int NB>
__device__ inline void write_omega_tile(const QrPass& q, uint64_t tile, const f64x4 (&om)[NB], double* stg) {
  constexpr int LD = NB * 16 + 1;
  const int lane = threadIdx.x & 63, kr = lane >> 4, ci = lane & 15;
  const uint32_t l = q.l;
#pragma unroll
  for (int c2 = 0; c2 < NB; ++c2)
#pragma unroll
    for (int v = 0; v < 4; ++v) stg[ci * LD + c2 * 16 + kr + 4 * v] = om[c2][v];
  __builtin_amdgcn_wave_barrier();
  const uint64_t r0 = tile * 16;
  const uint32_t nr = (uint32_t)std::min<uint64_t>(16, q.n - r0);
  const uint32_t ne = nr * l;
  for (uint32_t e = lane; e < ne; e += 64) {
    const uint32_t r = e / l, c = e - r * l;
    const double x = stg[r * LD + c];
    const uint64_t gi = r0 * l + e;
    q.omega[gi] = x;
    if (q.zero_acc) q.zero_acc[gi] = 0.0;
    if (q.omega32) q.omega32[gi] = (float)x;
  }
  if (q.bprec == PREC_SPLIT_BF16) {
    __bf16* bg = reinterpret_cast<__bf16*>(q.bg);
    for (uint32_t e = lane; e < l * 4; e += 64) {
      const uint32_t c = e >> 2, g = e & 3;
      __bf16 hi[4], lo[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {  // position 4 g + j holds row kperm16(4 g + j) = 4 j + g of the tile
        const uint32_t r = q.kperm ? 4 * j + g : 4 * g + j;
        // rows past n are the operand's zero padding, whatever the tile holds
        // there (a broken-down factor leaves 0 * inf = NaN)
        const double x = r < nr ? stg[r * LD + c] : 0.0;
        hi[j] = (__bf16)(float)x;
        lo[j] = (__bf16)(float)(x - (double)(float)hi[j]);
      }
      uint2 h2, l2;
      __builtin_memcpy(&h2, hi, 8);
      __builtin_memcpy(&l2, lo, 8);
      *reinterpret_cast<uint2*>(bg + (uint64_t)c * q.kpad + r0 + 4 * g) = h2;
      *reinterpret_cast<uint2*>(bg + (uint64_t)(q.lo_off + c) * q.kpad + r0 + 4 * g) = l2;
    }
  } else {
    for (uint32_t e = lane; e < ne; e += 64) {
      const uint32_t r = e / l, c = e - r * l;
      const double x = stg[r * LD + c];
      const uint64_t kpos = q.kperm ? kperm16(r0 + r) : r0 + r;
      if (q.bprec == PREC_FP32)
        reinterpret_cast<float*>(q.bg)[(uint64_t)c * q.kpad + kpos] = (float)x;
      else
        reinterpret_cast<double*>(q.bg)[(uint64_t)c * q.kpad + kpos] = x;
    }
  }
  __builtin_amdgcn_wave_barrier();
}

template <int NB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_qr_flip(QrPass q) {
  using S = Shape<NB>;
  constexpr int W = S::W;
  constexpr int PRO = 2 * W * W + W;
  __shared__ double r1[W * W], r2[W * W];
  __shared__ double sh[PRO > kWaves * S::STAT ? PRO : kWaves * S::STAT];
  // the top block of Omega for the one-wave sign LU and its row / column exchange
  __shared__ double btw[wave_v<NB>() > 0 ? W * W : 1];
  __shared__ double lus[wave_v<NB>() > 0 ? 2 * W : 1];
  constexpr bool kStage = NB <= 3;  // LDS room for the staged writes
  __shared__ double stg[kStage ? kWaves * 16 * (W + 1) : 1];  // each wave's tile on its way out
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kr = lane >> 4, ci = lane & 15;
  const uint32_t l = q.l, t = threadIdx.x, nt = blockDim.x;
  constexpr int V = wave_v<NB>();
  const bool wave_lu = l <= 8 * V;
  // the previous Omega's column signs (read here, before the last block
  // replaces them; the prologue's barriers order them before their use)
  __shared__ double fps[W];
  for (uint32_t c = t; c < (uint32_t)W; c += nt) fps[c] = c < l ? q.fsign[c] : 1.0;
  // the wave's first tile of H is in flight during the prologue (for l <= 48;
  // above, the registers go to the prologue)
  constexpr bool kPre = NB <= 3;
  uint64_t t0, t1;
  wave_tiles(q.n, wave_lu, t0, t1);
  double x[NB * 4], o[NB][4];
  if (kPre && t0 < t1) load_tile<NB>(q.H, q.n, l, t0, x);
  // prologue: R2^-1 from the Gram of Q1
  stage_padded(q.R1i, l, W, r1);
  {
    double* a = sh;
    double* E = a + W * W;
    double* isd = E + W * W;
    gram_to_upper<NB>(q.gram, l, a);
    const int bad = factor<NB>(a, E, isd, l, r2);
    if (blockIdx.x == 0) {
      store_compact(r2, l, W, q.R2i);
      if (t == 0 && bad && q.status[1] == 0) q.status[1] = bad;
    }
  }
  // the Householder signs of the unsigned Omega's top l x l block: one wave
  // beside the pass when l fits wave_v, else the whole of block 0 before it
  if (blockIdx.x == 0) {
    if constexpr (V > 0) {
      if (wave_lu) {  // waves 0 .. NB-1 compute the top tiles; wave 0 then runs the LU
        omega_top<NB>(q, r1, r2, NB, btw);
        __syncthreads();
        if (wv == 0) householder_signs_2d<W, V>(btw, l, q.sign, lus);
      }
    }
    if (!wave_lu) {
      omega_top<NB>(q, r1, r2, kWaves, sh);
      householder_signs_block<W>(sh, l, q.sign);
    }
  }
  double dm[NB][4], dp[NB][4], cs[NB][4];
#pragma unroll
  for (int c2 = 0; c2 < NB; ++c2)
#pragma unroll
    for (int v = 0; v < 4; ++v) dm[c2][v] = dp[c2][v] = cs[c2][v] = 0.0;
  for (uint64_t tile = t0; tile < t1; ++tile) {
    if (!kPre || tile != t0) load_tile<NB>(q.H, q.n, l, tile, x);
    // the previous Omega's rows of this tile are read before the tile's new
    // rows replace them (same wave, same addresses)
    if (kPre) load_prev<NB>(q.omega, q.n, l, tile, o);  // waited for after the products
    f64x4 om[NB];
    tile_omega<NB>(x, l, r1, r2, om);
    if (!kPre) load_prev<NB>(q.omega, q.n, l, tile, o);
    const uint64_t row = tile * 16 + ci;
#pragma unroll
    for (int c2 = 0; c2 < NB; ++c2)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t c = c2 * 16 + kr + 4 * v;
        if (row < q.n && c < l) {
          const double y = om[c2][v], p = fps[c] * o[c2][v];
          dm[c2][v] += fabs(p - y);
          dp[c2][v] += fabs(p + y);
          cs[c2][v] += y;
        }
      }
    if constexpr (kStage)
      write_omega_tile<NB>(q, tile, om, stg + wv * 16 * (W + 1));
    else
      write_omega_tile_direct<NB>(q, tile, om);
  }
  // rows are spread over lane % 16: fold them, then the waves
#pragma unroll
  for (int c2 = 0; c2 < NB; ++c2)
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        dm[c2][v] += __shfl_xor(dm[c2][v], off);
        dp[c2][v] += __shfl_xor(dp[c2][v], off);
        cs[c2][v] += __shfl_xor(cs[c2][v], off);
      }
  double* red = sh;
  if (ci == 0)
#pragma unroll
    for (int c2 = 0; c2 < NB; ++c2)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int c = c2 * 16 + kr + 4 * v;
        red[wv * S::STAT + c] = dm[c2][v];
        red[wv * S::STAT + W + c] = dp[c2][v];
        red[wv * S::STAT + 2 * W + c] = cs[c2][v];
      }
  __syncthreads();
  double* const stats = q.stats;
  for (int e = t; e < S::STAT; e += nt) {
    double s = 0.0;
    for (int w = 0; w < kWaves; ++w) s += red[w * S::STAT + e];
    stats[(uint64_t)blockIdx.x * S::STAT + e] = s;
  }
}

// out = the sum of nparts partials of count doubles (stride apart), in block
// order; with square (Gram partials in tile layout), out is the full symmetric
// l x l matrix (row-major) instead.  Each thread has 16 partials' loads in
// flight per batch.
template <int NB>
__global__ __launch_bounds__(256) void k_gram_sum(const double* __restrict__ part, int nparts, uint64_t stride,
                                                  int count, uint32_t l, int square, double* __restrict__ out) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (e < count)
    for (int b0 = wv * 16; b0 < nparts; b0 += 64) {  // wave w takes batches 4 j + w of 16 partials
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
  int ca = 0, rem = p;
  while (rem >= NB - ca) {
    rem -= NB - ca;
    ++ca;
  }
  const int cb = ca + rem;
  const uint32_t r = ca * 16 + (ln >> 4) + 4 * v, c = cb * 16 + (ln & 15);
  if (r < l && c < l && r <= c) {
    out[r * l + c] = tot;
    out[c * l + r] = tot;
  }
}

template <int NB>
void launch_parts_sum(const double* part, int nparts, uint64_t stride, int count, uint32_t l, int square,
                      double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_gram_sum<NB>, dim3((uint32_t)ceil_div(count, 64)), dim3(256), 0, s, part, nparts, stride,
                     count, l, square, out);
}

// blocks of the reduction passes over n rows: about one 16-row tile per wave,
// at most kMaxBlocks partials
int qr_nparts(uint64_t n) {
  return (int)std::min<uint64_t>(std::max<uint64_t>(ceil_div(ceil_div(n, 16), kWaves), 1), kMaxBlocks);
}

template <int NB>
void launch_nb(QrPass q, int pass, hipStream_t s) {
  q.nparts = qr_nparts(q.n);
  using S = Shape<NB>;
  switch (pass) {
    case 0:
      hipLaunchKernelGGL(k_qr_gram_h<NB>, dim3(q.nparts), dim3(kThreads), 0, s, q);
      launch_parts_sum<NB>(q.part, q.nparts, S::GRAM, S::GRAM, q.l, 0, q.gram, s);
      break;
    case 1:
      hipLaunchKernelGGL(k_qr_gram_q<NB>, dim3(q.nparts), dim3(kThreads), 0, s, q);
      launch_parts_sum<NB>(q.part + (uint64_t)kMaxBlocks * S::GRAM, q.nparts, S::GRAM, S::GRAM, q.l, 0, q.gram, s);
      break;
    default: hipLaunchKernelGGL(k_qr_flip<NB>, dim3(q.nparts), dim3(kThreads), 0, s, q); break;
  }
}

// ---------------------------------------------------------------------------
// computeUSV's CholeskyQR2 of G (M x l) and B^T = H R^-1 (src/Halko.cpp:53-68)
// ---------------------------------------------------------------------------
constexpr int kUsvBlocks = 256;  // one block per CU for the M-row passes

// Gram partials of G's rows.
template <int NB>
__global__ __launch_bounds__(kThreads) void k_usv_gram(UsvPass u) {
  using S = Shape<NB>;
  __shared__ double slots[S::SLOTS];
  gram_rows_pass<NB>(u.G, nullptr, nullptr, u.m, u.l, slots, u.part + (uint64_t)blockIdx.x * S::GRAM);
}

// prologue: R1^-1 from the reduced Gram of G; pass: Gram partials of
// Q1 = G R1^-1 (Q1 is not stored).
template <int NB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_usv_q1(UsvPass u) {
  using S = Shape<NB>;
  constexpr int W = S::W;
  constexpr int PRO = 2 * W * W + W;
  __shared__ double r1[W * W];
  __shared__ double sh[S::SLOTS > PRO ? S::SLOTS : PRO];
  const uint32_t l = u.l;
  uint64_t t0, t1;
  wave_tiles(u.m, false, t0, t1);
  double x[NB * 4];
  {
    double* a = sh;
    double* E = a + W * W;
    double* isd = E + W * W;
    gram_to_upper<NB>(u.gram, l, a);
    const int bad = factor<NB>(a, E, isd, l, r1);
    if (blockIdx.x == 0) {
      store_compact(r1, l, W, u.R1i);
      if (threadIdx.x == 0 && bad && u.status[0] == 0) u.status[0] = bad;
    }
  }
  apply_gram_pass<NB, false>(u.G, u.m, l, r1, nullptr, sh, u.part + (uint64_t)blockIdx.x * S::GRAM, t0, t1, x);
}

// prologue: R2^-1 from the reduced Gram of Q1 and Rt^-1 = R1^-1 R2^-1 (the
// inverse of R = R2 R1); pass: B^T = H Rt^-1 (written) and its Gram partials.
template <int NB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_usv_bt(UsvPass u) {
  using S = Shape<NB>;
  constexpr int W = S::W;
  constexpr int PRO = 2 * W * W + W;
  __shared__ double r1[W * W], rt[W * W];
  __shared__ double sh[S::SLOTS > PRO ? S::SLOTS : PRO];
  const uint32_t l = u.l, t = threadIdx.x;
  uint64_t t0, t1;
  wave_tiles(u.n, false, t0, t1);
  double x[NB * 4];
  stage_padded(u.R1i, l, W, r1);
  {
    double* a = sh;
    double* E = a + W * W;
    double* isd = E + W * W;
    gram_to_upper<NB>(u.gram, l, a);
    const int bad = factor<NB>(a, E, isd, l, rt);  // rt holds R2^-1 for now
    if (blockIdx.x == 0) {
      store_compact(rt, l, W, u.R2i);
      if (t == 0 && bad && u.status[1] == 0) u.status[1] = bad;
    }
    // Rt^-1 = R1^-1 R2^-1 (both upper triangular) into a, then back to rt
    for (uint32_t e = t; e < W * W; e += blockDim.x) {
      const uint32_t i = e / W, j = e % W;
      double v = 0.0;
      for (uint32_t k = i; k <= j && k < l; ++k) v = fma(r1[i * W + k], rt[k * W + j], v);
      a[e] = v;
    }
    __syncthreads();
    for (uint32_t e = t; e < W * W; e += blockDim.x) rt[e] = a[e];
    __syncthreads();
    if (blockIdx.x == 0) store_compact(rt, l, W, u.Rti);
  }
  apply_gram_pass<NB, false>(u.H, u.n, l, rt, u.Bt, sh, u.part + (uint64_t)blockIdx.x * S::GRAM, t0, t1, x);
}

template <int NB>
void launch_usv_nb(UsvPass u, hipStream_t s, int step) {
  using S = Shape<NB>;
  const uint32_t mt = (uint32_t)ceil_div(u.m, 16), nt = (uint32_t)ceil_div(u.n, 16);
  const int gm = (int)std::min<uint64_t>(std::max<uint64_t>(ceil_div(mt, kWaves), 1), kUsvBlocks);
  const int gn = (int)std::min<uint64_t>(std::max<uint64_t>(ceil_div(nt, kWaves), 1), kMaxBlocks);
  switch (step) {
    case 0:  // Gram of G, reduced
      hipLaunchKernelGGL(k_usv_gram<NB>, dim3(gm), dim3(kThreads), 0, s, u);
      launch_parts_sum<NB>(u.part, gm, S::GRAM, S::GRAM, u.l, 0, u.gram, s);
      break;
    case 1:  // R1^-1, Gram of Q1, reduced
      hipLaunchKernelGGL(k_usv_q1<NB>, dim3(gm), dim3(kThreads), 0, s, u);
      launch_parts_sum<NB>(u.part, gm, S::GRAM, S::GRAM, u.l, 0, u.gram, s);
      break;
    default:  // R2^-1, Rt^-1, B^T and B B^T (square)
      hipLaunchKernelGGL(k_usv_bt<NB>, dim3(gn), dim3(kThreads), 0, s, u);
      launch_parts_sum<NB>(u.part, gn, S::GRAM, S::GRAM, u.l, 1, u.bbt, s);
      break;
  }
}

}  // namespace

uint64_t qr_stat_doubles() { return (uint64_t)kMaxBlocks * 3 * 64; }

uint64_t qr_part_doubles(uint32_t l) {
  const uint64_t nb = ceil_div(l, 16), np = nb * (nb + 1) / 2;
  return (uint64_t)kMaxBlocks * (2 * np * 256 + 3 * nb * 16);
}

uint64_t usv_part_doubles(uint32_t l) {
  const uint64_t nb = ceil_div(l, 16), np = nb * (nb + 1) / 2;
  return (uint64_t)kUsvBlocks * np * 256;
}

uint64_t usv_gram_doubles(uint32_t l) {
  const uint64_t nb = ceil_div(l, 16), np = nb * (nb + 1) / 2;
  return np * 256;
}

void launch_usv_qr(const UsvPass& u, int step, hipStream_t s) {
  PCAONE_REQUIRE(u.l >= 1 && u.l <= 64, -1, "l must be in [1, 64]");
  PCAONE_REQUIRE(u.m >= u.l && u.n >= 1, -1, "computeUSV: G needs at least l rows");
  switch ((u.l + 15) / 16) {
    case 1: launch_usv_nb<1>(u, s, step); break;
    case 2: launch_usv_nb<2>(u, s, step); break;
    case 3: launch_usv_nb<3>(u, s, step); break;
    default: launch_usv_nb<4>(u, s, step); break;
  }
}

SideJobs qr_sign_job(const QrPass& q) {
  const uint32_t nb = (q.l + 15) / 16, np = nb * (nb + 1) / 2;
  (void)np;
  SideJobs j;
  j.stats = q.stats;
  j.nparts = (uint32_t)qr_nparts(q.n);
  j.W = nb * 16;
  j.l = q.l;
  j.sign = q.sign;
  j.fsign = q.fsign;
  j.colsum = q.colsum;
  return j;
}

void launch_qr_pass(const QrPass& q, int pass, hipStream_t s) {
  PCAONE_REQUIRE(q.l >= 1 && q.l <= 64, -1, "l must be in [1, 64]");
  PCAONE_REQUIRE(q.n >= 2ull * q.l, -1, "the Omega update needs at least 2 l rows");
  PCAONE_REQUIRE(!q.kperm || q.kpad % 16 == 0, -1, "omega operand: permuted K needs 16-code groups");
  PCAONE_REQUIRE(q.bprec != PREC_SPLIT_BF16 || (q.kpad % 16 == 0 && q.kpad >= ceil_div(q.n, 16) * 16), -1,
                 "omega operand: the split operand is written in whole 16-position groups");
  switch ((q.l + 15) / 16) {
    case 1: launch_nb<1>(q, pass, s); break;
    case 2: launch_nb<2>(q, pass, s); break;
    case 3: launch_nb<3>(q, pass, s); break;
    default: launch_nb<4>(q, pass, s); break;
  }
}

}  // namespace pcaone
