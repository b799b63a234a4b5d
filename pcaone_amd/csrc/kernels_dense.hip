// Dense FP32 feature-major matrices (FileBin's .bin / .residuals layout,
// src/FileBinary.cpp:21-46): HBM layout and the contractions over them.
//
// X[j,i] = D[i][j] - mu_i (reference: cast to double, centre each feature).
//   G_i = sum_j X_ji Omega_j = sum_j D_ij Omega_j - mu_i colsum(Omega)
//   H_j = sum_i X_ji G_i     = sum_i D_ij G_i     - sum_i mu_i G_i
// The dense sums run on MFMA (FP32 arm: v_mfma_f32_16x16x4f32; split arm: D and
// the operand both split into BF16 hi + lo, three products hi*hi + hi*lo +
// lo*hi accumulated in FP32) or on FP64 FMA (reference arm); the centring
// terms are applied in the FP64 epilogues shared with the packed path.
#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// layout
// ---------------------------------------------------------------------------
// rows (laid-out order) of D, padded to `ld` floats with zeros
__global__ void k_dense_rows(const float* __restrict__ raw, uint32_t nsamples, uint32_t nfeat,
                             const int32_t* __restrict__ perm, float* __restrict__ out, uint32_t ld) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (uint64_t)nfeat * ld) return;
  const uint32_t r = (uint32_t)(gid / ld), j = (uint32_t)(gid % ld);
  const uint32_t src = perm ? (uint32_t)perm[r] : r;
  out[gid] = (j < nsamples) ? raw[(uint64_t)src * nsamples + j] : 0.0f;
}

void launch_dense_rows(const float* raw, uint32_t nsamples, uint32_t nfeat, const int32_t* perm, float* out,
                       uint32_t ld, hipStream_t s) {
  const uint64_t total = (uint64_t)nfeat * ld;
  hipLaunchKernelGGL(k_dense_rows, dim3((uint32_t)ceil_div(total, 256)), dim3(256), 0, s, raw, nsamples, nfeat,
                     perm, out, ld);
}

// per-feature mean over samples, FP64 accumulation (one wave per feature)
__global__ void k_dense_mean(const float* __restrict__ raw, uint32_t nsamples, uint32_t nfeat,
                             double* __restrict__ mu) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (r >= nfeat) return;
  const float* p = raw + r * nsamples;
  double s = 0.0;
  for (uint32_t j = lane; j < nsamples; j += 64) s += (double)p[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) mu[r] = s / (double)nsamples;
}

void launch_dense_mean(const float* raw, uint32_t nsamples, uint32_t nfeat, double* mu, hipStream_t s) {
  hipLaunchKernelGGL(k_dense_mean, dim3((uint32_t)ceil_div((uint64_t)nfeat * 64, 256)), dim3(256), 0, s, raw,
                     nsamples, nfeat, mu);
}

// windowed transpose: dt[w][j][k] = rows[w*bs + k][j] (zero past the window)
__global__ __launch_bounds__(256) void k_dense_transpose(const float* __restrict__ rows, uint32_t ldr,
                                                         uint32_t nfeat, uint32_t bs, uint32_t nsamples,
                                                         float* __restrict__ dt, uint32_t ldt) {
  __shared__ float tile[64][65];
  const uint32_t w = blockIdx.z, tk = blockIdx.x, tj = blockIdx.y;
  const uint32_t f0 = w * bs + tk * 64, fend = min(w * bs + bs, nfeat);
  const uint32_t j0 = tj * 64;
  const uint32_t t = threadIdx.x;
  for (uint32_t e = t; e < 64 * 64; e += 256) {
    const uint32_t rr = e / 64, cc = e % 64;
    const uint32_t f = f0 + rr, j = j0 + cc;
    tile[rr][cc] = (f < fend && j < nsamples) ? rows[(uint64_t)f * ldr + j] : 0.0f;
  }
  __syncthreads();
  for (uint32_t e = t; e < 64 * 64; e += 256) {
    const uint32_t jj = e / 64, kk = e % 64;
    const uint32_t j = j0 + jj;
    if (j < nsamples) dt[((uint64_t)w * nsamples + j) * ldt + tk * 64 + kk] = tile[kk][jj];
  }
}

void launch_dense_transpose(const float* rows, uint32_t ldr, uint32_t nfeat, uint32_t bs, uint32_t nwin,
                            uint32_t nsamples, float* dt, uint32_t ldt, hipStream_t s) {
  dim3 grid((uint32_t)ceil_div(bs, 64), (uint32_t)ceil_div(nsamples, 64), nwin);
  hipLaunchKernelGGL(k_dense_transpose, grid, dim3(256), 0, s, rows, ldr, nfeat, bs, nsamples, dt, ldt);
}

// ---------------------------------------------------------------------------
// FP32 MFMA contraction: out[z][split][row][col] = sum_k A[row][k] B[col][k]
// ---------------------------------------------------------------------------
// Block = 4 waves x 64 rows; K advances 64 floats per chunk.  Lane (r, g)
// loads float4 q (q = 0..3) at k = 16q + 4g; MFMA step (q, s) consumes element
// s, i.e. k = 16q + 4g + s, and B is read with the same mapping from LDS.
constexpr int kFStride = 68;  // floats per LDS B row (64 + pad)

template <int NT>
__global__ __launch_bounds__(256) void k_dense_gemm_f32(DenseGemm g) {
  __shared__ __attribute__((aligned(16))) float bl[NT * 16 * kFStride];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t z = blockIdx.z, split = blockIdx.y;
  const float* A = g.A + (uint64_t)z * g.a_zstride;
  const float* B = reinterpret_cast<const float*>(g.B) + (uint64_t)z * g.b_zstride;
  float* out = reinterpret_cast<float*>(g.out) + (uint64_t)z * g.o_zstride;
  const uint32_t row0 = blockIdx.x * 256 + wave * 64;
  const uint32_t c0 = split * g.chunks_per_split, c1 = min(c0 + g.chunks_per_split, g.kchunks);
  const int rl = lane & 15, gl = lane >> 4;
  const float* arow[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) arow[f] = A + (uint64_t)min(row0 + 16 * f + rl, g.nrows - 1) * g.lda + 4 * gl;
  f32x4 acc[4][NT];
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[f][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (uint32_t c = c0; c < c1; ++c) {
    f32x4 a[4][4];
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int q = 0; q < 4; ++q) a[f][q] = *reinterpret_cast<const f32x4*>(arow[f] + (uint64_t)c * 64 + 16 * q);
    __syncthreads();
    for (int it = threadIdx.x; it < NT * 16 * 16; it += 256) {
      const int n = it >> 4, seg = it & 15;
      *reinterpret_cast<f32x4*>(bl + n * kFStride + seg * 4) =
          *reinterpret_cast<const f32x4*>(B + (uint64_t)n * g.ldb + (uint64_t)c * 64 + seg * 4);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      f32x4 b[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) b[t] = *reinterpret_cast<const f32x4*>(bl + (t * 16 + rl) * kFStride + 16 * q + 4 * gl);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[f][q][s], b[t][s], acc[f][t], 0, 0, 0);
    }
  }
  float* o = out + (uint64_t)split * g.nrows * g.ldo;
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t r = row0 + 16 * f + 4 * gl + i;
      if (r < g.nrows) {
#pragma unroll
        for (int t = 0; t < NT; ++t) o[(uint64_t)r * g.ldo + t * 16 + rl] = acc[f][t][i];
      }
    }
}

// ---------------------------------------------------------------------------
// split-BF16 contraction with three products
// ---------------------------------------------------------------------------
// B holds 2*NTH tiles of 16 columns: tiles [0, NTH) the BF16 hi parts of the
// operand columns, tiles [NTH, 2 NTH) their lo parts.  A (FP32) is split in
// registers; hi*hi + hi*lo + lo*hi accumulate in one FP32 accumulator per
// output tile.  K-step s (0, 1) of a 64-float chunk feeds codes 32s + 8g .. +7
// of lane (r, g).
constexpr int kHStride = 72;  // bf16 per LDS B row (64 + pad)

__device__ inline void split8(const f32x4 x0, const f32x4 x1, bf16x8& hi, bf16x8& lo) {
  const float v[8] = {x0[0], x0[1], x0[2], x0[3], x1[0], x1[1], x1[2], x1[3]};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const __bf16 h = (__bf16)v[i];
    hi[i] = h;
    lo[i] = (__bf16)(v[i] - (float)h);
  }
}

template <int NTH>
__global__ __launch_bounds__(256) void k_dense_gemm_bf16x3(DenseGemm g) {
  __shared__ __attribute__((aligned(16))) __bf16 bl[2 * NTH * 16 * kHStride];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t z = blockIdx.z, split = blockIdx.y;
  const float* A = g.A + (uint64_t)z * g.a_zstride;
  const __bf16* B = reinterpret_cast<const __bf16*>(g.B) + (uint64_t)z * g.b_zstride;
  float* out = reinterpret_cast<float*>(g.out) + (uint64_t)z * g.o_zstride;
  const uint32_t row0 = blockIdx.x * 256 + wave * 64;
  const uint32_t c0 = split * g.chunks_per_split, c1 = min(c0 + g.chunks_per_split, g.kchunks);
  const int rl = lane & 15, gl = lane >> 4;
  const float* arow[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) arow[f] = A + (uint64_t)min(row0 + 16 * f + rl, g.nrows - 1) * g.lda + 8 * gl;
  f32x4 acc[4][NTH];
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int t = 0; t < NTH; ++t) acc[f][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (uint32_t c = c0; c < c1; ++c) {
    f32x4 a[4][4];
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        a[f][2 * q] = *reinterpret_cast<const f32x4*>(arow[f] + (uint64_t)c * 64 + 32 * q);
        a[f][2 * q + 1] = *reinterpret_cast<const f32x4*>(arow[f] + (uint64_t)c * 64 + 32 * q + 4);
      }
    __syncthreads();
    for (int it = threadIdx.x; it < 2 * NTH * 16 * 8; it += 256) {
      const int n = it >> 3, seg = it & 7;
      *reinterpret_cast<uint4*>(bl + n * kHStride + seg * 8) =
          *reinterpret_cast<const uint4*>(B + (uint64_t)n * g.ldb + (uint64_t)c * 64 + seg * 8);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      bf16x8 bh[NTH], blo[NTH];
#pragma unroll
      for (int t = 0; t < NTH; ++t) {
        bh[t] = *reinterpret_cast<const bf16x8*>(bl + (t * 16 + rl) * kHStride + 32 * s + 8 * gl);
        blo[t] = *reinterpret_cast<const bf16x8*>(bl + ((NTH + t) * 16 + rl) * kHStride + 32 * s + 8 * gl);
      }
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        bf16x8 ah, al;
        split8(a[f][2 * s], a[f][2 * s + 1], ah, al);
#pragma unroll
        for (int t = 0; t < NTH; ++t) {
          acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[t], acc[f][t], 0, 0, 0);
          acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, blo[t], acc[f][t], 0, 0, 0);
          acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[t], acc[f][t], 0, 0, 0);
        }
      }
    }
  }
  float* o = out + (uint64_t)split * g.nrows * g.ldo;
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t r = row0 + 16 * f + 4 * gl + i;
      if (r < g.nrows) {
#pragma unroll
        for (int t = 0; t < NTH; ++t) o[(uint64_t)r * g.ldo + t * 16 + rl] = acc[f][t][i];
      }
    }
}

// ---------------------------------------------------------------------------
// FP64 vector-FMA contraction (reference arm)
// ---------------------------------------------------------------------------
// Block = 64 * ncg threads; thread (tr, tc) owns rows tr + 64 i (i < 4) and
// columns 8 tc .. 8 tc + 7; K advances 32 at a time with B staged in LDS.
__global__ __launch_bounds__(1024) void k_dense_gemm_f64(DenseGemm g, int ncols) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  double* bl = reinterpret_cast<double*>(smem_raw);  // [32][lc]
  const int ncg = blockDim.x >> 6, lc = ncg * 8;
  const int tr = threadIdx.x & 63, tc = threadIdx.x >> 6;
  const uint32_t z = blockIdx.z, split = blockIdx.y;
  const float* A = g.A + (uint64_t)z * g.a_zstride;
  const double* B = reinterpret_cast<const double*>(g.B) + (uint64_t)z * g.b_zstride;
  double* out = reinterpret_cast<double*>(g.out) + (uint64_t)z * g.o_zstride;
  const uint32_t rbase = blockIdx.x * 256 + tr;
  const uint32_t c0 = split * g.chunks_per_split, c1 = min(c0 + g.chunks_per_split, g.kchunks);
  const float* arow[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) arow[i] = A + (uint64_t)min(rbase + 64 * i, g.nrows - 1) * g.lda;
  double acc[4][8];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;
  for (uint32_t c = c0; c < c1; ++c) {
    for (int sub = 0; sub < 2; ++sub) {
      const uint64_t kb = (uint64_t)c * 64 + sub * 32;
      __syncthreads();
      for (int it = threadIdx.x; it < 32 * lc; it += blockDim.x) {
        const int kk = it & 31, col = it >> 5;
        bl[kk * lc + col] = (col < ncols) ? B[(uint64_t)col * g.ldb + kb + kk] : 0.0;
      }
      __syncthreads();
      for (int k4 = 0; k4 < 32; k4 += 4) {
        f32x4 av[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) av[i] = *reinterpret_cast<const f32x4*>(arow[i] + kb + k4);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          double bv[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) bv[j] = bl[(k4 + s) * lc + tc * 8 + j];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const double x = (double)av[i][s];
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[i][j] = fma(x, bv[j], acc[i][j]);
          }
        }
      }
    }
  }
  double* o = out + (uint64_t)split * g.nrows * g.ldo;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t r = rbase + 64 * i;
    if (r < g.nrows) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int col = tc * 8 + j;
        if (col < ncols) o[(uint64_t)r * g.ldo + col] = acc[i][j];
      }
    }
  }
}

void launch_dense_gemm(const DenseGemm& g, int arm, int ncols, int ntiles, hipStream_t s) {
  dim3 grid((uint32_t)ceil_div(g.nrows, 256), g.nsplit, g.nz);
  if (arm == PREC_FP64) {
    const int ncg = (ncols + 7) / 8;
    PCAONE_REQUIRE(ncg >= 1 && ncg <= 16, -1, "dense FP64 contraction supports 1..128 columns");
    hipLaunchKernelGGL(k_dense_gemm_f64, grid, dim3(64 * ncg), (size_t)32 * ncg * 8 * sizeof(double), s, g,
                       ncols);
    return;
  }
  if (arm == PREC_FP32) {
    switch (ntiles) {
#define PCAONE_T(TT) \
  case TT: hipLaunchKernelGGL(k_dense_gemm_f32<TT>, grid, dim3(256), 0, s, g); break;
      PCAONE_T(1) PCAONE_T(2) PCAONE_T(3) PCAONE_T(4) PCAONE_T(5) PCAONE_T(6) PCAONE_T(7) PCAONE_T(8)
#undef PCAONE_T
      default: throw Error(-1, "dense FP32 contraction supports at most 128 columns");
    }
    return;
  }
  switch (ntiles) {
#define PCAONE_T(TT) \
  case TT: hipLaunchKernelGGL(k_dense_gemm_bf16x3<TT>, grid, dim3(256), 0, s, g); break;
    PCAONE_T(1) PCAONE_T(2) PCAONE_T(3) PCAONE_T(4) PCAONE_T(5) PCAONE_T(6) PCAONE_T(7) PCAONE_T(8)
#undef PCAONE_T
    default: throw Error(-1, "dense split-BF16 contraction supports at most 128 columns");
  }
}

}  // namespace pcaone
