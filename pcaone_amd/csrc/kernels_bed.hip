// Packed PLINK genotype kernels for gfx950: HBM layout, code counts, missing
// lists, the two contraction families (split-BF16 MFMA and FP32/FP64 vector
// FMA) and their epilogues.
//
// The algebra (reference layout X is N x M, column i = SNP i):
//   X[j,i] = d_i * (BED2GENO[c_ij] - F_i), missing -> 0     (src/FilePlink.cpp:106-114,
//                                                            src/Data.cpp:290-299)
// With y = A1 dosage in {0,1,2} (BED2GENO = y/2; y = 0 for missing),
//   G_i = (X^T Omega)_i = 1/2 d_i ( sum_j y_ij Omega_j - 2 F_i ( colsum(Omega) - sum_{j miss} Omega_j ) )
//   H_j = (X G)_j       = 1/2 ( sum_i y_ij g'_i - ( S'' - sum_{i miss} g''_i ) ),
//   g' = d G, g'' = 2 F g', S'' = sum_i g''_i.
// The dense parts sum_j y_ij Omega_j and sum_i y_ij g'_i run on the packed codes;
// the missing entries (sparse) are corrected in the epilogues from CSR lists.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"
#include "side_jobs.hpp"

namespace pcaone {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// layout
// ---------------------------------------------------------------------------
__global__ void k_bed_to_snpmajor(const uint8_t* __restrict__ raw, uint32_t nsamples, uint32_t nsnps,
                                  const int32_t* __restrict__ perm, uint8_t* __restrict__ out,
                                  uint32_t row_bytes) {
  const uint32_t nb = (nsamples + 3) >> 2;
  const uint32_t dw_per_row = row_bytes >> 2;
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (uint64_t)nsnps * dw_per_row) return;
  const uint32_t r = (uint32_t)(gid / dw_per_row), q = (uint32_t)(gid % dw_per_row);
  const uint32_t src = perm ? (uint32_t)perm[r] : r;
  const uint8_t* in = raw + (uint64_t)src * nb;
  uint32_t word = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const uint32_t bi = q * 4 + t;
    uint32_t b = 0x55u;  // padding: missing codes
    if (bi < nb) {
      b = in[bi];
      const uint32_t first = bi * 4;
      if (first + 4 > nsamples) {  // last byte: codes past N become missing
        const uint32_t keep = nsamples - first;
        const uint32_t mask = (1u << (2 * keep)) - 1u;
        b = (b & mask) | (0x55u & ~mask);
      }
    }
    word |= b << (8 * t);
  }
  reinterpret_cast<uint32_t*>(out)[gid] = word;
}

void launch_bed_to_snpmajor(const uint8_t* raw, uint32_t nsamples, uint32_t nsnps, const int32_t* perm,
                            uint8_t* bed_snp, uint32_t row_bytes, hipStream_t s) {
  const uint64_t total = (uint64_t)nsnps * (row_bytes / 4);
  const uint32_t blocks = (uint32_t)ceil_div(total, 256);
  hipLaunchKernelGGL(k_bed_to_snpmajor, dim3(blocks), dim3(256), 0, s, raw, nsamples, nsnps, perm, bed_snp,
                     row_bytes);
}

// one wave per row: counts of the four codes among the first nvalid codes
__global__ void k_count_codes(const uint8_t* __restrict__ bed, uint32_t nrows, uint32_t row_bytes,
                              uint32_t nvalid, int32_t* __restrict__ counts) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t row = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= nrows) return;
  const uint32_t* p = reinterpret_cast<const uint32_t*>(bed + row * row_bytes);
  const uint32_t ndw = row_bytes >> 2;
  int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  for (uint32_t q = lane; q < ndw; q += 64) {
    const uint32_t x = p[q];
    const uint32_t lo = x & 0x55555555u, hi = (x >> 1) & 0x55555555u;
    c0 += __popc(~lo & ~hi & 0x55555555u);
    c1 += __popc(lo & ~hi);
    c2 += __popc(~lo & hi);
    c3 += __popc(lo & hi);
  }
  for (int off = 32; off > 0; off >>= 1) {
    c0 += __shfl_down(c0, off);
    c1 += __shfl_down(c1, off);
    c2 += __shfl_down(c2, off);
    c3 += __shfl_down(c3, off);
  }
  if (lane == 0) {
    counts[4 * row + 0] = c0;
    counts[4 * row + 1] = c1 - (int)(row_bytes * 4 - nvalid);  // padding codes are 01
    counts[4 * row + 2] = c2;
    counts[4 * row + 3] = c3;
  }
}

void launch_count_codes(const uint8_t* bed_snp, uint32_t nrows, uint32_t row_bytes, uint32_t nvalid,
                        int32_t* counts4, hipStream_t s) {
  const uint32_t blocks = (uint32_t)ceil_div((uint64_t)nrows * 64, 256);
  hipLaunchKernelGGL(k_count_codes, dim3(blocks), dim3(256), 0, s, bed_snp, nrows, row_bytes, nvalid, counts4);
}

// Tile transpose: 64 SNP rows x 256 samples of bed_snp -> 256 sample rows x 16
// bytes of one window of bed_smp.  Rows past the window stay missing (0x55,
// pre-filled by the caller).
__global__ __launch_bounds__(256) void k_transpose_windows(const uint8_t* __restrict__ bed_snp,
                                                           uint32_t row_bytes, uint32_t nsnps,
                                                           uint32_t blocksize, uint8_t* __restrict__ bed_smp,
                                                           uint32_t npad, uint32_t win_bytes) {
  __shared__ uint32_t tile[64][17];  // 64 rows x 64 bytes (+1 dword pad)
  const uint32_t w = blockIdx.z, tk = blockIdx.x, ts = blockIdx.y;
  const uint32_t ws = w * blocksize;
  const uint32_t we = min(ws + blocksize, nsnps);
  const uint32_t s0 = ws + tk * 64;
  if (s0 >= we) return;
  const uint32_t t = threadIdx.x;
  {
    const uint32_t row = t >> 2, seg = t & 3;
    uint4 v = make_uint4(0x55555555u, 0x55555555u, 0x55555555u, 0x55555555u);
    if (s0 + row < we)
      v = *reinterpret_cast<const uint4*>(bed_snp + (uint64_t)(s0 + row) * row_bytes + ts * 64 + 16 * seg);
    tile[row][4 * seg + 0] = v.x;
    tile[row][4 * seg + 1] = v.y;
    tile[row][4 * seg + 2] = v.z;
    tile[row][4 * seg + 3] = v.w;
  }
  __syncthreads();
  uint32_t o[4] = {0u, 0u, 0u, 0u};
  const uint32_t dw = t >> 4, sh = 2 * (t & 15);  // sample t lives in dword t/16 at bit 2*(t%16)
#pragma unroll
  for (int sp = 0; sp < 64; ++sp) {
    const uint32_t code = (tile[sp][dw] >> sh) & 3u;
    o[sp >> 4] |= code << (2 * (sp & 15));
  }
  const uint64_t j = (uint64_t)ts * 256 + t;
  uint8_t* dst = bed_smp + ((uint64_t)w * npad + j) * win_bytes + tk * 16;
  *reinterpret_cast<uint4*>(dst) = make_uint4(o[0], o[1], o[2], o[3]);
}

void launch_transpose_windows(const uint8_t* bed_snp, uint32_t row_bytes, uint32_t nsnps, uint32_t blocksize,
                              uint32_t nwin, uint8_t* bed_smp, uint32_t npad, uint32_t win_bytes,
                              hipStream_t s) {
  dim3 grid((uint32_t)ceil_div(blocksize, 64), npad / 256, nwin);
  hipLaunchKernelGGL(k_transpose_windows, grid, dim3(256), 0, s, bed_snp, row_bytes, nsnps, blocksize, bed_smp,
                     npad, win_bytes);
}

// ---------------------------------------------------------------------------
// missing lists (code 01)
// ---------------------------------------------------------------------------
__device__ inline bool pr_row(const PackedRows& pr, uint64_t r, uint32_t& nvalid, uint32_t& base) {
  const uint32_t g = (uint32_t)(r / pr.rows_per_group), jr = (uint32_t)(r % pr.rows_per_group);
  if (jr >= pr.live_rows) return false;
  nvalid = (g + 1 == pr.ngroups) ? pr.valid_last : pr.valid;
  base = pr.base0 + g * pr.base_step;
  return true;
}

__device__ inline uint32_t missing_mask(uint32_t x, uint32_t q, uint32_t nvalid) {
  const uint32_t lo = x & 0x55555555u, hi = (x >> 1) & 0x55555555u;
  uint32_t m = lo & ~hi;
  const uint32_t first = q * 16;
  if (first >= nvalid) return 0u;
  if (first + 16 > nvalid) m &= (1u << (2 * (nvalid - first))) - 1u;
  return m;
}

__global__ void k_count_missing(PackedRows pr, int32_t* __restrict__ counts) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (r >= pr.nrows) return;
  uint32_t nvalid, base;
  int c = 0;
  if (pr_row(pr, r, nvalid, base)) {
    const uint32_t* p = reinterpret_cast<const uint32_t*>(pr.data + r * pr.row_bytes);
    const uint32_t ndw = (nvalid + 15) / 16;
    for (uint32_t q = lane; q < ndw; q += 64) c += __popc(missing_mask(p[q], q, nvalid));
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
  if (lane == 0) counts[r] = c;
}

__global__ void k_fill_missing(PackedRows pr, const int64_t* __restrict__ ptr, int32_t* __restrict__ idx) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (r >= pr.nrows) return;
  uint32_t nvalid, base;
  if (!pr_row(pr, r, nvalid, base)) return;
  const uint32_t* p = reinterpret_cast<const uint32_t*>(pr.data + r * pr.row_bytes);
  const uint32_t ndw = (nvalid + 15) / 16;
  int64_t pos = ptr[r];
  for (uint32_t q0 = 0; q0 < ndw; q0 += 64) {
    const uint32_t q = q0 + lane;
    uint32_t m = (q < ndw) ? missing_mask(p[q], q, nvalid) : 0u;
    const int cnt = __popc(m);
    int incl = cnt;  // inclusive prefix over lanes
    for (int off = 1; off < 64; off <<= 1) {
      const int v = __shfl_up(incl, off);
      if ((int)lane >= off) incl += v;
    }
    int64_t o = pos + (incl - cnt);
    while (m) {
      const uint32_t b = __ffs(m) - 1;  // even bit position
      idx[o++] = (int32_t)(base + q * 16 + (b >> 1));
      m &= m - 1;
    }
    pos += __shfl(incl, 63);
  }
}

void launch_count_missing(const PackedRows& pr, int32_t* counts, hipStream_t s) {
  const uint32_t blocks = (uint32_t)ceil_div((uint64_t)pr.nrows * 64, 256);
  hipLaunchKernelGGL(k_count_missing, dim3(blocks), dim3(256), 0, s, pr, counts);
}

void launch_fill_missing(const PackedRows& pr, const int64_t* ptr, int32_t* idx, hipStream_t s) {
  const uint32_t blocks = (uint32_t)ceil_div((uint64_t)pr.nrows * 64, 256);
  hipLaunchKernelGGL(k_fill_missing, dim3(blocks), dim3(256), 0, s, pr, ptr, idx);
}

// ---------------------------------------------------------------------------
// split-BF16 MFMA contraction
// ---------------------------------------------------------------------------
// Spread four 2-bit codes (one byte) into the low 2 bits of four bytes.
__device__ inline uint32_t spread4(uint32_t b) {
  uint32_t t = (b | (b << 12)) & 0x000F000Fu;
  return (t | (t << 6)) & 0x03030303u;
}

// Eight consecutive codes (16 bits) -> eight BF16 dosages {00:2, 10:1, 11:0, 01:0}.
// v_perm picks each BF16 byte from two 4-entry byte tables indexed by the code.
__device__ inline bf16x8 decode8(uint32_t h) {
  constexpr uint32_t LUT_HI = 0x003F0040u;  // high bytes of bf16 {2.0, 0, 1.0, 0}
  constexpr uint32_t LUT_LO = 0x00800000u;  // low bytes
  const uint32_t e0 = spread4(h & 0xFFu), e1 = spread4((h >> 8) & 0xFFu);
  const uint32_t f0 = e0 + 0x04040404u, f1 = e1 + 0x04040404u;
  uint32_t o[4];
  o[0] = __builtin_amdgcn_perm(LUT_LO, LUT_HI, __builtin_amdgcn_perm(f0, e0, 0x01050004u));
  o[1] = __builtin_amdgcn_perm(LUT_LO, LUT_HI, __builtin_amdgcn_perm(f0, e0, 0x03070206u));
  o[2] = __builtin_amdgcn_perm(LUT_LO, LUT_HI, __builtin_amdgcn_perm(f1, e1, 0x01050004u));
  o[3] = __builtin_amdgcn_perm(LUT_LO, LUT_HI, __builtin_amdgcn_perm(f1, e1, 0x03070206u));
  bf16x8 r;
  __builtin_memcpy(&r, o, sizeof(r));
  return r;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16 codes (one 32-bit word) -> the A operands of two MFMA K steps, in kperm16
// order.  The operand is not the dosage y in {2, 1, 0} but y' = 2 - 2 y in
// {-2, 0, 2}: every such BF16 value has a zero low byte, so a code becomes its
// BF16 value with one table lookup (v_perm over the high bytes) and one v_perm
// that interleaves zero bytes.  The epilogues recover X B = colsum(B) - X' B / 2
// exactly (their `affine` flag).  Spread word x_j = (w >> 2j) & 0x03030303 holds
// codes j, j+4, j+8, j+12, one per byte.
__device__ inline void decode16(uint32_t w, bf16x8& a0, bf16x8& a1) {
  // high bytes of bf16(2 - 2 y) by code: 00 (y 2) -> -2, 01 (missing, y 0) -> 2,
  // 10 (y 1) -> 0, 11 (y 0) -> 2
  constexpr uint32_t LUT = 0x400040C0u;
  uint32_t o[8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t x = (w >> (2 * j)) & 0x03030303u;
    const uint32_t h = __builtin_amdgcn_perm(0u, LUT, x);   // high bytes of codes j, j+4, j+8, j+12
    o[2 * j] = __builtin_amdgcn_perm(0u, h, 0x010c000cu);       // [0, h0, 0, h1]
    o[2 * j + 1] = __builtin_amdgcn_perm(0u, h, 0x030c020cu);   // [0, h2, 0, h3]
  }
  __builtin_memcpy(&a0, o, sizeof(a0));
  __builtin_memcpy(&a1, o + 4, sizeof(a1));
}

// Block = 4 waves, 256 rows (64 per wave = 4 fragments of 16).  K advances in
// chunks of 256 codes (64 bytes per row); lane (r, g) holds bytes [16g, 16g+16)
// of its row's chunk, i.e. codes [64g, 64g+64), and K-step s (0..7) feeds the
// MFMA with codes 64g+8s..64g+8s+7.  The B operand is staged per chunk in LDS,
// rows padded to 264 bf16 so the 16 column-lanes of a ds_read_b128 hit
// distinct banks.
// Accumulators to the split-K partial: every column, or with g.fold the hi and
// lo columns of each output column summed (the lo column of c sits fold
// columns on, i.e. in tile (c + fold) / 16 at lane (c + fold) % 16: for
// fold % 16 == 8 a rotation by 8 lanes inside each 16-lane DPP row).
__device__ inline float ror8_row(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x128, 0xf, 0xf, false));
}

// The folded store with the tile offset TB = fold / 16 and HALF = fold % 16
// == 8 as constants, so every accumulator index is static (a runtime index
// would move the accumulators to scratch memory).
template <int NT, int TB, bool HALF>
__device__ inline void store_acc_fold(const PackedGemm& g, float* o, uint32_t row0, const f32x4 (&acc)[4][NT]) {
  constexpr int NTO = HALF ? TB + 1 : TB;  // output tiles: ceil(fold / 16)
  const int lane = threadIdx.x & 63, L = lane & 15;
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t r = row0 + 16 * f + 4 * (lane >> 4) + i;
#pragma unroll
      for (int t = 0; t < NTO && t < NT; ++t) {
        const int t1 = t + TB < NT ? t + TB : NT - 1, t2 = t + TB + 1 < NT ? t + TB + 1 : NT - 1;
        float lo;
        if constexpr (HALF) {  // every lane of the wave runs the rotations
          const float a = ror8_row(acc[f][t1][i]), b = ror8_row(acc[f][t2][i]);
          lo = L < 8 ? a : b;
        } else {
          lo = acc[f][t1][i];
        }
        if (r < g.nrows) o[(uint64_t)r * g.ldo + t * 16 + L] = acc[f][t][i] + lo;
      }
    }
}

template <int NT, bool FOLD>
__device__ inline void store_acc(const PackedGemm& g, float* o, uint32_t row0, const f32x4 (&acc)[4][NT]) {
  const int lane = threadIdx.x & 63, L = lane & 15;
  if constexpr (FOLD) {
    if (g.fold) {
      switch (g.fold >> 3) {  // fold / 8: TB = fold / 16, HALF = odd
#define PCAONE_FOLD(K)                                                                  \
  case K:                                                                               \
    if constexpr (K / 2 + (K & 1) <= NT) store_acc_fold<NT, K / 2, (K & 1) != 0>(g, o, row0, acc); \
    return;
        PCAONE_FOLD(1) PCAONE_FOLD(2) PCAONE_FOLD(3) PCAONE_FOLD(4) PCAONE_FOLD(5) PCAONE_FOLD(6) PCAONE_FOLD(7)
        PCAONE_FOLD(8)
#undef PCAONE_FOLD
        default: break;
      }
    }
  }
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t r = row0 + 16 * f + 4 * (lane >> 4) + i;
      if (r < g.nrows) {
#pragma unroll
        for (int t = 0; t < NT; ++t) o[(uint64_t)r * g.ldo + t * 16 + L] = acc[f][t][i];
      }
    }
}

constexpr int kBStride = 264;

template <int NT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_packed_gemm_bf16(PackedGemm g) {
  constexpr int NTH = 64 * WAVES;
  // uint4 of the B chunk per thread; held in registers one chunk ahead when small
  constexpr int BV = NT * 16 * 32 / NTH;
  constexpr bool BPRE = BV <= 10;
  static_assert(NT * 16 * 32 % NTH == 0, "B chunk must split evenly over the block");
  __shared__ __attribute__((aligned(16))) __bf16 bl[NT * 16 * kBStride];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t z = blockIdx.z, split = blockIdx.y;
  const uint8_t* A = g.A + (uint64_t)z * g.a_zstride;
  const __bf16* B = reinterpret_cast<const __bf16*>(g.B) + (uint64_t)z * g.b_zstride;
  float* out = reinterpret_cast<float*>(g.out) + (uint64_t)z * g.o_zstride;
  const uint32_t row0 = blockIdx.x * NTH + wave * 64;
  const uint32_t c0 = split * g.chunks_per_split;
  const uint32_t c1 = min(c0 + g.chunks_per_split, g.kchunks);

  const uint8_t* arow[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    uint32_t r = row0 + 16 * f + (lane & 15);
    r = min(r, g.nrows - 1);
    arow[f] = A + (uint64_t)r * g.lda + 16 * (lane >> 4);
  }
  f32x4 acc[4][NT];
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[f][t] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int bn = lane & 15, bg = lane >> 4;
  // the next chunk's A bytes (and B tile, when BPRE) travel in registers while
  // the current chunk computes; B goes through LDS (two barriers per chunk)
  uint4 an[4];
  u32x4 bn4[BPRE ? BV : 1];
#define PCAONE_LOAD_A(CH)                                                                                \
  do {                                                                                                   \
    _Pragma("unroll") for (int f = 0; f < 4; ++f) an[f] =                                                \
        *reinterpret_cast<const uint4*>(arow[f] + (uint64_t)(CH) * 64);                                  \
  } while (0)
#define PCAONE_B_SRC(CH, V)                                                                              \
  (*reinterpret_cast<const u32x4*>(B + (uint64_t)((threadIdx.x + NTH * (V)) >> 5) * g.ldb +              \
                                   (uint64_t)(CH) * 256 + ((threadIdx.x + NTH * (V)) & 31) * 8))
#define PCAONE_B_DST(V) \
  (*reinterpret_cast<u32x4*>(bl + ((threadIdx.x + NTH * (V)) >> 5) * kBStride + ((threadIdx.x + NTH * (V)) & 31) * 8))
  if (c0 < c1) {
    PCAONE_LOAD_A(c0);
    if constexpr (BPRE) {
#pragma unroll
      for (int v = 0; v < BV; ++v) bn4[v] = PCAONE_B_SRC(c0, v);
    }
  }
  for (uint32_t c = c0; c < c1; ++c) {
    __syncthreads();
    if constexpr (BPRE) {
#pragma unroll
      for (int v = 0; v < BV; ++v) PCAONE_B_DST(v) = bn4[v];
    } else {
#pragma unroll
      for (int v = 0; v < BV; ++v) PCAONE_B_DST(v) = PCAONE_B_SRC(c, v);
    }
    uint4 ar[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) ar[f] = an[f];
    __syncthreads();
    if (c + 1 < c1) {
      PCAONE_LOAD_A(c + 1);
      if constexpr (BPRE) {
#pragma unroll
        for (int v = 0; v < BV; ++v) bn4[v] = PCAONE_B_SRC(c + 1, v);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // 32-bit word q of each fragment: K steps 2q, 2q + 1
      bf16x8 b0[NT], b1[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        b0[t] = *reinterpret_cast<const bf16x8*>(bl + (t * 16 + bn) * kBStride + 64 * bg + 16 * q);
        b1[t] = *reinterpret_cast<const bf16x8*>(bl + (t * 16 + bn) * kBStride + 64 * bg + 16 * q + 8);
      }
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const uint32_t w = (q == 0) ? ar[f].x : (q == 1) ? ar[f].y : (q == 2) ? ar[f].z : ar[f].w;
        bf16x8 a0, a1;
        decode16(w, a0, a1);
        // NT independent accumulators between two updates of the same one
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0[t], acc[f][t], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1[t], acc[f][t], 0, 0, 0);
      }
    }
  }
#undef PCAONE_LOAD_A
#undef PCAONE_B_SRC
#undef PCAONE_B_DST
  store_acc<NT, false>(g, out + (uint64_t)split * g.nrows * g.ldo, row0, acc);
}

// Long row ranges: 8 waves x 64 rows per block, the B chunk streamed straight
// from global memory into a double-buffered LDS tile (global_load_lds_dwordx4:
// lane L of a load fills bytes [16 L, 16 L + 16) of a 1 KB span, two B rows),
// so chunk c + 1 lands while chunk c computes and one barrier per chunk
// suffices.  B rows are unpadded (512 bytes); the 16-byte segment s of row n
// sits at slot s ^ (n & 7), which spreads the 16 rows a ds_read_b128 touches
// over distinct banks.
template <int NT>
__global__ __launch_bounds__(512) void k_packed_gemm_bf16_stream(PackedGemm g) {
  constexpr int ROWS_B = NT * 16;  // B rows per chunk (512 bytes each)
  __shared__ __attribute__((aligned(16))) __bf16 bs[2][ROWS_B * 256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t z = blockIdx.z, split = blockIdx.y;
  const uint8_t* A = g.A + (uint64_t)z * g.a_zstride;
  const __bf16* B = reinterpret_cast<const __bf16*>(g.B) + (uint64_t)z * g.b_zstride;
  float* out = reinterpret_cast<float*>(g.out) + (uint64_t)z * g.o_zstride;
  const uint32_t row0 = blockIdx.x * 512 + wave * 64;
  const uint32_t c0 = split * g.chunks_per_split;
  const uint32_t c1 = min(c0 + g.chunks_per_split, g.kchunks);

  const uint8_t* arow[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    uint32_t r = row0 + 16 * f + (lane & 15);
    r = min(r, g.nrows - 1);
    arow[f] = A + (uint64_t)r * g.lda + 16 * (lane >> 4);
  }
  f32x4 acc[4][NT];
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[f][t] = f32x4{0.f, 0.f, 0.f, 0.f};

  // this lane's share of a B chunk: load j fills rows 2 i, 2 i + 1 (i = wave + 8 j)
  const int half = lane >> 5, sp = lane & 31;
#define PCAONE_ISSUE_B(CH, BUF)                                                                          \
  do {                                                                                                   \
    _Pragma("unroll") for (int j = 0; j < NT; ++j) {                                                     \
      const int i = wave + 8 * j;                                                                        \
      const int row = 2 * i + half;                                                                      \
      const __bf16* src = B + (uint64_t)row * g.ldb + (uint64_t)(CH) * 256 + (sp ^ (row & 7)) * 8;       \
      __builtin_amdgcn_global_load_lds((const void*)src,                                                 \
                                       (__attribute__((address_space(3))) void*)(&bs[BUF][2 * i * 256]), \
                                       16, 0, 0);                                                        \
    }                                                                                                    \
  } while (0)
  uint4 an[4];
  if (c0 < c1) {
    PCAONE_ISSUE_B(c0, 0);
#pragma unroll
    for (int f = 0; f < 4; ++f) an[f] = *reinterpret_cast<const uint4*>(arow[f] + (uint64_t)c0 * 64);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  const int bn = lane & 15, bg = lane >> 4;
  for (uint32_t c = c0; c < c1; ++c) {
    const int cur = (int)((c - c0) & 1);
    uint4 ar[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) ar[f] = an[f];
    if (c + 1 < c1) {
      PCAONE_ISSUE_B(c + 1, cur ^ 1);
#pragma unroll
      for (int f = 0; f < 4; ++f) an[f] = *reinterpret_cast<const uint4*>(arow[f] + (uint64_t)(c + 1) * 64);
    }
    const __bf16* bl = bs[cur];
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // 32-bit word q of each fragment: K steps 2q, 2q + 1
      const int sg = 8 * bg + 2 * q;  // 16-byte segment of positions 64 bg + 16 q
      bf16x8 b0[NT], b1[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int row = t * 16 + bn;
        b0[t] = *reinterpret_cast<const bf16x8*>(bl + row * 256 + ((sg ^ (row & 7)) * 8));
        b1[t] = *reinterpret_cast<const bf16x8*>(bl + row * 256 + (((sg + 1) ^ (row & 7)) * 8));
      }
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const uint32_t w = (q == 0) ? ar[f].x : (q == 1) ? ar[f].y : (q == 2) ? ar[f].z : ar[f].w;
        bf16x8 a0, a1;
        decode16(w, a0, a1);
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0[t], acc[f][t], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1[t], acc[f][t], 0, 0, 0);
      }
    }
    // this wave's loads of chunk c + 1 are in LDS; after the barrier every
    // wave's are, and nobody reads buffer `cur` any more
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
#undef PCAONE_ISSUE_B
  store_acc<NT, true>(g, out + (uint64_t)split * g.nrows * g.ldo, row0, acc);
}

// Streamed 8-wave blocks (512 rows, half the B traffic per row) wherever the
// rows fill a few blocks; measured on the 10,000 x 500,000 winSVD, they beat
// the 4-wave blocks from one window's H product (10,000 rows, 25 -> 21 us) up,
// with the split-K chosen by resident rounds.  4-wave blocks for smaller ones.
bool packed_bf16_streams(uint64_t nrows, uint32_t nz) { return nrows * nz >= 4096; }

// One launch holds at most 8 column tiles in its accumulators (128 FP32
// registers per lane at two waves per SIMD); wider operands (l > 64) run as
// column blocks of at most 8 tiles, each reading the packed rows again, with B
// and the output offset to the block's first column.
constexpr int kMaxTilesPerLaunch = 8;

void launch_packed_gemm_bf16(const PackedGemm& g, int nt16, hipStream_t s) {
  const bool big = packed_bf16_streams(g.nrows, g.nz);
  PCAONE_REQUIRE(nt16 >= 1 && nt16 <= 2 * kMaxTilesPerLaunch, -1,
                 "split-BF16 contraction supports at most 256 operand columns (l <= 128)");
  PCAONE_REQUIRE(!g.fold || (big && nt16 <= kMaxTilesPerLaunch && (g.fold % 16 == 0 || g.fold % 16 == 8) &&
                             2 * g.fold <= 16u * (uint32_t)nt16 && g.ldo >= (g.fold + 15) / 16 * 16),
                 -1, "split contraction: fold needs the streamed kernel and a multiple of 8 within the operand");
  PCAONE_REQUIRE(g.ldo >= 16ull * (uint64_t)nt16 || g.fold, -1, "split contraction: output rows too short");
  dim3 grid((uint32_t)ceil_div(g.nrows, big ? 512 : 256), g.nsplit, g.nz);
  const int nblk = (nt16 + kMaxTilesPerLaunch - 1) / kMaxTilesPerLaunch;
  const int per = (nt16 + nblk - 1) / nblk;
  for (int t0 = 0; t0 < nt16; t0 += per) {
    PackedGemm gb = g;
    const int nt = std::min(per, nt16 - t0);
    gb.B = reinterpret_cast<const __bf16*>(g.B) + (uint64_t)t0 * 16 * g.ldb;
    gb.out = reinterpret_cast<float*>(g.out) + (uint64_t)t0 * 16;
    switch (nt) {
#define PCAONE_NT(n)                                                                     \
  case n:                                                                                \
    if (big)                                                                             \
      hipLaunchKernelGGL((k_packed_gemm_bf16_stream<n>), grid, dim3(512), 0, s, gb);     \
    else                                                                                 \
      hipLaunchKernelGGL((k_packed_gemm_bf16<n, 4>), grid, dim3(256), 0, s, gb);         \
    break;
      PCAONE_NT(1) PCAONE_NT(2) PCAONE_NT(3) PCAONE_NT(4) PCAONE_NT(5) PCAONE_NT(6) PCAONE_NT(7) PCAONE_NT(8)
#undef PCAONE_NT
      default: throw Error(-1, "split-BF16 contraction: bad column block");
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 / FP64 vector-FMA contraction (plain drop-in and FP64 arm)
// ---------------------------------------------------------------------------
// Block = 64 * ncg threads (wave w owns columns [8w, 8w+8)); each thread owns 8
// rows (tr + 64 i) x 8 columns.  K advances 64 codes at a time; B is staged in
// LDS transposed to [64][8*ncg].
template <class T>
__global__ __launch_bounds__(1024) void k_packed_gemm_valu(PackedGemm g, int ncols) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  T* bl = reinterpret_cast<T*>(smem_raw);
  const int ncg = blockDim.x >> 6;
  const int lc = ncg * 8;
  const int tr = threadIdx.x & 63, tc = threadIdx.x >> 6;
  const uint32_t z = blockIdx.z, split = blockIdx.y;
  const uint8_t* A = g.A + (uint64_t)z * g.a_zstride;
  const T* B = reinterpret_cast<const T*>(g.B) + (uint64_t)z * g.b_zstride;
  T* out = reinterpret_cast<T*>(g.out) + (uint64_t)z * g.o_zstride;
  const uint32_t rbase = blockIdx.x * 512 + tr;
  const uint32_t c0 = split * g.chunks_per_split;
  const uint32_t c1 = min(c0 + g.chunks_per_split, g.kchunks);
  const uint8_t* arow[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) arow[i] = A + (uint64_t)min(rbase + 64 * i, g.nrows - 1) * g.lda;
  T acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = T(0);

  for (uint32_t c = c0; c < c1; ++c) {
    for (int sub = 0; sub < 4; ++sub) {
      const uint64_t kb = (uint64_t)c * 256 + sub * 64;
      uint32_t a[8][4];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint4 v = *reinterpret_cast<const uint4*>(arow[i] + (uint64_t)c * 64 + sub * 16);
        const uint32_t xs[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {  // 2-bit codes -> 2-bit dosages
          const uint32_t lo = xs[q] & 0x55555555u, hi = (xs[q] >> 1) & 0x55555555u;
          const uint32_t y1 = ~(lo | hi) & 0x55555555u;  // code 00 -> 2
          const uint32_t y0 = hi & ~lo;                  // code 10 -> 1
          a[i][q] = y0 | (y1 << 1);
        }
      }
      __syncthreads();
      for (int it = threadIdx.x; it < 64 * lc; it += blockDim.x) {
        const int kk = it & 63, col = it >> 6;
        bl[kk * lc + col] = (col < ncols) ? B[(uint64_t)col * g.ldb + kb + kk] : T(0);
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < 64; ++kk) {
        T bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = bl[kk * lc + tc * 8 + j];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const T y = (T)((a[i][kk >> 4] >> (2 * (kk & 15))) & 3u);
#pragma unroll
          for (int j = 0; j < 8; ++j) acc[i][j] = fma(y, bv[j], acc[i][j]);
        }
      }
    }
  }
  T* o = out + (uint64_t)split * g.nrows * g.ldo;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
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

template <class T>
static void launch_valu(const PackedGemm& g, int ncols, hipStream_t s) {
  const int ncg = (ncols + 7) / 8;
  PCAONE_REQUIRE(ncg >= 1 && ncg <= 16, -1, "vector contraction supports 1..128 columns");
  dim3 grid((uint32_t)ceil_div(g.nrows, 512), g.nsplit, g.nz);
  const size_t shm = (size_t)64 * ncg * 8 * sizeof(T);
  hipLaunchKernelGGL(k_packed_gemm_valu<T>, grid, dim3(64 * ncg), shm, s, g, ncols);
}

void launch_packed_gemm_f32(const PackedGemm& g, int ncols, hipStream_t s) { launch_valu<float>(g, ncols, s); }
void launch_packed_gemm_f64(const PackedGemm& g, int ncols, hipStream_t s) { launch_valu<double>(g, ncols, s); }

// ---------------------------------------------------------------------------
// epilogues
// ---------------------------------------------------------------------------

// Grid (row chunk of 32, window).  Besides G, g'' and the H operand the
// epilogue writes the chunk's partial column sums of g'' (s2p), which the
// window-sum job (side_jobs.hpp) reduces into S''.  The operand (column-major
// per window) is staged through LDS so its stores run along the window's SNPs.
constexpr uint32_t kFinRows = 32;  // two 512-thread finalize blocks fit one CU

// bh[w][c][k0 + p] for the operand positions p < 64 of a chunk, from
// tile[q][c] (q < nq; rows past nq are window padding and get 0).  Each item
// is four consecutive positions of one column, stored as one 8-byte (BF16
// halves), 16-byte (FP32) or 32-byte (FP64) write; with kperm, position p
// holds row kperm16(p) (a 4 x 4 transpose inside each 16-position group).
template <int LD>
__device__ inline void store_operand_tile(const double* tile, uint32_t nq, uint32_t l, uint32_t w, uint32_t k0,
                                          int bprec, void* bh, uint32_t bcols, uint32_t lo_off, uint64_t kw,
                                          int kperm) {
  for (uint32_t it = threadIdx.x; it < l * (kFinRows / 4); it += blockDim.x) {
    const uint32_t c = it / (kFinRows / 4), p0 = 4 * (it % (kFinRows / 4));
    double v[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const uint32_t q = kperm ? (uint32_t)kperm16(p0 + b) : p0 + b;
      v[b] = q < nq ? tile[q * LD + c] : 0.0;
    }
    const uint64_t o = ((uint64_t)w * bcols + c) * kw + k0 + p0;  // k0 is a multiple of kFinRows
    if (bprec == PREC_FP64) {
      *reinterpret_cast<double4*>(reinterpret_cast<double*>(bh) + o) = make_double4(v[0], v[1], v[2], v[3]);
    } else if (bprec == PREC_FP32) {
      *reinterpret_cast<float4*>(reinterpret_cast<float*>(bh) + o) =
          make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
    } else {
      __bf16 hi[4], lo[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        hi[b] = (__bf16)(float)v[b];
        lo[b] = (__bf16)(float)(v[b] - (double)(float)hi[b]);
      }
      __bf16* bb = reinterpret_cast<__bf16*>(bh);
      uint2 h2, l2;
      __builtin_memcpy(&h2, hi, 8);
      __builtin_memcpy(&l2, lo, 8);
      *reinterpret_cast<uint2*>(bb + o) = h2;
      *reinterpret_cast<uint2*>(bb + ((uint64_t)w * bcols + lo_off + c) * kw + k0 + p0) = l2;
    }
  }
}

// sum over a missing-entry list of v[idx[p] * l + c] (and of val[p] times it),
// four independent loads at a time
__device__ inline void miss_sums(const int64_t* ptr, const int32_t* idx, uint64_t li, const double* v, uint32_t l,
                                 uint32_t c, const double* val, double& corr, double& corr_e) {
  int64_t p = ptr[li];
  const int64_t pe = ptr[li + 1];
  for (; p + 4 <= pe; p += 4) {
    const double o0 = v[(uint64_t)idx[p] * l + c], o1 = v[(uint64_t)idx[p + 1] * l + c];
    const double o2 = v[(uint64_t)idx[p + 2] * l + c], o3 = v[(uint64_t)idx[p + 3] * l + c];
    corr += (o0 + o1) + (o2 + o3);
    if (val) corr_e += (val[p] * o0 + val[p + 1] * o1) + (val[p + 2] * o2 + val[p + 3] * o3);
  }
  for (; p < pe; ++p) {
    const double o = v[(uint64_t)idx[p] * l + c];
    corr += o;
    if (val) corr_e += val[p] * o;
  }
}

__device__ inline void store_g2(const FinalizeG& f, uint64_t i, double v) {
  if (f.g2_f32)
    reinterpret_cast<float*>(f.g2)[i] = (float)v;
  else
    reinterpret_cast<double*>(f.g2)[i] = v;
}

// 512 threads over the chunk's rows x l columns in flat (row, column) order,
// so every lane has an element and the partial, correction, G and g''
// accesses run along rows; each thread's elements (at most 4 for l <= 64, 8
// for l <= 128: LW is the width class) have their loads in flight together,
// four split-K partials at a time.
constexpr int kFinThreads = 512;
static_assert(kFinThreads / 64 == 8, "finalize_G sums its column slices eight at a time");

// KS split-K slices per round (1 when the launch has one, so no load repeats);
// LO: the partials hold separate hi and lo columns.
template <typename P, int KS, bool LO, int LW>
__global__ __launch_bounds__(kFinThreads) void k_finalize_G(FinalizeG f) {
  constexpr int kTileLd = LW + 1;
  constexpr int kFinEl = kFinRows * LW / kFinThreads;
  __shared__ double tile[kFinRows * kTileLd];  // g' = d G, the H operand
  __shared__ double acct[kFinRows * kTileLd];  // the column-sum term (g'', or g'' - g' when affine)
  __shared__ double sh[kFinThreads / 64][LW];
  const uint32_t t = threadIdx.x, l = f.l;
  const uint32_t w = f.w0 + blockIdx.y, chunk = blockIdx.x;
  const uint32_t wstart = w * f.blocksize;
  const uint32_t wend = min(wstart + f.blocksize, f.nsnps);
  const uint32_t cb = wstart + chunk * kFinRows;
  const uint32_t nq = cb < wend ? min(kFinRows, wend - cb) : 0;
  const uint32_t n = nq * l;
  const uint64_t nrow = f.r1 - f.r0;
  const P* __restrict__ part = reinterpret_cast<const P*>(f.part);
  const double* __restrict__ colsum = f.colsum;
  const P* __restrict__ corrp = reinterpret_cast<const P*>(f.corr);
  const double* __restrict__ F = f.F;
  const double* __restrict__ d = f.d;
  const uint64_t sstride = nrow * f.ldp;
  // Element offsets inside one split's partial (32 bits: nrow * ldp < 2^32 is
  // checked at launch).  Every load is unconditional, from a row and split
  // clamped into range, and its value is dropped afterwards when out of range:
  // the loads then issue together instead of one branch at a time.
  uint32_t qe[kFinEl], ce[kFinEl], off[kFinEl];
  double raw[kFinEl], cs[kFinEl], Fi[kFinEl], di[kFinEl], sg[kFinEl];
  P corr[kFinEl];  // widened where used
  const uint32_t rlast = f.r1 - 1;
#pragma unroll
  for (int u = 0; u < kFinEl; ++u) {
    const uint32_t e = t + kFinThreads * u;
    const bool ok = e < n;
    qe[u] = ok ? e / l : 0;
    ce[u] = ok ? e - qe[u] * l : 0;
    const uint32_t r = min(cb + qe[u], rlast);
    off[u] = (r - f.r0) * (uint32_t)f.ldp + ce[u];
    raw[u] = 0.0;
    cs[u] = colsum[ce[u]];
    sg[u] = f.fsign ? f.fsign[ce[u]] : 1.0;
    Fi[u] = F[r];
    di[u] = d[r];
  }
  if (corrp) {
#pragma unroll
    for (int u = 0; u < kFinEl; ++u) corr[u] = corrp[(uint64_t)min(cb + qe[u], rlast) * l + ce[u]];
  } else {
#pragma unroll
    for (int u = 0; u < kFinEl; ++u) corr[u] = P(0);
  }
  const uint32_t lo = LO ? l : 0;
  for (uint32_t s0 = 0; s0 < f.nsplit; s0 += KS) {
    P x[KS][kFinEl], y[KS][kFinEl];
#pragma unroll
    for (int j = 0; j < KS; ++j) {
      const P* __restrict__ ps = part + (uint64_t)min(s0 + j, f.nsplit - 1) * sstride;
#pragma unroll
      for (int u = 0; u < kFinEl; ++u) {
        x[j][u] = ps[off[u]];
        y[j][u] = LO ? ps[off[u] + lo] : P(0);
      }
    }
#pragma unroll
    for (int j = 0; j < KS; ++j)
      if (s0 + j < f.nsplit)
#pragma unroll
        for (int u = 0; u < kFinEl; ++u) raw[u] += (double)x[j][u] + (LO ? (double)y[j][u] : 0.0);
  }
#pragma unroll
  for (int u = 0; u < kFinEl; ++u) {
    if (t + kFinThreads * u >= n) continue;
    const uint32_t q = qe[u], c = ce[u], r = cb + q;
    const double rv = f.affine ? cs[u] - 0.5 * raw[u] : raw[u];  // the contraction ran on 2 - 2 y
    double corr_e = 0.0;
    if (f.mval) {
      double unused = 0.0;
      miss_sums(f.miss_ptr, f.miss_idx, r, f.omega, l, c, f.mval, unused, corr_e);
    }
    // every term is linear in Omega's column c: its sign applies to the whole
    const double Gv = sg[u] * (di[u] * (f.hf * (rv - 2.0 * Fi[u] * (cs[u] - (double)corr[u])) + corr_e));
    f.G[(uint64_t)r * l + c] = Gv;
    const double gp = di[u] * Gv;
    const double g2 = 2.0 * Fi[u] * gp;
    store_g2(f, (uint64_t)r * l + c, g2);
    tile[q * kTileLd + c] = gp;
    acct[q * kTileLd + c] = f.affine ? g2 - gp : g2;
  }
  __syncthreads();
  {  // column sums of the chunk: 8 row slices, then the slices in a fixed order
    const uint32_t sl = t >> 6;
    for (uint32_t c = t & 63; c < (uint32_t)LW; c += 64) {
      double a = 0.0;
      if (c < l)
        for (uint32_t q = sl; q < nq; q += kFinThreads / 64) a += acct[q * kTileLd + c];
      sh[sl][c] = a;
    }
  }
  __syncthreads();
  if (t < l) {
    double v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = sh[i][t];
    f.s2p[((uint64_t)w * gridDim.x + chunk) * l + t] = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
  }
  store_operand_tile<kTileLd>(tile, nq, l, w, cb - wstart, f.bprec, f.bh, f.bcols, f.lo_off, f.kw, f.kperm);
}

template <typename P, int LW>
void launch_finalize_G_t(const FinalizeG& f, dim3 grid, hipStream_t s) {
  const bool one = f.nsplit == 1, lo = f.split_cols != 0;
  if (one && !lo)
    hipLaunchKernelGGL((k_finalize_G<P, 1, false, LW>), grid, dim3(kFinThreads), 0, s, f);
  else if (one)
    hipLaunchKernelGGL((k_finalize_G<P, 1, true, LW>), grid, dim3(kFinThreads), 0, s, f);
  else if (!lo)
    hipLaunchKernelGGL((k_finalize_G<P, 4, false, LW>), grid, dim3(kFinThreads), 0, s, f);
  else
    hipLaunchKernelGGL((k_finalize_G<P, 4, true, LW>), grid, dim3(kFinThreads), 0, s, f);
}

void launch_finalize_G(const FinalizeG& f, hipStream_t s) {
  PCAONE_REQUIRE(f.l >= 1 && f.l <= 128, -1, "finalize: l must be in [1, 128]");
  PCAONE_REQUIRE(f.kw % kFinRows == 0, -1, "finalize: window operand rows must be whole chunks");
  PCAONE_REQUIRE((uint64_t)(f.r1 - f.r0) * f.ldp < (1ull << 32), -1, "finalize: partial block too large");
  if (f.w1 <= f.w0) return;
  dim3 grid((uint32_t)ceil_div(f.blocksize, kFinRows), f.w1 - f.w0);
  if (f.l <= 64) {
    if (f.part_is_f64)
      launch_finalize_G_t<double, 64>(f, grid, s);
    else
      launch_finalize_G_t<float, 64>(f, grid, s);
  } else {
    if (f.part_is_f64)
      launch_finalize_G_t<double, 128>(f, grid, s);
    else
      launch_finalize_G_t<float, 128>(f, grid, s);
  }
}

uint32_t finalize_chunks(uint32_t blocksize) { return (uint32_t)ceil_div(blocksize, kFinRows); }

// H-contraction operand from a given G (rows [r0, r1) of windows [w0, w1)): the
// output half of finalize_G for an externally supplied G, so X G can run
// without a preceding X^T Omega (Gram mat-vec X^T X x starts here).
template <int LW>
__global__ __launch_bounds__(256) void k_operand_from_G(FinalizeG f) {
  constexpr int kTileLd = LW + 1;
  __shared__ double sh[4][LW];
  __shared__ double tile[kFinRows * kTileLd];
  const uint32_t t = threadIdx.x, rs = t >> 6;
  const uint32_t w = f.w0 + blockIdx.y, chunk = blockIdx.x;
  const uint32_t wstart = w * f.blocksize;
  const uint32_t wend = min(wstart + f.blocksize, f.nsnps);
  const uint32_t cb = wstart + chunk * kFinRows;
  const uint32_t nq = cb < wend ? min(kFinRows, wend - cb) : 0;
  for (uint32_t c = t & 63; c < (uint32_t)LW; c += 64) {
    double acc2 = 0.0;
    if (c < f.l) {
      for (uint32_t q = rs; q < nq; q += 4) {
        const uint32_t r = cb + q;
        const double gp = f.d[r] * f.G[(uint64_t)r * f.l + c];
        const double g2 = 2.0 * f.F[r] * gp;
        store_g2(f, (uint64_t)r * f.l + c, g2);
        acc2 += f.affine ? g2 - gp : g2;
        tile[q * kTileLd + c] = gp;
      }
    }
    sh[rs][c] = acc2;
  }
  __syncthreads();
  if (t < f.l)
    f.s2p[((uint64_t)w * gridDim.x + chunk) * f.l + t] = (sh[0][t] + sh[1][t]) + (sh[2][t] + sh[3][t]);
  store_operand_tile<kTileLd>(tile, nq, f.l, w, cb - wstart, f.bprec, f.bh, f.bcols, f.lo_off, f.kw, f.kperm);
}

void launch_operand_from_G(const FinalizeG& f, hipStream_t s) {
  PCAONE_REQUIRE(f.l >= 1 && f.l <= 128, -1, "operand: l must be in [1, 128]");
  if (f.w1 <= f.w0) return;
  dim3 grid((uint32_t)ceil_div(f.blocksize, kFinRows), f.w1 - f.w0);
  if (f.l <= 64)
    hipLaunchKernelGGL(k_operand_from_G<64>, grid, dim3(256), 0, s, f);
  else
    hipLaunchKernelGGL(k_operand_from_G<128>, grid, dim3(256), 0, s, f);
}

// S2[w][c] = sum of the chunk partials of window w: 16 slices of chunks (all
// of a slice's loads in flight), combined in a fixed order
void launch_window_colsum(const double* s2p, uint32_t l, uint32_t blocksize, uint32_t w0, uint32_t w1, double* S2,
                          hipStream_t s) {
  PCAONE_REQUIRE(l >= 1 && l <= 128, -1, "l out of range");
  if (w1 <= w0) return;
  launch_side_jobs(window_colsum_job(s2p, l, blocksize, w0, w1, S2), s);
}

SideJobs window_colsum_job(const double* s2p, uint32_t l, uint32_t blocksize, uint32_t w0, uint32_t w1, double* S2) {
  SideJobs j;
  j.s2p = s2p;
  j.nchunk = finalize_chunks(blocksize);
  j.w0 = w0;
  j.w1 = w1;
  j.l2 = l;
  j.S2 = S2;
  return j;
}

// thread (sample j, column c); the windows' contributions are summed in
// registers per destination and added once.  Four windows' partials, KS
// split-K slices each (KS = 1 when the launch has one slice, so no load is
// repeated), and their corrections load together from clamped indices
// (unconditional loads issue together; out-of-range values drop out).  LO:
// the partials hold separate hi and lo columns.
template <typename P, int KS, bool LO>
__global__ __launch_bounds__(256) void k_finalize_H(FinalizeH f) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (uint64_t)f.nsamples * f.l) return;
  const uint32_t j = (uint32_t)(gid / f.l), c = (uint32_t)(gid % f.l);
  const P* __restrict__ part = reinterpret_cast<const P*>(f.part);
  const P* __restrict__ corrp = reinterpret_cast<const P*>(f.corr);
  const double* __restrict__ S2 = f.S2;
  const uint64_t lo = LO ? f.l : 0;
  double ha = 0.0, hb = 0.0;
  bool any_a = false, any_b = false;
  for (uint32_t w0 = f.w0; w0 < f.w1; w0 += 4) {
    double raw[4], s2[4];
    P corr[4];  // widened where used: a conversion here would wait for each load in turn
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t w = min(w0 + u, f.w1 - 1);
      raw[u] = 0.0;
      s2[u] = S2[(uint64_t)w * f.l + c];
    }
    if (corrp) {
#pragma unroll
      for (int u = 0; u < 4; ++u) corr[u] = corrp[((uint64_t)min(w0 + u, f.w1 - 1) * f.nsamples + j) * f.l + c];
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u) corr[u] = P(0);
    }
    for (uint32_t s0 = 0; s0 < f.nsplit; s0 += KS) {
      P x[4][KS], y[4][KS];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int k = 0; k < KS; ++k) {
          const uint32_t wz = min(w0 + u, f.w1 - 1) - f.wpart0, sk = min(s0 + k, f.nsplit - 1);
          const uint64_t base = (((uint64_t)wz * f.nsplit + sk) * f.part_rows + j) * f.ldp + c;
          x[u][k] = part[base];
          y[u][k] = LO ? part[base + lo] : P(0);
        }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int k = 0; k < KS; ++k)
          if (s0 + k < f.nsplit) raw[u] += (double)x[u][k] + (LO ? (double)y[u][k] : 0.0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t w = w0 + u;
      if (w >= f.w1) continue;
      // the contraction ran on 2 - 2 y: X^T g' = sum g' - part / 2, and S2 already
      // holds sum (g'' - g') for this layout
      const double rv = f.affine ? -0.5 * raw[u] : raw[u];
      double corr_e = 0.0;
      if (f.mval) {
        const uint64_t li = (uint64_t)w * f.miss_rows + j;
        for (int64_t p = f.miss_ptr[li]; p < f.miss_ptr[li + 1]; ++p) {
          const uint64_t i = (uint64_t)f.miss_idx[p];
          corr_e += f.mval[f.mpos[p]] * f.d[i] * f.G[i * f.l + c];
        }
      }
      const double h = f.hf * (rv - (s2[u] - (double)corr[u])) + corr_e;
      if ((f.route_b >> (w - f.w0)) & 1ull) {
        hb += h;
        any_b = true;
      } else {
        ha += h;
        any_a = true;
      }
    }
  }
  if (any_a) f.Ha[(uint64_t)j * f.l + c] += ha;
  if (any_b) f.Hb[(uint64_t)j * f.l + c] += hb;
}

template <typename P>
void launch_finalize_H_t(const FinalizeH& f, dim3 grid, hipStream_t s) {
  const bool one = f.nsplit == 1, lo = f.split_cols != 0;
  if (one && !lo)
    hipLaunchKernelGGL((k_finalize_H<P, 1, false>), grid, dim3(256), 0, s, f);
  else if (one)
    hipLaunchKernelGGL((k_finalize_H<P, 1, true>), grid, dim3(256), 0, s, f);
  else if (!lo)
    hipLaunchKernelGGL((k_finalize_H<P, 4, false>), grid, dim3(256), 0, s, f);
  else
    hipLaunchKernelGGL((k_finalize_H<P, 4, true>), grid, dim3(256), 0, s, f);
}

void launch_finalize_H(const FinalizeH& f, hipStream_t s) {
  const uint64_t total = (uint64_t)f.nsamples * f.l;
  const dim3 grid((uint32_t)ceil_div(total, 256));
  if (f.part_is_f64)
    launch_finalize_H_t<double>(f, grid, s);
  else
    launch_finalize_H_t<float>(f, grid, s);
}

// one thread per SNP row: EMU value of each of its missing samples
__global__ void k_emu_update(const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx, uint32_t nsnps,
                             const double* __restrict__ U, const double* __restrict__ S,
                             const double* __restrict__ V, uint32_t k, const double* __restrict__ F,
                             double* __restrict__ val) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsnps) return;
  const double lo = -F[i], hi = 1.0 - F[i];
  for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
    const uint64_t j = (uint64_t)idx[p];
    double e = 0.0;
    for (uint32_t c = 0; c < k; ++c) e += U[j * k + c] * (S[c] * V[i * k + c]);
    val[p] = fmin(fmax(e, lo), hi);
  }
}

void launch_emu_update(const int64_t* miss_ptr, const int32_t* miss_idx, uint32_t nsnps, const double* U,
                       const double* S, const double* V, uint32_t k, const double* F, double* val, hipStream_t s) {
  PCAONE_REQUIRE(k >= 1 && k <= 128, -1, "EMU update supports k <= 128");
  hipLaunchKernelGGL(k_emu_update, dim3((uint32_t)ceil_div(nsnps, 256)), dim3(256), 0, s, miss_ptr, miss_idx,
                     nsnps, U, S, V, k, F, val);
}

// one thread per (window, sample) row of the window lists
__global__ void k_map_missing(const int64_t* __restrict__ wptr, const int32_t* __restrict__ widx, uint32_t nwin,
                              uint32_t npad, uint32_t nsamples, const int64_t* __restrict__ sptr,
                              const int32_t* __restrict__ sidx, int64_t* __restrict__ pos) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (uint64_t)nwin * npad) return;
  const int32_t j = (int32_t)(r % npad);
  for (int64_t p = wptr[r]; p < wptr[r + 1]; ++p) {
    const uint64_t i = (uint64_t)widx[p];
    int64_t lo = sptr[i], hi = sptr[i + 1] - 1;  // sample lists are ascending
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (sidx[mid] < j)
        lo = mid + 1;
      else
        hi = mid;
    }
    pos[p] = lo;
  }
  (void)nsamples;
}

void launch_map_missing(const int64_t* wptr, const int32_t* widx, uint32_t nwin, uint32_t npad,
                        uint32_t nsamples, const int64_t* sptr, const int32_t* sidx, int64_t* pos,
                        hipStream_t s) {
  const uint64_t rows = (uint64_t)nwin * npad;
  hipLaunchKernelGGL(k_map_missing, dim3((uint32_t)ceil_div(rows, 256)), dim3(256), 0, s, wptr, widx, nwin, npad,
                     nsamples, sptr, sidx, pos);
}

// One wave per list row, grid-stride.  The row's indices are loaded 64 at a
// time (one per lane) and handed out with lane shuffles.  Lane group grp
// (lpg = ceil(l / V) lanes, V consecutive columns per lane read as one vector
// load) takes entries grp, grp + ng, ..., two per round.  The sums run in the
// table's own type (FP32 tables: FP32 sums of FP32 values, a relative error
// of a few 1e-7 on each correction; FP64 tables: FP64), the groups' sums
// combine in a fixed order, and the row is stored in that type (one vector
// store per lane), which is the type the epilogues' partials have.  Measured on the 500,000-row missing lists of the
// 10,000-sample workload: the FP32 sums (half the shuffles, fewer registers,
// more waves resident), two entries per round and the grid stride take a
// launch from 387 to 257 us; four entries per round in flight take a
// 25 M-entry list from 296 to 273 us (500,000 rows) and 324 to 292 us
// (640,000 window rows).
// The first side_blocks(g.side) blocks run the side jobs instead.
template <class T, int V>
__global__ __launch_bounds__(256) void k_csr_gather(CsrGather g) {
  typedef T vec __attribute__((ext_vector_type(V)));
  constexpr int D = 4;  // entries per group per round
  const uint32_t nside = side_blocks(g.side);
  if (blockIdx.x < nside) {
    __shared__ double red[kSideLds];
    side_block(g.side, blockIdx.x, red);
    return;
  }
  const uint32_t bid = blockIdx.x - nside;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int lpg = (int)((g.ncols + V - 1) / V);
  const int ng = 64 / lpg;
  const int grp = lane / lpg, li = lane - grp * lpg;
  const bool act = grp < ng && (uint32_t)(li * V) < g.ncols;
  const uint32_t coff = act ? g.col0 + (uint32_t)(li * V) : g.col0;
  const T* __restrict__ tab = reinterpret_cast<const T*>(g.table);
  const uint64_t total = (uint64_t)g.per * g.ngroups;
  for (uint64_t o = (uint64_t)bid * 4 + wv; o < total; o += (uint64_t)g.gblocks * 4) {
    const uint64_t q = o / g.per;
    const uint64_t row = g.in_base + q * g.stride_in + (o - q * g.per);
    const int64_t p0 = g.ptr[row], p1 = g.ptr[row + 1];
    T s[V];
#pragma unroll
    for (int v = 0; v < V; ++v) s[v] = T(0);
    for (int64_t b = p0; b < p1; b += 64) {
      const int cnt = (int)(p1 - b < 64 ? p1 - b : 64);
      const int32_t mine = lane < cnt ? g.idx[b + lane] : 0;
      for (int e0 = 0; e0 < cnt; e0 += D * ng) {
        vec x[D];
        bool ok[D];
#pragma unroll
        for (int u = 0; u < D; ++u) {
          const int e = e0 + u * ng + grp;
          ok[u] = act && e < cnt;
          const int32_t j = __shfl(mine, (e < cnt) ? e : 0);
          // clamped, unconditional loads (row 0 stands in for idle lanes)
          x[u] = *reinterpret_cast<const vec*>(tab + (uint64_t)(ok[u] ? j : 0) * g.l + coff);
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
          T acc = T(0);
#pragma unroll
          for (int u = 0; u < D; ++u) acc += ok[u] ? x[u][v] : T(0);
          s[v] += acc;
        }
      }
    }
    T t[V];
#pragma unroll
    for (int v = 0; v < V; ++v) t[v] = s[v];
    for (int gg = 1; gg < ng; ++gg)
#pragma unroll
      for (int v = 0; v < V; ++v) t[v] += __shfl(s[v], gg * lpg + li);
    if (grp == 0 && act) {  // act: the lane's V columns are inside the row (l % V == 0)
      vec y;
#pragma unroll
      for (int v = 0; v < V; ++v) y[v] = t[v];
      *reinterpret_cast<vec*>(reinterpret_cast<T*>(g.out) + (g.out_base + o) * g.l + coff) = y;
    }
  }
}

void launch_csr_gather(const CsrGather& g0, hipStream_t s) {
  PCAONE_REQUIRE(g0.l >= 1 && g0.l <= 128, -1, "gather: l must be in [1, 128]");
  const uint64_t total = (uint64_t)g0.per * g0.ngroups;
  if (!total) {
    launch_side_jobs(g0.side, s);
    return;
  }
  const uint32_t gblocks = (uint32_t)std::min<uint64_t>(ceil_div(total, 4), 32768);
  // 16-byte vector loads need rows that are whole vectors (tables are hipMalloc'd);
  // a lane group spans at most the 64 lanes of a wave, so rows of single
  // columns wider than 64 run as two column windows
  const int V = g0.table_f32 ? (g0.l % 4 == 0 ? 4 : g0.l % 2 == 0 ? 2 : 1) : (g0.l % 2 == 0 ? 2 : 1);
  const uint32_t wmax = 64u * (uint32_t)V;
  for (uint32_t c0 = 0; c0 < g0.l; c0 += wmax) {
    CsrGather g = g0;
    g.col0 = c0;
    g.ncols = std::min(wmax, g0.l - c0);
    g.gblocks = gblocks;
    if (c0 > 0) g.side = SideJobs{};  // the jobs ride on the first launch only
    const uint32_t blocks = gblocks + side_blocks(g.side);
    if (g.table_f32) {
      if (V == 4)
        hipLaunchKernelGGL((k_csr_gather<float, 4>), dim3(blocks), dim3(256), 0, s, g);
      else if (V == 2)
        hipLaunchKernelGGL((k_csr_gather<float, 2>), dim3(blocks), dim3(256), 0, s, g);
      else
        hipLaunchKernelGGL((k_csr_gather<float, 1>), dim3(blocks), dim3(256), 0, s, g);
    } else {
      if (V == 2)
        hipLaunchKernelGGL((k_csr_gather<double, 2>), dim3(blocks), dim3(256), 0, s, g);
      else
        hipLaunchKernelGGL((k_csr_gather<double, 1>), dim3(blocks), dim3(256), 0, s, g);
    }
  }
}

__global__ __launch_bounds__(kSideThreads) void k_side_jobs(SideJobs j) {
  __shared__ double red[kSideLds];
  side_block(j, blockIdx.x, red);
}

void launch_side_jobs(const SideJobs& j, hipStream_t s) {
  PCAONE_REQUIRE(!j.stats || (j.W >= 1 && j.W <= 64 && j.l <= j.W), -1, "side jobs: bad sign job");
  PCAONE_REQUIRE(!j.s2p || j.l2 >= 1, -1, "side jobs: bad column-sum job");
  const uint32_t nb = side_blocks(j);
  if (nb) hipLaunchKernelGGL(k_side_jobs, dim3(nb), dim3(kSideThreads), 0, s, j);
}

__global__ void k_cast_f32(const double* __restrict__ src, float* __restrict__ dst, uint64_t n) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < n) dst[gid] = (float)src[gid];
}

void launch_cast_f32(const double* src, float* dst, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_cast_f32, dim3((uint32_t)ceil_div(n, 256)), dim3(256), 0, s, src, dst, n);
}

__global__ void k_omega_operand(const double* __restrict__ omega, uint32_t n, uint32_t l, int bprec, void* bg,
                                uint32_t bcols, uint32_t lo_off, uint64_t kpad, int kperm) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (uint64_t)bcols * kpad) return;
  const uint32_t col = (uint32_t)(gid / kpad);
  const uint64_t k = kperm ? kperm16(gid % kpad) : gid % kpad;  // the sample this position holds
  if (bprec == PREC_SPLIT_BF16) {
    __bf16 v = (__bf16)0.0f;
    const bool is_hi = col < l, is_lo = col >= lo_off && col < lo_off + l;
    if (k < n && (is_hi || is_lo)) {
      const double x = omega[k * l + (is_hi ? col : col - lo_off)];
      const __bf16 hi = (__bf16)(float)x;
      v = is_hi ? hi : (__bf16)(float)(x - (double)(float)hi);
    }
    reinterpret_cast<__bf16*>(bg)[gid] = v;
  } else {
    const double x = (k < n && col < l) ? omega[k * l + col] : 0.0;
    if (bprec == PREC_FP64)
      reinterpret_cast<double*>(bg)[gid] = x;
    else
      reinterpret_cast<float*>(bg)[gid] = (float)x;
  }
}

void launch_omega_operand(const double* omega, uint32_t n, uint32_t l, int bprec, void* bg, uint32_t bcols,
                          uint32_t lo_off, uint64_t kpad, int kperm, hipStream_t s) {
  PCAONE_REQUIRE(!kperm || kpad % 16 == 0, -1, "omega operand: permuted K needs 16-code groups");
  const uint64_t total = (uint64_t)bcols * kpad;
  hipLaunchKernelGGL(k_omega_operand, dim3((uint32_t)ceil_div(total, 256)), dim3(256), 0, s, omega, n, l, bprec,
                     bg, bcols, lo_off, kpad, kperm);
}

__global__ void k_fill_f64(double* p, uint64_t n, double v) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < n) p[gid] = v;
}

void launch_fill_f64(double* p, uint64_t n, double v, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_f64, dim3((uint32_t)ceil_div(n, 256)), dim3(256), 0, s, p, n, v);
}

}  // namespace pcaone
