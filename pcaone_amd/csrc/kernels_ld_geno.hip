// LD r^2 straight from the packed PLINK genotypes, with the semantics of the
// reference's --ld --ld-stats 1 route: Data::write_residuals writes the
// centred genotypes BED2GENO - F (missing 0, each SNP re-centred by its mean,
// src/Data.cpp:196-221) and ld_r2_big correlates them over the windows of
// divide_pos_by_window (src/LD.cpp:125-139, 419-442).
//
// With y the A1 dosage (0, 1, 2; missing 0) and w the non-missing mask, the r^2
// of a pair follows exactly from four integer sums over samples,
//   S_yy = sum y_a y_b,  S_yw = sum y_a w_b,  S_wy = sum w_a y_b,  S_ww = sum w_a w_b,
// and per-SNP F, sum y, sum y^2, n (see k_ld_geno's epilogue).  The sums come
// from int8 MFMA with int32 accumulation, so they are exact: the operand of
// 8 SNPs is 16 stacked int8 rows [y of the 8 SNPs; w of the 8 SNPs], and one
// v_mfma_i32_16x16x64_i8 of two such fragments yields all four sums for 8 x 8
// SNP pairs (its four 8 x 8 quadrants).
//
// One persistent launch covers every band tile of every window.  Per K chunk
// the block decodes the packed codes of its TM + TN SNP rows once into LDS
// (one 32-bit word = 16 samples -> 16 y bytes and 16 w bytes by byte-table
// v_perm), then every wave runs its FA x FB fragments from LDS; the next
// chunk's packed words are loaded while the current chunk computes.  The
// epilogue gathers each pair's four sums into one lane with three lane
// shuffles and writes r^2 (FP64) at the pair's place in the window-ordered
// output.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {
namespace {

typedef int32_t i32x4 __attribute__((ext_vector_type(4)));

// byte tables indexed by the 2-bit code (00 hom A1, 01 missing, 10 het, 11 hom A2)
constexpr uint32_t kLutY = 0x00010002u;  // dosage y: 2, 0, 1, 0
constexpr uint32_t kLutW = 0x01010001u;  // non-missing w: 1, 0, 1, 1

// 16 codes of one 32-bit word -> 16 bytes (word j holds codes j, j+4, j+8, j+12:
// a fixed permutation of the samples, the same for every operand, so the sums
// are unchanged)
__device__ inline uint4 decode_bytes(uint32_t x, uint32_t lut) {
  uint32_t o[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) o[j] = __builtin_amdgcn_perm(0u, lut, (x >> (2 * j)) & 0x03030303u);
  return make_uint4(o[0], o[1], o[2], o[3]);
}

template <int TM, int TN, int WR, int WC, int KS>
struct LdShape {
  static constexpr int NT = 64 * WR * WC;           // threads
  static constexpr int NFA = TM / 8, NFB = TN / 8;  // 8-SNP fragments per side
  static constexpr int NF = NFA + NFB;
  static constexpr int FA = NFA / WR, FB = NFB / WC;  // fragments per wave
  static constexpr int ROWS = TM + TN;
  static constexpr int LOADS = ROWS * KS;           // 16-byte (64-sample) loads per chunk
  static constexpr int LPT = (LOADS + NT - 1) / NT;
  static constexpr int LDS_U4 = KS * NF * 4 * 16;   // uint4 slots: [kstep][fragment][group][row]
  static_assert(NFA % WR == 0 && NFB % WC == 0, "fragments must split evenly over the waves");
};

template <int TM, int TN, int WR, int WC, int KS>
__global__ __launch_bounds__(64 * WR * WC) void k_ld_geno(LdGeno g) {
  using S = LdShape<TM, TN, WR, WC, KS>;
  __shared__ uint4 dec[S::LDS_U4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wr = wv / WC, wc = wv % WC;
  const int lg = lane >> 4, li = lane & 15;
  const uint32_t nchunks = g.row_bytes / (16 * KS);
  // tiles of this block: a contiguous share of the tile list per XCD (blocks
  // are dealt to the 8 XCDs round-robin), neighbouring tiles share SNP rows in
  // that XCD's L2
  const uint32_t xcd = blockIdx.x % 8, slot = blockIdx.x / 8, per_xcd_blocks = gridDim.x / 8;
  const uint32_t share = (g.ntiles + 7) / 8;
  const uint32_t t_begin = xcd * share, t_end = min(g.ntiles, t_begin + share);
  for (uint32_t tix = t_begin + slot; tix < t_end; tix += per_xcd_blocks) {
    const int2 tile = g.tiles[tix];
    const uint32_t a0 = (uint32_t)tile.x * TM, b0 = (uint32_t)tile.y * TN;
    // this thread's loads: load e = (row s, kstep t), t fastest (64 contiguous
    // bytes of a row per chunk for KS = 4)
    const uint8_t* src[S::LPT];
    int dst_f[S::LPT], dst_i[S::LPT], dst_t[S::LPT];
    bool live[S::LPT];
#pragma unroll
    for (int u = 0; u < S::LPT; ++u) {
      const int e = threadIdx.x + S::NT * u;
      live[u] = e < S::LOADS;
      const int s = live[u] ? e / KS : 0, t = e % KS;
      uint32_t snp = s < TM ? a0 + s : b0 + (s - TM);
      snp = min(snp, g.nsnps - 1);  // rows past the last SNP: any row, their pairs are never written
      src[u] = g.bed + (uint64_t)snp * g.row_bytes + 16 * t;
      dst_f[u] = s / 8;
      dst_i[u] = s % 8;
      dst_t[u] = t;
    }
    i32x4 acc[S::FA][S::FB];
#pragma unroll
    for (int fa = 0; fa < S::FA; ++fa)
#pragma unroll
      for (int fb = 0; fb < S::FB; ++fb) acc[fa][fb] = i32x4{0, 0, 0, 0};
    uint4 reg[S::LPT];
#pragma unroll
    for (int u = 0; u < S::LPT; ++u) reg[u] = live[u] ? *reinterpret_cast<const uint4*>(src[u]) : make_uint4(0, 0, 0, 0);
    for (uint32_t c = 0; c < nchunks; ++c) {
      __syncthreads();  // the previous chunk's operands are no longer read
#pragma unroll
      for (int u = 0; u < S::LPT; ++u) {
        if (!live[u]) continue;
        const uint32_t x[4] = {reg[u].x, reg[u].y, reg[u].z, reg[u].w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {  // 16-sample group q of this 64-sample step
          const int base = ((dst_t[u] * S::NF + dst_f[u]) * 4 + q) * 16;
          dec[base + dst_i[u]] = decode_bytes(x[q], kLutY);
          dec[base + 8 + dst_i[u]] = decode_bytes(x[q], kLutW);
        }
      }
      if (c + 1 < nchunks) {
#pragma unroll
        for (int u = 0; u < S::LPT; ++u)
          if (live[u]) reg[u] = *reinterpret_cast<const uint4*>(src[u] + (uint64_t)(c + 1) * 16 * KS);
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < KS; ++t) {
        i32x4 a[S::FA], b[S::FB];
#pragma unroll
        for (int fa = 0; fa < S::FA; ++fa) {
          const uint4 v = dec[((t * S::NF + wr * S::FA + fa) * 4 + lg) * 16 + li];
          a[fa] = i32x4{(int)v.x, (int)v.y, (int)v.z, (int)v.w};
        }
#pragma unroll
        for (int fb = 0; fb < S::FB; ++fb) {
          const uint4 v = dec[((t * S::NF + S::NFA + wc * S::FB + fb) * 4 + lg) * 16 + li];
          b[fb] = i32x4{(int)v.x, (int)v.y, (int)v.z, (int)v.w};
        }
#pragma unroll
        for (int fa = 0; fa < S::FA; ++fa)
#pragma unroll
          for (int fb = 0; fb < S::FB; ++fb)
            acc[fa][fb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[fa], b[fb], acc[fa][fb], 0, 0, 0);
      }
    }
    // Epilogue.  Accumulator element v of lane (j, q) = (lane % 16, lane / 16)
    // is D[4 q + v][j]: rows 0-7 are the y of A's 8 SNPs, rows 8-15 their w;
    // columns likewise for B.  Lane (jj, q) with jj < 8, q < 2 holds S_yy of
    // SNPs (4 q + v, jj); S_yw, S_wy and S_ww sit 8, 32 and 40 lanes up.
    const bool act = li < 8 && lg < 2;
#pragma unroll
    for (int fa = 0; fa < S::FA; ++fa)
#pragma unroll
      for (int fb = 0; fb < S::FB; ++fb) {
        int yw[4], wy[4], ww[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          yw[v] = __shfl(acc[fa][fb][v], (lane + 8) & 63);
          wy[v] = __shfl(acc[fa][fb][v], (lane + 32) & 63);
          ww[v] = __shfl(acc[fa][fb][v], (lane + 40) & 63);
        }
        if (!act) continue;
        const uint32_t b = b0 + (uint32_t)(wc * S::FB + fb) * 8 + li;
        if (b >= g.nsnps) continue;
        const double Fb = g.F[b], mb = g.mun[b], sb = g.ss[b];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const uint32_t a = a0 + (uint32_t)(wr * S::FA + fa) * 8 + 4 * lg + v;
          if (a >= b || a < g.win_lo || a - g.win_lo >= g.win_rows) continue;
          const int32_t w = g.win_of_row[a - g.win_lo];
          if (w < 0) continue;
          const uint32_t d = b - a;
          if (d >= (uint32_t)g.we[w]) continue;
          const uint64_t o = (uint64_t)g.off[w] + d - 1;
          const int yy = acc[fa][fb][v];
          const double Fa = g.F[a];
          const double cov = 0.25 * (double)yy - 0.5 * Fb * (double)yw[v] - 0.5 * Fa * (double)wy[v] +
                             Fa * Fb * (double)ww[v] - g.mun[a] * mb * g.inv_n;
          // a SNP without variation (monomorphic or all missing: ss is then
          // exactly 0) has a zero residual column, and the reference's
          // r = 0 * (1 / 0) is NaN (calc_sds / ld_r2_big, src/LD.cpp:20-23, 427)
          const double sa = g.ss[a];
          g.r2[o] = (sa > 0.0 && sb > 0.0) ? cov * cov / (sa * sb) : __builtin_nan("");
          if (g.sums) reinterpret_cast<int4*>(g.sums)[o] = make_int4(yy, yw[v], wy[v], ww[v]);
        }
      }
  }
}

struct LdConfig {
  int tm, tn;
  void (*launch)(const LdGeno&, uint32_t, hipStream_t);
  int (*blocks_per_cu)();
};

template <int TM, int TN, int WR, int WC, int KS>
void launch_cfg(const LdGeno& g, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((k_ld_geno<TM, TN, WR, WC, KS>), dim3(grid), dim3(64 * WR * WC), 0, s, g);
}

template <int TM, int TN, int WR, int WC, int KS>
int occupancy_cfg() {
  int n = 0;
  PCAONE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(k_ld_geno<TM, TN, WR, WC, KS>),
                                                          64 * WR * WC, 0));
  return std::max(n, 1);
}

#define PCAONE_LDCFG(TM, TN, WR, WC, KS) \
  LdConfig { TM, TN, launch_cfg<TM, TN, WR, WC, KS>, occupancy_cfg<TM, TN, WR, WC, KS> }

// The candidates the autotuner times (tile rows x tile columns; waves as
// rows x columns of the tile; 64-sample K steps per LDS chunk):
const LdConfig kConfigs[] = {
    PCAONE_LDCFG(32, 32, 1, 2, 4),    // 2 waves, 32 x 16 SNPs each: short windows
    PCAONE_LDCFG(64, 64, 2, 2, 4),    // 4 waves, 32 x 32 each
    PCAONE_LDCFG(64, 64, 1, 2, 4),    // 2 waves, 64 x 32 each: more reuse per LDS read
    PCAONE_LDCFG(128, 64, 2, 2, 2),   // 4 waves, 64 x 32 each
    PCAONE_LDCFG(128, 128, 2, 4, 2),  // 8 waves, 64 x 32 each: long windows
};
#undef PCAONE_LDCFG

}  // namespace

int ld_geno_nconfigs() { return (int)(sizeof(kConfigs) / sizeof(kConfigs[0])); }

void ld_geno_tile(int config, int* tm, int* tn) {
  PCAONE_REQUIRE(config >= 0 && config < ld_geno_nconfigs(), -1, "ld: no such tile configuration");
  *tm = kConfigs[config].tm;
  *tn = kConfigs[config].tn;
}

uint32_t ld_geno_grid(int config, uint32_t ntiles, uint32_t ncu) {
  const uint32_t slots = ncu * (uint32_t)kConfigs[config].blocks_per_cu();
  // a multiple of 8 (one share per XCD), at most one block per tile and share
  const uint32_t want = std::min<uint32_t>(slots, std::max<uint32_t>(8, (ntiles + 7) / 8 * 8));
  return std::max<uint32_t>(8, want / 8 * 8);
}

void launch_ld_geno(const LdGeno& g, int config, uint32_t grid, hipStream_t s) {
  PCAONE_REQUIRE(config >= 0 && config < ld_geno_nconfigs(), -1, "ld: no such tile configuration");
  PCAONE_REQUIRE(g.row_bytes % 64 == 0 && g.nsnps > 0, -1, "ld: packed rows must be whole 64-byte chunks");
  PCAONE_REQUIRE(grid % 8 == 0, -1, "ld: grid must be a multiple of 8");
  if (!g.ntiles) return;
  kConfigs[config].launch(g, grid, s);
}

}  // namespace pcaone
