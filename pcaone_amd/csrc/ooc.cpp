// Out-of-core (-m) planning on the host: the block plan of Data::prepare and
// the SNP order of permute_plink.  Both decide which SNPs form each window of
// the winSVD, so they follow the reference's integer arithmetic step by step
// (32-bit unsigned products included) to reproduce its windows exactly.
#include <cmath>
#include <vector>

#include "engine.hpp"

namespace pcaone {

// src/Data.cpp:27-71.  svd: 0 IRAM, 1 Halko, 2 winSVD.  memory in GB
// (2^27 doubles per GB).
void ooc_plan(uint32_t nsamples, uint32_t nsnps, int svd, int k, int os, int bands, double memory, uint32_t& bs,
              uint32_t& nblocks, uint32_t& factor) {
  PCAONE_REQUIRE(nsamples > 0 && nsnps > 0, -1, "ooc_plan: empty matrix");
  PCAONE_REQUIRE(memory > 0.0, -1, "ooc_plan: -m must be positive");
  PCAONE_REQUIRE(svd >= 0 && svd <= 2, -1, "ooc_plan: svd must be 0 (IRAM), 1 or 2");
  PCAONE_REQUIRE(svd != 2 || bands >= 4, -1, "ooc_plan: bands must be at least 4");
  constexpr double kDoublesPerGB = 134217728.0;
  if (svd == 0) {
    bs = (uint32_t)std::ceil(memory * kDoublesPerGB / nsamples);
  } else {
    const uint32_t l = (uint32_t)(k + os);
    const uint32_t a = 3u * nsamples * l, b = 2u * nsnps * l, c = 5u * nsnps;  // 32-bit, as the reference
    double m = (double)(a + b + c) / kDoublesPerGB;
    if (memory > 1.1 * m) m = 0.0;
    bs = (uint32_t)std::ceil((double)((m + memory) * kDoublesPerGB - a - b - c) / nsamples);
  }
  PCAONE_REQUIRE(bs > 0, -1, "ooc_plan: -m leaves no room for a block");
  nblocks = (uint32_t)std::ceil((double)nsnps / bs);
  PCAONE_REQUIRE(nblocks != 1, -1, "only one block exists. please remove -m option");
  factor = 1;
  if (svd == 2) {
    if (nblocks < (uint32_t)bands) {
      bs = (uint32_t)std::ceil((double)nsnps / bands);
    } else {
      factor = (uint32_t)std::ceil((double)nblocks / bands);
      bs = (uint32_t)std::ceil((double)nsnps / ((double)bands * factor));
    }
    nblocks = (uint32_t)std::ceil((double)nsnps / bs);
  }
}

// src/FilePlink.cpp:298-405: perm[ib] = ia, input SNP ia at position ib of
// the permuted file.  Buffers of two_snps SNPs (a multiple of nbands) are
// dealt round-robin into the bands, band b of buffer i starting at
// i * bufsize + bandidx[b]; the last buffer may be short.
void plink_perm(uint32_t nsamples, uint32_t nsnps, uint32_t gb, uint32_t nbands, int32_t* perm) {
  PCAONE_REQUIRE(nsamples > 0 && nsnps > 0 && nbands > 0 && gb > 0, -1, "plink_perm: bad sizes");
  const uint64_t bps = (nsamples + 3) >> 2;
  uint64_t two = (uint64_t)std::floor(1073741824.0 * gb / (double)bps);
  if (two > nsnps) two = nsnps;
  uint64_t bufsize = two / nbands;
  two = bufsize * nbands;
  PCAONE_REQUIRE(two > 0, -1, "plink_perm: fewer SNPs than bands");
  const uint64_t nblocks = (nsnps + two - 1) / two;
  uint64_t modr2 = nsnps % two;
  const uint64_t modr = nsnps % nbands, bandsize = (nsnps + nbands - 1) / nbands;
  std::vector<uint64_t> bandidx(nbands);
  for (uint32_t i = 0; i < nbands; ++i)
    bandidx[i] = (modr == 0 || i < modr) ? i * bandsize : modr * bandsize + (bandsize - 1) * (i - modr);
  const uint64_t bufidx = bufsize;
  std::vector<char> seen(nsnps, 0);
  uint64_t placed = 0;
  auto put = [&](uint64_t ib, uint64_t ia) {
    PCAONE_REQUIRE(ib < nsnps && ia < nsnps && !seen[ib], -2,
                   "plink_perm: the reference's interleave is not a permutation for these sizes");
    seen[ib] = 1;
    perm[ib] = (int32_t)ia;
    ++placed;
  };
  for (uint64_t i = 0; i < nblocks; ++i) {
    if (i == nblocks - 1 && modr2 != 0) {
      const uint64_t two2 = nsnps - (nblocks - 1) * two;
      bufsize = (two2 + nbands - 1) / nbands;
      modr2 = two2 % nbands;
    }
    for (uint64_t b = 0; b < nbands; ++b) {
      uint64_t j = 0;
      for (; j + 1 < bufsize; ++j) put(i * bufidx + bandidx[b] + j, i * two + j * nbands + b);
      if (i != nblocks - 1 || b < modr2 || modr2 == 0) put(i * bufidx + bandidx[b] + j, i * two + j * nbands + b);
    }
  }
  PCAONE_REQUIRE(placed == nsnps, -2, "plink_perm: the reference's interleave is not a permutation for these sizes");
}

}  // namespace pcaone
