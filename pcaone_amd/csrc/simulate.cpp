// Seeded synthetic genotype generator: Balding-Nichols populations written
// straight into the PLINK .bed payload (SNP-major, 4 samples per byte,
// LSB first; code 00 = two A1 copies, 10 = one, 11 = none, 01 = missing).
//
// Each SNP draws from its own counter-seeded stream, so the output depends only
// on (seed, sizes, parameters) and not on the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <zlib.h>

#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pcaone_gpu.h"
#include "zstd_dl.hpp"

namespace {

inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Xoshiro256 {
  uint64_t s[4];
  explicit Xoshiro256(uint64_t seed) {
    for (auto& v : s) v = splitmix64(seed);
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  inline uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  inline double uniform() { return (next() >> 11) * 0x1.0p-53; }  // [0,1)
  inline double normal() {  // Box-Muller, one value per call
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // Marsaglia-Tsang gamma(shape, 1)
  double gamma(double a) {
    if (a < 1.0) {
      double u = uniform();
      if (u < 1e-300) u = 1e-300;
      return gamma(a + 1.0) * std::pow(u, 1.0 / a);
    }
    const double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
      double x = normal(), v = 1.0 + c * x;
      if (v <= 0) continue;
      v = v * v * v;
      double u = uniform();
      if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
      if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
    }
  }
  double beta(double a, double b) {
    double x = gamma(a), y = gamma(b);
    return x / (x + y);
  }
};

uint64_t snp_seed(uint64_t seed, uint64_t snp, uint64_t attempt) {
  uint64_t x = seed ^ (snp * 0xD1B54A32D192ED03ull) ^ (attempt * 0x8CB92BA72F3D8DD7ull);
  return splitmix64(x);
}

// One SNP row.  Returns true when polymorphic among non-missing genotypes.
bool gen_row(uint64_t seed, uint32_t snp, uint32_t attempt, uint32_t n, uint32_t npops, double fst, double pmin,
             double pmax, double miss, uint8_t* row, uint32_t nbytes, double* p_out) {
  Xoshiro256 rng(snp_seed(seed, snp, attempt));
  const double p = pmin + (pmax - pmin) * rng.uniform();
  if (p_out) *p_out = p;
  std::vector<uint32_t> t2(npops), t1(npops);
  for (uint32_t k = 0; k < npops; ++k) {
    double pk = p;
    if (fst > 0) {
      const double a = p * (1.0 - fst) / fst, b = (1.0 - p) * (1.0 - fst) / fst;
      pk = rng.beta(a, b);
    }
    const double f2 = pk * pk, f1 = f2 + 2.0 * pk * (1.0 - pk);
    t2[k] = (uint32_t)std::min(4294967295.0, std::floor(f2 * 4294967296.0));
    t1[k] = (uint32_t)std::min(4294967295.0, std::floor(f1 * 4294967296.0));
  }
  const uint32_t tm = (uint32_t)std::min(4294967295.0, std::floor(miss * 4294967296.0));
  std::memset(row, 0, nbytes);
  uint64_t c[4] = {0, 0, 0, 0};
  for (uint32_t j = 0; j < n; ++j) {
    const uint32_t k = (uint32_t)(((uint64_t)j * npops) / n);
    const uint64_t r = rng.next();
    const uint32_t ug = (uint32_t)(r >> 32), um = (uint32_t)r;
    uint32_t code;
    if (miss > 0 && um < tm)
      code = 1;
    else if (ug < t2[k])
      code = 0;  // two A1 copies
    else if (ug < t1[k])
      code = 2;  // het
    else
      code = 3;  // no A1 copy
    c[code]++;
    row[j >> 2] |= (uint8_t)(code << (2 * (j & 3)));
  }
  const uint64_t nonmiss = c[0] + c[2] + c[3];
  if (nonmiss == 0) return false;
  // F = (n00 + n10/2) / n in {0, 1} means monomorphic
  if (c[0] + c[2] == 0) return false;
  if (c[2] + c[3] == 0) return false;
  return true;
}

}  // namespace

namespace {

// Poisson(lambda) by inversion for small lambda, normal approximation with
// continuity correction above 64 (never reached by the count model below at
// the sizes used, kept for completeness).
uint32_t poisson(Xoshiro256& rng, double lam) {
  if (lam <= 0) return 0;
  if (lam > 64) {
    const double x = std::floor(lam + std::sqrt(lam) * rng.normal() + 0.5);
    return x < 0 ? 0u : (uint32_t)x;
  }
  const double L = std::exp(-lam);
  double p = 1.0;
  uint32_t k = 0;
  do {
    ++k;
    p *= rng.uniform();
  } while (p > L);
  return k - 1;
}

}  // namespace

// scRNA-like dense matrix, feature-major (ngenes rows of ncells float32):
// gene means m_g = base * LogNormal(0, sdlog), cell size factors
// s_c = LogNormal(0, 0.3), counts ~ NegBinomial(mean m_g s_c, dispersion phi)
// (Gamma-Poisson), value = log1p(1e6 * count / library size of the cell).
extern "C" int pcaone_sim_counts(uint64_t seed, uint32_t ncells, uint32_t ngenes, double base, double sdlog,
                                 double phi, float* out, int nthreads) {
  if (!out || ncells == 0 || ngenes == 0 || !(base > 0) || !(phi > 0)) return -1;
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  unsigned nt = nthreads > 0 ? (unsigned)nthreads : std::min(hw, 16u);
  nt = std::max(1u, std::min<unsigned>(nt, ngenes));
  std::vector<double> sc(ncells);
  {
    Xoshiro256 rng(snp_seed(seed, 0xFFFFFFFFull, 7));
    for (auto& v : sc) v = std::exp(0.3 * rng.normal());
  }
  std::vector<std::vector<double>> lib(nt, std::vector<double>(ncells, 0.0));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t) {
    pool.emplace_back([&, t]() {
      const uint64_t lo = (uint64_t)ngenes * t / nt, hi = (uint64_t)ngenes * (t + 1) / nt;
      for (uint64_t g = lo; g < hi; ++g) {
        Xoshiro256 rng(snp_seed(seed, g, 0));
        const double mg = base * std::exp(sdlog * rng.normal());
        const double shape = 1.0 / phi;
        float* row = out + g * ncells;
        for (uint32_t c = 0; c < ncells; ++c) {
          const double mean = mg * sc[c];
          const double lam = rng.gamma(shape) * mean / shape;
          const uint32_t k = poisson(rng, lam);
          row[c] = (float)k;
          lib[t][c] += k;
        }
      }
    });
  }
  for (auto& th : pool) th.join();
  std::vector<double> L(ncells, 0.0);
  for (unsigned t = 0; t < nt; ++t)
    for (uint32_t c = 0; c < ncells; ++c) L[c] += lib[t][c];
  pool.clear();
  for (unsigned t = 0; t < nt; ++t) {
    pool.emplace_back([&, t]() {
      const uint64_t lo = (uint64_t)ngenes * t / nt, hi = (uint64_t)ngenes * (t + 1) / nt;
      for (uint64_t g = lo; g < hi; ++g) {
        float* row = out + g * ncells;
        for (uint32_t c = 0; c < ncells; ++c)
          row[c] = (float)std::log1p(L[c] > 0 ? 1e6 * (double)row[c] / L[c] : 0.0);
      }
    });
  }
  for (auto& th : pool) th.join();
  return 0;
}

// Genotype likelihoods (BEAGLE order, 3 per sample and site) from
// Balding-Nichols genotypes: depth ~ Poisson(depth), each read carries one of
// the genotype's two alleles, flipped with probability err; GL(g) =
// prod_reads P(read | g), normalised to sum 1 (depth 0 gives 1/3 each).
// out: nsites x nsamples x 3 float32; g counts copies of the allele whose
// frequency is modelled (the third GL is that homozygote).
extern "C" int pcaone_sim_gl(uint64_t seed, uint32_t nsamples, uint32_t nsites, uint32_t npops, double fst,
                             double pmin, double pmax, double depth, double err, float* out, int nthreads) {
  if (!out || nsamples == 0 || nsites == 0 || npops == 0 || npops > nsamples) return -1;
  if (!(fst >= 0 && fst < 1) || !(pmin > 0 && pmax < 1 && pmin <= pmax) || !(depth >= 0) || !(err >= 0 && err < 0.5))
    return -1;
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  unsigned nt = nthreads > 0 ? (unsigned)nthreads : std::min(hw, 16u);
  nt = std::max(1u, std::min<unsigned>(nt, nsites));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t) {
    pool.emplace_back([&, t]() {
      const uint64_t lo = (uint64_t)nsites * t / nt, hi = (uint64_t)nsites * (t + 1) / nt;
      std::vector<double> pk(npops);
      for (uint64_t s = lo; s < hi; ++s) {
        Xoshiro256 rng(snp_seed(seed ^ 0x5851F42D4C957F2Dull, s, 0));
        const double p = pmin + (pmax - pmin) * rng.uniform();
        for (uint32_t k = 0; k < npops; ++k) {
          pk[k] = p;
          if (fst > 0) pk[k] = rng.beta(p * (1.0 - fst) / fst, (1.0 - p) * (1.0 - fst) / fst);
        }
        float* row = out + s * (uint64_t)nsamples * 3;
        for (uint32_t j = 0; j < nsamples; ++j) {
          const double f = pk[(uint64_t)j * npops / nsamples];
          const int g = (rng.uniform() < f) + (rng.uniform() < f);
          const uint32_t d = poisson(rng, depth);
          double L[3] = {1.0, 1.0, 1.0};
          for (uint32_t r = 0; r < d; ++r) {
            int b = rng.uniform() < g / 2.0;  // allele carried by the read
            if (rng.uniform() < err) b ^= 1;
            for (int gg = 0; gg < 3; ++gg) {
              const double q = gg / 2.0;  // fraction of the modelled allele
              L[gg] *= b ? q * (1 - err) + (1 - q) * err : (1 - q) * (1 - err) + q * err;
            }
          }
          const double sum = L[0] + L[1] + L[2];
          for (int gg = 0; gg < 3; ++gg) row[3 * j + gg] = (float)(L[gg] / sum);
        }
      }
    });
  }
  for (auto& th : pool) th.join();
  return 0;
}

extern "C" int pcaone_sim_bed(uint64_t seed, uint32_t nsamples, uint32_t nsnps, uint32_t npops, double fst,
                              double pmin, double pmax, double miss_rate, uint8_t* payload, double* p_anc,
                              int nthreads) {
  if (!payload || nsamples == 0 || nsnps == 0 || npops == 0 || npops > nsamples) return -1;
  if (!(fst >= 0 && fst < 1) || !(pmin > 0 && pmax < 1 && pmin <= pmax) || !(miss_rate >= 0 && miss_rate < 1))
    return -1;
  const uint32_t nbytes = (nsamples + 3) >> 2;
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  unsigned nt = nthreads > 0 ? (unsigned)nthreads : std::min(hw, 16u);
  nt = std::max(1u, std::min<unsigned>(nt, nsnps));
  std::vector<std::thread> pool;
  std::vector<int> fail(nt, 0);
  for (unsigned t = 0; t < nt; ++t) {
    pool.emplace_back([&, t]() {
      const uint64_t lo = (uint64_t)nsnps * t / nt, hi = (uint64_t)nsnps * (t + 1) / nt;
      for (uint64_t i = lo; i < hi; ++i) {
        uint8_t* row = payload + i * nbytes;
        bool ok = false;
        for (uint32_t a = 0; a < 1000 && !ok; ++a)
          ok = gen_row(seed, (uint32_t)i, a, nsamples, npops, fst, pmin, pmax, miss_rate, row, nbytes,
                       p_anc ? p_anc + i : nullptr);
        if (!ok) fail[t] = 1;
      }
    });
  }
  for (auto& th : pool) th.join();
  for (int f : fail)
    if (f) return -3;
  return 0;
}

// ---- BGEN writer ------------------------------------------------------------
// Genotype probabilities of Balding-Nichols genotypes in a BGEN v1.2 file
// (layout 2) or a v1.1 one (layout 1).  The container: u32 offset of the first
// variant (counted from byte 4), the header block (its length, variant and
// sample counts, "bgen", flags = compression | layout << 2 | ids << 31), the
// optional sample identifier block, then per variant its identifiers, position,
// alleles and genotype block.

namespace {

void put16(std::vector<uint8_t>& o, uint32_t v) {
  o.push_back((uint8_t)v);
  o.push_back((uint8_t)(v >> 8));
}
void put32(std::vector<uint8_t>& o, uint32_t v) {
  for (int i = 0; i < 4; ++i) o.push_back((uint8_t)(v >> (8 * i)));
}
void put_str16(std::vector<uint8_t>& o, const std::string& s) {
  put16(o, (uint32_t)s.size());
  o.insert(o.end(), s.begin(), s.end());
}

// appends bits [0, nbits) of v to a little-endian bit stream
struct BitWriter {
  std::vector<uint8_t>& o;
  uint64_t acc = 0;
  int have = 0;
  explicit BitWriter(std::vector<uint8_t>& out) : o(out) {}
  void put(uint64_t v, int nbits) {
    acc |= v << have;
    have += nbits;
    while (have >= 8) {
      o.push_back((uint8_t)acc);
      acc >>= 8;
      have -= 8;
    }
  }
  void flush() {
    if (have > 0) o.push_back((uint8_t)acc);
    acc = 0;
    have = 0;
  }
};

// probabilities q (summing to 1) quantised to integers summing to scale: floors,
// then the remainder to the largest fractional parts
void quantise(const double* q, int n, uint64_t scale, uint64_t* v) {
  double frac[3];
  uint64_t sum = 0;
  for (int i = 0; i < n; ++i) {
    const double x = q[i] * (double)scale;
    v[i] = (uint64_t)std::floor(x);
    frac[i] = x - (double)v[i];
    sum += v[i];
  }
  while (sum < scale) {
    int best = 0;
    for (int i = 1; i < n; ++i)
      if (frac[i] > frac[best]) best = i;
    ++v[best];
    frac[best] = -1.0;
    ++sum;
  }
}

}  // namespace

extern "C" int pcaone_sim_bgen(const char* path, uint64_t seed, uint32_t nsamples, uint32_t nsnps, uint32_t npops,
                               double fst, double pmin, double pmax, double miss_rate, int layout, int compression,
                               int bits, int pmin_declared, int sample_ids, int nthreads) {
  if (!path || nsamples == 0 || nsnps == 0 || npops == 0 || npops > nsamples) return -1;
  if (!(fst >= 0 && fst < 1) || !(pmin > 0 && pmax < 1 && pmin <= pmax) || !(miss_rate >= 0 && miss_rate < 1))
    return -1;
  if ((layout != 1 && layout != 2) || compression < 0 || compression > 2 || (layout == 1 && compression == 2))
    return -1;
  if (layout == 2 && (bits < 1 || bits > 32)) return -1;
  std::unique_ptr<pcaone::Zstd> z;
  try {
    if (compression == 2) z.reset(new pcaone::Zstd("sim_bgen"));
  } catch (...) {
    return -1;
  }
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  unsigned nt = nthreads > 0 ? (unsigned)nthreads : std::min(hw, 16u);
  nt = std::max(1u, std::min<unsigned>(nt, nsnps));
  std::vector<std::vector<uint8_t>> part(nt);
  std::vector<int> fail(nt, 0);
  const uint64_t scale = layout == 2 ? ((1ull << bits) - 1) : 32768;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t) {
    pool.emplace_back([&, t]() {
      const uint64_t lo = (uint64_t)nsnps * t / nt, hi = (uint64_t)nsnps * (t + 1) / nt;
      std::vector<double> pk(npops);
      std::vector<uint8_t> raw, comp;
      std::vector<double> q(3ull * nsamples);
      std::vector<uint8_t> miss(nsamples);
      auto& o = part[t];
      for (uint64_t s = lo; s < hi && !fail[t]; ++s) {
        bool ok = false;
        for (uint32_t a = 0; a < 1000 && !ok; ++a) {  // redraw monomorphic variants
          Xoshiro256 rng(snp_seed(seed ^ 0xB6E0A1F4C3D2E5F7ull, s, a));
          const double p = pmin + (pmax - pmin) * rng.uniform();
          for (uint32_t k = 0; k < npops; ++k) {
            pk[k] = p;
            if (fst > 0) pk[k] = rng.beta(p * (1.0 - fst) / fst, (1.0 - p) * (1.0 - fst) / fst);
          }
          uint64_t gsum = 0, nobs = 0;
          for (uint32_t j = 0; j < nsamples; ++j) {
            const double f = pk[(uint64_t)j * npops / nsamples];
            const int g = (rng.uniform() < f) + (rng.uniform() < f);  // copies of the modelled allele
            double* qq = &q[3ull * j];
            qq[0] = qq[1] = qq[2] = 0.0;
            if (rng.uniform() < 0.6) {
              qq[g] = 1.0;
            } else {
              const double w = 0.5 + 0.5 * rng.uniform(), r = rng.uniform();
              qq[g] = w;
              qq[(g + 1) % 3] = (1.0 - w) * r;
              qq[(g + 2) % 3] = (1.0 - w) * (1.0 - r);
            }
            miss[j] = rng.uniform() < miss_rate;
            if (!miss[j]) {
              gsum += g;
              ++nobs;
            }
          }
          ok = nobs > 0 && gsum > 0 && gsum < 2 * nobs;
        }
        if (!ok) {
          fail[t] = 1;
          break;
        }
        // genotype order (AA, AB, BB) of the first allele A: the modelled one
        // for even variants, the other for odd ones
        const bool first_modelled = (s % 2) == 0;
        raw.clear();
        if (layout == 2) {
          put32(raw, nsamples);
          put16(raw, 2);
          raw.push_back((uint8_t)(pmin_declared ? 1 : 2));
          raw.push_back(2);
          for (uint32_t j = 0; j < nsamples; ++j) raw.push_back((uint8_t)(miss[j] ? 0x82 : 0x02));
          raw.push_back(0);  // unphased
          raw.push_back((uint8_t)bits);
          BitWriter bw(raw);
          for (uint32_t j = 0; j < nsamples; ++j) {
            const double* qq = &q[3ull * j];
            const double g3[3] = {first_modelled ? qq[2] : qq[0], qq[1], first_modelled ? qq[0] : qq[2]};
            uint64_t v[3] = {0, 0, 0};
            if (!miss[j]) quantise(g3, 3, scale, v);
            bw.put(v[0], bits);
            bw.put(v[1], bits);
          }
          bw.flush();
        } else {
          for (uint32_t j = 0; j < nsamples; ++j) {
            const double* qq = &q[3ull * j];
            const double g3[3] = {first_modelled ? qq[2] : qq[0], qq[1], first_modelled ? qq[0] : qq[2]};
            for (int i = 0; i < 3; ++i) put16(raw, miss[j] ? 0u : (uint32_t)std::lround(g3[i] * 32768.0));
          }
        }
        // variant identifying data
        if (layout == 1) put32(o, nsamples);
        put_str16(o, "snp" + std::to_string(s + 1));
        put_str16(o, "rs" + std::to_string(s + 1));
        put_str16(o, "1");
        put32(o, (uint32_t)(1000 * (s + 1)));
        if (layout == 2) put16(o, 2);
        for (const char* al : {"A", "G"}) {
          put32(o, 1);
          o.push_back((uint8_t)al[0]);
        }
        // genotype block
        if (compression == 0) {
          if (layout == 2) put32(o, (uint32_t)raw.size());
          o.insert(o.end(), raw.begin(), raw.end());
          continue;
        }
        size_t clen;
        if (compression == 1) {
          uLongf cap = compressBound(raw.size());
          comp.resize(cap);
          if (compress2(comp.data(), &cap, raw.data(), raw.size(), 1) != Z_OK) {
            fail[t] = 1;
            break;
          }
          clen = cap;
        } else {
          comp.resize(z->bound(raw.size()));
          clen = z->compress(comp.data(), comp.size(), raw.data(), raw.size(), 1);
          if (z->is_error(clen)) {
            fail[t] = 1;
            break;
          }
        }
        if (layout == 2) {
          put32(o, (uint32_t)(clen + 4));
          put32(o, (uint32_t)raw.size());
        } else {
          put32(o, (uint32_t)clen);
        }
        o.insert(o.end(), comp.begin(), comp.begin() + clen);
      }
    });
  }
  for (auto& th : pool) th.join();
  for (int f : fail)
    if (f) return -3;
  std::vector<uint8_t> head;
  std::vector<uint8_t> ids;
  if (sample_ids) {
    std::vector<uint8_t> body;
    for (uint32_t j = 0; j < nsamples; ++j) put_str16(body, "ind" + std::to_string(j + 1));
    put32(ids, (uint32_t)(8 + body.size()));
    put32(ids, nsamples);
    ids.insert(ids.end(), body.begin(), body.end());
  }
  const uint32_t hlen = 20;
  put32(head, hlen + (uint32_t)ids.size());
  put32(head, hlen);
  put32(head, nsnps);
  put32(head, nsamples);
  for (char c : std::string("bgen")) head.push_back((uint8_t)c);
  put32(head, (uint32_t)compression | ((uint32_t)layout << 2) | (sample_ids ? 0x80000000u : 0u));
  head.insert(head.end(), ids.begin(), ids.end());
  FILE* fp = std::fopen(path, "wb");
  if (!fp) return -1;
  bool good = std::fwrite(head.data(), 1, head.size(), fp) == head.size();
  for (auto& pv : part) good = good && std::fwrite(pv.data(), 1, pv.size(), fp) == pv.size();
  good = (std::fclose(fp) == 0) && good;
  return good ? 0 : -1;
}
