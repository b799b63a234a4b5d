// BGEN input (FileBgen, src/FileBgen.cpp:11-107) over the BGEN v1.1 / v1.2
// container as the reference's vendored bgen library reads it
// (external/bgen: header.cpp, variant.cpp, genotypes.cpp, utils.cpp).
//
// Host: the file is mapped read-only and one pass walks the variant blocks
// (identifiers, alleles, genotype-block length).  The genotype blocks of a
// chunk of variants (about 128 MB decompressed) are decompressed on worker
// threads (zlib from libz, zstd from libzstd.so.1) into page-locked memory,
// each block 16-byte aligned and followed by 8 zero bytes, and each block's
// own header (sample count, allele count, ploidy range, phasing, bit depth,
// length) is checked there.  The chunk is copied to one of two device slots
// and decoded while the host decompresses the next chunk into the other.
//
// Device, one workgroup per variant:
//   mode 0 (FileBgen::read_all without --pcangsd): the first allele's dosage
//     of every sample (Genotypes::ref_dosage_fast / ref_dosage_slow), the
//     minor allele by the library's sampled test (find_minor_allele,
//     minor_certain), the minor allele's dosage (swap_allele_dosage),
//     F = mean of dose / 2 over the non-missing samples, and the row
//     dose / 2 - F (missing 0), divided by sqrt(ploidy F (1 - F)) when
//     standardising (Data::standardize_E, src/Data.cpp:290-299), stored in
//     FP32 as the dense operator's matrix;
//   mode 1 (--pcangsd, src/FileBgen.cpp:72-106): the first two genotype
//     probabilities of every sample (Genotypes::probabilities) as the
//     likelihood pair of the PCAngsd path.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "zstd_dl.hpp"

namespace pcaone {

// ---- container --------------------------------------------------------------

BgenIndex::~BgenIndex() {
  if (map) munmap(map, len);
}

namespace {

uint16_t rd16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }

}  // namespace

// Header (header.cpp: offset, header length, variant and sample counts, magic,
// free data, flags) and every variant block (variant.cpp: identifiers, position,
// alleles, genotype-block length), in file order.
void bgen_index(const char* path, BgenIndex& ix) {
  PCAONE_REQUIRE(path, -1, "load_bgen: no file name");
  const int fd = ::open(path, O_RDONLY);
  PCAONE_REQUIRE(fd >= 0, -1, std::string("error reading from '") + path + "'");
  struct stat st {};
  if (fstat(fd, &st) != 0 || st.st_size < 24) {
    ::close(fd);
    throw Error(-1, std::string("load_bgen: ") + path + " is too short for a BGEN header");
  }
  ix.len = (size_t)st.st_size;
  void* m = mmap(nullptr, ix.len, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  PCAONE_REQUIRE(m != MAP_FAILED, -2, std::string("load_bgen: cannot map ") + path);
  ix.map = m;
  const uint8_t* b = static_cast<const uint8_t*>(m);
  const uint64_t size = ix.len;
  const uint32_t first = rd32(b), hlen = rd32(b + 4);
  ix.nvariants = rd32(b + 8);
  ix.nsamples = rd32(b + 12);
  const uint8_t* mg = b + 16;
  const bool bgen_magic = mg[0] == 'b' && mg[1] == 'g' && mg[2] == 'e' && mg[3] == 'n';
  PCAONE_REQUIRE(bgen_magic || (mg[0] & mg[1] & mg[2] & mg[3]) == 0, -1, "doesn't appear to be a bgen file");
  const uint64_t flag_at = std::max<uint64_t>(20, hlen);
  PCAONE_REQUIRE(flag_at + 4 <= size, -1, "load_bgen: the header runs past the end of the file");
  const uint32_t flags = rd32(b + flag_at);
  ix.compression = (int)(flags & 3u);
  ix.layout = (int)((flags >> 2) & 15u);
  PCAONE_REQUIRE(ix.layout == 1 || ix.layout == 2, -1, "load_bgen: layout " + std::to_string(ix.layout) +
                                                            " (only layouts 1 and 2 exist)");
  PCAONE_REQUIRE(ix.compression <= 2, -1, "load_bgen: unknown compression 3");
  PCAONE_REQUIRE(ix.nsamples > 0 && ix.nvariants > 0, -1, "load_bgen: no samples or no variants");
  const uint32_t N = ix.nsamples, V = ix.nvariants;
  ix.data.resize(V);
  ix.clen.resize(V);
  ix.dlen.resize(V);
  ix.nalleles.resize(V);
  uint64_t off = (uint64_t)first + 4;
  auto need = [&](uint64_t p, uint64_t n) {
    PCAONE_REQUIRE(p + n <= size, -1, "load_bgen: a variant block runs past the end of the file");
  };
  for (uint32_t v = 0; v < V; ++v) {
    PCAONE_REQUIRE(off < size, -1, "reached end of file");
    uint64_t p = off;
    if (ix.layout == 1) {
      need(p, 4);
      PCAONE_REQUIRE(rd32(b + p) == N, -1, "number of samples doesn't match");
      p += 4;
    }
    for (int s = 0; s < 3; ++s) {  // variant id, rsid, chromosome
      need(p, 2);
      p += 2 + rd16(b + p);
    }
    p += 4;  // position
    uint32_t K = 2;
    if (ix.layout == 2) {
      need(p, 2);
      K = rd16(b + p);
      p += 2;
    }
    for (uint32_t a = 0; a < K; ++a) {
      need(p, 4);
      p += 4 + (uint64_t)rd32(b + p);
    }
    uint64_t length;
    if (ix.layout == 1 && ix.compression == 0) {
      length = 6ull * N;
    } else {
      need(p, 4);
      length = rd32(b + p);
      p += 4;
    }
    need(p, length);
    ix.nalleles[v] = (uint16_t)K;
    if (ix.compression == 0) {
      ix.data[v] = p;
      ix.clen[v] = (uint32_t)length;
      ix.dlen[v] = (uint32_t)length;
    } else if (ix.layout == 2) {
      PCAONE_REQUIRE(length >= 4, -1, "couldn't read the compressed length");
      ix.dlen[v] = rd32(b + p);
      ix.data[v] = p + 4;
      ix.clen[v] = (uint32_t)(length - 4);
    } else {
      ix.data[v] = p;
      ix.clen[v] = (uint32_t)length;
      ix.dlen[v] = 6u * N;
    }
    off = p + length;
  }
}

namespace {

// ---- device ----------------------------------------------------------------

constexpr uint8_t kConst = 1;    // one ploidy for every sample (Pmin == Pmax)
constexpr uint8_t kLayout1 = 2;  // layout 1: three 16-bit probabilities per sample
constexpr uint8_t kFast8 = 4;    // diploid, unphased, 8 bits: the library's byte-pair path

struct BgenMeta {
  uint64_t off;   // block offset in the chunk
  uint32_t prob;  // probability data offset within the block
  float factor;   // dosage scale of ref_dosage_slow: 1 / (2^bits - 1), layout 1: 1 / 32768
  float pfactor;  // probability scale of probabilities_layout2: 1 / ((float)2^bits - 1)
  uint8_t bits, maxp, flags, pad;
};

// the bit field [bit, bit + B) of a little-endian bit stream, B <= 32 (the
// library reads it as a 64-bit word at byte bit / 8, shifted by bit % 8)
__device__ inline uint32_t bits_at(const uint8_t* p, uint64_t bit, uint32_t B) {
  const uint8_t* q = p + (bit >> 3);
  const uint64_t w = (uint64_t)q[0] | ((uint64_t)q[1] << 8) | ((uint64_t)q[2] << 16) | ((uint64_t)q[3] << 24) |
                     ((uint64_t)q[4] << 32);
  return (uint32_t)((w >> (bit & 7)) & (0xFFFFFFFFFFFFFFFFull >> (64 - B)));
}

// dosage of the first allele of sample n, and whether n is missing.  The
// byte-pair path is the library's: 8-lane integer sums times 1/255 in its
// vector loop (samples below vec_end), its 511-entry table after it.
__device__ inline float ref_dose(const uint8_t* blk, const BgenMeta& m, uint32_t n, uint32_t vec_end,
                                 const float* lut, bool& miss) {
  const uint8_t* q = blk + m.prob;
  if (m.flags & kLayout1) {
    const uint8_t* s = q + 6ull * n;
    const uint32_t hom = s[0] | (s[1] << 8), het = s[2] | (s[3] << 8), alt = s[4] | (s[5] << 8);
    miss = (hom | het | alt) == 0;
    return (float)(hom * 2u + het) * m.factor;
  }
  const uint8_t pl = blk[8 + n];
  miss = (pl & 0x80) != 0;
  if (m.flags & kFast8) {
    const uint32_t i = 2u * q[2ull * n] + q[2ull * n + 1];
    return n < vec_end ? (float)i * (1.0f / 255.0f) : lut[i];
  }
  const uint32_t cp = (m.flags & kConst) ? m.maxp : (pl & 63u);
  const uint64_t bit = 2ull * m.bits * n;
  const uint32_t hom = bits_at(q, bit, m.bits), het = bits_at(q, bit + m.bits, m.bits);
  return (float)(hom * cp + het * (cp / 2u)) * m.factor;
}

// utils.cpp minor_certain: does a 5-sigma interval around freq exclude 0.5?
__device__ inline bool minor_certain(double freq, int n_checked, double z) {
#pragma clang fp contract(off)
  const double delta = z * sqrt((freq * (1.0 - freq)) / (double)n_checked);
  return !((freq - delta < 0.5) & (freq + delta > 0.5));
}

constexpr int kBT = 256;

__device__ inline double block_sum_d(double x, double* red) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// mode 0: one workgroup per variant of the chunk; row, F and minor index go to
// the variant's global index row0 + blockIdx.x
__global__ __launch_bounds__(kBT) void k_bgen_dosage(const uint8_t* __restrict__ buf,
                                                     const BgenMeta* __restrict__ meta, uint32_t nsamples,
                                                     uint32_t vec_end, const float* __restrict__ lut,
                                                     int standardize, double ploidy, uint64_t row0,
                                                     float* __restrict__ out, double* __restrict__ F,
                                                     uint8_t* __restrict__ minor) {
#pragma clang fp contract(off)
  __shared__ double red[4];
  __shared__ double scan[kBT];
  __shared__ uint32_t first;
  const BgenMeta m = meta[blockIdx.x];
  const uint8_t* blk = buf + m.off;
  const uint32_t t = threadIdx.x;
  const uint64_t gv = row0 + blockIdx.x;
  bool miss;
  // find_minor_allele: residue classes n = r (mod inc) in order, each summed
  // over its samples; the running total after class r gives
  // freq = total / (200 (r + 1)), checked with n_checked = 100 (r + 1).  The
  // dosages of 8-bit data are multiples of 2^-32 below 2, so the double sums
  // are exact below 2^20 samples and the order of the additions is immaterial.
  const uint32_t inc = max(nsamples / 100u, 1u);
  double base = 0.0, freq = 0.0;
  for (uint32_t c0 = 0; c0 < inc; c0 += kBT) {
    const uint32_t r = c0 + t;
    double s = 0.0;
    if (r < inc)
      for (uint32_t n = r; n < nsamples; n += inc) s += (double)ref_dose(blk, m, n, vec_end, lut, miss);
    scan[t] = s;
    if (t == 0) first = 0xffffffffu;
    __syncthreads();
    for (uint32_t o = 1; o < kBT; o <<= 1) {  // inclusive scan of the class sums
      const double a = t >= o ? scan[t - o] : 0.0;
      __syncthreads();
      scan[t] += a;
      __syncthreads();
    }
    if (r < inc) {
      const double tot = base + scan[t];
      if (minor_certain(tot / (double)(200u * (r + 1)), (int)(100u * (r + 1)), 5.0)) atomicMin(&first, r);
    }
    __syncthreads();
    const uint32_t last = min(inc, c0 + kBT) - 1;
    if (first != 0xffffffffu || last + 1 == inc) {
      const uint32_t r1 = first != 0xffffffffu ? first : last;
      freq = (base + scan[r1 - c0]) / (double)(200u * (r1 + 1));
      break;
    }
    base += scan[last - c0];
    __syncthreads();  // scan is rewritten
  }
  const bool swap = !(freq <= 0.5);
  // F over the non-missing samples of the minor allele dosage
  double gs = 0.0, gc = 0.0;
  for (uint32_t n = t; n < nsamples; n += kBT) {
    float d = ref_dose(blk, m, n, vec_end, lut, miss);
    if (swap) d = 2.0f - d;
    if (!miss) {
      gs += (double)d / 2.0;
      gc += 1.0;
    }
  }
  gs = block_sum_d(gs, red);
  gc = block_sum_d(gc, red);
  const double af = gc == 0.0 ? 0.0 : gs / gc;
  const double sd = sqrt(ploidy * af * (1.0 - af));
  const bool div = standardize && sd > 1e-9;
  float* row = out + gv * nsamples;
  for (uint32_t n = t; n < nsamples; n += kBT) {
    float d = ref_dose(blk, m, n, vec_end, lut, miss);
    if (swap) d = 2.0f - d;
    double v = miss ? 0.0 : (double)d / 2.0 - af;
    if (div) v /= sd;
    row[n] = (float)v;
  }
  if (t == 0) {
    F[gv] = af;
    minor[gv] = swap ? 1 : 0;
  }
}

// mode 1: P[variant][sample][0..1] (probabilities_layout1 / _layout2 for
// diploid unphased biallelic data; missing samples NaN)
__global__ __launch_bounds__(kBT) void k_bgen_probs(const uint8_t* __restrict__ buf, const BgenMeta* __restrict__ meta,
                                                    uint32_t nsamples, const float* __restrict__ lut, uint64_t row0,
                                                    float* __restrict__ P) {
  const BgenMeta m = meta[blockIdx.x];
  const uint8_t* blk = buf + m.off;
  const uint8_t* q = blk + m.prob;
  float2* row = reinterpret_cast<float2*>(P) + (row0 + blockIdx.x) * nsamples;
  const float qnan = __builtin_nanf("");
  for (uint32_t n = threadIdx.x; n < nsamples; n += kBT) {
    float p0, p1;
    bool miss;
    if (m.flags & kLayout1) {
      const uint8_t* s = q + 6ull * n;
      const uint32_t a = s[0] | (s[1] << 8), b = s[2] | (s[3] << 8), c = s[4] | (s[5] << 8);
      p0 = (float)a * m.factor;
      p1 = (float)b * m.factor;
      miss = (a | b | c) == 0;
    } else {
      miss = (blk[8 + n] & 0x80) != 0;
      if (m.flags & kFast8) {
        p0 = lut[q[2ull * n]];
        p1 = lut[q[2ull * n + 1]];
      } else {
        const uint64_t bit = 2ull * m.bits * n;
        p0 = (float)bits_at(q, bit, m.bits) * m.pfactor;
        p1 = (float)bits_at(q, bit + m.bits, m.bits) * m.pfactor;
      }
    }
    row[n] = miss ? make_float2(qnan, qnan) : make_float2(p0, p1);
  }
}

// rows idx[r] of src to row r of dst
__global__ __launch_bounds__(kBT) void k_gather_rows(const float* __restrict__ src, const int32_t* __restrict__ idx,
                                                     uint32_t ncol, float* __restrict__ dst) {
  const float* s = src + (uint64_t)idx[blockIdx.x] * ncol;
  float* d = dst + (uint64_t)blockIdx.x * ncol;
  for (uint32_t j = threadIdx.x; j < ncol; j += kBT) d[j] = s[j];
}

// ---- host ------------------------------------------------------------------

// genotypes.cpp's lut8: i / 255 rounded to 7 decimals, read as float
std::vector<float> bgen_lut8() {
  std::vector<float> lut(511);
  char txt[32];
  for (int i = 0; i < 511; ++i) {
    std::snprintf(txt, sizeof txt, "%.7f", i / 255.0);
    lut[i] = std::strtof(txt, nullptr);
  }
  return lut;
}

struct PinnedBuf {
  uint8_t* p = nullptr;
  size_t n = 0;
  void reserve(size_t bytes) {
    if (bytes <= n) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    PCAONE_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), bytes, 0));
    n = bytes;
  }
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
};

struct EventPair {
  hipEvent_t e[2] = {nullptr, nullptr};
  bool rec[2] = {false, false};
  EventPair() {
    for (auto& x : e) PCAONE_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  }
  ~EventPair() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

constexpr uint64_t kChunkBytes = 128ull << 20;

// one variant's genotype block decompressed to dst (dlen bytes + 8 zero bytes),
// its header checked and its decode parameters returned
BgenMeta inflate_block(const BgenIndex& ix, uint32_t v, uint8_t* dst, const Zstd* z, int mode) {
  const uint8_t* src = static_cast<const uint8_t*>(ix.map) + ix.data[v];
  const uint32_t dlen = ix.dlen[v], clen = ix.clen[v];
  if (ix.compression == 0) {
    std::memcpy(dst, src, dlen);
  } else if (ix.compression == 1) {
    uLongf out = dlen;
    const int r = uncompress(dst, &out, src, clen);
    PCAONE_REQUIRE(r == Z_OK && out == dlen, -1, "zlib decompression gave data of wrong length");
  } else {
    const size_t r = z->decompress(dst, dlen, src, clen);
    PCAONE_REQUIRE(!z->is_error(r) && r == dlen, -1, "zstd decompression gave data of wrong length");
  }
  std::memset(dst + dlen, 0, 8);
  const uint32_t N = ix.nsamples;
  BgenMeta m{};
  if (ix.layout == 1) {
    PCAONE_REQUIRE(dlen == 6ull * N, -1, "load_bgen: a layout 1 block is not 6 bytes per sample");
    m.prob = 0;
    m.bits = 16;
    m.maxp = 2;
    m.flags = kConst | kLayout1;
    m.factor = 1.0f / 32768;
    m.pfactor = m.factor;
    return m;
  }
  PCAONE_REQUIRE(dlen >= 10ull + N, -1, "load_bgen: a genotype block is shorter than its header");
  PCAONE_REQUIRE(rd32(dst) == N, -1, "number of samples doesn't match!");
  const uint32_t K = rd16(dst + 4);
  PCAONE_REQUIRE(K == ix.nalleles[v], -1, "number of alleles doesn't match!");
  PCAONE_REQUIRE(K == 2, -1, "can't get allele dosages for non-biallelic var.");
  const uint32_t pmin = dst[6], pmax = dst[7], phased = dst[8 + N], B = dst[9 + N];
  PCAONE_REQUIRE(B >= 1 && B <= 32, -1, "probabilities bit depth out of bounds");
  const bool cons = pmin == pmax;
  if (mode == 1)
    PCAONE_REQUIRE(cons && pmax == 2 && !phased, -1,
                   "load_bgen: --pcangsd reads diploid unphased probabilities with one ploidy for all samples");
  // both paths read two B-bit values per sample
  PCAONE_REQUIRE(dlen >= 10ull + N + ceil_div(2ull * N * B, 8), -1,
                 "load_bgen: a genotype block is shorter than two probabilities per sample");
  m.prob = 10 + N;
  m.bits = (uint8_t)B;
  m.maxp = (uint8_t)pmax;
  m.flags = (uint8_t)((cons ? kConst : 0) | ((cons && !phased && pmax == 2 && B == 8) ? kFast8 : 0));
  const uint32_t maxval = (uint32_t)(std::pow(2.0, (double)B) - 1.0);
  m.factor = 1.0f / (float)maxval;
  const float den = (float)std::pow(2.0, (double)B) - 1;
  m.pfactor = (float)(1.0 / den);
  return m;
}

}  // namespace

void Engine::load_bgen(const char* path, int mode, double maf, int standardize, int ploidy, int nthreads, double* F,
                       uint8_t* minor) {
  PCAONE_REQUIRE(mode == 0 || mode == 1, -1, "load_bgen: mode is 0 (dosages) or 1 (--pcangsd)");
  PCAONE_REQUIRE(ploidy == 1 || ploidy == 2, -1, "load_bgen: ploidy is 1 or 2");
  if (maf > 0.5) maf = 1.0 - maf;  // src/Cmd.cpp:196-199
  PCAONE_REQUIRE(mode == 0 || maf <= 0.0, -1, "load_bgen: --maf with --pcangsd is not supported on BGEN input");
  BgenIndex ix;
  bgen_index(path, ix);
  const uint32_t N = ix.nsamples, V = ix.nvariants;
  std::unique_ptr<Zstd> z;
  if (ix.compression == 2) z.reset(new Zstd("load_bgen"));
  PCAONE_HIP(hipSetDevice(device_));
  end_stream();
  have_op_ = have_mv_ = false;
  // chunks of consecutive variants, each block 16-byte aligned plus 8 pad bytes
  std::vector<uint64_t> boff(V);
  std::vector<uint32_t> cstart{0};
  uint64_t acc = 0, slot_bytes = 0;
  for (uint32_t v = 0; v < V; ++v) {
    const uint64_t b = round_up((uint64_t)ix.dlen[v] + 8, 16);
    if (acc > 0 && acc + b > kChunkBytes) {
      slot_bytes = std::max(slot_bytes, acc);
      cstart.push_back(v);
      acc = 0;
    }
    boff[v] = acc;
    acc += b;
  }
  slot_bytes = std::max(slot_bytes, acc);
  cstart.push_back(V);
  const uint32_t nchunk = (uint32_t)cstart.size() - 1;
  uint32_t maxvar = 0;
  for (uint32_t c = 0; c < nchunk; ++c) maxvar = std::max(maxvar, cstart[c + 1] - cstart[c]);
  const uint64_t meta_at = round_up(slot_bytes, 256);
  const uint64_t pin_bytes = meta_at + (uint64_t)maxvar * sizeof(BgenMeta);

  PinnedBuf pin[2];
  DevBuf<uint8_t> dbuf[2];
  for (int s = 0; s < 2 && s < (int)nchunk; ++s) {
    pin[s].reserve(pin_bytes);
    dbuf[s].reserve(pin_bytes);
  }
  const std::vector<float> lut_h = bgen_lut8();
  DevBuf<float> lut;
  lut.reserve(lut_h.size());
  PCAONE_HIP(hipMemcpyAsync(lut.p, lut_h.data(), lut_h.size() * 4, hipMemcpyHostToDevice, stream_));
  DevBuf<double> dF;
  DevBuf<uint8_t> dminor;
  if (mode == 0) {
    rawf_.reserve((uint64_t)V * N);
    dF.reserve(V);
    dminor.reserve(V);
  } else {
    glP_.reserve((uint64_t)V * N * 2);
  }
  const uint32_t vec_end = N > 32 ? (uint32_t)round_up(N - 32, 16) : 0;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const unsigned nt = nthreads > 0 ? (unsigned)nthreads : std::min(hw, 16u);
  EventPair ev;
  // on any exit the stream drains before the staging buffers are freed
  struct Drain {
    hipStream_t s;
    ~Drain() { (void)hipStreamSynchronize(s); }
  } drain{stream_};
  for (uint32_t c = 0; c < nchunk; ++c) {
    const int s = (int)(c & 1);
    if (ev.rec[s]) PCAONE_HIP(hipEventSynchronize(ev.e[s]));  // slot s is free again
    const uint32_t v0 = cstart[c], nv = cstart[c + 1] - v0;
    BgenMeta* meta = reinterpret_cast<BgenMeta*>(pin[s].p + meta_at);
    std::atomic<uint32_t> next{0};
    std::vector<std::string> err(nt);
    auto work = [&](unsigned t) {
      try {
        for (uint32_t i = next.fetch_add(1); i < nv; i = next.fetch_add(1)) {
          BgenMeta m = inflate_block(ix, v0 + i, pin[s].p + boff[v0 + i], z.get(), mode);
          m.off = boff[v0 + i];
          meta[i] = m;
        }
      } catch (const std::exception& e) {
        err[t] = e.what();
        next.store(nv);
      }
    };
    const unsigned ntc = std::max(1u, std::min<unsigned>(nt, nv));
    std::vector<std::thread> th;
    for (unsigned t = 1; t < ntc; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    for (auto& e : err) PCAONE_REQUIRE(e.empty(), -1, e);
    const uint64_t used = boff[v0 + nv - 1] + round_up((uint64_t)ix.dlen[v0 + nv - 1] + 8, 16);
    PCAONE_HIP(hipMemcpyAsync(dbuf[s].p, pin[s].p, used, hipMemcpyHostToDevice, stream_));
    PCAONE_HIP(hipMemcpyAsync(dbuf[s].p + meta_at, meta, (size_t)nv * sizeof(BgenMeta), hipMemcpyHostToDevice,
                              stream_));
    const BgenMeta* dmeta = reinterpret_cast<const BgenMeta*>(dbuf[s].p + meta_at);
    if (mode == 0) {
      timed("bgen_dosage", [&] {
        hipLaunchKernelGGL(k_bgen_dosage, dim3(nv), dim3(kBT), 0, stream_, dbuf[s].p, dmeta, N, vec_end, lut.p,
                           standardize, (double)ploidy, (uint64_t)v0, rawf_.p, dF.p, dminor.p);
      });
    } else {
      timed("bgen_probs", [&] {
        hipLaunchKernelGGL(k_bgen_probs, dim3(nv), dim3(kBT), 0, stream_, dbuf[s].p, dmeta, N, lut.p, (uint64_t)v0,
                           glP_.p);
      });
    }
    PCAONE_HIP(hipEventRecord(ev.e[s], stream_));
    ev.rec[s] = true;
  }
  N_ = N;
  if (mode == 1) {
    sync();
    M_ = V;
    mu_orig_.assign(M_, 0.0);
    F_orig_.assign(M_, 0.0);
    counts_orig_.clear();
    have_bed_ = true;
    src_ = SRC_DENSE;
    gl_mode_ = true;
    have_glF_ = gl_bed_ = have_resid_ = have_ldg_ = have_grm_ = false;
    return;
  }
  std::vector<double> Fh(V);
  std::vector<uint8_t> mh(V);
  PCAONE_HIP(hipMemcpyAsync(Fh.data(), dF.p, (size_t)V * 8, hipMemcpyDeviceToHost, stream_));
  PCAONE_HIP(hipMemcpyAsync(mh.data(), dminor.p, V, hipMemcpyDeviceToHost, stream_));
  sync();
  if (F) std::memcpy(F, Fh.data(), (size_t)V * 8);
  if (minor) std::memcpy(minor, mh.data(), V);
  // MAF filter (src/FileBgen.cpp:39-42): a variant stays when F > maf
  std::vector<int32_t> keep;
  keep.reserve(V);
  for (uint32_t v = 0; v < V; ++v)
    if (Fh[v] > maf) keep.push_back((int32_t)v);
  PCAONE_REQUIRE(!keep.empty(), -1, "the number of SNPs after filtering should be 0!");
  if (keep.size() < V) {
    DevBuf<int32_t> didx;
    didx.reserve(keep.size());
    DevBuf<float> kept;
    kept.reserve((uint64_t)keep.size() * N);
    PCAONE_HIP(hipMemcpyAsync(didx.p, keep.data(), keep.size() * 4, hipMemcpyHostToDevice, stream_));
    timed("bgen_keep", [&] {
      hipLaunchKernelGGL(k_gather_rows, dim3((uint32_t)keep.size()), dim3(kBT), 0, stream_, rawf_.p, didx.p, N,
                         kept.p);
    });
    sync();
    rawf_ = std::move(kept);
  }
  M_ = (uint32_t)keep.size();
  load_dense_finish(false);
}

}  // namespace pcaone
