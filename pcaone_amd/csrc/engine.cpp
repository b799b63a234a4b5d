// Host side of the RSVD hot path: HBM layout, the Normal / Fancy computeGandH
// schedules and computeUSV, all driving the gfx950 kernels on one stream.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

namespace pcaone {

// ---------------------------------------------------------------------------
// kernel timing
// ---------------------------------------------------------------------------
void KernelTimer::resolve() {
  for (auto& r : pending) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto& e = agg[r.name];
      e.first += 1;
      e.second += ms;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  pending.clear();
}

void KernelTimer::clear() {
  resolve();
  agg.clear();
}

KernelTimer::~KernelTimer() {
  for (auto& r : pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
}

// ---------------------------------------------------------------------------
// winSVD window schedule, src/Halko.cpp:159-223 (in core: bands windows);
// out of core (src/Halko.cpp:227-269) the bands are cut into nblocks blocks
// and the schedule counts blocks, starting at the band factor
// ---------------------------------------------------------------------------
FancyPlan fancy_plan(int bands, int pi, uint64_t& bandsize, uint32_t nblocks, uint32_t factor) {
  FancyPlan pl;
  if (nblocks == 0) {
    nblocks = (uint32_t)bands;
    factor = 1;
  }
  pl.reset_both = std::ldexp(1.0, pi) >= (double)bands;
  if (pi == 0) bandsize = factor;
  bandsize = std::min<uint64_t>(bandsize * 2, (uint64_t)nblocks);
  uint64_t i = 1;
  for (uint32_t b = 0; b < nblocks; ++b, ++i) {
    WindowStep st{(i <= bandsize / 2) ? 1 : 2, false, 0};
    const bool adjacent = pi > 0 && (double)(b + 1) == std::ldexp(1.0, pi - 1) * factor && std::ldexp(1.0, pi) < bands;
    if (!((uint64_t)(b + 1) < bandsize && !adjacent)) {
      if (i == bandsize || i == bandsize / 2 || adjacent) {
        st.update = true;
        if (i == bandsize) {
          st.zero = 1;
          i = 0;
        } else {
          st.zero = 2;
        }
      }
    }
    pl.steps.push_back(st);
  }
  return pl;
}

// ---------------------------------------------------------------------------
// lifecycle
// ---------------------------------------------------------------------------
Engine::Engine(int device) : device_(device) {
  int n = 0;
  PCAONE_HIP(hipGetDeviceCount(&n));
  PCAONE_REQUIRE(device >= 0 && device < n, -1, "no such HIP device");
  PCAONE_HIP(hipSetDevice(device));
  PCAONE_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  PCAONE_HIP(hipStreamCreateWithFlags(&ustream_, hipStreamNonBlocking));
  PCAONE_HIP(hipEventCreateWithFlags(&ev_epoch_, hipEventDisableTiming));
  PCAONE_HIP(hipHostMalloc(&hpipe_, 64, 0));
  PCAONE_HIP(hipDeviceGetAttribute(&ncu_, hipDeviceAttributeMultiprocessorCount, device));
  (void)ld_scratch(device);  // the allocation-free entry point's scratch address, once per device
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  try {
    end_stream();
  } catch (...) {
  }
  if (ustream_) (void)hipStreamSynchronize(ustream_);
  if (copy_stream_) (void)hipStreamDestroy(copy_stream_);
  destroy_blas(blas_);
  timer.resolve();
  if (ev_epoch_) (void)hipEventDestroy(ev_epoch_);
  if (hpipe_) (void)hipHostFree(hpipe_);
  if (ustream_) (void)hipStreamDestroy(ustream_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::sync() { PCAONE_HIP(hipStreamSynchronize(stream_)); }

// ---------------------------------------------------------------------------
// data residency, src/FilePlink.cpp:29-131
// ---------------------------------------------------------------------------
void Engine::load_bed(const uint8_t* payload, uint32_t nsamples, uint32_t nsnps) {
  PCAONE_REQUIRE(payload && nsamples > 0 && nsnps > 0, -1, "load_bed: empty input");
  PCAONE_HIP(hipSetDevice(device_));
  have_op_ = false;
  end_stream();
  N_ = nsamples;
  M_ = nsnps;
  const uint64_t nb = (N_ + 3) >> 2;
  raw_.reserve(nb * M_);
  PCAONE_HIP(hipMemcpyAsync(raw_.p, payload, nb * M_, hipMemcpyHostToDevice, stream_));
  // identity SNP-major layout to count codes
  const uint32_t rb = (uint32_t)round_up(nb, 64);
  DevBuf<uint8_t> tmp;
  tmp.reserve((uint64_t)rb * M_);
  DevBuf<int32_t> cnt;
  cnt.reserve((uint64_t)M_ * 4);
  timed("bed_layout", [&] { launch_bed_to_snpmajor(raw_.p, N_, M_, nullptr, tmp.p, rb, stream_); });
  timed("count_codes", [&] { launch_count_codes(tmp.p, M_, rb, N_, cnt.p, stream_); });
  std::vector<int32_t> h((size_t)M_ * 4);
  PCAONE_HIP(hipMemcpyAsync(h.data(), cnt.p, h.size() * 4, hipMemcpyDeviceToHost, stream_));
  sync();
  set_code_counts(h);
}

// F from the per-SNP code counts (00, 01, 10, 11), src/FilePlink.cpp:36-62
void Engine::set_code_counts(const std::vector<int32_t>& h) {
  F_orig_.assign(M_, 0.0);
  counts_orig_.assign((size_t)M_ * 4, 0);
  for (uint32_t i = 0; i < M_; ++i) {
    for (int c = 0; c < 4; ++c) counts_orig_[4 * (size_t)i + c] = h[4 * (size_t)i + c];
    const int64_t n00 = h[4 * (size_t)i], n10 = h[4 * (size_t)i + 2], n11 = h[4 * (size_t)i + 3];
    const int64_t n = n00 + n10 + n11;
    // sum of BED2GENO over non-missing (exact in double) / count
    F_orig_[i] = n ? ((double)n00 + 0.5 * (double)n10) / (double)n : 0.0;
  }
  have_bed_ = true;
  src_ = SRC_BED;
  gl_mode_ = have_glF_ = gl_bed_ = have_resid_ = have_ldg_ = have_grm_ = false;
}

// Replaces FileBin::read_all (src/FileBinary.cpp:21-30): nfeatures rows of
// nsamples float32 stay in HBM; each feature's mean (FP64) is subtracted in
// the contraction epilogues.
void Engine::load_dense(const float* featmajor, uint32_t nsamples, uint32_t nfeatures) {
  PCAONE_REQUIRE(featmajor && nsamples > 0 && nfeatures > 0, -1, "load_dense: empty input");
  PCAONE_HIP(hipSetDevice(device_));
  have_op_ = false;
  N_ = nsamples;
  M_ = nfeatures;
  rawf_.reserve((uint64_t)N_ * M_);
  PCAONE_HIP(hipMemcpyAsync(rawf_.p, featmajor, (uint64_t)N_ * M_ * 4, hipMemcpyHostToDevice, stream_));
  load_dense_finish(true);
}

// The features in rawf_ become the operator's matrix: with `centre`, each
// feature's FP64 mean is what the epilogues subtract; without, nothing is.
void Engine::load_dense_finish(bool centre) {
  end_stream();
  mu_orig_.assign(M_, 0.0);
  if (centre) {
    DevBuf<double> mu;
    mu.reserve(M_);
    timed("dense_mean", [&] { launch_dense_mean(rawf_.p, N_, M_, mu.p, stream_); });
    PCAONE_HIP(hipMemcpyAsync(mu_orig_.data(), mu.p, (size_t)M_ * 8, hipMemcpyDeviceToHost, stream_));
  }
  sync();
  // the epilogues subtract 2 F_i colsum, so F carries mu / 2
  F_orig_.resize(M_);
  for (uint32_t i = 0; i < M_; ++i) F_orig_[i] = 0.5 * mu_orig_[i];
  counts_orig_.clear();
  have_bed_ = true;
  src_ = SRC_DENSE;
  gl_mode_ = have_glF_ = gl_bed_ = have_resid_ = have_ldg_ = have_grm_ = false;
}

void Engine::feature_means(double* mu) const {
  PCAONE_REQUIRE(have_bed_ && src_ == SRC_DENSE, -3, "feature_means: no dense matrix loaded");
  std::memcpy(mu, mu_orig_.data(), mu_orig_.size() * sizeof(double));
}

void Engine::dense_rows(uint32_t r0, uint32_t nr, float* out) {
  PCAONE_REQUIRE(have_bed_ && src_ == SRC_DENSE && !gl_mode_, -3, "dense_rows: no dense matrix loaded");
  PCAONE_REQUIRE(out && (uint64_t)r0 + nr <= M_, -1, "dense_rows: rows out of range");
  PCAONE_HIP(hipSetDevice(device_));
  PCAONE_HIP(hipMemcpyAsync(out, rawf_.p + (uint64_t)r0 * N_, (size_t)nr * N_ * sizeof(float),
                            hipMemcpyDeviceToHost, stream_));
  sync();
}

void Engine::allele_freq(double* F, int64_t* counts) const {
  PCAONE_REQUIRE(have_bed_ && src_ == SRC_BED, -3, "allele_freq: no genotypes loaded");
  if (F) std::memcpy(F, F_orig_.data(), F_orig_.size() * sizeof(double));
  if (counts) std::memcpy(counts, counts_orig_.data(), counts_orig_.size() * sizeof(int64_t));
}

void Engine::dims(int64_t* rows, int64_t* cols, int64_t* size) const {
  PCAONE_REQUIRE(have_op_, -3, "rsvd_dims: operator not initialised");
  if (rows) *rows = M_;
  if (cols) *cols = N_;
  if (size) *size = l_;
}

void Engine::rsvd_perm(int32_t* perm) const {
  PCAONE_REQUIRE(have_op_, -3, "rsvd_perm: operator not initialised");
  std::memcpy(perm, perm_.data(), perm_.size() * sizeof(int32_t));
}

// ---------------------------------------------------------------------------
// operator set-up
// ---------------------------------------------------------------------------
// split-K factor: about 4,096 workgroups (16 per CU, so each CU holds several
// blocks whose K-chunk loads overlap), at least four K chunks per slice, at most
// 128 slices.  A window segment of a few hundred rows against a long K (200,000
// samples) needs a deep split; the all-window dense G product (79 row blocks)
// moved its 16 GB at 1.9 TB/s on 553 blocks (a target of 512) and takes 18%
// less time at 4,096.
static uint32_t split_for(uint64_t blocks, uint32_t kchunks, uint32_t& cps) {
  uint64_t want = ceil_div(4096, std::max<uint64_t>(blocks, 1));
  want = std::min<uint64_t>(std::max<uint64_t>(want, 1), std::min<uint32_t>(std::max<uint32_t>(kchunks / 4, 1), 128));
  cps = (uint32_t)ceil_div(kchunks, want);
  return (uint32_t)ceil_div(kchunks, cps);
}

// Split-K factor of the split-BF16 packed contraction: `blocks` row blocks (all
// windows) over kchunks K chunks with `slots` blocks resident on the chip at
// once.  The split (at most 8) minimises rounds of resident blocks times the
// chunks each block runs plus one chunk's worth for its epilogue, the smallest
// split on a tie.  (Counting only "enough blocks" put 320 blocks on 256 slots
// for the 16-window H product: two rounds of whole-K blocks.)
static uint32_t split_rounds(uint64_t blocks, uint32_t kchunks, uint32_t slots, uint32_t& cps) {
  uint32_t best = 1;
  uint64_t best_t = ~0ull;
  for (uint32_t s = 1; s <= std::min<uint32_t>(kchunks, 8); ++s) {
    const uint32_t c = (uint32_t)ceil_div(kchunks, s), ns = (uint32_t)ceil_div(kchunks, c);
    const uint64_t t = ceil_div(blocks * ns, slots) * (uint64_t)(c + 1);
    if (t < best_t) {
      best_t = t;
      best = ns;
    }
  }
  cps = (uint32_t)ceil_div(kchunks, best);
  return (uint32_t)ceil_div(kchunks, cps);
}

// split-K of one packed contraction launch: rows per window group, nz groups
uint32_t Engine::packed_split(int prec, uint64_t rows, uint32_t nz, uint32_t kchunks, uint32_t& cps) const {
  if (prec == PREC_SPLIT_BF16) {
    // the streamed kernel: 512-row blocks, one resident per CU; the 4-wave one:
    // 256-row blocks, two per CU
    const bool big = packed_bf16_streams(rows, nz);
    return split_rounds(ceil_div(rows, big ? 512 : 256) * nz, kchunks, big ? 256 : 512, cps);
  }
  return split_for(ceil_div(rows, rows_per_block_(prec)) * nz, kchunks, cps);
}

void Engine::rsvd_init(const RsvdParams& p, const double* omega, const int32_t* perm) {
  PCAONE_REQUIRE(have_bed_, -3, "rsvd_init: no data loaded");
  PCAONE_HIP(hipSetDevice(device_));
  side_ = SideJobs{};  // pending jobs point into buffers this may reallocate
  have_mv_ = false;
  PCAONE_REQUIRE(p.variant == 1 || p.variant == 2, -1, "rsvd_init: variant must be 1 (normal) or 2 (fancy)");
  PCAONE_REQUIRE(p.precision >= 0 && p.precision <= 2, -1, "rsvd_init: bad precision");
  PCAONE_REQUIRE(p.k >= 1 && p.os >= 0, -1, "rsvd_init: bad k / oversamples");
  // the reference's rule (src/Cmd.cpp:213-214): at least 4 and even
  PCAONE_REQUIRE(p.bands >= 4 && p.bands % 2 == 0, -1, "the -w/--batches must be a power of 2 and the minimun is 4.");
  PCAONE_REQUIRE(p.ploidy == 1 || p.ploidy == 2, -1, "rsvd_init: ploidy must be 1 or 2");
  const uint32_t l = (uint32_t)(p.k + p.os);
  PCAONE_REQUIRE(l <= kWideMaxL, -1, "rsvd_init: k + oversamples must be <= 128");
  PCAONE_REQUIRE(p.k <= 64, -1, "rsvd_init: k must be <= 64");
  PCAONE_REQUIRE(l <= N_ && l <= M_, -1, "rsvd_init: k + oversamples exceeds the matrix dimensions");
  p_ = p;
  if (src_ == SRC_DENSE) p_.standardize = 0;  // dense inputs are only centred (setFlags(false, false))
  l_ = l;
  // windows: winSVD uses blocksize = ceil(M / bands) (src/Halko.cpp:178); the
  // plain RSVD never reads -w (src/Halko.cpp:97-151) and cuts the SNPs into
  // its own windows, only as a tiling of the H contraction: 64, fewer when
  // there are fewer SNPs.
  // An out-of-core plan (-m, src/Data.cpp:27-71) gives the windows itself:
  // blocks of p.blocksize SNPs, and for the winSVD a schedule that counts
  // blocks from the band factor.
  if (p.blocksize > 0) {
    PCAONE_REQUIRE(p.band_factor >= 1, -1, "rsvd_init: band_factor must be >= 1");
    bs_ = p.blocksize;
    nwin_ = (uint32_t)ceil_div(M_, bs_);
    PCAONE_REQUIRE(nwin_ >= 2, -1, "only one block exists. please remove -m option");
  } else {
    if (p.variant == 1) {
      p_.bands = 64;
      while (p_.bands > 4 && (uint64_t)p_.bands > M_) p_.bands /= 2;
    }
    p_.band_factor = 1;
    bs_ = (uint32_t)ceil_div(M_, (uint64_t)p_.bands);
    nwin_ = (uint32_t)ceil_div(M_, bs_);
    if (p.variant == 2)
      PCAONE_REQUIRE(nwin_ == (uint32_t)p.bands, -1,
                     "winSVD needs every window non-empty: nsnps too small for the number of windows");
  }
  PCAONE_REQUIRE(N_ >= 2 * l_, -1, "the GPU operator needs nsamples >= 2 (k + oversamples)");
  // permutation (src/RSVD.hpp:63-79): only the winSVD path shuffles
  perm_.resize(M_);
  std::iota(perm_.begin(), perm_.end(), 0);
  if (p.variant == 2 && p.perm) {
    if (perm) {
      std::vector<char> seen(M_, 0);
      for (uint32_t i = 0; i < M_; ++i) {
        PCAONE_REQUIRE(perm[i] >= 0 && (uint32_t)perm[i] < M_ && !seen[perm[i]], -1, "rsvd_init: invalid perm");
        seen[perm[i]] = 1;
        perm_[i] = perm[i];
      }
    } else {
      auto rng = std::default_random_engine{};
      std::shuffle(perm_.data(), perm_.data() + perm_.size(), rng);
    }
  }
  omega_given_.clear();
  if (omega) omega_given_.assign(omega, omega + (size_t)N_ * l_);
  build_layout();

  // operands and partials
  const int prec = p.precision;
  const bool dense = src_ == SRC_DENSE;
  opsz_ = prec == PREC_SPLIT_BF16 ? 2 : prec == PREC_FP32 ? 4 : 8;
  part_f64_ = prec == PREC_FP64;
  const size_t partsz = part_f64_ ? 8 : 4;
  if (!dense) {
    // packed: split operand = [hi | lo | 0-pad] in nt16 tiles; the partials
    // keep both halves, or hi + lo per column where the contraction folds
    // them (l % 16 in {0, 8}; the streamed kernel, launches of 65536 rows or
    // more)
    nt16_ = (int)ceil_div(2 * l_, 16);
    bcols_ = prec == PREC_SPLIT_BF16 ? (uint32_t)nt16_ * 16 : l_;
    lo_off_ = l_;
    fold_ = (prec == PREC_SPLIT_BF16 && l_ <= 64 && (l_ % 16 == 0 || l_ % 16 == 8)) ? l_ : 0;
    ldp_ = prec == PREC_SPLIT_BF16 ? (uint64_t)nt16_ * 16 : l_;
    split_cols_ = prec == PREC_SPLIT_BF16;
    hf_ = 0.5;
  } else {
    // dense: FP32 arm in 16-column tiles; split arm = hi tiles | lo tiles, the
    // three products land in one accumulator per tile
    ntiles_ = (int)ceil_div(l_, 16);
    fold_ = 0;
    bcols_ = prec == PREC_FP64 ? l_ : prec == PREC_FP32 ? (uint32_t)ntiles_ * 16 : 2u * ntiles_ * 16;
    lo_off_ = (uint32_t)ntiles_ * 16;
    ldp_ = prec == PREC_FP64 ? l_ : (uint64_t)ntiles_ * 16;
    split_cols_ = false;
    hf_ = 1.0;
  }
  const uint64_t ldp = ldp_;
  omega_.reserve((uint64_t)N_ * l_);
  colsum_.reserve(l_);
  bg_.reserve((uint64_t)bcols_ * kg_ * opsz_);
  bh_.reserve((uint64_t)nwin_ * bcols_ * kw_ * opsz_);
  PCAONE_HIP(hipMemsetAsync(bh_.p, 0, bh_.bytes(), stream_));
  G_.reserve((uint64_t)M_ * l_);
  g2_.reserve((uint64_t)M_ * l_ * (prec == PREC_FP64 ? 8 : 4));
  if (nnz_snp_ > 0) {
    const uint64_t csz = prec == PREC_FP64 ? 8 : 4;  // the partials' type
    corrG_.reserve((uint64_t)M_ * l_ * csz);
    corrH_.reserve((uint64_t)nwin_ * N_ * l_ * csz);
    if (prec != PREC_FP64) omega32_.reserve((uint64_t)N_ * l_);
  }
  S2_.reserve((uint64_t)nwin_ * l_);
  s2p_.reserve((uint64_t)nwin_ * finalize_chunks(bs_) * l_);
  for (auto* h : {&H1_, &H2_, &H_}) {
    h->reserve((uint64_t)N_ * l_);
    PCAONE_HIP(hipMemsetAsync(h->p, 0, h->bytes(), stream_));
  }
  uint64_t gmax = 0, hmax = 0;
  partial_rows(prec, gmax, hmax);
  gpart_.reserve(gmax * ldp * partsz);
  hpart_.reserve(hmax * ldp * partsz);
  // linear algebra scratch
  wide_ = l_ > 64;
  robust_ = false;
  part_.reserve(std::max<uint64_t>({(uint64_t)1024 * l_ * l_, qr_part_doubles(l_), usv_part_doubles(l_),
                                    wide_ ? wide_part_doubles() : 0}));
  // reduced Grams between passes: tile layout (l <= 64) or full l x l (wide);
  // qred_ also holds the flip statistics
  const uint64_t red = std::max<uint64_t>({usv_gram_doubles(l_), (uint64_t)l_ * l_, 3ull * kWideMaxL});
  gred_.reserve(red);
  qred_.reserve(red);
  if (wide_) {
    wq_.reserve((uint64_t)std::max(N_, M_) * l_);
    eigtmp_.reserve((uint64_t)l_ * l_);
  }
  snapH1_.reserve((uint64_t)N_ * l_);
  snapH2_.reserve((uint64_t)N_ * l_);
  bbt_.reserve((uint64_t)l_ * l_);
  for (auto* b : {&R1_, &R1i_, &R2_, &R2i_, &Rt_, &Rti_, &W_}) b->reserve((uint64_t)l_ * l_);
  sign_.reserve(l_);
  fsign_.reserve(l_);
  qstat_.reserve(qr_stat_doubles());
  evals_.reserve(l_);
  sinv_.reserve(l_);
  Bt_.reserve((uint64_t)N_ * l_);
  Ucur_.reserve((uint64_t)N_ * p.k);
  Upre_.reserve((uint64_t)N_ * p.k);
  V_.reserve((uint64_t)M_ * p.k);
  scal_.reserve(4);
  miss_val_.reserve(std::max<int64_t>(nnz_snp_, 1));
  Sdev_.reserve(p.k);
  Vpre_.reserve((uint64_t)M_ * p.k);
  update_ = have_usv_ = have_vals_ = false;
  status_.reserve(2);
  PCAONE_HIP(hipMemsetAsync(status_.p, 0, 2 * sizeof(int), stream_));
  bandsize_ = 1;
  have_op_ = true;
  draw_omega();
  init_omega();
  sync();
}

void Engine::build_layout() {
  if (src_ == SRC_DENSE) {
    // rows (laid-out order, zero padded to 64 samples) for the G contraction and
    // a windowed transpose [window][sample][feature] for the H contraction
    kg_ = round_up(N_, 64);
    kw_ = round_up(bs_, 64);
    miss_rows_ = N_;
    d_perm_.reserve(M_);
    PCAONE_HIP(hipMemcpyAsync(d_perm_.p, perm_.data(), (size_t)M_ * 4, hipMemcpyHostToDevice, stream_));
    dns_.reserve((uint64_t)M_ * kg_);
    dnt_.reserve((uint64_t)nwin_ * N_ * kw_);
    if (gl_mode_) {
      // PCAngsd: the initial E (posterior at pt = F) straight into the laid-out rows
      PCAONE_REQUIRE(have_glF_, -3, "PCAngsd: estimate the allele frequencies (gl_em_maf) before rsvd_init");
      PCAONE_HIP(hipMemsetAsync(dns_.p, 0, dns_.bytes(), stream_));
      gl_fill_E(nullptr, nullptr, nullptr, 0, false, nullptr);
    } else {
      timed("dense_layout",
            [&] { launch_dense_rows(rawf_.p, N_, M_, d_perm_.p, dns_.p, (uint32_t)kg_, stream_); });
      timed("dense_layout", [&] {
        launch_dense_transpose(dns_.p, (uint32_t)kg_, M_, bs_, nwin_, N_, dnt_.p, (uint32_t)kw_, stream_);
      });
    }
    F_lay_.resize(M_);
    for (uint32_t r = 0; r < M_; ++r) F_lay_[r] = F_orig_[perm_[r]];
    F_.reserve(M_);
    PCAONE_HIP(hipMemcpyAsync(F_.p, F_lay_.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
    upload_scale();
    // no missing entries
    nnz_snp_ = 0;
    miss_snp_ptr_.reserve((uint64_t)M_ + 1);
    miss_win_ptr_.reserve((uint64_t)nwin_ * N_ + 1);
    PCAONE_HIP(hipMemsetAsync(miss_snp_ptr_.p, 0, miss_snp_ptr_.bytes(), stream_));
    PCAONE_HIP(hipMemsetAsync(miss_win_ptr_.p, 0, miss_win_ptr_.bytes(), stream_));
    miss_snp_idx_.reserve(1);
    miss_win_idx_.reserve(1);
    miss_win_pos_.reserve(1);
    sync();
    return;
  }
  const uint64_t nb = (N_ + 3) >> 2;
  row_bytes_ = (uint32_t)round_up(nb, 64);
  npad_ = row_bytes_ * 4;
  win_bytes_ = (uint32_t)round_up(ceil_div(bs_, 4), 64);
  kw_ = (uint64_t)win_bytes_ * 4;
  kg_ = npad_;
  miss_rows_ = npad_;
  if (streamed_) {
    // the windows are laid out in the ring at each use; F, d, the permutation
    // (V's rows come back through it) and the missing lists (one pass over the
    // rows) stay resident
    d_perm_.reserve(M_);
    PCAONE_HIP(hipMemcpyAsync(d_perm_.p, perm_.data(), (size_t)M_ * 4, hipMemcpyHostToDevice, stream_));
    F_lay_.resize(M_);
    for (uint32_t r = 0; r < M_; ++r) F_lay_[r] = F_orig_[perm_[r]];
    F_.reserve(M_);
    PCAONE_HIP(hipMemcpyAsync(F_.p, F_lay_.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
    upload_scale();
    stream_setup();
    stream_missing_lists();
    return;
  }
  d_perm_.reserve(M_);
  PCAONE_HIP(hipMemcpyAsync(d_perm_.p, perm_.data(), (size_t)M_ * 4, hipMemcpyHostToDevice, stream_));
  bed_snp_.reserve((uint64_t)row_bytes_ * M_);
  timed("bed_layout", [&] { launch_bed_to_snpmajor(raw_.p, N_, M_, d_perm_.p, bed_snp_.p, row_bytes_, stream_); });
  bed_smp_.reserve((uint64_t)nwin_ * npad_ * win_bytes_);
  PCAONE_HIP(hipMemsetAsync(bed_smp_.p, 0x55, bed_smp_.bytes(), stream_));
  timed("bed_transpose", [&] {
    launch_transpose_windows(bed_snp_.p, row_bytes_, M_, bs_, nwin_, bed_smp_.p, npad_, win_bytes_, stream_);
  });
  F_lay_.resize(M_);
  for (uint32_t r = 0; r < M_; ++r) F_lay_[r] = F_orig_[perm_[r]];
  F_.reserve(M_);
  PCAONE_HIP(hipMemcpyAsync(F_.p, F_lay_.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
  upload_scale();
  // missing lists
  auto build = [&](const PackedRows& pr, DevBuf<int64_t>& ptr, DevBuf<int32_t>& idx, const char* nm) -> int64_t {
    DevBuf<int32_t> cnt;
    cnt.reserve(pr.nrows);
    timed(nm, [&] { launch_count_missing(pr, cnt.p, stream_); });
    std::vector<int32_t> hc(pr.nrows);
    PCAONE_HIP(hipMemcpyAsync(hc.data(), cnt.p, (size_t)pr.nrows * 4, hipMemcpyDeviceToHost, stream_));
    sync();
    std::vector<int64_t> hp((size_t)pr.nrows + 1, 0);
    for (uint32_t r = 0; r < pr.nrows; ++r) hp[r + 1] = hp[r] + hc[r];
    ptr.reserve(hp.size());
    PCAONE_HIP(hipMemcpyAsync(ptr.p, hp.data(), hp.size() * 8, hipMemcpyHostToDevice, stream_));
    idx.reserve(std::max<int64_t>(hp.back(), 1));
    timed(nm, [&] { launch_fill_missing(pr, ptr.p, idx.p, stream_); });
    sync();  // hp must outlive the copy
    return hp.back();
  };
  PackedRows ps{bed_snp_.p, row_bytes_, M_, M_, M_, N_, N_, 1, 0};
  nnz_snp_ = build(ps, miss_snp_ptr_, miss_snp_idx_, "missing_lists");
  PackedRows pw{bed_smp_.p, win_bytes_, nwin_ * npad_, npad_, N_, bs_, M_ - (nwin_ - 1) * bs_, nwin_, bs_};
  const int64_t nnz_win = build(pw, miss_win_ptr_, miss_win_idx_, "missing_lists");
  PCAONE_REQUIRE(nnz_win == nnz_snp_, -2, "missing lists disagree between the two layouts");
  miss_win_pos_.reserve(std::max<int64_t>(nnz_win, 1));
  timed("missing_lists", [&] {
    launch_map_missing(miss_win_ptr_.p, miss_win_idx_.p, nwin_, npad_, N_, miss_snp_ptr_.p, miss_snp_idx_.p,
                       miss_win_pos_.p, stream_);
  });
}

// 1/sd of each laid-out SNP, or 1 when not standardising or sd <= VAR_TOL
// (src/Data.cpp:290-299, VAR_TOL = 1e-9 at src/Data.hpp:7)
void Engine::upload_scale() {
  std::vector<double> d(M_);
  for (uint32_t r = 0; r < M_; ++r) {
    const double f = F_lay_[r];
    const double sd = std::sqrt((double)p_.ploidy * f * (1 - f));
    d[r] = (p_.standardize && sd > 1e-9) ? 1.0 / sd : 1.0;
  }
  d_.reserve(M_);
  PCAONE_HIP(hipMemcpyAsync(d_.p, d.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
  sync();
}

// RsvdOpData::setFlags (src/Halko.hpp:32-35): `update` re-imputes the missing
// genotypes from the operator's U, S, V at the next pi == 0 (EMU), and
// `standardize` switches the 1/sd scaling.
void Engine::set_flags(bool update, bool standardize) {
  PCAONE_REQUIRE(have_op_, -3, "setFlags: operator not initialised");
  PCAONE_REQUIRE(!update || have_usv_, -3, "setFlags: update needs a previous computeUSV");
  update_ = update;
  if (!update) have_vals_ = false;
  if ((int)standardize != p_.standardize) {
    p_.standardize = standardize;
    upload_scale();
  }
}

// Largest partial-product row counts over every possible segment (split-K
// slices x rows) of the G and H contractions.
void Engine::partial_rows(int prec, uint64_t& gmax, uint64_t& hmax) const {
  const bool dense = src_ == SRC_DENSE;
  const uint32_t rpb = rows_per_block_(prec);
  const uint32_t kcg = (uint32_t)(kg_ / (dense ? 64 : 256)), kch = (uint32_t)(kw_ / (dense ? 64 : 256));
  gmax = hmax = 0;
  // streamed: every launch covers one window
  for (uint32_t nw = 1; nw <= (streamed_ ? 1u : nwin_); ++nw) {
    const uint64_t rows = std::min<uint64_t>((uint64_t)nw * bs_, M_);
    uint32_t cps;
    const uint32_t sg = dense ? split_for(ceil_div(rows, rpb), kcg, cps) : packed_split(prec, rows, 1, kcg, cps);
    gmax = std::max<uint64_t>(gmax, sg * rows);
    const uint32_t sh = dense ? split_for(ceil_div(N_, rpb) * nw, kch, cps) : packed_split(prec, N_, nw, kch, cps);
    hmax = std::max<uint64_t>(hmax, (uint64_t)sh * N_ * nw);
  }
}

// ---------------------------------------------------------------------------
// Gram mat-vec: the Spectra op of IRAM (ArnoldiOpData::perform_op,
// src/Arnoldi.cpp:18-46; SVDTallMatOp / SVDWideMatOp::perform_op,
// external/Spectra/contrib/PartialSVDSolver.h:65-71, 104-110), on the same
// contractions as the RSVD operator with one operand column.
// ---------------------------------------------------------------------------
void Engine::gram_init(int precision, bool standardize, int ploidy) {
  PCAONE_REQUIRE(have_bed_, -3, "gram_init: no data loaded");
  side_ = SideJobs{};  // pending jobs point into buffers this may reallocate
  PCAONE_REQUIRE(precision >= 0 && precision <= 2, -1, "gram_init: bad precision");
  RsvdParams p;
  p.variant = 1;  // identity SNP order, windows only tile the H contraction
  p.k = 1;
  p.os = 1;
  p.bands = 64;
  while (p.bands > 4 && (uint64_t)p.bands > M_) p.bands /= 2;
  p.precision = precision;
  p.standardize = standardize;
  p.ploidy = ploidy;
  rsvd_init(p, nullptr, nullptr);
  const bool dense = src_ == SRC_DENSE;
  SegBufs& b = mv_;
  b = SegBufs{};
  b.prec = precision;
  b.l = 1;
  b.opsz = precision == PREC_SPLIT_BF16 ? 2 : precision == PREC_FP32 ? 4 : 8;
  b.part_f64 = precision == PREC_FP64;
  if (!dense) {
    b.nt16 = 1;
    b.bcols = precision == PREC_SPLIT_BF16 ? 16 : 1;
    b.lo_off = 1;
    b.ldp = precision == PREC_SPLIT_BF16 ? 16 : 1;
    b.split_cols = precision == PREC_SPLIT_BF16;
    b.hf = 0.5;
  } else {
    b.ntiles = 1;
    b.bcols = precision == PREC_FP64 ? 1 : precision == PREC_FP32 ? 16 : 32;
    b.lo_off = 16;
    b.ldp = precision == PREC_FP64 ? 1 : 16;
    b.hf = 1.0;
  }
  uint64_t gmax, hmax;
  partial_rows(precision, gmax, hmax);
  const size_t partsz = b.part_f64 ? 8 : 4;
  mv_omega_.reserve(N_);
  mv_colsum_.reserve(1);
  mv_G_.reserve(M_);
  mv_g2_.reserve((uint64_t)M_ * (precision == PREC_FP64 ? 8 : 4));
  b.g2_f32 = precision != PREC_FP64;
  if (nnz_snp_ > 0) {
    mv_corrG_.reserve((uint64_t)M_ * partsz);
    mv_corrH_.reserve((uint64_t)nwin_ * N_ * partsz);
    if (b.g2_f32) mv_omega32_.reserve(N_);
    b.corrG = mv_corrG_.p;
    b.corrH = mv_corrH_.p;
    b.omega32 = mv_omega32_.p;
  }
  mv_S2_.reserve(nwin_);
  mv_s2p_.reserve((uint64_t)nwin_ * finalize_chunks(bs_));
  mv_H_.reserve(N_);
  mv_bg_.reserve((uint64_t)b.bcols * kg_ * b.opsz);
  mv_bh_.reserve((uint64_t)nwin_ * b.bcols * kw_ * b.opsz);
  PCAONE_HIP(hipMemsetAsync(mv_bh_.p, 0, mv_bh_.bytes(), stream_));
  mv_gpart_.reserve(gmax * b.ldp * partsz);
  mv_hpart_.reserve(hmax * b.ldp * partsz);
  b.omega = mv_omega_.p;
  b.colsum = mv_colsum_.p;
  b.G = mv_G_.p;
  b.g2 = mv_g2_.p;
  b.S2 = mv_S2_.p;
  b.s2p = mv_s2p_.p;
  b.affine = src_ == SRC_BED && precision == PREC_SPLIT_BF16;
  b.bg = mv_bg_.p;
  b.bh = mv_bh_.p;
  b.gpart = mv_gpart_.p;
  b.hpart = mv_hpart_.p;
  b.Ha = b.Hb = mv_H_.p;
  b.mval = nullptr;
  have_mv_ = true;
  nops_ = 0;
  sync();
}

void Engine::set_omega_vec(const double* dev_v, const double* host_v) {
  std::vector<double> h;
  if (!host_v) {
    h.resize(N_);
    PCAONE_HIP(hipMemcpyAsync(h.data(), dev_v, (size_t)N_ * 8, hipMemcpyDeviceToHost, stream_));
    sync();
    host_v = h.data();
  } else {
    PCAONE_HIP(hipMemcpyAsync(mv_omega_.p, host_v, (size_t)N_ * 8, hipMemcpyHostToDevice, stream_));
  }
  if (dev_v) PCAONE_HIP(hipMemcpyAsync(mv_omega_.p, dev_v, (size_t)N_ * 8, hipMemcpyDeviceToDevice, stream_));
  double cs = 0.0;
  for (uint32_t j = 0; j < N_; ++j) cs += host_v[j];
  PCAONE_HIP(hipMemcpyAsync(mv_colsum_.p, &cs, 8, hipMemcpyHostToDevice, stream_));
  timed("omega_operand", [&] {
    launch_omega_operand(mv_omega_.p, N_, 1, mv_.prec, mv_bg_.p, mv_.bcols, mv_.lo_off, kg_, kperm_(mv_.prec),
                         stream_);
  });
  refresh_omega32(mv_);
  sync();  // cs lives on this stack frame
}

// v = X x (x: M, original feature order; v: N)
void Engine::apply_X(const double* x, double* v) {
  const std::vector<int> route(nwin_, 1);
  std::vector<double> xl(M_);
  for (uint32_t r = 0; r < M_; ++r) xl[r] = x[perm_[r]];
  PCAONE_HIP(hipMemcpyAsync(mv_G_.p, xl.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
  FinalizeG fd = finalize_desc(0, nwin_, mv_);
  timed("finalize_G", [&] { launch_operand_from_G(fd, stream_); });
  timed("window_colsum", [&] { launch_window_colsum(mv_s2p_.p, 1, bs_, 0, nwin_, mv_S2_.p, stream_); });
  timed("fill", [&] { launch_fill_f64(mv_H_.p, N_, 0.0, stream_); });
  seg_H(0, nwin_, route, mv_);
  if (v) {
    PCAONE_HIP(hipMemcpyAsync(v, mv_H_.p, (size_t)N_ * 8, hipMemcpyDeviceToHost, stream_));
  }
  sync();
}

// y = X^T v (v: N; y: M, original feature order); v == nullptr uses mv_H_
void Engine::apply_Xt(const double* v, double* y) {
  if (v)
    set_omega_vec(nullptr, v);
  else
    set_omega_vec(mv_H_.p, nullptr);
  seg_G(0, nwin_, mv_);
  std::vector<double> gl(M_);
  PCAONE_HIP(hipMemcpyAsync(gl.data(), mv_G_.p, (size_t)M_ * 8, hipMemcpyDeviceToHost, stream_));
  sync();
  for (uint32_t r = 0; r < M_; ++r) y[perm_[r]] = gl[r];
}

void Engine::gram_matvec(int side, const double* x, double* y) {
  PCAONE_REQUIRE(have_mv_, -3, "gram_matvec: call gram_init first");
  PCAONE_REQUIRE(side == 0 || side == 1, -1, "gram_matvec: side must be 0 (X X^T x) or 1 (X^T X x)");
  PCAONE_REQUIRE(x && y, -1, "gram_matvec: null vector");
  PCAONE_HIP(hipSetDevice(device_));
  if (side == 0) {
    // y = X (X^T x): x plays Omega (N x 1)
    const std::vector<int> route(nwin_, 1);
    set_omega_vec(nullptr, x);
    timed("fill", [&] { launch_fill_f64(mv_H_.p, N_, 0.0, stream_); });
    run_segment(0, nwin_, route, mv_);
    PCAONE_HIP(hipMemcpyAsync(y, mv_H_.p, (size_t)N_ * 8, hipMemcpyDeviceToHost, stream_));
    sync();
  } else {
    // y = X^T (X x)
    apply_X(x, nullptr);
    apply_Xt(nullptr, y);
  }
  ++nops_;
}

// RsvdOpData::initOmg, src/Halko.cpp:14-22: default_random_engine seeded with
// --seed, an N x l NullaryExpr filled in column-major order.  The draw only
// depends on (seed, rand, N, l), so it is made once per operator and every
// pi == 0 restores it from a pristine device copy.
void Engine::draw_omega() {
  std::vector<double> cm((size_t)N_ * l_);
  if (!omega_given_.empty()) {
    cm = omega_given_;
  } else {
    auto rng = std::default_random_engine{};
    rng.seed(p_.seed);
    if (p_.rand) {
      std::normal_distribution<double> dist{0, 1};
      for (auto& v : cm) v = dist(rng);
    } else {
      std::uniform_real_distribution<double> dist{-1, 1};
      for (auto& v : cm) v = dist(rng);
    }
  }
  std::vector<double> rm((size_t)N_ * l_), cs(l_, 0.0);
  for (uint32_t c = 0; c < l_; ++c)
    for (uint32_t r = 0; r < N_; ++r) {
      const double v = cm[(size_t)c * N_ + r];
      rm[(size_t)r * l_ + c] = v;
      cs[c] += v;
    }
  omega0_.reserve((uint64_t)N_ * l_);
  colsum0_.reserve(l_);
  PCAONE_HIP(hipMemcpyAsync(omega0_.p, rm.data(), rm.size() * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(colsum0_.p, cs.data(), cs.size() * 8, hipMemcpyHostToDevice, stream_));
  sync();  // host vectors go out of scope
}

void Engine::init_omega() {
  const size_t nb = (size_t)N_ * l_ * 8;
  PCAONE_HIP(hipMemcpyAsync(omega_.p, omega0_.p, nb, hipMemcpyDeviceToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(colsum_.p, colsum0_.p, (size_t)l_ * 8, hipMemcpyDeviceToDevice, stream_));
  side_.stats = nullptr;  // a pending sign job of an earlier refresh is stale
  launch_fill_f64(fsign_.p, l_, 1.0, stream_);  // the drawn Omega carries its signs
  timed("omega_operand",
        [&] { launch_omega_operand(omega_.p, N_, l_, p_.precision, bg_.p, bcols_, lo_off_, kg_, kperm_(p_.precision),
                                   stream_); });
  refresh_omega32(op_bufs());
}

// Omega <- Householder Q of H (H = H1 + H2 for winSVD), then flipOmg
// (src/Halko.cpp:118-121, 213-216): CholeskyQR2 in FP64 with the Householder
// R-diagonal signs reproduced, so Q equals Eigen's householderQ() columns,
// fused into three passes, the last of which also writes the contraction
// operand and the FP32 gather table.  Omega is stored unsigned; the column
// signs (fsign_) are applied where G is finalised.
void Engine::omega_refresh(bool from_two, double* zero_acc) {
  if (wide_ || robust_) {
    omega_refresh_wide(from_two, zero_acc);
    return;
  }
  flush_side(true, true);  // the previous refresh's signs before this one's flip reads them
  const SegBufs b = op_bufs();
  QrPass q{};
  q.H1 = from_two ? H1_.p : H_.p;
  q.H2 = from_two ? H2_.p : nullptr;
  q.H = H_.p;
  q.n = N_;
  q.l = l_;
  q.part = part_.p;
  q.gram = qred_.p;
  q.status = status_.p;
  q.R1i = R1i_.p;
  q.R2i = R2i_.p;
  q.sign = sign_.p;
  q.fsign = fsign_.p;
  q.stats = qstat_.p;
  q.colsum = colsum_.p;
  q.omega = omega_.p;
  q.zero_acc = zero_acc;
  q.omega32 = (b.corrG && b.g2_f32) ? b.omega32 : nullptr;
  q.bg = bg_.p;
  q.bprec = p_.precision;
  q.lo_off = lo_off_;
  q.kpad = kg_;
  q.kperm = kperm_(p_.precision);
  timed("qr_gram_h", [&] { launch_qr_pass(q, 0, stream_); });
  timed("qr_gram_q", [&] { launch_qr_pass(q, 1, stream_); });
  timed("qr_flip", [&] { launch_qr_pass(q, 2, stream_); });
  // the column signs settle in the next gather's first block (or on their own)
  const SideJobs sj = qr_sign_job(q);
  side_.stats = sj.stats;
  side_.nparts = sj.nparts;
  side_.W = sj.W;
  side_.l = sj.l;
  side_.sign = sj.sign;
  side_.fsign = sj.fsign;
  side_.colsum = sj.colsum;
}

// The same refresh for l > 64 in general form: Gram of H (= H1 + H2, written),
// R1^-1, Q1 = H R1^-1 (stored), Gram of Q1, R2^-1 with the Householder signs of
// the top block of Q1 R2^-1, Omega_u = Q1 R2^-1 with flipOmg's statistics
// against the previous Omega, then the signed, flipped Omega into every
// consumer.
void Engine::omega_refresh_wide(bool from_two, double* zero_acc) {
  flush_side(true, true);
  const SegBufs b = op_bufs();
  const double* H1 = from_two ? H1_.p : H_.p;
  const double* H2 = from_two ? H2_.p : nullptr;
  if (!wq_.p) wq_.reserve((uint64_t)std::max(N_, M_) * l_);  // robust mode at l <= 64
  if (part_.n < wide_part_doubles()) part_.reserve(wide_part_doubles());
  timed("qr_gram_h", [&] { launch_wide_gram(H1, H2, H_.p, N_, l_, part_.p, qred_.p, stream_); });
  const double* Hq = H_.p;  // the matrix orthogonalised (H itself stays for computeUSV)
  if (robust_) {
    hrep_.reserve((uint64_t)N_ * l_);
    rmask_.reserve(l_);
    rscale_.reserve(1);
    timed("qr_repair", [&] { launch_wide_repair(H_.p, qred_.p, N_, l_, rmask_.p, rscale_.p, hrep_.p, stream_); });
    timed("qr_gram_h", [&] { launch_wide_gram(hrep_.p, nullptr, nullptr, N_, l_, part_.p, qred_.p, stream_); });
    Hq = hrep_.p;
  }
  timed("qr_chol", [&] { launch_wide_chol(qred_.p, l_, R1i_.p, status_.p, stream_); });
  WideApply a1{Hq, N_, l_, R1i_.p, wq_.p, nullptr, nullptr};
  timed("qr_apply", [&] { launch_wide_apply(a1, nullptr, stream_); });
  timed("qr_gram_q", [&] { launch_wide_gram(wq_.p, nullptr, nullptr, N_, l_, part_.p, qred_.p, stream_); });
  timed("qr_chol", [&] {
    launch_wide_r2(qred_.p, l_, R1i_.p, R2i_.p, Rti_.p, wq_.p, sign_.p, status_.p + 1, stream_);
  });
  WideApply a2{wq_.p, N_, l_, R2i_.p, Bt_.p, omega_.p, part_.p};
  timed("qr_flip", [&] { launch_wide_apply(a2, qred_.p, stream_); });
  WideWrite w{};
  w.omu = Bt_.p;
  w.n = N_;
  w.l = l_;
  w.stat = qred_.p;
  w.sign = sign_.p;
  w.omega = omega_.p;
  w.colsum = colsum_.p;
  w.zero_acc = zero_acc;
  w.omega32 = (b.corrG && b.g2_f32) ? b.omega32 : nullptr;
  w.bg = bg_.p;
  w.bprec = p_.precision;
  w.lo_off = lo_off_;
  w.kpad = kg_;
  w.kperm = kperm_(p_.precision);
  timed("qr_write", [&] { launch_wide_write(w, stream_); });
  launch_fill_f64(fsign_.p, l_, 1.0, stream_);  // written signed
}

// The pending side jobs selected, taken out of side_.
SideJobs Engine::take_side(bool signs, bool colsum) {
  SideJobs j;
  if (signs && side_.stats) {
    j.stats = side_.stats;
    j.nparts = side_.nparts;
    j.W = side_.W;
    j.l = side_.l;
    j.sign = side_.sign;
    j.fsign = side_.fsign;
    j.colsum = side_.colsum;
    side_.stats = nullptr;
  }
  if (colsum && side_.s2p) {
    j.s2p = side_.s2p;
    j.nchunk = side_.nchunk;
    j.w0 = side_.w0;
    j.w1 = side_.w1;
    j.l2 = side_.l2;
    j.S2 = side_.S2;
    side_.s2p = nullptr;
  }
  return j;
}

void Engine::flush_side(bool signs, bool colsum) {
  if (signs && side_.stats) {
    const SideJobs j = take_side(true, false);
    timed("qr_signs", [&] { launch_side_jobs(j, stream_); });
  }
  if (colsum && side_.s2p) {
    const SideJobs j = take_side(false, true);
    timed("window_colsum", [&] { launch_side_jobs(j, stream_); });
  }
}

// the FP32 table the missing-entry gathers of the G contraction read
void Engine::refresh_omega32(const SegBufs& b) {
  if (!b.corrG || !b.g2_f32) return;
  timed("omega_operand", [&] { launch_cast_f32(b.omega, b.omega32, (uint64_t)N_ * b.l, stream_); });
}

// G rows of windows [w0, w1) and their H contributions, routed per window.
SegBufs Engine::op_bufs() {
  SegBufs b;
  b.prec = p_.precision;
  b.l = l_;
  b.nt16 = nt16_;
  b.ntiles = ntiles_;
  b.bcols = bcols_;
  b.lo_off = lo_off_;
  b.ldp = ldp_;
  b.opsz = opsz_;
  b.part_f64 = part_f64_;
  b.split_cols = split_cols_;
  b.fold = fold_;
  b.hf = hf_;
  b.omega = omega_.p;
  b.colsum = colsum_.p;
  b.fsign = fsign_.p;
  b.G = G_.p;
  b.g2 = g2_.p;
  b.g2_f32 = p_.precision != PREC_FP64;
  b.omega32 = omega32_.p;
  b.corrG = nnz_snp_ > 0 ? corrG_.p : nullptr;
  b.corrH = nnz_snp_ > 0 ? corrH_.p : nullptr;
  b.S2 = S2_.p;
  b.s2p = s2p_.p;
  b.affine = src_ == SRC_BED && p_.precision == PREC_SPLIT_BF16;
  b.bg = bg_.p;
  b.bh = bh_.p;
  b.gpart = gpart_.p;
  b.hpart = hpart_.p;
  b.Ha = (p_.variant == 2) ? H1_.p : H_.p;
  b.Hb = (p_.variant == 2) ? H2_.p : H_.p;
  b.mval = have_vals_ ? miss_val_.p : nullptr;
  return b;
}

// G rows of windows [w0, w1): G = X_seg^T Omega, its epilogue (centring,
// missing-entry corrections, H operand) and the per-window sums S''.
void Engine::seg_G(uint32_t w0, uint32_t w1, const SegBufs& b) {
  if (!streamed_) {
    seg_G_on(w0, w1, b, src_ == SRC_BED ? bed_snp_.p + (uint64_t)w0 * bs_ * row_bytes_ : nullptr);
    return;
  }
  for (uint32_t w = w0; w < w1; ++w) {
    const int s = stream_acquire(w);
    seg_G_on(w, w + 1, b, slots_[s].snp.p);
    stream_release(s);
    stream_prefetch((w + 1) % nwin_);
  }
}

void Engine::seg_G_on(uint32_t w0, uint32_t w1, const SegBufs& b, const uint8_t* A) {
  const int prec = b.prec;
  const bool dense = src_ == SRC_DENSE;
  const uint32_t r0 = w0 * bs_, r1 = std::min<uint32_t>(w1 * bs_, M_);
  const uint32_t rows = r1 - r0;
  const uint32_t rpb = rows_per_block_(prec);
  uint32_t g_nsplit = 0, g_fold = 0;
  if (b.corrG) {
    CsrGather cg{};
    cg.ptr = miss_snp_ptr_.p;
    cg.idx = miss_snp_idx_.p;
    cg.table = b.g2_f32 ? static_cast<const void*>(b.omega32) : static_cast<const void*>(b.omega);
    cg.table_f32 = b.g2_f32;
    cg.l = b.l;
    cg.in_base = r0;
    cg.stride_in = 0;
    cg.out_base = r0;
    cg.per = rows;
    cg.ngroups = 1;
    cg.out = b.corrG;
    cg.side = take_side(true, false);  // the last refresh's column signs
    // In line before the contraction: run beside it on a second stream, the
    // gather's random table reads evict the contraction's operand tiles from
    // L2 and cost it more time than the gather takes alone.
    timed("miss_gather_G", [&] { launch_csr_gather(cg, stream_); });
  } else {
    flush_side(true, false);
  }
  if (!dense) {
    PackedGemm g{};
    g.A = A;
    g.lda = row_bytes_;
    g.nrows = rows;
    g.kchunks = row_bytes_ / 64;
    g.nsplit = packed_split(prec, rows, 1, g.kchunks, g.chunks_per_split);
    g.B = b.bg;
    g.ldb = kg_;
    g.out = b.gpart;
    g.nz = 1;
    g.fold = packed_bf16_streams(rows, 1) ? b.fold : 0;
    g.ldo = g.fold ? ceil_div(g.fold, 16) * 16 : b.ldp;
    g_nsplit = g.nsplit;
    g_fold = g.fold;
    timed("gemm_G", [&] {
      if (prec == PREC_SPLIT_BF16)
        launch_packed_gemm_bf16(g, b.nt16, stream_);
      else if (prec == PREC_FP32)
        launch_packed_gemm_f32(g, (int)b.l, stream_);
      else
        launch_packed_gemm_f64(g, (int)b.l, stream_);
    });
  } else {
    DenseGemm g{};
    g.A = dns_.p + (uint64_t)r0 * kg_;
    g.lda = kg_;
    g.nrows = rows;
    g.kchunks = (uint32_t)(kg_ / 64);
    g.nsplit = split_for(ceil_div(rows, rpb), g.kchunks, g.chunks_per_split);
    g.B = b.bg;
    g.ldb = kg_;
    g.out = b.gpart;
    g.ldo = b.ldp;
    g.nz = 1;
    g_nsplit = g.nsplit;
    timed("gemm_G", [&] { launch_dense_gemm(g, prec, (int)b.l, b.ntiles, stream_); });
  }
  FinalizeG fg = finalize_desc(w0, w1, b);
  fg.part = b.gpart;
  fg.nsplit = g_nsplit;
  if (g_fold) {  // hi + lo already summed per column
    fg.ldp = ceil_div(g_fold, 16) * 16;
    fg.split_cols = 0;
  }
  timed("finalize_G", [&] { launch_finalize_G(fg, stream_); });
  // the window sums of g'' are first read by finalize_H: they ride on the H
  // product's gather when there is one (resident windows)
  flush_side(false, true);
  const SideJobs cj = window_colsum_job(b.s2p, b.l, bs_, w0, w1, b.S2);
  side_.s2p = cj.s2p;
  side_.nchunk = cj.nchunk;
  side_.w0 = cj.w0;
  side_.w1 = cj.w1;
  side_.l2 = cj.l2;
  side_.S2 = cj.S2;
  // the operator's own windows in HBM only (the mat-vec's buffers and the
  // streamed windows sum at once)
  if (streamed_ || b.S2 != S2_.p) flush_side(false, true);
}

FinalizeG Engine::finalize_desc(uint32_t w0, uint32_t w1, const SegBufs& b) const {
  FinalizeG fg{};
  fg.part_is_f64 = b.part_f64;
  fg.ldp = b.ldp;
  fg.split_cols = b.split_cols;
  fg.r0 = w0 * bs_;
  fg.r1 = std::min<uint32_t>(w1 * bs_, M_);
  fg.l = b.l;
  fg.F = F_.p;
  fg.d = d_.p;
  fg.colsum = b.colsum;
  fg.fsign = b.fsign;
  fg.miss_ptr = miss_snp_ptr_.p;
  fg.miss_idx = miss_snp_idx_.p;
  fg.omega = b.omega;
  fg.corr = b.corrG;
  fg.G = b.G;
  fg.g2 = b.g2;
  fg.g2_f32 = b.g2_f32;
  fg.bprec = b.prec;
  fg.bh = b.bh;
  fg.bcols = b.bcols;
  fg.lo_off = b.lo_off;
  fg.kw = kw_;
  fg.blocksize = bs_;
  fg.mval = b.mval;
  fg.hf = b.hf;
  fg.w0 = w0;
  fg.w1 = w1;
  fg.nsnps = M_;
  fg.s2p = b.s2p;
  fg.affine = b.affine ? 1 : 0;
  fg.kperm = kperm_(b.prec);
  return fg;
}

// H contributions of windows [w0, w1) (X_w G_w), routed per window to Ha / Hb.
void Engine::seg_H(uint32_t w0, uint32_t w1, const std::vector<int>& route, const SegBufs& b) {
  if (!streamed_) {
    seg_H_on(w0, w1, route, b, src_ == SRC_BED ? bed_smp_.p + (uint64_t)w0 * npad_ * win_bytes_ : nullptr);
    return;
  }
  for (uint32_t w = w0; w < w1; ++w) {
    const int s = stream_acquire(w);
    seg_H_on(w, w + 1, route, b, slots_[s].smp.p);
    stream_release(s);
    stream_prefetch((w + 1) % nwin_);
  }
}

void Engine::seg_H_on(uint32_t w0, uint32_t w1, const std::vector<int>& route, const SegBufs& b,
                      const uint8_t* A) {
  const int prec = b.prec;
  const bool dense = src_ == SRC_DENSE;
  const uint32_t rpb = rows_per_block_(prec);
  const uint32_t nz = w1 - w0;
  uint32_t h_nsplit = 0, h_fold = 0;
  if (b.corrH) {
    CsrGather cg{};
    cg.ptr = miss_win_ptr_.p;
    cg.idx = miss_win_idx_.p;
    cg.table = b.g2;
    cg.table_f32 = b.g2_f32;
    cg.l = b.l;
    cg.in_base = (uint64_t)w0 * miss_rows_;
    cg.stride_in = miss_rows_;
    cg.out_base = (uint64_t)w0 * N_;
    cg.per = N_;
    cg.ngroups = nz;
    cg.out = b.corrH;
    cg.side = take_side(false, true);  // finalize_G's window sums
    timed("miss_gather_H", [&] { launch_csr_gather(cg, stream_); });
  }
  if (!dense) {
    PackedGemm h{};
    h.A = A;
    h.lda = win_bytes_;
    h.a_zstride = (uint64_t)npad_ * win_bytes_;
    h.nrows = N_;
    h.kchunks = win_bytes_ / 64;
    h.nsplit = packed_split(prec, N_, nz, h.kchunks, h.chunks_per_split);
    h.B = b.bh + (uint64_t)w0 * b.bcols * kw_ * b.opsz;
    h.ldb = kw_;
    h.b_zstride = (uint64_t)b.bcols * kw_;
    h.out = b.hpart;
    h.nz = nz;
    h.fold = packed_bf16_streams(N_, nz) ? b.fold : 0;
    h.ldo = h.fold ? ceil_div(h.fold, 16) * 16 : b.ldp;
    h.o_zstride = (uint64_t)h.nsplit * N_ * h.ldo;
    h_nsplit = h.nsplit;
    h_fold = h.fold;
    timed("gemm_H", [&] {
      if (prec == PREC_SPLIT_BF16)
        launch_packed_gemm_bf16(h, b.nt16, stream_);
      else if (prec == PREC_FP32)
        launch_packed_gemm_f32(h, (int)b.l, stream_);
      else
        launch_packed_gemm_f64(h, (int)b.l, stream_);
    });
  } else {
    DenseGemm h{};
    h.A = dnt_.p + (uint64_t)w0 * N_ * kw_;
    h.lda = kw_;
    h.a_zstride = (uint64_t)N_ * kw_;
    h.nrows = N_;
    h.kchunks = (uint32_t)(kw_ / 64);
    h.nsplit = split_for(ceil_div(N_, rpb) * nz, h.kchunks, h.chunks_per_split);
    h.B = b.bh + (uint64_t)w0 * b.bcols * kw_ * b.opsz;
    h.ldb = kw_;
    h.b_zstride = (uint64_t)b.bcols * kw_;
    h.out = b.hpart;
    h.ldo = b.ldp;
    h.o_zstride = (uint64_t)h.nsplit * N_ * b.ldp;
    h.nz = nz;
    h_nsplit = h.nsplit;
    timed("gemm_H", [&] { launch_dense_gemm(h, prec, (int)b.l, b.ntiles, stream_); });
  }
  flush_side(false, true);  // S2 when no gather carried it
  FinalizeH fh{};
  fh.part = b.hpart;
  fh.part_is_f64 = b.part_f64;
  fh.nsplit = h_nsplit;
  fh.ldp = h_fold ? ceil_div(h_fold, 16) * 16 : b.ldp;  // folded: hi + lo already summed per column
  fh.part_rows = N_;
  fh.miss_rows = miss_rows_;
  fh.split_cols = h_fold ? 0 : b.split_cols;
  fh.nsamples = N_;
  fh.l = b.l;
  fh.wpart0 = w0;
  fh.Ha = b.Ha;
  fh.Hb = b.Hb;
  fh.S2 = b.S2;
  fh.affine = b.affine ? 1 : 0;
  fh.miss_ptr = miss_win_ptr_.p;
  fh.miss_idx = miss_win_idx_.p;
  fh.corr = b.corrH;
  fh.mval = b.mval;
  fh.mpos = miss_win_pos_.p;
  fh.G = b.G;
  fh.d = d_.p;
  fh.hf = b.hf;
  for (uint32_t a = w0; a < w1; a += 64) {  // route travels as a 64-bit mask
    fh.w0 = a;
    fh.w1 = std::min<uint32_t>(a + 64, w1);
    fh.route_b = 0;
    for (uint32_t w = fh.w0; w < fh.w1; ++w)
      if (route[w] == 2) fh.route_b |= 1ull << (w - a);
    timed("finalize_H", [&] { launch_finalize_H(fh, stream_); });
  }
}

void Engine::run_segment(uint32_t w0, uint32_t w1, const std::vector<int>& route) {
  run_segment(w0, w1, route, op_bufs());
}

// G rows and H contributions of windows [w0, w1).  Streamed, each window
// travels once: its G rows, then its H contribution, then the next window
// (the reference's out-of-core loop, src/Halko.cpp:138-149, 232-268).
void Engine::run_segment(uint32_t w0, uint32_t w1, const std::vector<int>& route, const SegBufs& b) {
  if (!streamed_) {
    seg_G(w0, w1, b);
    if (spec_gate_) {
      const std::function<bool()> gate = std::move(spec_gate_);
      spec_gate_ = nullptr;
      if (!gate()) {
        spec_abort_ = true;
        return;
      }
    }
    seg_H(w0, w1, route, b);
    return;
  }
  for (uint32_t w = w0; w < w1; ++w) {
    const int s = stream_acquire(w);
    seg_G_on(w, w + 1, b, slots_[s].snp.p);
    seg_H_on(w, w + 1, route, b, slots_[s].smp.p);
    stream_release(s);
    stream_prefetch((w + 1) % nwin_);
  }
}

void Engine::rsvd_gandh(int pi, double* Gout, double* Hout, bool check) {
  PCAONE_REQUIRE(have_op_, -3, "computeGandH: operator not initialised");
  PCAONE_REQUIRE(pi >= 0, -1, "computeGandH: pi must be >= 0");
  PCAONE_HIP(hipSetDevice(device_));
  if (pi == 0 && update_) {
    // fit_with_pi, EMU branch (src/Data.cpp:272-287; src/FilePlink.cpp:253-261)
    PCAONE_HIP(hipMemcpyAsync(Sdev_.p, last_S_.data(), last_S_.size() * 8, hipMemcpyHostToDevice, stream_));
    timed("emu_update", [&] {
      launch_emu_update(miss_snp_ptr_.p, miss_snp_idx_.p, M_, Ucur_.p, Sdev_.p, V_.p, (uint32_t)p_.k, F_.p,
                        miss_val_.p, stream_);
    });
    have_vals_ = true;
  }
  bool refreshed = false;
  if (p_.variant == 1) {
    // NormalRsvdOpData in-core, src/Halko.cpp:105-126
    if (pi == 0) {
      init_omega();
    } else {
      omega_refresh(false, nullptr);
      refreshed = true;
    }
    timed("fill", [&] { launch_fill_f64(H_.p, (uint64_t)N_ * l_, 0.0, stream_); });
    run_segment(0, nwin_, std::vector<int>(nwin_, 1));
    if (spec_abort_) return;
  } else {
    // FancyRsvdOpData in-core, src/Halko.cpp:153-226
    if (pi == 0) init_omega();
    const FancyPlan pl = fancy_plan(p_.bands, pi, bandsize_, nwin_, (uint32_t)p_.band_factor);
    if (pl.reset_both) {
      timed("fill", [&] { launch_fill_f64(H1_.p, (uint64_t)N_ * l_, 0.0, stream_); });
      timed("fill", [&] { launch_fill_f64(H2_.p, (uint64_t)N_ * l_, 0.0, stream_); });
    }
    std::vector<int> route(nwin_, 1);
    for (uint32_t b = 0; b < nwin_; ++b) route[b] = pl.steps[b].acc;
    uint32_t start = 0;
    for (uint32_t b = 0; b < nwin_; ++b) {
      const WindowStep& st = pl.steps[b];
      if (!st.update && b + 1 < nwin_) continue;
      run_segment(start, b + 1, route);
      if (spec_abort_) return;
      start = b + 1;
      if (st.update) {
        omega_refresh(true, st.zero == 1 ? H1_.p : H2_.p);
        refreshed = true;
      }
    }
  }
  if (check && refreshed) check_status();
  if (Gout || Hout) {
    sync();
    std::vector<double> tmp;
    if (Gout) {
      tmp.resize((size_t)M_ * l_);
      PCAONE_HIP(hipMemcpy(tmp.data(), G_.p, tmp.size() * 8, hipMemcpyDeviceToHost));
      for (uint32_t r = 0; r < M_; ++r)
        for (uint32_t c = 0; c < l_; ++c) Gout[(size_t)c * M_ + r] = tmp[(size_t)r * l_ + c];
    }
    if (Hout) {
      tmp.resize((size_t)N_ * l_);
      PCAONE_HIP(hipMemcpy(tmp.data(), H_.p, tmp.size() * 8, hipMemcpyDeviceToHost));
      for (uint32_t r = 0; r < N_; ++r)
        for (uint32_t c = 0; c < l_; ++c) Hout[(size_t)c * N_ + r] = tmp[(size_t)r * l_ + c];
    }
  }
}

void Engine::check_status(const char* context) {
  int st[2] = {0, 0};
  PCAONE_HIP(hipMemcpyAsync(st, status_.p, sizeof(st), hipMemcpyDeviceToHost, stream_));
  sync();
  if (st[0] || st[1]) {
    PCAONE_HIP(hipMemsetAsync(status_.p, 0, 2 * sizeof(int), stream_));
    throw Error(-4, "CholeskyQR breakdown (pivot " + std::to_string(st[0] ? st[0] : st[1]) +
                        "): the sketch is numerically rank deficient" + (context ? context : ""));
  }
}

// One pass of computeUSV's linear algebra (src/Halko.cpp:53-68): the
// CholeskyQR2 factors of G (R = R2 R1; Q = G R^-1 is never stored: only the
// last pass needs it, for V), B^T = H R^-1, B B^T = U_B S^2 U_B^T,
// Ucur = B^T U_B S^-1.
void Engine::usv_step() {
  if (robust_) {
    usv_step_robust();
    return;
  }
  if (wide_) {
    usv_step_wide();
    return;
  }
  usv_step_on(stream_, G_.p, H_.p, part_.p, R1i_.p, R2i_.p, status_.p);
  Gfin_ = G_.p;
}

void Engine::usv_step_on(hipStream_t s, const double* G, const double* H, double* part, double* R1i, double* R2i,
                         int* status) {
  const uint32_t l = l_;
  UsvPass u{};
  u.G = G;
  u.m = M_;
  u.H = H;
  u.n = N_;
  u.l = l;
  u.part = part;
  u.gram = gred_.p;
  u.bbt = bbt_.p;
  u.R1i = R1i;
  u.R2i = R2i;
  u.Rti = Rti_.p;
  u.Bt = Bt_.p;
  u.status = status;
  timed_on(s, "usv_gram", [&] { launch_usv_qr(u, 0, s); });
  timed_on(s, "usv_q1", [&] { launch_usv_qr(u, 1, s); });
  timed_on(s, "usv_bt", [&] { launch_usv_qr(u, 2, s); });
  // later epochs start from the previous epoch's eigenvectors
  timed_on(s, "usv_eig", [&] {
    launch_sym_eig(bbt_.p, 1, l, evals_.p, W_.p, sinv_.p, (uint32_t)p_.k, have_W_ ? W_.p : nullptr, s);
  });
  have_W_ = true;
  timed_on(s, "usv_rows", [&] {
    launch_rows_times_small(Bt_.p, N_, l, W_.p, l, (uint32_t)p_.k, sinv_.p, Ucur_.p, s);
  });
}

// usv_step for l > 64: CholeskyQR2 factors of G (Q1 = G R1^-1 stored),
// B^T = H Rt^-1, the Gram B B^T and its eigensolve, Ucur = B^T U_B S^-1.
void Engine::usv_step_wide() {
  const uint32_t l = l_;
  Gfin_ = G_.p;
  timed("usv_gram", [&] { launch_wide_gram(G_.p, nullptr, nullptr, M_, l, part_.p, gred_.p, stream_); });
  timed("usv_chol", [&] { launch_wide_chol(gred_.p, l, R1i_.p, status_.p, stream_); });
  WideApply a1{G_.p, M_, l, R1i_.p, wq_.p, nullptr, nullptr};
  timed("usv_q1", [&] { launch_wide_apply(a1, nullptr, stream_); });
  timed("usv_gram", [&] { launch_wide_gram(wq_.p, nullptr, nullptr, M_, l, part_.p, gred_.p, stream_); });
  timed("usv_chol", [&] {
    launch_wide_r2(gred_.p, l, R1i_.p, R2i_.p, Rti_.p, nullptr, nullptr, status_.p + 1, stream_);
  });
  WideApply a2{H_.p, N_, l, Rti_.p, Bt_.p, nullptr, nullptr};
  timed("usv_bt", [&] { launch_wide_apply(a2, nullptr, stream_); });
  timed("usv_gram", [&] { launch_wide_gram(Bt_.p, nullptr, nullptr, N_, l, part_.p, bbt_.p, stream_); });
  timed("usv_eig", [&] {
    launch_sym_eig_wide(bbt_.p, l, evals_.p, W_.p, sinv_.p, (uint32_t)p_.k, have_W_ ? W_.p : nullptr, eigtmp_.p,
                        stream_);
  });
  have_W_ = true;
  timed("usv_rows", [&] {
    launch_rows_times_small(Bt_.p, N_, l, W_.p, l, (uint32_t)p_.k, sinv_.p, Ucur_.p, stream_);
  });
}

// The epochs of computeUSV (src/Halko.cpp:51-93); `done` is false when the
// loop ran out of epochs (the last Ucur then sits in Upre).
int Engine::usv_epochs(int maxp, double tol, double* diffs, bool& done) {
  if (robust_ || wide_) return usv_epochs_serial(maxp, tol, diffs, done);
  // Pipelined: computeUSV of epoch pi (its QR, eigensolve and mev) runs on
  // ustream_ from a copy of H and the G buffer just written, while epoch
  // pi + 1's computeGandH, which needs only Omega and the accumulators, runs on
  // stream_.  That next epoch is speculative: when epoch pi ends the solve, the
  // accumulators (which carry over between solves) and the window schedule
  // state are put back as they were, so every result equals the serial order's.
  const uint32_t k = (uint32_t)p_.k;
  const size_t hb = (size_t)N_ * l_ * 8;
  Gu_.reserve((uint64_t)M_ * l_);
  Hu_.reserve((uint64_t)N_ * l_);
  sH1_.reserve((uint64_t)N_ * l_);
  sH2_.reserve((uint64_t)N_ * l_);
  upart_.reserve(part_.n);
  uR1i_.reserve((uint64_t)l_ * l_);
  uR2i_.reserve((uint64_t)l_ * l_);
  ustat_.reserve(4);
  PCAONE_HIP(hipMemsetAsync(ustat_.p, 0, 4 * sizeof(int), stream_));
  int p = maxp, ep = 0;
  have_W_ = false;
  if (diffs)
    for (int i = 0; i <= maxp; ++i) diffs[i] = 0.0;
  done = false;
  const bool winsvd = p_.variant == 2;
  const int pwin = (int)std::log2((double)p_.bands);
  double* hm = static_cast<double*>(hpipe_);
  int* hst = reinterpret_cast<int*>(hm + 1);
  rsvd_gandh(0, nullptr, nullptr, false);
  ep = 1;
  for (int pi = 0; pi <= p; ++pi) {
    // epoch pi's G and H for its computeUSV
    std::swap(G_, Gu_);
    PCAONE_HIP(hipMemcpyAsync(Hu_.p, H_.p, hb, hipMemcpyDeviceToDevice, stream_));
    PCAONE_HIP(hipMemcpyAsync(ustat_.p, status_.p, 2 * sizeof(int), hipMemcpyDeviceToDevice, stream_));
    PCAONE_HIP(hipEventRecord(ev_epoch_, stream_));
    PCAONE_HIP(hipStreamWaitEvent(ustream_, ev_epoch_, 0));
    usv_step_on(ustream_, Gu_.p, Hu_.p, upart_.p, uR1i_.p, uR2i_.p, ustat_.p + 2);
    if (pi > 0) {
      const int nbN = gram_blocks(N_);
      timed_on(ustream_, "mev", [&] { launch_xty_partial(Ucur_.p, Upre_.p, N_, k, upart_.p, nbN, ustream_); });
      timed_on(ustream_, "mev", [&] { launch_mev_finish(upart_.p, nbN, k, scal_.p, ustream_); });
      PCAONE_HIP(hipMemcpyAsync(hm, scal_.p, 8, hipMemcpyDeviceToHost, ustream_));
    }
    PCAONE_HIP(hipMemcpyAsync(hst, ustat_.p, 4 * sizeof(int), hipMemcpyDeviceToHost, ustream_));
    // the next epoch, if the schedule may still run one
    const int pcap = (winsvd && pi > 0) ? std::max(p, pwin) : p;
    const bool spec = pi + 1 <= pcap;
    const uint64_t bandsize_before = bandsize_;
    bool decided = false, stop = false, broke = false;
    // epoch pi's verdict from its computeUSV (waits for it): true to go on
    auto decide = [&]() -> bool {
      if (decided) return !stop;
      decided = true;
      PCAONE_HIP(hipStreamSynchronize(ustream_));
      if (hst[0] || hst[1] || hst[2] || hst[3]) {
        broke = stop = true;
        return false;
      }
      if (pi > 0) {
        const double diff = 1.0 - hm[0];
        if (diffs && pi <= maxp) diffs[pi] = diff;
        if (diff < tol || pi == p) {
          if (winsvd && std::ldexp(1.0, pi) < (double)p_.bands) {
            p = pwin;  // winSVD runs log2(bands)+1 epochs at least
          } else {
            done = true;
            stop = true;
          }
        } else {
          std::swap(Ucur_.p, Upre_.p);
        }
      } else {
        std::swap(Ucur_.p, Upre_.p);
      }
      if (!stop && pi + 1 > p) stop = true;  // out of epochs
      return !stop;
    };
    if (spec) {
      PCAONE_HIP(hipMemcpyAsync(sH1_.p, H1_.p, hb, hipMemcpyDeviceToDevice, stream_));
      PCAONE_HIP(hipMemcpyAsync(sH2_.p, H2_.p, hb, hipMemcpyDeviceToDevice, stream_));
      // where epoch pi may end the solve, the next epoch stops after its first
      // G product unless epoch pi's verdict is to go on (the winSVD's first
      // log2(bands) epochs always go on: no wait there)
      const bool may_stop = pi > 0 && !(winsvd && std::ldexp(1.0, pi) < (double)p_.bands && pwin > pi);
      spec_abort_ = false;
      if (may_stop) spec_gate_ = decide;
      try {
        rsvd_gandh(pi + 1, nullptr, nullptr, false);
      } catch (...) {
        spec_gate_ = nullptr;
        spec_abort_ = false;
        throw;
      }
      spec_gate_ = nullptr;
      spec_abort_ = false;
    }
    decide();
    auto undo_spec = [&]() {
      if (!spec) return;
      PCAONE_HIP(hipMemcpyAsync(H1_.p, sH1_.p, hb, hipMemcpyDeviceToDevice, stream_));
      PCAONE_HIP(hipMemcpyAsync(H2_.p, sH2_.p, hb, hipMemcpyDeviceToDevice, stream_));
      PCAONE_HIP(hipMemsetAsync(status_.p, 0, 2 * sizeof(int), stream_));
      bandsize_ = bandsize_before;
      sync();
    };
    if (broke) {
      const int piv = hst[0] ? hst[0] : hst[1] ? hst[1] : hst[2] ? hst[2] : hst[3];
      undo_spec();
      PCAONE_HIP(hipMemsetAsync(status_.p, 0, 2 * sizeof(int), stream_));
      PCAONE_HIP(hipMemsetAsync(ustat_.p, 0, 4 * sizeof(int), stream_));
      sync();
      throw Error(-4, "CholeskyQR breakdown (pivot " + std::to_string(piv) +
                          "): the sketch is numerically rank deficient");
    }
    if (stop) {
      undo_spec();
      break;
    }
    ++ep;  // epoch pi + 1, running
  }
  // G_ holds the last computeUSV step's G again (any speculative work is done)
  if (Gu_.p != G_.p) std::swap(G_, Gu_);
  Gfin_ = G_.p;
  return ep;
}

int Engine::usv_epochs_serial(int maxp, double tol, double* diffs, bool& done) {
  const uint32_t k = (uint32_t)p_.k;
  int p = maxp, ep = 0;
  have_W_ = false;
  if (diffs)
    for (int i = 0; i <= maxp; ++i) diffs[i] = 0.0;
  done = false;
  for (int pi = 0; pi <= p; ++pi) {
    rsvd_gandh(pi, nullptr, nullptr, false);
    ++ep;
    usv_step();
    check_status(robust_ ? " even in the rank-robust mode" : nullptr);
    if (pi > 0) {
      const int nbN = gram_blocks(N_);
      timed("mev", [&] { launch_xty_partial(Ucur_.p, Upre_.p, N_, k, part_.p, nbN, stream_); });
      timed("mev", [&] { launch_mev_finish(part_.p, nbN, k, scal_.p, stream_); });
      double m = 0.0;
      PCAONE_HIP(hipMemcpyAsync(&m, scal_.p, 8, hipMemcpyDeviceToHost, stream_));
      sync();
      const double diff = 1.0 - m;
      if (diffs && pi <= maxp) diffs[pi] = diff;
      if (diff < tol || pi == p) {
        if (p_.variant == 2 && std::ldexp(1.0, pi) < (double)p_.bands) {
          p = (int)std::log2((double)p_.bands);  // winSVD runs log2(bands)+1 epochs at least
        } else {
          done = true;
          break;
        }
      } else {
        std::swap(Ucur_.p, Upre_.p);
      }
    } else {
      std::swap(Ucur_.p, Upre_.p);
    }
  }
  return ep;
}

// usv_step in robust mode, for a G of numerical rank below l (the data
// matrix itself of rank below l, where Householder QR -- the reference's --
// fills Q's trailing columns from rounding and CholeskyQR breaks down): the
// eigenpairs of G^T G give T = W_r Lambda_r^-1/2 over its numerical range
// (eigenvalues above 1e-12 of the largest), so Q = G T has orthonormal
// columns and B^T = X^T Q = H T; the rest is usv_step's: B B^T, its
// eigenvectors, Ucur = B^T U_B S^-1, and at the end V = G (T U_B).
void Engine::usv_step_robust() {
  const uint32_t l = l_;
  Gfin_ = G_.p;
  if (part_.n < wide_part_doubles()) part_.reserve(wide_part_doubles());
  if (!eigtmp_.p) eigtmp_.reserve((uint64_t)l * l);
  Wg_.reserve((uint64_t)l * l);
  timed("usv_gram", [&] { launch_wide_gram(G_.p, nullptr, nullptr, M_, l, part_.p, gred_.p, stream_); });
  timed("usv_eig", [&] {
    launch_sym_eig_wide(gred_.p, l, evals_.p, Wg_.p, sinv_.p, 0, nullptr, eigtmp_.p, stream_);
  });
  timed("usv_rows", [&] { launch_whiten(Wg_.p, evals_.p, l, 1e-12, Rti_.p, stream_); });
  WideApply a2{H_.p, N_, l, Rti_.p, Bt_.p, nullptr, nullptr};
  timed("usv_bt", [&] { launch_wide_apply(a2, nullptr, stream_); });
  timed("usv_gram", [&] { launch_wide_gram(Bt_.p, nullptr, nullptr, N_, l, part_.p, bbt_.p, stream_); });
  timed("usv_eig", [&] {
    launch_sym_eig_wide(bbt_.p, l, evals_.p, W_.p, sinv_.p, (uint32_t)p_.k, have_W_ ? W_.p : nullptr, eigtmp_.p,
                        stream_);
  });
  have_W_ = true;
  timed("usv_rows", [&] {
    launch_rows_times_small(Bt_.p, N_, l, W_.p, l, (uint32_t)p_.k, sinv_.p, Ucur_.p, stream_);
  });
}

void Engine::rsvd_usv(int maxp, double tol, double* U, double* S, double* V, int* epochs, double* diffs) {
  PCAONE_REQUIRE(have_op_, -3, "computeUSV: operator not initialised");
  PCAONE_REQUIRE(maxp >= 0, -1, "computeUSV: p must be >= 0");
  PCAONE_HIP(hipSetDevice(device_));
  const uint32_t k = (uint32_t)p_.k;
  // the winSVD accumulators carry over between solves (the reference resets
  // them only when 2^pi >= bands): keep them for a restart
  const size_t hb = (size_t)N_ * l_ * 8;
  PCAONE_HIP(hipMemcpyAsync(snapH1_.p, H1_.p, hb, hipMemcpyDeviceToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(snapH2_.p, H2_.p, hb, hipMemcpyDeviceToDevice, stream_));
  bool done = false;
  int ep = 0;
  try {
    ep = usv_epochs(maxp, tol, diffs, done);
  } catch (const Error& e) {
    if (e.code != -4 || robust_) throw;
    // A partial H or a G of lower rank than l (a window of fewer SNPs than l,
    // or a data matrix of rank below l): Householder QR, as the reference
    // uses, carries on with trailing columns set by rounding; CholeskyQR
    // breaks down.  The solve restarts in the rank-robust mode (rank repair
    // in every refresh, G through its Gram eigenpairs) from the same state
    // (Omega is redrawn at pi == 0; the EMU values of this solve are kept).
    robust_ = true;
    PCAONE_HIP(hipMemcpyAsync(H1_.p, snapH1_.p, hb, hipMemcpyDeviceToDevice, stream_));
    PCAONE_HIP(hipMemcpyAsync(H2_.p, snapH2_.p, hb, hipMemcpyDeviceToDevice, stream_));
    const bool upd = update_;
    update_ = false;
    try {
      ep = usv_epochs(maxp, tol, diffs, done);
    } catch (...) {
      update_ = upd;
      throw;
    }
    update_ = upd;
  }
  if (!done) std::swap(Ucur_.p, Upre_.p);  // the last Ucur sits in Upre after the swap
  // V = Q U_B[:, :k] = G (R^-1 U_B[:, :k]) (rows in laid-out order), S = sqrt(eig(B B^T))[:k]
  timed("usv_rows", [&] { launch_rows_times_small(Rti_.p, l_, l_, W_.p, l_, k, nullptr, Rt_.p, stream_); });
  timed("usv_rows", [&] { launch_rows_times_small(Gfin_, M_, l_, Rt_.p, k, k, nullptr, V_.p, stream_); });
  std::vector<double> ev(l_);
  PCAONE_HIP(hipMemcpyAsync(ev.data(), evals_.p, l_ * 8, hipMemcpyDeviceToHost, stream_));
  sync();
  last_S_.assign(k, 0.0);
  for (uint32_t c = 0; c < k; ++c) last_S_[c] = std::sqrt(std::max(ev[c], 0.0));
  have_usv_ = true;
  copy_usv(U, S, V);
  if (epochs) *epochs = ep;
}

// U (N x k) and V (M x k, original feature order) column-major, S (k) of the
// last computeUSV
void Engine::copy_usv(double* U, double* S, double* V) {
  const uint32_t k = (uint32_t)p_.k;
  // the transposes and the un-shuffle of V's rows run on the device, so each
  // output is one contiguous copy into the caller's buffer
  if (U || V) out_.reserve((uint64_t)std::max(N_, M_) * k);
  if (U) {
    timed("usv_out", [&] { launch_rows_to_cols(Ucur_.p, N_, k, nullptr, out_.p, stream_); });
    PCAONE_HIP(hipMemcpyAsync(U, out_.p, (size_t)N_ * k * 8, hipMemcpyDeviceToHost, stream_));
  }
  if (V) {
    timed("usv_out", [&] { launch_rows_to_cols(V_.p, M_, k, d_perm_.p, out_.p, stream_); });
    PCAONE_HIP(hipMemcpyAsync(V, out_.p, (size_t)M_ * k * 8, hipMemcpyDeviceToHost, stream_));
  }
  sync();
  if (S) std::memcpy(S, last_S_.data(), k * sizeof(double));
}

// The EM branch of run_pca_with_halko for --emu (src/Halko.cpp:294-336):
// one plain solve, then solves with the missing genotypes re-imputed from the
// previous U, S, V until 1 - mev(V, Vpre) < tolem, then a final standardised
// solve.  `diffs` (optional, maxiter entries) receives 1 - mev per iteration.
void Engine::rsvd_em(int maxp, double tol, int maxiter, double tolem, double* U, double* S, double* V, int* iters,
                     double* diffs) {
  PCAONE_REQUIRE(have_op_, -3, "EM: operator not initialised");
  PCAONE_REQUIRE(maxiter >= 0, -1, "EM: maxiter must be >= 0");
  const uint32_t k = (uint32_t)p_.k;
  set_flags(false, false);
  rsvd_usv(maxp, tol, nullptr, nullptr, nullptr, nullptr, nullptr);
  int it = 0;
  for (int i = 0; i < maxiter; ++i) {
    PCAONE_HIP(hipMemcpyAsync(Vpre_.p, V_.p, (size_t)M_ * k * 8, hipMemcpyDeviceToDevice, stream_));
    set_flags(true, false);
    rsvd_usv(maxp, tol, nullptr, nullptr, nullptr, nullptr, nullptr);
    const int nb = gram_blocks(M_);
    timed("mev", [&] { launch_xty_partial(V_.p, Vpre_.p, M_, k, part_.p, nb, stream_); });
    timed("mev", [&] { launch_mev_finish(part_.p, nb, k, scal_.p, stream_); });
    double m = 0.0;
    PCAONE_HIP(hipMemcpyAsync(&m, scal_.p, 8, hipMemcpyDeviceToHost, stream_));
    sync();
    it = i + 1;
    if (diffs) diffs[i] = 1.0 - m;
    if (1.0 - m < tolem) break;
  }
  set_flags(true, true);  // EMU always ends with a standardised solve (src/Halko.cpp:332-336)
  rsvd_usv(maxp, tol, U, S, V, nullptr, nullptr);
  if (iters) *iters = it;
}

}  // namespace pcaone
