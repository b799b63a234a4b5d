// Host engine behind the C ABI: owns the HIP stream, the HBM layout of the
// packed genotypes and the RSVD operator state (the members of RsvdOpData /
// NormalRsvdOpData / FancyRsvdOpData, src/Halko.hpp:7-85).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace pcaone {

struct KernelTimer {
  bool on = false;
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::map<std::string, std::pair<int64_t, double>> agg;  // name -> (count, total ms)
  void resolve();
  void clear();
  ~KernelTimer();
};

// One step of the winSVD window loop (src/Halko.cpp:189-223).
struct WindowStep {
  int acc;     // 1: H1 += window, 2: H2 += window
  bool update; // Omega refreshed from H1 + H2 after this window
  int zero;    // accumulator cleared after the refresh (1 or 2), 0 if none
};

// Plan of one FancyRsvdOpData::computeGandH call; `bandsize` is the state the
// operator carries between calls (reset to 1 at pi == 0).
struct FancyPlan {
  bool reset_both;
  std::vector<WindowStep> steps;
};
// nblocks = 0: the in-core plan (bands windows, factor 1)
FancyPlan fancy_plan(int bands, int pi, uint64_t& bandsize, uint32_t nblocks = 0, uint32_t factor = 1);

struct RsvdParams {
  int variant = 2, k = 10, os = 10, bands = 64, precision = 2, seed = 101, rand = 1, perm = 1, standardize = 1,
      ploidy = 2;
  // out-of-core (-m) block plan: windows of `blocksize` SNPs and the winSVD band
  // factor (0: the in-core windows)
  uint32_t blocksize = 0;
  int band_factor = 1;
};

// -m block plan (src/Data.cpp:27-71) and permute_plink's SNP order
// (src/FilePlink.cpp:298-405), host only
void ooc_plan(uint32_t nsamples, uint32_t nsnps, int svd, int k, int os, int bands, double memory, uint32_t& bs,
              uint32_t& nblocks, uint32_t& factor);
void plink_perm(uint32_t nsamples, uint32_t nsnps, uint32_t gb, uint32_t nbands, int32_t* perm);

// Buffers and operand format of one set of contractions (the RSVD operator's
// l columns, or the single column of the Gram mat-vec).
struct SegBufs {
  int prec = 0;
  uint32_t l = 0;
  int nt16 = 0, ntiles = 0;
  uint32_t bcols = 0, lo_off = 0;
  uint64_t ldp = 0;
  size_t opsz = 8;
  bool part_f64 = false, split_cols = false;
  uint32_t fold = 0;  // split contraction folds lo onto hi columns (PackedGemm::fold)
  double hf = 0.5;
  const double* omega = nullptr;
  float* omega32 = nullptr;  // FP32 copy of Omega: the missing-entry gather table of the non-FP64 arms
  const double* colsum = nullptr;
  const double* fsign = nullptr;  // Omega's column signs (null: all +1)
  double *G = nullptr, *S2 = nullptr, *s2p = nullptr;
  bool affine = false;  // the packed split-BF16 contraction runs on 2 - 2 y
  void* g2 = nullptr;        // 2 F d G, float when g2_f32
  bool g2_f32 = false;
  void *corrG = nullptr, *corrH = nullptr;  // missing-entry sums in the partials' type (null: nothing missing)
  unsigned char *bg = nullptr, *bh = nullptr, *gpart = nullptr, *hpart = nullptr;
  double *Ha = nullptr, *Hb = nullptr;
  const double* mval = nullptr;
};

// window pairs -> output offsets (from 0) and the lead-SNP map over [0, m);
// returns the pair count
uint64_t window_offsets(const int32_t* ws, const int32_t* we, uint32_t nwin, uint32_t m, std::vector<int64_t>& off,
                        std::vector<int32_t>& win_of_row);
// band tiles (row block of tm SNPs, column block of tn SNPs) of the windows
// .mbim: the .bim lines with each SNP's allele frequency appended (ld_tools.hip)
void write_mbim(const char* fileout, const char* filebim, const double* F, uint32_t nsnps);
// the CSV reader alone (csv.hip): counts feature-major, library sizes per sample
void read_csv_host(const char* path, int nthreads, float* values, uint64_t capacity, uint32_t* nsamples,
                   uint64_t* nfeatures, int64_t* libsize);
// the .hwe table (inbreed.hip)
void write_hwe(const char* fileout, const char* filebim, const double* hwe_p, const double* lrt, const double* coef,
               uint32_t nsnps);
// rocBLAS handle teardown of grm_eig (grm_eig.cpp); null is a no-op
void destroy_blas(void* handle);
// the .ld.gz text of the window pairs (ld_write.cpp)
void write_ld_gz(const char* fileout, const char* filebim, const int32_t* ws, const int32_t* we, uint32_t nwin,
                 const double* r2, uint64_t npairs, int level, int nthreads);
std::vector<int2> band_tiles(const int32_t* ws, const int32_t* we, uint32_t nwin, uint32_t m, int tm, int tn);
// the per-device scratch of the allocation-free LD entry point (address taken
// once per device); kLdScratchBytes is its fixed size
constexpr uint64_t kLdScratchBytes = 32ull << 20;
void* ld_scratch(int device);

// A BGEN file mapped read-only with its variant blocks located (bgen.hip):
// per variant the genotype block's data offset, stored and decompressed
// lengths and allele count
struct BgenIndex {
  void* map = nullptr;
  size_t len = 0;
  uint32_t nsamples = 0, nvariants = 0;
  int layout = 0, compression = 0;
  std::vector<uint64_t> data;
  std::vector<uint32_t> clen, dlen;
  std::vector<uint16_t> nalleles;
  BgenIndex() = default;
  BgenIndex(const BgenIndex&) = delete;
  BgenIndex& operator=(const BgenIndex&) = delete;
  ~BgenIndex();
};
void bgen_index(const char* path, BgenIndex& ix);

// divide_pos_by_window (src/LD.cpp:125-139); returns the number of windows
uint32_t ld_windows(const int32_t* pos, uint32_t nsnps, const int32_t* chr_end, uint32_t nchr, int32_t ld_bp,
                    int32_t* ws, int32_t* we);

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();

  void load_bed(const uint8_t* payload, uint32_t nsamples, uint32_t nsnps);
  // a .bed file: read into HBM, or (out_of_core) mapped and streamed window by
  // window through a two-slot device ring at every use
  void load_bed_file(const char* filebed, uint32_t nsamples, uint32_t nsnps, bool out_of_core);
  bool streamed() const { return streamed_; }
  void load_dense(const float* featmajor, uint32_t nsamples, uint32_t nfeatures);
  // FileCsv::read_all: a (zstd-compressed) CSV of counts, --scale on the device
  void load_csv(const char* path, int scale, double scale_factor, int nthreads);
  // FileBgen: mode 0 minor-allele dosages (MAF filter, centring, optional
  // standardisation) into the dense path; mode 1 the genotype probabilities as
  // PCAngsd likelihoods.  F and minor: every variant of the file (optional)
  void load_bgen(const char* path, int mode, double maf, int standardize, int ploidy, int nthreads, double* F,
                 uint8_t* minor);
  // genotype likelihoods (PCAngsd): the E matrix is built in HBM from them
  void load_gl(const float* gl, uint32_t nsamples, uint32_t nsites, int per_site);
  // PCAngsd on PLINK input: the loaded genotypes become hard-call likelihoods
  void gl_from_bed();
  void gl_em_maf(int maxiter, double tol, double* F, int* iters);
  void pcangsd_update(const double* U, const double* S, const double* V, int k, int standardize, float* E,
                      double* Dc);
  // .residuals of the loaded PLINK genotypes (ld_stats 0: minus U S V^T)
  void write_residuals(const char* fileout, const double* U, const double* S, const double* V, uint32_t k,
                       int ld_stats);
  // r^2 of window pairs on the LD matrix the handle holds (dense or PLINK)
  std::vector<double> ld_pairs(const std::vector<int32_t>& ws, const std::vector<int32_t>& we);
  void ld_prune(const char* filebim, int ld_bp, double r2_tol, const char* prefix);
  void ld_clump(const char* fileassoc, const char* colnames, const char* filebim, int clump_bp, double clump_r2,
                double clump_p1, double clump_p2, const char* fileout);
  // projection of the loaded PLINK samples onto a reference PCA; returns the matched SNP count
  uint32_t project(const char* filebim, const char* filembim, const double* S, const double* Vref, uint32_t k,
                   uint32_t ref_nsnps, int method, int standardize, double* U, double* Wout);
  // projection of the loaded BEAGLE likelihoods onto a reference PCA (the EM of
  // run_projection_beagle_gl); returns the matched site count
  uint32_t project_gl(const char* filebeagle, const char* filembim, const double* S, const double* Vref, uint32_t k,
                      uint32_t ref_nsnps, int standardize, int maxiter, double tolem, double* U, double* VSout,
                      int* iters, double* diffs);
  uint32_t project_gl_matched(const int32_t* ref_idx, const uint8_t* flip, const double* F, const double* S,
                              const double* Vref, uint32_t k, uint32_t ref_nsnps, int standardize, int maxiter,
                              double tolem, double* U, double* VSout, int* iters, double* diffs);
  // per-site inbreeding coefficients and HWE LRT on the loaded PLINK hard calls
  int inbreeding(const double* U, const double* S, const double* V, uint32_t k, const double* F, int maxiter,
                 double tolem, double* coef, double* lrt, double* hwe_p);
  // eigenpairs of the kept GRM as JacobiSVD(C) gives them (src/Halko.cpp:326-330):
  // |lambda| descending, U2 row-major N x N (row i = sample i)
  void grm_eig(double* evals, double* U2);
  // run_pca_with_arnoldi's in-core PCAngsd EM (IRAM on E each iteration)
  void pcangsd_iram(int k, int ncv, int maxit, double tol, int maxiter, double tolem, double* U, double* S,
                    double* V, int* iters, double* diffs, double* Dc, double* C);
  void pcangsd(int maxp, double tol, int maxiter, double tolem, double* U, double* S, double* V, int* iters,
               double* diffs, double* Dc, double* C);
  // LD r^2 over windows on the loaded dense matrix, or on the --USV-adjusted genotypes
  uint64_t ld_r2(const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2, uint64_t capacity);
  // --USV on PLINK input: the LD matrix becomes (I - U U^T)(BED2GENO - F)
  void ld_usv(const double* U, uint32_t k);
  // LD r^2 on the loaded PLINK genotypes as the --ld --ld-stats 1 route gives
  // it, from exact integer sums (int8 MFMA); config < 0 autotunes the tile
  // shape (the choice is written back)
  uint64_t ld_r2_geno(const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2, uint64_t capacity,
                      int32_t* sums, int* config);
  // the same without device allocation: metadata, tiles and result staging in
  // the per-device scratch (needs the layout of an earlier ld_r2_geno)
  uint64_t ld_r2_geno_noalloc(const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2, uint64_t capacity,
                              int32_t* sums, int config);
  void allele_freq(double* F, int64_t* counts) const;
  void feature_means(double* mu) const;
  // rows [r0, r0 + nr) of the loaded dense matrix as held in HBM (FP32)
  void dense_rows(uint32_t r0, uint32_t nr, float* out);
  void load_dense_finish(bool centre);
  void loaded_dims(uint32_t* nsamples, uint32_t* nfeatures) const {
    if (nsamples) *nsamples = N_;
    if (nfeatures) *nfeatures = M_;
  }

  void rsvd_init(const RsvdParams& p, const double* omega, const int32_t* perm);
  // computeGandH; `check` reads the CholeskyQR breakdown flag back when an
  // Omega refresh ran (the public entry point), so a breakdown is reported by
  // the call that caused it (computeUSV checks once per epoch instead)
  void rsvd_gandh(int pi, double* G, double* H, bool check = true);
  void rsvd_usv(int maxp, double tol, double* U, double* S, double* V, int* epochs, double* diffs);
  void rsvd_perm(int32_t* perm) const;
  void set_flags(bool update, bool standardize);
  void rsvd_em(int maxp, double tol, int maxiter, double tolem, double* U, double* S, double* V, int* iters,
               double* diffs);
  void dims(int64_t* rows, int64_t* cols, int64_t* size) const;
  void gram_init(int precision, bool standardize, int ploidy);
  void gram_matvec(int side, const double* x, double* y);
  void eigs(int side, int k, int ncv, int maxit, double tol, double* evals, double* evecs, int* nconv,
            int* nops);
  void iram(int k, int ncv, int maxit, double tol, double* U, double* S, double* V, int* nconv, int* nops);

  KernelTimer timer;
  void sync();
  std::string last_error;

 private:
  template <class F>
  void timed(const char* name, F&& launch) {
    timed_on(stream_, name, std::forward<F>(launch));
  }
  template <class F>
  void timed_on(hipStream_t s, const char* name, F&& launch);

  void build_layout();
  void draw_omega();
  void init_omega();
  void omega_refresh(bool from_two, double* zero_acc);
  void omega_refresh_wide(bool from_two, double* zero_acc);  // l > 64, and every l in robust mode
  void run_segment(uint32_t w0, uint32_t w1, const std::vector<int>& route);
  void run_segment(uint32_t w0, uint32_t w1, const std::vector<int>& route, const SegBufs& b);
  void usv_step();
  // usv_step's normal path on stream s with the given G, H and scratch
  void usv_step_on(hipStream_t s, const double* G, const double* H, double* part, double* R1i, double* R2i,
                   int* status);
  // computeUSV's epochs one at a time (robust and l > 64 solves)
  int usv_epochs_serial(int maxp, double tol, double* diffs, bool& done);
  void usv_step_wide();  // l > 64
  void usv_step_robust();  // robust mode: G orthogonalised through its Gram eigenpairs
  // computeUSV's epochs; returns the number of computeGandH calls
  int usv_epochs(int maxp, double tol, double* diffs, bool& done);
  void check_status(const char* context = nullptr);
  void upload_scale();
  uint32_t rows_per_block_(int prec) const {
    return (src_ == SRC_DENSE || prec == PREC_SPLIT_BF16) ? 256u : 512u;
  }
  // the split packed contraction reads its operands in the kperm16 K order
  int kperm_(int prec) const { return (src_ == SRC_BED && prec == PREC_SPLIT_BF16) ? 1 : 0; }
  SegBufs op_bufs();
  void refresh_omega32(const SegBufs& b);
  FinalizeG finalize_desc(uint32_t w0, uint32_t w1, const SegBufs& b) const;
  // windows [w0, w1): resident, or one window at a time through the ring
  void seg_G(uint32_t w0, uint32_t w1, const SegBufs& b);
  void seg_H(uint32_t w0, uint32_t w1, const std::vector<int>& route, const SegBufs& b);
  // A: SNP-major rows of window w0 on (seg_G); sample-major block of window w0 on (seg_H)
  void seg_G_on(uint32_t w0, uint32_t w1, const SegBufs& b, const uint8_t* A);
  void seg_H_on(uint32_t w0, uint32_t w1, const std::vector<int>& route, const SegBufs& b, const uint8_t* A);
  void partial_rows(int prec, uint64_t& gmax, uint64_t& hmax) const;
  uint32_t packed_split(int prec, uint64_t rows, uint32_t nz, uint32_t kchunks, uint32_t& cps) const;
  void set_omega_vec(const double* dev_v, const double* host_v);
  void apply_X(const double* x, double* v);
  void apply_Xt(const double* v, double* y);
  // U, S, V (column-major) from the top-k Gram eigenpairs of the smaller side
  void usv_from_eigvecs(int k, const double* lam, const double* vecs, double* U, double* S, double* V);
  void copy_usv(double* U, double* S, double* V);
  void gl_fill_E(const double* U, const double* S, const double* V, uint32_t k, bool standardize, double* Dc);

  int device_;
  hipStream_t stream_ = nullptr;  // every launch, in order, at the default priority
  // computeUSV's linear algebra of epoch pi runs here, beside epoch pi + 1's
  // computeGandH on stream_ (default priority as well)
  hipStream_t ustream_ = nullptr;
  hipEvent_t ev_epoch_ = nullptr;   // stream_: epoch pi's G and H are ready
  DevBuf<double> Gu_, Hu_;          // epoch pi's G (swapped with G_) and H (a copy)
  DevBuf<double> sH1_, sH2_;        // the accumulators before a speculative epoch
  DevBuf<double> upart_, uR1i_, uR2i_;
  DevBuf<int> ustat_;               // [0..1] the refreshes' flags of epoch pi, [2..3] its usv's
  void* hpipe_ = nullptr;           // page-locked: 1 - mev and the four flags
  const double* Gfin_ = nullptr;    // the G of the last computeUSV step
  // A speculative epoch asks spec_gate_ (once, after its first segment's G
  // product) whether to go on; a false answer sets spec_abort_ and the rest of
  // the epoch is never enqueued.
  std::function<bool()> spec_gate_;
  bool spec_abort_ = false;

  // ---- out of core: the payload in host memory, windows through a ring ---
  struct StreamSlot {
    DevBuf<uint8_t> raw, snp, smp;  // window rows as read, SNP-major rows, sample-major block
    uint8_t* stage = nullptr;       // pinned host staging of the rows
    size_t stage_bytes = 0;
    hipEvent_t copied = nullptr;    // the staging's copy to raw is done (copy stream)
    hipEvent_t used = nullptr;      // the last work on this slot is done (stream_)
    bool copy_rec = false, used_rec = false, laid = false;
    int64_t win = -1;
  };
  bool streamed_ = false;
  const uint8_t* host_bed_ = nullptr;  // M rows of ceil(N/4) bytes
  void* map_base_ = nullptr;
  size_t map_len_ = 0;
  StreamSlot slots_[2];
  int last_slot_ = -1;
  hipStream_t copy_stream_ = nullptr;  // host-to-device copies of the ring (default priority)
  void end_stream();                   // back to resident data: unmap, free the ring
  void set_code_counts(const std::vector<int32_t>& h);  // F and counts from (00, 01, 10, 11) per SNP
  void stream_counts();                // code counts and F by one pass over the mapped rows
  void stream_setup();                 // ring buffers for the current windows
  void stream_missing_lists();
  void stream_fill(int s, uint32_t w);
  int stream_acquire(uint32_t w);      // slot holding window w, laid out, ordered on stream_
  void stream_release(int s);
  void stream_prefetch(uint32_t w);
  void gather_rows(uint8_t* dst, uint32_t r0, uint32_t nr, bool permuted) const;
  uint32_t win_rows(uint32_t w) const { return std::min<uint32_t>(bs_, M_ - w * bs_); }
  // the payload in HBM, for the paths that need all of it at once
  const uint8_t* resident_raw(const char* what) const {
    PCAONE_REQUIRE(!streamed_, -3,
                   std::string(what) + ": needs the genotypes resident in HBM (load without out_of_core)");
    return raw_.p;
  }

  // ---- data --------------------------------------------------------------
  enum Source { SRC_NONE, SRC_BED, SRC_DENSE };
  Source src_ = SRC_NONE;
  uint32_t N_ = 0, M_ = 0;
  bool have_bed_ = false;
  DevBuf<uint8_t> raw_;            // payload as loaded (M x ceil(N/4))
  DevBuf<float> rawf_;             // dense features as loaded (M x N)
  DevBuf<float> dns_, dnt_;        // dense rows (laid-out, padded) and windowed transpose
  std::vector<double> mu_orig_;    // dense: per-feature mean (FileBin centring)
  std::vector<double> F_orig_;     // allele frequency, original order
  std::vector<int64_t> counts_orig_;
  // genotype likelihoods: P [site][sample][2] FP32, F from emMAF_with_GL
  bool gl_mode_ = false, have_glF_ = false;
  bool gl_bed_ = false;  // likelihoods are the hard calls of a loaded PLINK set
  DevBuf<float> glP_;
  DevBuf<double> glF_, gl_part_, gl_Dc_;
  std::vector<double> glF_host_;
  // --USV LD input: (I - U U^T) G, [SNP][sample] FP64, original SNP order
  bool have_resid_ = false;
  DevBuf<double> resid_;
  // integer LD on the packed genotypes: identity SNP-major copy and per-SNP
  // terms, built on the first call after a load; tuned tile shape per
  // (window-size class)
  bool have_ldg_ = false;
  // the GRM of the last pcangsd call that returned it (N x N, kept for grm_eig)
  DevBuf<double> grm_;
  bool have_grm_ = false;
  void* blas_ = nullptr;  // rocBLAS handle of the GRM eigensolve, made on first use
  uint32_t ldg_row_bytes_ = 0;
  DevBuf<uint8_t> ldg_bed_;
  DevBuf<double> ldg_F_, ldg_mun_, ldg_ss_;
  std::map<int, int> ldg_tuned_;
  int ncu_ = 0;

  // ---- layout (depends on the permutation and the windowing) -------------
  bool have_op_ = false;
  RsvdParams p_;
  uint32_t l_ = 0, nwin_ = 0, bs_ = 0;
  uint32_t row_bytes_ = 0, npad_ = 0, win_bytes_ = 0;
  uint64_t kw_ = 0;
  std::vector<int32_t> perm_;      // laid-out row r holds SNP perm_[r]
  std::vector<double> F_lay_;      // allele frequency in laid-out order
  DevBuf<int32_t> d_perm_;
  DevBuf<uint8_t> bed_snp_, bed_smp_;
  DevBuf<double> F_, d_;
  DevBuf<int64_t> miss_snp_ptr_, miss_win_ptr_;
  DevBuf<int32_t> miss_snp_idx_, miss_win_idx_;
  DevBuf<int64_t> miss_win_pos_;   // window entry -> SNP-CSR position
  int64_t nnz_snp_ = 0;

  // ---- operator state ----------------------------------------------------
  int nt16_ = 0;        // 16-column tiles of the split operand
  uint32_t bcols_ = 0;  // columns of the contraction operands
  size_t opsz_ = 8;     // bytes per operand element
  uint32_t lo_off_ = 0; // split operands: row of the lo part of column c is lo_off_ + c
  uint64_t ldp_ = 0;    // columns of the partial products
  bool part_f64_ = false, split_cols_ = false;
  uint32_t fold_ = 0;
  int ntiles_ = 0;      // dense MFMA: 16-column output tiles
  uint64_t kg_ = 0;     // K extent of the G contraction (padded samples)
  uint32_t miss_rows_ = 0;
  double hf_ = 0.5;
  std::vector<double> omega_given_;
  DevBuf<double> omega_, colsum_, omega0_, colsum0_;
  DevBuf<unsigned char> bg_, bh_;   // operands of the G and H contractions
  DevBuf<double> G_, S2_, s2p_, H1_, H2_, H_;
  DevBuf<unsigned char> g2_;
  DevBuf<float> omega32_;
  DevBuf<unsigned char> corrG_, corrH_;  // [M][l] and [nwin][N][l] missing-entry sums (partials' type)
  DevBuf<unsigned char> gpart_, hpart_;
  uint64_t bandsize_ = 1;
  // RsvdOpData::update / U, S, V of the last computeUSV (src/Halko.hpp:11-14)
  bool update_ = false, have_usv_ = false, have_vals_ = false;
  bool have_W_ = false;  // W_ holds this solve's previous eigenvectors
  std::vector<double> last_S_;
  DevBuf<double> miss_val_, Sdev_, Vpre_;
  // linear algebra scratch
  DevBuf<double> part_, R1_, R1i_, R2_, R2i_, Rt_, Rti_, sign_;
  DevBuf<double> fsign_;           // Omega's column signs (omega_ and colsum_ are unsigned)
  DevBuf<double> qstat_;           // the flip pass's statistics, kept until the signs settle
  // Jobs deferred to the next missing-entry gather (which does not need them):
  // the sign settle after an Omega refresh (stats != null) and the window sums
  // after finalize_G (s2p != null).  flush_side launches them on their own.
  SideJobs side_;
  SideJobs take_side(bool signs, bool colsum);
  void flush_side(bool signs, bool colsum);
  DevBuf<double> gred_, bbt_;  // computeUSV: reduced Gram between passes, B B^T
  DevBuf<double> qred_;        // Omega update: reduced partials between passes
  DevBuf<double> Bt_, evals_, W_, sinv_, Ucur_, Upre_, V_, scal_;
  DevBuf<double> out_;  // column-major U / V staged for the copy to the host
  // l > 64: the tall-skinny steps run in general form (kernels_wide.hip)
  bool wide_ = false;
  DevBuf<double> wq_;      // max(N, M) x l: Q1 of the current factorisation
  DevBuf<double> eigtmp_;  // l x l eigensolve scratch
  // Robust mode (after a CholeskyQR breakdown of an Omega refresh, the solve
  // restarts in it): every refresh marks the columns of H numerically in the
  // span of the earlier ones and orthogonalises a copy of H with those columns
  // replaced by pseudo-random ones.  The accumulators are kept as they were at
  // the start of the solve, for the restart.
  bool robust_ = false;
  DevBuf<double> hrep_, snapH1_, snapH2_, rscale_, Wg_;
  DevBuf<int> rmask_;
  DevBuf<int> status_;
  // Gram mat-vec state (one operand column)
  bool have_mv_ = false;
  int64_t nops_ = 0;
  SegBufs mv_;
  DevBuf<double> mv_omega_, mv_colsum_, mv_G_, mv_S2_, mv_s2p_, mv_H_;
  DevBuf<unsigned char> mv_corrG_, mv_corrH_;
  DevBuf<unsigned char> mv_g2_;
  DevBuf<float> mv_omega32_;
  DevBuf<unsigned char> mv_bg_, mv_bh_, mv_gpart_, mv_hpart_;
};

// Launch wrapper: with profiling on, HIP events on the handle's stream bracket
// the launch and are resolved into per-name totals later.
template <class F>
void Engine::timed_on(hipStream_t s, const char* name, F&& launch) {
  if (!timer.on) {
    launch();
    PCAONE_HIP(hipGetLastError());
    return;
  }
  KernelTimer::Rec r{name, nullptr, nullptr};
  PCAONE_HIP(hipEventCreate(&r.a));
  PCAONE_HIP(hipEventCreate(&r.b));
  PCAONE_HIP(hipEventRecord(r.a, s));
  launch();
  PCAONE_HIP(hipGetLastError());
  PCAONE_HIP(hipEventRecord(r.b, s));
  timer.pending.push_back(r);
  if (timer.pending.size() > 4096) timer.resolve();
}

}  // namespace pcaone
