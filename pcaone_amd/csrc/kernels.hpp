// Host-side launch wrappers for the gfx950 kernels.  Every wrapper enqueues on
// the given stream and returns immediately; argument shapes are checked by the
// callers in engine.cpp before anything is launched.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pcaone {

// ---- layout of the packed genotype matrix in HBM --------------------------
// bed_snp : [M][row_bytes]     SNP-major, row_bytes = round_up(ceil(N/4), 64);
//           code 01 (missing) in every padding position, so padding decodes
//           to genotype 0 and is never counted as data.
// bed_smp : [nwin][Npad][win_bytes] sample-major copy cut into SNP windows,
//           win_bytes = round_up(ceil(max_window/4), 64), Npad = 4*row_bytes.

void launch_bed_to_snpmajor(const uint8_t* raw, uint32_t nsamples, uint32_t nsnps, const int32_t* perm,
                            uint8_t* bed_snp, uint32_t row_bytes, hipStream_t s);

void launch_count_codes(const uint8_t* bed_snp, uint32_t nrows, uint32_t row_bytes, uint32_t nvalid,
                        int32_t* counts4, hipStream_t s);

void launch_transpose_windows(const uint8_t* bed_snp, uint32_t row_bytes, uint32_t nsnps, uint32_t blocksize,
                              uint32_t nwin, uint8_t* bed_smp, uint32_t npad, uint32_t win_bytes,
                              hipStream_t s);

// Missing-code lists.  Row r of a packed matrix holds `nvalid(r)` codes; the
// kernels either count the 01 codes per row or write their positions
// (+ index_base(r)) into idx[ptr[r] ..].
struct PackedRows {
  const uint8_t* data;
  uint32_t row_bytes;
  uint32_t nrows;
  // rows are grouped: group g = r / rows_per_group covers codes
  // [0, valid_last) when g is the last group, else [0, valid); rows whose
  // in-group index >= live_rows contribute nothing.
  uint32_t rows_per_group;
  uint32_t live_rows;
  uint32_t valid;
  uint32_t valid_last;
  uint32_t ngroups;
  uint32_t base_step;  // index_base(r) = base0 + group * base_step
  uint32_t base0;      // 0 unless the rows are one window cut out of many
};
void launch_count_missing(const PackedRows& pr, int32_t* counts, hipStream_t s);
void launch_fill_missing(const PackedRows& pr, const int64_t* ptr, int32_t* idx, hipStream_t s);

// ---- contractions over packed rows ----------------------------------------
// out[z][split][row][col] = sum_{k in split} y(row,k) * B[z][col][k],
// y = A1 dosage of the 2-bit code (00->2, 10->1, 11->0, 01->0).
struct PackedGemm {
  const uint8_t* A;
  uint64_t lda;        // bytes per packed row (multiple of 64)
  uint64_t a_zstride;  // bytes between z-slices of A
  uint32_t nrows;      // valid rows per z-slice
  uint32_t kchunks;    // number of 256-code chunks in a row
  uint32_t chunks_per_split;
  uint32_t nsplit;
  const void* B;       // [z][ncols_b][ldb]
  uint64_t ldb;        // elements
  uint64_t b_zstride;  // elements
  void* out;           // [z][nsplit][nrows][ldo]
  uint64_t ldo;
  uint64_t o_zstride;  // elements between z-slices
  uint32_t nz;
  uint32_t fold;       // split-BF16, streamed kernel only: out column c < fold = column c +
                       // column fold + c (hi + lo; fold % 16 in {0, 8}; ldo >= round_up(fold, 16));
                       // 0: every column stored
};
// split-BF16 MFMA: B is bf16 with nt16*16 columns ([hi | lo | 0-pad]); out fp32.
// The A operand is y' = 2 - 2 y (y the dosage), so out = sum_k y'(row,k) B[col][k]
// and X B = colsum(B) - out / 2 (the finalize kernels' `affine` flag).
void launch_packed_gemm_bf16(const PackedGemm& g, int nt16, hipStream_t s);
// whether a launch of nrows x nz rows runs the streamed kernel (the one that folds)
bool packed_bf16_streams(uint64_t nrows, uint32_t nz);
// FP32 / FP64 vector-FMA: B has `ncols` columns of float / double; out same type
void launch_packed_gemm_f32(const PackedGemm& g, int ncols, hipStream_t s);
void launch_packed_gemm_f64(const PackedGemm& g, int ncols, hipStream_t s);

// ---- epilogues -------------------------------------------------------------
// G rows [r0, r1): reduce split partials, add the missing-mask correction and
// the centring/scaling, then emit the H-contraction operand for each row.
struct FinalizeG {
  const void* part;     // [nsplit][nrows][ldp]  (float or double)
  int part_is_f64;
  uint32_t nsplit;
  uint64_t ldp;
  int split_cols;       // 1: value = part[c] + part[l + c]
  uint32_t r0, r1;      // global (laid-out) SNP rows
  uint32_t l;
  const double* F;      // [M] allele frequency, laid-out order
  const double* d;      // [M] 1/sd or 1
  const double* colsum; // [l]  sum_j Omega[j][c]
  const double* fsign;  // [l] sign of Omega's column c (omega, colsum and corr hold the unsigned
                        // columns; G takes the sign), or null (all +1)
  const int64_t* miss_ptr;  // [M+1]  (EMU values only)
  const int32_t* miss_idx;  // sample indices
  const double* omega;  // [N][l]
  const void* corr;     // [M][l] sum of Omega over each row's missing samples, in the partials' type, or
                        // null (none missing)
  double* G;            // [M][l]
  void* g2;             // [M][l]  2 f d G, float when g2_f32 else double
  int g2_f32;
  // H operand, per window: [nwin][bcols][kw]
  int bprec;            // 0 fp64, 1 fp32, 2 bf16 split
  void* bh;
  uint32_t bcols;       // rows of the operand (l, or the padded tile count * 16)
  uint32_t lo_off;      // split operand: row of the lo part of column c is lo_off + c
  uint64_t kw;          // codes per window row (win_bytes*4)
  uint32_t blocksize;   // window length (last may be shorter)
  const double* mval;   // EMU values of the missing entries (CSR order) or null
  double hf;            // 0.5 for A1 dosages (BED2GENO = y/2), 1 for dense values
  uint32_t w0, w1;      // windows of the segment ([r0, r1) are their rows)
  uint32_t nsnps;
  double* s2p;          // [nwin][finalize_chunks(blocksize)][l] partial sums of g2 (affine: of g2 - g')
  int kperm;            // operand positions in the split packed contraction's K order
  int affine;           // part holds the 2 - 2 y contraction: X Omega = colsum - part / 2
};
// grid = (chunks of 64 rows, windows of the segment)
void launch_finalize_G(const FinalizeG& f, hipStream_t s);
uint32_t finalize_chunks(uint32_t blocksize);
// the H operand, g2 and its chunk sums from a given f.G (no contraction before)
void launch_operand_from_G(const FinalizeG& f, hipStream_t s);

// EMU imputation of every missing entry (SNP-CSR order):
// val = clamp(sum_c U[j][c] S[c] V[i][c], -F_i, 1 - F_i)   (src/FilePlink.cpp:253-261)
void launch_emu_update(const int64_t* miss_ptr, const int32_t* miss_idx, uint32_t nsnps, const double* U,
                       const double* S, const double* V, uint32_t k, const double* F, double* val, hipStream_t s);
// pos[p] = position in the SNP-CSR list of the p-th entry of the window lists
void launch_map_missing(const int64_t* wptr, const int32_t* widx, uint32_t nwin, uint32_t npad,
                        uint32_t nsamples, const int64_t* sptr, const int32_t* sidx, int64_t* pos,
                        hipStream_t s);

// S2[w][c] = sum over window w's rows of g2[r][c] for windows [w0, w1), from
// the chunk partials finalize_G wrote
void launch_window_colsum(const double* s2p, uint32_t l, uint32_t blocksize, uint32_t w0, uint32_t w1, double* S2,
                          hipStream_t s);

struct FinalizeH {
  const void* part;     // [nwseg][nsplit][part_rows][ldp]
  int part_is_f64;
  uint32_t nsplit;
  uint64_t ldp;
  uint32_t part_rows;   // rows per split slice of the partials
  uint32_t miss_rows;   // row stride of the per-window missing pointers (Npad)
  int split_cols;
  uint32_t nsamples, l;
  uint32_t w0, w1;      // windows of this launch (w1 - w0 <= 64)
  uint32_t wpart0;      // window whose partials sit at index 0 of `part`
  uint64_t route_b;     // bit (w - w0) set: window w adds into Hb, else Ha
  double* Ha;
  double* Hb;
  const double* S2;     // [nwin][l] window sums of g2 (affine: of g2 - g')
  int affine;           // part holds the 2 - 2 y contraction: X^T g' = sum g' - part / 2
  const int64_t* miss_ptr;  // [nwin*miss_rows+1]  (EMU values only)
  const int32_t* miss_idx;  // SNP rows
  const void* corr;     // [nwin][nsamples][l] sum of g2 over each (window, sample) missing list, in the
                        // partials' type, or null
  // EMU: H_j += sum over missing (i, j) of val * d_i G_i
  const double* mval;   // SNP-CSR values or null
  const int64_t* mpos;  // window entry -> SNP-CSR position
  const double* G;      // [M][l]
  const double* d;      // [M]
  double hf;            // 0.5 for A1 dosages, 1 for dense values
};
void launch_finalize_H(const FinalizeH& f, hipStream_t s);

// Sparse corrections for the missing genotypes: for each list row o of a
// CSR (ptr, idx), out[o][c] = sum_p table[idx[p]][c] (summed and stored in
// the table's type, FP32 or FP64).  List rows come in ngroups groups of `per` rows: row o reads
// the list at in_base + (o / per) * stride_in + o % per and writes out row
// out_base + o.
// Jobs for extra blocks of a launch that does not need their results (see
// side_jobs.hpp); a null stats / s2p pointer leaves that job out.
struct SideJobs {
  // flipOmg's column signs: the flip pass's statistics [nparts][3 W]
  const double* stats = nullptr;
  uint32_t nparts = 0, W = 0, l = 0;
  const double* sign = nullptr;  // [l] Householder signs
  double* fsign = nullptr;       // [l] out: column signs
  double* colsum = nullptr;      // [l] out: unsigned column sums
  // window sums S2[w][c] = sum over chunks of s2p[w][chunk][c], windows [w0, w1)
  const double* s2p = nullptr;
  uint32_t nchunk = 0, w0 = 0, w1 = 0, l2 = 0;
  double* S2 = nullptr;
};
// the jobs in a launch of their own
void launch_side_jobs(const SideJobs& j, hipStream_t s);
// the window sums of finalize_G's chunk partials (k_finalize_G's s2p) as a job
SideJobs window_colsum_job(const double* s2p, uint32_t l, uint32_t blocksize, uint32_t w0, uint32_t w1, double* S2);

struct CsrGather {
  const int64_t* ptr;
  const int32_t* idx;
  const void* table;
  int table_f32;
  uint32_t l;                // columns of the table and of out (row stride)
  uint32_t col0, ncols;      // set by the launcher: the column window of one launch
  uint64_t in_base, stride_in, out_base;
  uint32_t per, ngroups;
  void* out;  // [.][l] of the table's type
  SideJobs side;     // run by the first blocks of the launch
  uint32_t gblocks;  // set by the launcher: the blocks that gather
};
void launch_csr_gather(const CsrGather& g, hipStream_t s);
// FP64 -> FP32 copy
void launch_cast_f32(const double* src, float* dst, uint64_t n, hipStream_t s);

// Omega (FP64 [N][l]) -> G-contraction operand [bcols][kpad] in the arm's type
// (split: hi at row c, lo at row lo_off + c).
// kperm: positions in the split packed contraction's K order (kperm16)
void launch_omega_operand(const double* omega, uint32_t n, uint32_t l, int bprec, void* bg, uint32_t bcols,
                          uint32_t lo_off, uint64_t kpad, int kperm, hipStream_t s);

// ---- dense FP32 feature-major matrices (src/FileBinary.cpp:21-46) ----------
struct DenseGemm {
  const float* A;      // [z][nrows][lda], K contiguous, zero padded to kchunks*64
  uint64_t lda;
  uint64_t a_zstride;
  uint32_t nrows;
  uint32_t kchunks;    // 64-float chunks
  uint32_t chunks_per_split;
  uint32_t nsplit;
  const void* B;       // [z][cols][ldb]: fp64 (arm 0), fp32 (1), bf16 hi tiles | lo tiles (2)
  uint64_t ldb;
  uint64_t b_zstride;
  void* out;           // [z][nsplit][nrows][ldo], fp64 for arm 0 else fp32
  uint64_t ldo;
  uint64_t o_zstride;
  uint32_t nz;
};
void launch_dense_rows(const float* raw, uint32_t nsamples, uint32_t nfeat, const int32_t* perm, float* out,
                       uint32_t ld, hipStream_t s);
void launch_dense_mean(const float* raw, uint32_t nsamples, uint32_t nfeat, double* mu, hipStream_t s);
void launch_dense_transpose(const float* rows, uint32_t ldr, uint32_t nfeat, uint32_t bs, uint32_t nwin,
                            uint32_t nsamples, float* dt, uint32_t ldt, hipStream_t s);

// ---- genotype likelihoods, PCAngsd (src/FileBeagle.cpp, src/Data.cpp:249-343)
// P: [site][sample][2] FP32, the first two of each sample's three GLs.
void launch_gl_em_maf(const float* P, uint32_t nsamples, uint32_t nsites, const double* Fin, double* Fout,
                      double* dpart, hipStream_t s);
uint32_t gl_em_maf_blocks(uint32_t nsites);
struct GlE {
  const float* P;
  uint32_t nsamples, nsites;
  const int32_t* perm;   // laid-out row -> site (null: identity)
  const double* F;       // per site (original order)
  const double* U;       // [nsamples][k] or null (initial E, prior = F)
  const double* S;       // [k]
  const double* V;       // [nsites][k], laid-out rows
  uint32_t k;
  int standardize;       // divide by sqrt(2F(1-F)) and emit Dc partials
  int bed_init;          // U == null and P holds PLINK hard calls: E = dosage / 2 - F, missing 0
  float* out;            // [laid-out row][ld]
  uint64_t ld;
  uint32_t rows_per_block;
  double* dpart;         // [gridDim.y][nsamples] Dc partials, or null
};
void launch_gl_E(const GlE& e, hipStream_t s);
// PLINK hard calls -> the two likelihoods per sample, P [site][sample][2] FP32
void launch_bed_to_gl(const uint8_t* raw, uint32_t nsamples, uint32_t nsites, float* P, hipStream_t s);
// rows_per_block for launch_gl_E: at most 256 row blocks (Dc partials)
uint32_t gl_rows_per_block(uint32_t nsites);
// out[i] = sum_y part[y][i] (fixed order)
void launch_sum_rows(const double* part, uint32_t ny, uint32_t n, double* out, hipStream_t s);
// ---- LD r^2 on a dense feature-major matrix (src/LD.cpp:20-23, 419-442) -----
void launch_ld_sds(const void* X, int x_f64, uint32_t n, uint32_t m, const double* mu, double* sds,
                   hipStream_t s);
struct LdTiles {
  const void* X;           // [m][n] feature-major, FP32 or (x_f64) FP64
  int x_f64;
  const double* mu;        // [m] feature means
  const double* sds;       // [m] 1 / sd
  uint32_t n, m;
  const int2* tiles;       // (row block, column block) pairs, 64 x 64
  const int32_t* win_of_row;  // lead row -> window, or -1
  const int32_t* we;       // [nwin] window sizes (lead included)
  const int64_t* off;      // [nwin] first output of each window
  double* r2;              // window-ordered pair values
};
void launch_ld_tiles(const LdTiles& t, uint32_t ntiles, hipStream_t s);
// ---- LD r^2 from packed genotypes with int8 MFMA (kernels_ld_geno.hip) -----
// Band tiles (row block x column block of SNPs) of the window pairs; r^2 of
// pair (a, b) lands at off[w] + b - a - 1 (w the window led by a), the four
// integer sums (S_yy, S_yw, S_wy, S_ww) at the same index of `sums` when given.
struct LdGeno {
  const uint8_t* bed;        // [nsnps][row_bytes] SNP-major packed codes, padding codes 01
  uint32_t row_bytes;        // multiple of 64
  uint32_t nsnps;
  const int2* tiles;         // (row block, column block) in units of the tile shape
  uint32_t ntiles;
  const double* F;           // allele frequency
  const double* mun;         // sum y / 2 - F n: N times the mean of the centred genotype
  const double* ss;          // sum of squares of the re-centred genotype
  double inv_n;              // 1 / N
  const int32_t* win_of_row; // lead SNP win_lo + r -> window (r < win_rows), or -1
  uint32_t win_lo, win_rows;
  const int32_t* we;         // window sizes (lead included)
  const int64_t* off;        // first output of each window
  double* r2;
  int32_t* sums;             // optional, [npairs][4]
};
int ld_geno_nconfigs();
void ld_geno_tile(int config, int* tm, int* tn);
// persistent grid for a tile list on ncu compute units (a multiple of 8)
uint32_t ld_geno_grid(int config, uint32_t ntiles, uint32_t ncu);
void launch_ld_geno(const LdGeno& g, int config, uint32_t grid, hipStream_t s);

// --USV on PLINK input: R = G - U (U^T G), G = BED2GENO - F (missing 0), from
// the packed codes raw (m rows of ceil(n / 4) bytes), F (m), U ([n][k]); W
// ([m][k]) receives U^T G, R ([m][n], FP64) the adjusted genotypes.
void launch_usv_residual(const uint8_t* raw, uint32_t n, uint32_t m, const double* F, const double* U, uint32_t k,
                         double* W, double* R, hipStream_t s);

// GRM C = E^T E / nsites with diagonal Dc / nsites (E rows = sites, ld stride)
void launch_grm(const float* E, uint64_t ld, uint32_t nsamples, uint32_t nsites, const double* Dc, double* C,
                hipStream_t s);
// ntiles: 16-column tiles of the output (FP32 arm) or of each half (split arm)
void launch_dense_gemm(const DenseGemm& g, int arm, int ncols, int ntiles, hipStream_t s);

void launch_fill_f64(double* p, uint64_t n, double v, hipStream_t s);

// ---- FP64 tall-skinny linear algebra --------------------------------------
// workgroups of the partial-product passes over nrows rows (the mev products)
int gram_blocks(uint64_t nrows);
// ---- the winSVD Omega update in three passes (kernels_qr.hip) ---------------
// Omega <- Householder-signed CholeskyQR2 Q of H (H = H1 + H2 when H2 is
// given), flipped against the previous Omega, written unsigned to every
// consumer with the column signs in fsign.
struct QrPass {
  const double* H1;  // pass 0 input (the only input when H2 is null)
  const double* H2;  // optional second accumulator: H1 + H2 is written to H
  double* H;         // H (passes 1..3 read it; must equal H1 when H2 is null)
  uint64_t n;        // rows (samples), n >= 2 l
  uint32_t l;        // columns, 1..64
  double* part;      // qr_part_doubles(l) scratch, kept between the three passes
  double* gram;      // usv_gram_doubles(l): the reduced partials handed to the next pass
  int nparts;        // blocks of the reduction passes (set by launch_qr_pass)
  int* status;       // [2]: first failing pivot of the two factorisations (sticky, 0 = none)
  double* R1i;       // l x l outputs: R1^-1, R2^-1
  double* R2i;
  double* sign;      // l: Householder signs of the new Omega's columns
  double* fsign;     // l: column signs of the previous Omega in, of the new one out
                     // (Householder sign times flipOmg's flip)
  double* colsum;    // l: column sums of the new (unsigned) Omega
  double* stats;     // qr_stat_doubles(): the flip pass's block statistics (read by qr_sign_job's job)
  double* omega;     // n x l FP64, unsigned: the previous Omega in, the new one out
  double* zero_acc;  // optional n x l accumulator to clear
  float* omega32;    // optional n x l FP32 copy
  void* bg;          // contraction operand [bcols][kpad] of bprec
  int bprec;
  uint32_t lo_off;   // split BF16: column of the low halves
  uint64_t kpad;
  int kperm;         // operand positions in kperm16 order
};
uint64_t qr_part_doubles(uint32_t l);
uint64_t qr_stat_doubles();
// pass 0: H and the Gram of H; 1: R1^-1 and the Gram of Q1; 2: R2^-1, the
// Householder signs, the unsigned Omega into every consumer and the flip
// statistics.  qr_sign_job then settles fsign and colsum (on its own or as the
// side job of the next gather).
void launch_qr_pass(const QrPass& q, int pass, hipStream_t s);
SideJobs qr_sign_job(const QrPass& q);
// computeUSV's linear algebra without materialising Q: CholeskyQR2 factors of
// G (M x l) and B^T = H (R2 R1)^-1 with B B^T for the eigensolve.
struct UsvPass {
  const double* G;  // M x l
  uint64_t m;
  const double* H;  // N x l
  uint64_t n;
  uint32_t l;
  double* part;     // usv_part_doubles(l) scratch
  double* gram;     // usv_gram_doubles(l): reduced Gram between the steps
  double* bbt;      // l x l: B B^T (step 2 output, row-major, symmetric)
  double* R1i;      // l x l outputs: R1^-1, R2^-1, Rt^-1 = R1^-1 R2^-1
  double* R2i;
  double* Rti;
  double* Bt;       // N x l output: H Rt^-1
  int* status;      // [2]: first failing pivot of the two factorisations (sticky)
};
uint64_t usv_part_doubles(uint32_t l);
uint64_t usv_gram_doubles(uint32_t l);
// step 0: Gram of G; 1: R1^-1 and the Gram of G R1^-1; 2: R2^-1, Rt^-1, B^T, B B^T.
void launch_usv_qr(const UsvPass& u, int step, hipStream_t s);
// small dense: C[n][kc] = A[n][l] * W[l][ldw] (first kc columns) * diag(scale)
void launch_rows_times_small(const double* A, uint64_t n, uint32_t l, const double* W, uint32_t ldw,
                             uint32_t kc, const double* scale, double* C, hipStream_t s);
// out[c * n + perm[r]] = A[r * kc + c] (perm null: identity); perm must be a
// permutation of [0, n)
void launch_rows_to_cols(const double* A, uint64_t n, uint32_t kc, const int32_t* perm, double* out,
                         hipStream_t s);
// symmetric eigen-decomposition of sum(part) (l x l) by cyclic Jacobi in one
// workgroup; eigenvalues descending into evals, eigenvectors (columns) into W.
// W0 (optional, may alias W): start from this basis (the previous eigenvectors).
void launch_sym_eig(const double* part, int nparts, uint32_t l, double* evals, double* W, double* sinv,
                    uint32_t kc, const double* W0, hipStream_t s);
// ---- the same steps for l in (64, kWideMaxL] (kernels_wide.hip) -----------
constexpr uint32_t kWideMaxL = 128;
// scratch (doubles) the wide passes use for their partials
uint64_t wide_part_doubles();
int wide_gram_blocks(uint64_t n);
// gram (l x l, full symmetric, row-major) = (A1 + A2)^T (A1 + A2) over n rows
// of row-major n x l matrices; with A2 the sum is also written to Aout
void launch_wide_gram(const double* A1, const double* A2, double* Aout, uint64_t n, uint32_t l, double* part,
                      double* gram, hipStream_t s);
// out = A S (A: n x l, S: l x l, row-major); with prev, flipOmg's column
// statistics of out against prev (sum |prev - out|, sum |prev + out|, sum out,
// 3 x kWideMaxL doubles) land in stat
struct WideApply {
  const double* A;
  uint64_t n;
  uint32_t l;
  const double* S;
  double* out;
  const double* prev;  // optional
  double* stat_part;   // block partials of the statistics (with prev)
};
void launch_wide_apply(const WideApply& w, double* stat, hipStream_t s);
// R^-1 (l x l, upper) of the Cholesky factor of gram; first failing pivot into
// *status when it is still 0
void launch_wide_chol(const double* gram, uint32_t l, double* rinv, int* status, hipStream_t s);
// R2^-1 from gram, Rt^-1 = R1^-1 R2^-1 and, with q1 (the first l rows of Q1),
// the Householder signs of the top l x l block of Q1 R2^-1
void launch_wide_r2(const double* gram, uint32_t l, const double* r1i, double* r2i, double* rti, const double* q1,
                    double* sign, int* status, hipStream_t s);
// Omega = omu diag(flip * sign) written signed to every consumer (fsign stays +1)
struct WideWrite {
  const double* omu;   // n x l unsigned Omega (Q1 R2^-1)
  uint64_t n;
  uint32_t l;
  const double* stat;  // 3 x kWideMaxL flip statistics of omu against the previous Omega
  const double* sign;  // l Householder signs
  double* omega;
  double* colsum;
  double* zero_acc;    // optional accumulator to clear
  float* omega32;      // optional FP32 copy
  void* bg;
  int bprec;
  uint32_t lo_off;
  uint64_t kpad;
  int kperm;
};
void launch_wide_write(const WideWrite& q, hipStream_t s);
// Rank repair for the robust Omega refresh: from the Gram of the n x l A, mark
// the columns whose Cholesky pivot is below 1e-12 of their squared norm
// (numerically in the span of the earlier kept columns) and write out = A
// with those columns replaced by deterministic pseudo-random normal columns
// (RMS entry that of A); mask (l ints) and scale (1 double) are scratch
void launch_wide_repair(const double* A, const double* gram, uint64_t n, uint32_t l, int* mask, double* scale,
                        double* out, hipStream_t s);
// T = W diag(evals)^-1/2 over the eigenpairs above tau * evals[0] (descending
// evals), zero columns for the rest: G T has orthonormal columns spanning the
// numerical range of G when W, evals are the eigenpairs of G^T G
void launch_whiten(const double* W, const double* evals, uint32_t l, double tau, double* T, hipStream_t s);
// Jacobi eigensolve of the full symmetric l x l A (as launch_sym_eig); tmp: l x l scratch
void launch_sym_eig_wide(const double* A, uint32_t l, double* evals, double* W, double* sinv, uint32_t kc,
                         const double* W0, double* tmp, hipStream_t s);

// mev partials: M = X^T Y (k x k) per block
void launch_xty_partial(const double* X, const double* Y, uint64_t n, uint32_t k, double* part, int nblocks,
                        hipStream_t s);
void launch_mev_finish(const double* part, int nparts, uint32_t k, double* out, hipStream_t s);

}  // namespace pcaone
