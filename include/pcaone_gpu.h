/*
 * pcaone_gpu.h -- C ABI of the MI355X (gfx950) implementation of PCAone's
 * randomized-SVD hot path.
 *
 * Plain C: opaque handle, plain pointers and sizes, int status (0 = OK, < 0 =
 * error; the message is kept per handle, see pcaone_gpu_last_error).  All host
 * pointers belong to the caller; the handle owns its device buffers and one HIP
 * stream.  A handle is not thread-safe.  Calls that return host data return
 * after that data is written.  Matrices crossing the boundary are FP64
 * column-major (Eigen's default, the layout of the reference's Mat2D).
 *
 * Reference interfaces each entry point replaces are cited as path:line of
 * PCAone v0.6.0.
 */
#ifndef PCAONE_GPU_H
#define PCAONE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCAONE_OK 0
#define PCAONE_EINVAL (-1)  /* bad argument or size                          */
#define PCAONE_EHIP (-2)    /* HIP runtime error                             */
#define PCAONE_ESTATE (-3)  /* call out of order (e.g. no data loaded)       */
#define PCAONE_ENUM (-4)    /* numerical breakdown (e.g. Cholesky pivot <= 0) */

/* arithmetic arms of the genotype contractions */
#define PCAONE_PREC_FP64 0       /* decode -> FP64, FP64 FMA                         */
#define PCAONE_PREC_FP32 1       /* decode -> FP32, FP32 FMA (plain drop-in)         */
#define PCAONE_PREC_SPLIT_BF16 2 /* genotypes exact in BF16, operand = hi+lo BF16,
                                    FP32 MFMA accumulation                          */

/* RSVD variants: --svd 1 and --svd 2 of the reference (src/Cmd.cpp:129-138) */
#define PCAONE_RSVD_NORMAL 1 /* NormalRsvdOpData (src/Halko.hpp:44-62)  */
#define PCAONE_RSVD_FANCY 2  /* FancyRsvdOpData  (src/Halko.hpp:64-85)  */

typedef struct pcaone_gpu pcaone_gpu;

/* ---- lifecycle -------------------------------------------------------- */
int pcaone_gpu_create(int device, pcaone_gpu** out);
void pcaone_gpu_destroy(pcaone_gpu* h);
const char* pcaone_gpu_last_error(const pcaone_gpu* h);
/* Number of HIP devices visible (0 on a host without GPU); never fails. */
int pcaone_gpu_device_count(void);
/* Page-locked host memory for result buffers (no counterpart in the
 * reference, whose results stay in host memory): U, S and V written into such
 * a buffer travel at the full PCIe rate, where a pageable buffer is staged and
 * faulted in page by page (an 80 MB V: 1.4 ms against 7.8 ms).  A caller such
 * as the reference's Eigen::Map over its output matrices may use it; any
 * host pointer works with every entry point.  Returns PCAONE_EINVAL for a
 * null out pointer, a negative status when the memory cannot be had. */
int pcaone_gpu_host_alloc(size_t bytes, void** out);
int pcaone_gpu_host_free(void* p);

/* ---- data residency ----------------------------------------------------
 * Replaces FileBed::read_all (src/FilePlink.cpp:29-131): `payload` is the .bed
 * body after the 3-byte magic (src/FilePlink.hpp:22-26), nsnps rows of
 * (nsamples+3)>>2 bytes.  The bytes stay packed in HBM; allele frequencies
 * are computed on the device from exact code counts. */
int pcaone_gpu_load_bed(pcaone_gpu* h, const uint8_t* payload, uint32_t nsamples, uint32_t nsnps);
/* A PLINK .bed file (magic 6c 1b 01, then nsnps rows of ceil(nsamples / 4)
 * bytes): FileBed's reader (src/FilePlink.cpp:18-131).  out_of_core = 0 reads
 * it into HBM as pcaone_gpu_load_bed does.  out_of_core = 1 is -m
 * (Param::out_of_core, src/Cmd.cpp:192; read_block_initial,
 * src/FilePlink.cpp:133-224): the file stays mapped in host memory and every
 * pass of the operators streams it window by window (rows gathered into pinned
 * memory, copied to a two-slot device ring, laid out there), so a .bed larger
 * than HBM runs.  F comes from one streamed pass.  Paths that need the whole
 * matrix in HBM at once (LD, PCAngsd on PLINK, inbreeding, residuals,
 * projection) refuse a streamed handle with PCAONE_ESTATE. */
int pcaone_gpu_load_bed_file(pcaone_gpu* h, const char* filebed, uint32_t nsamples, uint32_t nsnps,
                             int out_of_core);
/* Replaces FileBin::read_all (src/FileBinary.cpp:21-30): `featmajor` holds
 * nfeatures rows of nsamples float32 (the .bin / .residuals body after the
 * u32 M, u32 N header, src/FileBinary.hpp:15-16).  The values stay FP32 in
 * HBM; each feature is centred by its FP64 mean inside the contractions.  The
 * RSVD operator on this input never standardises (src/Halko.cpp:290-291). */
int pcaone_gpu_load_dense_f32(pcaone_gpu* h, const float* featmajor, uint32_t nsamples, uint32_t nfeatures);
/* Per-feature means (nfeatures, original order) of the loaded dense matrix. */
int pcaone_gpu_feature_means(pcaone_gpu* h, double* mu);
/* F (nsnps) as estimated in src/FilePlink.cpp:36-62 (original SNP order) and,
 * optionally, counts (nsnps x 4, row-major: codes 00, 01, 10, 11). */
int pcaone_gpu_allele_freq(pcaone_gpu* h, double* F, int64_t* counts);

/* ---- RSVD operator (RsvdOpData, src/Halko.hpp:7-42) -------------------- */
typedef struct {
  int variant;     /* PCAONE_RSVD_NORMAL | PCAONE_RSVD_FANCY                        */
  int k;           /* -k                                                           */
  int oversamples; /* --oversamples, after Param's max(os, k) (src/Cmd.cpp:190)     */
  int bands;       /* -w/--batches: even, >= 4 (src/Cmd.cpp:213-214); read by the
                      winSVD variant only (the plain variant tiles its own)     */
  int precision;   /* PCAONE_PREC_*                                                */
  int seed;        /* --seed (src/Cmd.hpp:86)                                      */
  int rand;        /* --rand: 1 Gaussian, 0 uniform(-1, 1) (src/Cmd.hpp:50)        */
  int perm;        /* winSVD column shuffle, Param::perm (src/Cmd.cpp:216)         */
  int standardize; /* divide by sqrt(ploidy F (1-F)) (src/Data.cpp:290-299)        */
  int ploidy;      /* 2, or 1 with --haploid                                       */
  uint32_t blocksize; /* out of core (-m): window length of pcaone_ooc_plan's block
                         plan (src/Data.cpp:27-71); 0: the in-core windows      */
  int band_factor;    /* out of core winSVD: Data::bandFactor; the schedule counts
                         blocks from it (src/Halko.cpp:227-269); 0 or 1 in core */
} pcaone_rsvd_params;

/* Builds the HBM layout for the operator (SNP shuffle, SNP-major and windowed
 * sample-major packed copies, missing lists) and draws Omega with libstdc++'s
 * default_random_engine seeded by p->seed (RsvdOpData::initOmg,
 * src/Halko.cpp:14-22).  `omega` (nsamples x l, column-major) replaces that
 * draw when non-NULL; `perm` (nsnps) replaces the std::shuffle permutation
 * (src/RSVD.hpp:63-79) when non-NULL and p->perm is set. */
int pcaone_gpu_rsvd_init(pcaone_gpu* h, const pcaone_rsvd_params* p, const double* omega, const int32_t* perm);
/* rows() = nsnps, cols() = nsamples, size() = k + oversamples
 * (src/Halko.hpp:21-25). */
int pcaone_gpu_rsvd_dims(const pcaone_gpu* h, int64_t* rows, int64_t* cols, int64_t* size);
/* The SNP permutation in use: row j of G / V belongs to SNP perm[j]
 * (identity unless the winSVD shuffle is on). */
int pcaone_gpu_rsvd_perm(const pcaone_gpu* h, int32_t* perm);
/* computeGandH(G, H, pi) (src/Halko.hpp:27; Normal src/Halko.cpp:97-126,
 * Fancy src/Halko.cpp:153-226).  G and H stay in HBM; G (rows() x size(),
 * rows in operator order) and H (cols() x size()) are copied out when
 * non-NULL. */
int pcaone_gpu_rsvd_gandh(pcaone_gpu* h, int pi, double* G, double* H);
/* computeUSV(p, tol) (src/Halko.cpp:45-95): U (nsamples x k), S (k),
 * V (nsnps x k, ORIGINAL SNP order, i.e. data->perm * V as written by
 * src/Halko.cpp:339-340).  epochs: computeGandH calls made; diffs (optional,
 * maxp+1 entries): 1 - mev per epoch (diffs[0] = 0). */
int pcaone_gpu_rsvd_usv(pcaone_gpu* h, int maxp, double tol, double* U, double* S, double* V, int* epochs,
                        double* diffs);

/* RsvdOpData::setFlags(update, standardize) (src/Halko.hpp:32-35): with
 * update set, the next computeGandH at pi == 0 re-imputes every missing
 * genotype from the operator's current U, S, V (fit_with_pi, EMU branch,
 * src/Data.cpp:272-287); standardize switches the 1/sd scaling. */
int pcaone_gpu_rsvd_set_flags(pcaone_gpu* h, int update, int standardize);
/* run_pca_with_halko's EM branch for --emu (src/Halko.cpp:294-336): a plain
 * solve, then re-imputed solves until 1 - mev(V, Vpre) < tolem (at most
 * maxiter), then the final standardised solve whose U, S, V are returned.
 * diffs (optional, maxiter entries) receives 1 - mev per EM iteration. */
int pcaone_gpu_rsvd_em(pcaone_gpu* h, int maxp, double tol, int maxiter, double tolem, double* U, double* S,
                       double* V, int* iters, double* diffs);

/* ---- PCAngsd on genotype likelihoods ------------------------------------ */
/* Replaces FileBeagle::read_all's likelihood parse (src/FileBeagle.cpp:99-100,
 * 160-180): gl holds nsites x nsamples x per_site float32 likelihoods (the
 * BEAGLE body, per_site = 3; or only the first two, per_site = 2).  Only the
 * first two stay in HBM; the third is taken as 1 - P0 - P1. */
int pcaone_gpu_load_gl(pcaone_gpu* h, const float* gl, uint32_t nsamples, uint32_t nsites, int per_site);
/* --pcangsd on in-core PLINK input (FileBed::read_all, src/FilePlink.cpp:68,
 * 89-114): the genotypes of pcaone_gpu_load_bed become the hard-call
 * likelihoods RR (1, 0), AA (0, 0), RA (0, 1), missing (0.333333, 0.333333);
 * F stays the genotype-count allele frequency (no emMAF_with_GL for PLINK,
 * src/FilePlink.cpp:127-130).  pcaone_gpu_rsvd_init then builds the initial E
 * as the centred genotype BED2GENO - F (missing 0) and pcaone_gpu_pcangsd runs
 * the EM (pass Dc = C = NULL: the GRM step is for BEAGLE input only,
 * src/Halko.cpp:317). */
int pcaone_gpu_gl_from_bed(pcaone_gpu* h);
/* emMAF_with_GL (src/Utils.cpp:501-531) from F = 0.25 (src/FileBeagle.cpp:127)
 * until the RMS change of F is below tol (at most maxiter passes).  F (nsites,
 * optional) and iters (optional) receive the result.  Required before
 * pcaone_gpu_rsvd_init, which then builds the initial E of
 * src/FileBeagle.cpp:133-143 in HBM. */
int pcaone_gpu_gl_em_maf(pcaone_gpu* h, int maxiter, double tol, double* F, int* iters);
/* Data::fit_with_pi's PCAngsd branch (src/Data.cpp:249-270) or, with
 * standardize, Data::pcangsd_standardize_E (src/Data.cpp:301-343), from U
 * (nsamples x k), S (k) and V (nsites x k, original site order), column-major.
 * The result replaces the operator's matrix (as data->G in the reference).
 * E (optional): nsites x nsamples float32, site-major; Dc (optional, only with
 * standardize): the nsamples GRM diagonal sums. */
int pcaone_gpu_pcangsd_update(pcaone_gpu* h, const double* U, const double* S, const double* V, int k,
                              int standardize, float* E, double* Dc);
/* run_pca_with_halko's EM branch for PCAngsd on BEAGLE input
 * (src/Halko.cpp:294-331): a solve on the initial E, then E <- fit_with_pi(U,
 * S, V) and a new solve until 1 - mev(V, Vpre) < tolem (at most maxiter); then
 * pcangsd_standardize_E.  U, S, V: the last solve; iters, diffs (maxiter
 * entries) optional; Dc (nsamples, optional); C (nsamples x nsamples,
 * optional): the GRM E E^T / nsites with diagonal Dc / nsites written to the
 * reference's .cov. */
int pcaone_gpu_pcangsd(pcaone_gpu* h, int maxp, double tol, int maxiter, double tolem, double* U, double* S,
                       double* V, int* iters, double* diffs, double* Dc, double* C);

/* run_pca_with_arnoldi's in-core EM branch for PCAngsd (src/Arnoldi.cpp:57-104,
 * --svd 0 with -G): IRAM (k, ncv, maxit restarts, Ritz tolerance tol, as
 * pcaone_gpu_iram) on the initial E, flip_UV, then E <- fit_with_pi(U, S, V)
 * and a new IRAM until rmse(V_new, V) < tolem (src/Utils.cpp:116-167), at
 * most maxiter times; diffs[i - 1] is iteration i's RMSE.  As the reference's
 * solver object does (it caches the eigenvectors it first returns,
 * external/Spectra/contrib/PartialSVDSolver.h:177-206), the singular vectors of
 * the smaller side stay those of the first IRAM; later IRAMs give S, and the
 * other side is the current E times those vectors over S.  Then
 * pcangsd_standardize_E: Dc (nsamples) and, when C is given, the GRM
 * (nsamples x nsamples, diagonal Dc / M) as pcaone_gpu_pcangsd.  Needs
 * pcaone_gpu_load_gl, pcaone_gpu_gl_em_maf and pcaone_gpu_gram_init. */
int pcaone_gpu_pcangsd_iram(pcaone_gpu* h, int k, int ncv, int maxit, double tol, int maxiter, double tolem,
                            double* U, double* S, double* V, int* iters, double* diffs, double* Dc, double* C);

/* FileCsv::read_all (src/FileCsv.cpp:11-62, 82-142): a CSV of counts, one
 * feature per line and one sample per comma-separated field, zstd-compressed
 * (the reference's input) or plain text, parsed on nthreads host threads.
 * --scale on the device: 2 log10(count * median(libsize) / libsize + 1),
 * 3 log1p(count / libsize * scale_factor), 4 count / libsize * scale_factor,
 * 1 the counts; scales 1-4 then standardise each feature over samples (N - 1,
 * sd > 1e-10); 0 leaves the counts as they are, uncentred, as the reference's
 * reader does.  libsize: per sample, the sum of the counts read as integers
 * (std::stoi).  The matrix then serves as for pcaone_gpu_load_dense_f32;
 * *nsamples and *nfeatures (optional) receive its size. */
int pcaone_gpu_load_csv(pcaone_gpu* h, const char* path, int scale, double scale_factor, int nthreads,
                        uint32_t* nsamples, uint32_t* nfeatures);
/* FileBgen (src/FileBgen.hpp:13-46, src/FileBgen.cpp:11-107) with the
 * reference's vendored BGEN reader (external/bgen): a BGEN v1.1 / v1.2 file,
 * layout 1 or 2, uncompressed, zlib or zstd (libzstd.so.1 opened at run
 * time).  The host walks the variant blocks and decompresses them on
 * nthreads threads; the device decodes the probabilities.
 * mode 0 (FileBgen::read_all without --pcangsd): per variant the minor
 * allele's dosage (Genotypes::get_allele_dosage with use_minor: the 8-bit
 * byte-pair and any-depth paths, the sampled minor-allele test), F = mean of
 * dose / 2 over the non-missing samples, the variant kept when F > maf (maf
 * above 0.5 read as 1 - maf, src/Cmd.cpp:196-199), and the feature
 * dose / 2 - F (missing 0), divided by sqrt(ploidy F (1 - F)) when
 * standardize and that exceeds 1e-9 (Data::standardize_E, which the
 * reference's operator applies to BGEN input, src/Halko.cpp:285-289,
 * src/Arnoldi.cpp:60-62).  The matrix then serves as for
 * pcaone_gpu_load_dense_f32, without further centring.  F and minor
 * (optional, nvariants entries each, file order): every variant's estimate
 * and minor allele index (0: the first allele).
 * mode 1 (--pcangsd, src/FileBgen.cpp:72-106): the first two genotype
 * probabilities of every sample (Genotypes::probabilities; missing samples
 * NaN) become the likelihoods of pcaone_gpu_load_gl (diploid unphased data
 * only; maf must be 0).  nsamples, nsnps (optional): the loaded size. */
int pcaone_gpu_load_bgen(pcaone_gpu* h, const char* path, int mode, double maf, int standardize, int ploidy,
                         int nthreads, uint32_t* nsamples, uint32_t* nsnps, double* F, uint8_t* minor);
/* Rows [r0, r0 + nrows) of the dense matrix in HBM (FP32, nsamples per row):
 * what FileBin, FileCsv or FileBgen's read_all leave in data->G, column by
 * column (src/Data.hpp:49).  No counterpart entry point in the reference,
 * whose matrix is a host array; for callers that write it out (.residuals)
 * or check it. */
int pcaone_gpu_dense_rows(pcaone_gpu* h, uint32_t r0, uint32_t nrows, float* out);
/* Host only: the BGEN header (CppBgenReader's Header, external/bgen/header.cpp)
 * and a walk over every variant block. */
int pcaone_bgen_info(const char* path, uint32_t* nsamples, uint32_t* nvariants, int* layout, int* compression);
/* The reader of pcaone_gpu_load_csv alone, host only: the counts as read
 * (feature-major float32, capacity in floats; values NULL: sizes only) and
 * each sample's integer library size (libsize, nsamples, optional). */
int pcaone_read_csv(const char* path, int nthreads, float* values, uint64_t capacity, uint32_t* nsamples,
                    uint64_t* nfeatures, int64_t* libsize);

/* ---- projection (--project 1|2 with --USV) -------------------------------- */
/* run_projection (src/Projection.cpp:147-208) for the loaded PLINK samples:
 * target .bim lines matched to the reference .mbim (ID, then chromosome and
 * position; alleles equal, or swapped with F = 1 - freq and the loading row
 * negated; src/FilePlink.hpp:31-93), g = BED2GENO - F (missing 0), divided by
 * sqrt(2F(1-F)) when standardize and that exceeds 1e-9.  method 1: U = G V S^-1;
 * method 2: per sample least squares on its observed SNPs, V S as the design
 * (normal equations, Cholesky).  S (k), Vref (ref_nsnps x k column-major, the
 * .loadings), 1 <= k <= 32.  U (nsamples x k column-major); W (optional,
 * noverlap x k column-major): the matched loadings the reference writes
 * (V S^-1 or V S); *noverlap the matched SNP count (W needs at most nsnps x k). */
int pcaone_gpu_project(pcaone_gpu* h, const char* filebim, const char* filembim, const double* S,
                       const double* Vref, uint32_t k, uint32_t ref_nsnps, int method, int standardize, double* U,
                       double* W, uint32_t* noverlap);
/* run_projection_beagle_gl (src/Projection.cpp:32-137) on the likelihoods of
 * pcaone_gpu_load_gl (the BEAGLE body of `filebeagle`), with FileBeagle's
 * matching of its sites against the reference .mbim (src/FileBeagle.cpp:
 * 12-118: by ID, else chromosome_position; alleles 0-3 read as A, C, G, T;
 * equal alleles F = the .mbim frequency, swapped F = 1 - it and the loading
 * row negated).  S (k) and Vref (ref_nsnps x k, column-major) are the
 * reference's .sigvals and .loadings.  E step on the device (posterior
 * genotype under pt = F, then pt = (U_i S) L_j (times sqrt(2F(1-F)) when
 * standardize) + F, clamped to [1e-4, 1 - 1e-4]; standardised when
 * standardize), least squares of every sample on V S over its observed sites
 * (solve_projection_scores, normal equations by Cholesky), columns aligned
 * in sign with the previous scores, until |U - Uprev| / |Uprev| < tolem (at
 * most maxiter).  U: nsamples x k column-major; VS (optional, nsites x k
 * capacity): the matched rows of V S column-major, the loadings the
 * reference writes; noverlap: matched sites; iters, diffs (maxiter entries):
 * the iterations and their relative changes. */
int pcaone_gpu_project_gl(pcaone_gpu* h, const char* filebeagle, const char* filembim, const double* S,
                          const double* Vref, uint32_t k, uint32_t ref_nsnps, int standardize, int maxiter,
                          double tolem, double* U, double* VS, uint32_t* noverlap, int* iters, double* diffs);
/* The same EM with the matching given, as FileBeagle::read_all leaves it in
 * Data (project_ref_indices, project_flip, F; src/Data.hpp:56-60): per loaded
 * site ref_idx (-1 unmatched), flip (0/1) and F. */
int pcaone_gpu_project_gl_matched(pcaone_gpu* h, const int32_t* ref_idx, const uint8_t* flip, const double* F,
                                  const double* S, const double* Vref, uint32_t k, uint32_t ref_nsnps,
                                  int standardize, int maxiter, double tolem, double* U, double* VS,
                                  uint32_t* noverlap, int* iters, double* diffs);

/* ---- inbreeding / HWE (--inbreed 1 with --USV) ---------------------------- */
/* run_inbreeding_em (src/Inbreeding.cpp:16-190) on the loaded PLINK hard
 * calls: Pi = clamp((U S V^T + 2F) / 2, 1e-4, 1 - 1e-4) (src/FileUSV.cpp:11-22)
 * from a PCA's U (nsamples x k), S (k), V (nsnps x k, column-major; 1 <= k <=
 * 64) and the .mbim frequencies F (nsnps); the per-site inbreeding
 * coefficients by the SQUAREM-accelerated EM (maxiter, tolem as --maxiter,
 * --tol-em), then the per-site LRT and its chi-square(1) p-value.  Outputs
 * (nsnps each, optional): coef, lrt, hwe_p; *iters the EM iterations run. */
int pcaone_gpu_inbreeding(pcaone_gpu* h, const double* U, const double* S, const double* V, uint32_t k,
                          const double* F, int maxiter, double tolem, double* coef, double* lrt, double* hwe_p,
                          int* iters);
/* write_hwe_per_site (src/Inbreeding.cpp:129-144), host only: "#ID\tHWE_P\tLRT\t
 * Inbreeding_coefficient", then per .bim line its second field and the three
 * values (6 significant digits). */
int pcaone_write_hwe(const char* fileout, const char* filebim, const double* hwe_p, const double* lrt,
                     const double* coef, uint32_t nsnps);

/* ---- LD inputs and consumers ------------------------------------------- */
/* Data::write_residuals (src/Data.cpp:196-245) for the loaded PLINK genotypes
 * (the --ld route of a PLINK PCA, which runs unstandardised): the centred
 * genotypes BED2GENO - F (missing 0), minus U S V^T when ld_stats == 0 (U
 * nsamples x k, V nsnps x k column-major, S k; 1 <= k <= 64; ignored for
 * ld_stats == 1), each SNP re-centred by its mean over samples, written as a
 * uint32 nsnps, a uint32 nsamples and then one float32 row of nsamples per SNP
 * in the .bed's SNP order. */
int pcaone_gpu_write_residuals(pcaone_gpu* h, const char* fileout, const double* U, const double* S,
                               const double* V, uint32_t k, int ld_stats);
/* Data::save_snps_in_bim (src/Data.cpp:96-135), host only: each line of
 * filebim, a tab and F[j] (6 significant digits); the .bim must have nsnps
 * lines. */
int pcaone_write_mbim(const char* fileout, const char* filebim, const double* F, uint32_t nsnps);
/* ld_prune_big (src/LD.cpp:194-242) on the handle's LD matrix (a dense
 * .residuals matrix, or the loaded PLINK genotypes): windows of ld_bp from
 * filebim's positions (as pcaone_ld_windows), r^2 of their pairs on the
 * device, then the greedy pass: a window's lead, if still kept, drops each
 * kept partner with r^2 > r2_tol (the partner, or when filebim has a 7th
 * column of allele frequencies, the one of lower MAF).  Writes
 * prefix.ld.prune.in and prefix.ld.prune.out (the kept / dropped .bim lines). */
int pcaone_gpu_ld_prune(pcaone_gpu* h, const char* filebim, int ld_bp, double r2_tol, const char* prefix);
/* ld_clump_single_pheno with valid_assoc_file, map_index_snps and
 * get_target_snp_idx (src/LD.cpp:244-370, 83-122) on the handle's LD matrix:
 * fileassoc is tab-separated with a header holding the chromosome, position
 * and p-value columns (colnames "CHR,BP,P" when NULL or empty); index SNPs
 * (p <= clump_p1, in p order) absorb the SNPs with p <= clump_p2 within
 * clump_bp whose r^2 >= clump_r2.  Writes fileout: the header + "\tSP2", then
 * per index SNP its line, a tab and the clumped positions by p (or NONE).
 * Equal p-values are taken by position (the reference's order for ties
 * follows its hash map). */
int pcaone_gpu_ld_clump(pcaone_gpu* h, const char* fileassoc, const char* colnames, const char* filebim, int clump_bp,
                        double clump_r2, double clump_p1, double clump_p2, const char* fileout);

/* The GRM's eigenvectors for .eigvecs2: the PCAngsd branch of the EM driver
 * decomposes C with Eigen::JacobiSVD and writes its left singular vectors
 * (src/Halko.cpp:317-331; src/Arnoldi.cpp:94-103).  C is symmetric, so these
 * are its eigenvectors ordered by |eigenvalue| descending; column signs are
 * the solver's own.  Runs on the C kept by the last pcaone_gpu_pcangsd call
 * that was given C (FP64, on the device).  evals (nsamples, optional): the
 * singular values |lambda|; U2 (nsamples x nsamples row-major, optional): row i
 * = sample i, column j = vector j. */
int pcaone_gpu_grm_eig(pcaone_gpu* h, double* evals, double* U2);

/* write_eigvecs2_beagle (src/Utils.cpp:426-438), host only: "#FID\tIID\tPC1 ..
 * PCn", then per sample (names from the BEAGLE header, plain or gzipped, as
 * parse_beagle_samples, src/Utils.cpp:407-424) "name\tname\t" and row i of U2
 * as Eigen prints a row: 6 significant digits, right-aligned to the row's
 * widest value, separated by spaces. */
int pcaone_write_eigvecs2(const char* fileout, const char* filebeagle, uint32_t n, const double* U2);

/* ---- LD r^2 (run_ld_stuff --print-r2 on a .residuals matrix) ------------ */
/* divide_pos_by_window (src/LD.cpp:125-139) from the positions of nsnps SNPs
 * and each chromosome's last SNP index (0-based, increasing, the last one
 * nsnps - 1; SNPld::end_pos, src/LD.hpp:8-15).  ws and we (capacity nsnps)
 * receive each window's lead SNP and its size (lead included); *nwin the
 * number of windows.  Host only. */
int pcaone_ld_windows(const int32_t* pos, uint32_t nsnps, const int32_t* chr_end, uint32_t nchr, int32_t ld_bp,
                      int32_t* ws, int32_t* we, uint32_t* nwin);
/* run_ld_stuff's --USV branch for PLINK input (src/Main.cpp:64-75,
 * src/LD.cpp:460-466): after pcaone_gpu_load_bed, the LD matrix becomes
 * (I - U U^T) G with G = BED2GENO - F (missing 0, FileBed::read_all,
 * src/FilePlink.cpp:106-114) and U the nsamples x k matrix read_usv returns
 * for the .eigvecs (src/Utils.cpp:255-285; column-major, 1 <= k <= 64: the
 * caller passes that Mat2D's data as it is).  It stays in HBM in FP64,
 * not re-centred, and the following pcaone_gpu_ld_r2 calls run on it (until
 * the next load). */
int pcaone_gpu_ld_usv(pcaone_gpu* h, const double* U, uint32_t k);
/* ld_r2_big (src/LD.cpp:419-442) on the matrix loaded by
 * pcaone_gpu_load_dense_f32 (centred per feature as FileBin::read_all), or on
 * the --USV-adjusted genotypes of pcaone_gpu_ld_usv: r^2 of the pairs (ws[w], ws[w] + j), j = 1 .. we[w] - 1, window after
 * window, the line order of the reference's .ld.gz.  With r2 == NULL only
 * *npairs is set; otherwise r2 must hold capacity >= npairs doubles. */
int pcaone_gpu_ld_r2(pcaone_gpu* h, const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2,
                     uint64_t capacity, uint64_t* npairs);

/* LD r^2 straight from the genotypes of pcaone_gpu_load_bed, as the
 * reference's two-step route gives it: a run with --ld --ld-stats 1 writes the
 * centred genotypes BED2GENO - F (missing 0, unstandardised, each SNP
 * re-centred by its mean; Data::write_residuals, src/Data.cpp:196-221) to
 * .residuals, and -B .residuals --print-r2 runs ld_r2_big on them
 * (src/LD.cpp:419-442).  Here no matrix is materialised: each window pair's
 * integer sums over samples (S_yy = sum y_a y_b, S_yw = sum y_a w_b,
 * S_wy = sum w_a y_b, S_ww = sum w_a w_b; y the A1 dosage 0/1/2 with missing 0,
 * w the non-missing mask) come exactly from int8 MFMA with int32
 * accumulation, and r^2 follows from them in FP64.  Pairs and their order as
 * for pcaone_gpu_ld_r2 (windows from pcaone_ld_windows).  With r2 == NULL only
 * *npairs is set.  sums (optional, npairs x 4 int32, row per pair in the
 * order above) receives the integer sums.  tile_config (optional): on input a
 * tile-shape index >= 0 to use, or < 0 / NULL to autotune (every candidate
 * timed on a prefix of the windows, the choice kept per window-size class of
 * the handle); on return the index used.  A monomorphic or all-missing SNP
 * has r^2 NaN, as the reference's 0 * (1 / 0). */
int pcaone_gpu_ld_r2_geno(pcaone_gpu* h, const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2,
                          uint64_t capacity, uint64_t* npairs, int32_t* sums, int* tile_config);

/* Allocation-free variant of pcaone_gpu_ld_r2_geno.  It has no counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It allocates no device memory: the window metadata, the tile
 * list and the r^2 / sums staging live in one fixed array of
 * PCAONE_LD_SCRATCH_BYTES (32 MiB) per device, static device memory held from
 * the library's load until the process ends.  A call whose windows need more
 * (about 8 bytes per pair, 24 with sums, plus 4 per SNP from the first to the
 * last lead) returns PCAONE_EINVAL: pass the windows in batches (leads
 * increasing within a batch; results equal the full call's pairs of those
 * windows).  Needs the handle's LD layout, built by one pcaone_gpu_ld_r2_geno
 * call after the load.  tile_config < 0 takes the shape the handle tuned for
 * this window-size class, or the 64 x 64 default; no timing runs.  All calls
 * on one device share its scratch and each leaves its contents behind: calls
 * that overlap on one device race. */
#define PCAONE_LD_SCRATCH_BYTES (32ull << 20)
int pcaone_gpu_ld_r2_geno_noalloc(pcaone_gpu* h, const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2,
                                  uint64_t capacity, uint64_t* npairs, int32_t* sums, int tile_config);

/* ---- IRAM: the Spectra op and its driver ----------------------------------
 * Prepares the single-column Gram mat-vec on the loaded matrix (identity
 * SNP order; standardize / ploidy as for the RSVD operator).  Re-initialises
 * the RSVD operator of the handle. */
int pcaone_gpu_gram_init(pcaone_gpu* h, int precision, int standardize, int ploidy);
/* perform_op(x_in, y_out) (src/Arnoldi.hpp:18; Spectra op concept,
 * external/Spectra/contrib/PartialSVDSolver.h:16-33): side 0 computes
 * y = X X^T x (x, y of length nsamples; ArnoldiOpData and SVDWideMatOp),
 * side 1 computes y = X^T X x (length nfeatures; SVDTallMatOp).  X is the
 * standardised genotype matrix or the centred dense matrix. */
int pcaone_gpu_gram_matvec(pcaone_gpu* h, int side, const double* x, double* y);
/* run_pca_with_arnoldi in-core (src/Arnoldi.cpp:57-68): top-k singular
 * triplets by a restarted FP64 Lanczos on X^T X (nsamples > nfeatures) or
 * X X^T, ncv basis vectors, at most maxit restarts, Ritz residual tolerance
 * tol (--ncv, --imaxiter, --itol).  U (nsamples x k), S (k), V (nfeatures x k)
 * column-major; nconv = converged pairs, nops = mat-vec count.  Fewer than k
 * converged pairs is an error (PCAONE_ENUM, "the nconv is not equal to k.",
 * src/Arnoldi.cpp:64); so is a zero singular value among the top k. */
int pcaone_gpu_iram(pcaone_gpu* h, int k, int ncv, int maxit, double tol, double* U, double* S, double* V,
                    int* nconv, int* nops);

/* Host-only view of the winSVD window schedule of one computeGandH call
 * (in core src/Halko.cpp:159-223; out of core src/Halko.cpp:227-269, with
 * nblocks blocks and the band factor; nblocks = 0: in core, bands windows):
 * `bandsize` is the operator state carried between calls (reset at pi == 0).
 * For each window b: acc[b] in {1, 2} (H1 or H2), update[b] (Omega refreshed
 * after b), zero[b] (accumulator cleared by that refresh, 0 if none);
 * *reset_both = both accumulators cleared first. */
int pcaone_fancy_plan(int bands, uint32_t nblocks, int band_factor, int pi, uint64_t* bandsize, int* acc,
                      int* update, int* zero, int* reset_both);

/* ---- out of core (-m) ---------------------------------------------------- */
/* Data::prepare's block plan (src/Data.cpp:27-71) for -m memory_gb: svd 0 IRAM,
 * 1 Halko, 2 winSVD (then bands = -w).  Returns the block length, the number of
 * blocks and the winSVD band factor; a plan of one block is the reference's
 * error (PCAONE_EINVAL).  Host only. */
int pcaone_ooc_plan(uint32_t nsamples, uint32_t nsnps, int svd, int k, int oversamples, int bands, double memory_gb,
                    uint32_t* blocksize, uint32_t* nblocks, uint32_t* band_factor);
/* The SNP order permute_plink gives the out-of-core winSVD
 * (src/FilePlink.cpp:298-405, --buffer buffer_gb, nbands = -w): perm[ib] = the
 * input SNP at position ib, the `perm` to pass pcaone_gpu_rsvd_init in place of
 * the in-core shuffle.  Host only; no file is written. */
int pcaone_plink_perm(uint32_t nsamples, uint32_t nsnps, uint32_t buffer_gb, uint32_t nbands, int32_t* perm);

/* The .ld.gz file of --print-r2: ld_r2_big / ld_r2_small (src/LD.cpp:388-442)
 * write a header line "CHR_A\tBP_A\tSNP_A\tCHR_B\tBP_B\tSNP_B\tR2", then one line
 * per pair of the windows (ws, we as for pcaone_ld_windows, r2 in the order of
 * pcaone_gpu_ld_r2 / pcaone_gpu_ld_r2_geno): the .bim fields (chromosome, bp,
 * id) of both SNPs and std::to_string(r2).  Same text; the file is a series of
 * gzip members (level: zlib's -1..9; the reference uses -1, the default)
 * formatted and compressed by nthreads host threads.  The .bim must hold at
 * least the SNPs the windows reach; npairs must equal the windows' pair count.
 * Host only. */
int pcaone_write_ld_gz(const char* fileout, const char* filebim, const int32_t* ws, const int32_t* we, uint32_t nwin,
                       const double* r2, uint64_t npairs, int level, int nthreads);

/* ---- output files --------------------------------------------------------
 * Data::write_eigs_files (src/Data.cpp:177-194), host only: prefix.sigvals
 * ("#N,M" then S), prefix.eigvals (S^2 / nsnps * ploidy), prefix.eigvecs (U,
 * nsamples x k) and, when V != NULL, prefix.loadings (V, nsnps x k); U and V
 * column-major; 6 significant digits, tab-separated columns, one row per
 * line, as Eigen::IOFormat(6, DontAlignCols, "\t", "\n"). */
int pcaone_write_eigs(const char* prefix, uint32_t nsamples, uint32_t nsnps, int k, const double* S,
                      const double* U, const double* V, int ploidy);

/* ---- device timing ------------------------------------------------------
 * When on, every kernel the handle launches is bracketed by HIP events on the
 * handle's stream.  Stats are grouped by kernel family name. */
int pcaone_gpu_set_profiling(pcaone_gpu* h, int on);
int pcaone_gpu_reset_stats(pcaone_gpu* h);
/* Returns the number of families when idx < 0; else fills entry idx. */
int pcaone_gpu_kernel_stats(pcaone_gpu* h, int idx, char* name, size_t name_len, int64_t* count,
                            double* total_ms);
int pcaone_gpu_synchronize(pcaone_gpu* h);

/* ---- synthetic data (host only, no GPU needed) ---------------------------
 * Balding-Nichols genotypes straight into a .bed payload (nsnps rows of
 * (nsamples+3)>>2 bytes): ancestral p ~ U(pmin, pmax), population
 * frequencies ~ Beta(p(1-fst)/fst, (1-p)(1-fst)/fst), genotypes ~
 * Binomial(2, p_k), each genotype missing with probability miss_rate,
 * monomorphic SNPs redrawn; sample j is in population floor(j*npops/N).
 * p_anc (optional, nsnps) receives the ancestral frequencies. */
int pcaone_sim_bed(uint64_t seed, uint32_t nsamples, uint32_t nsnps, uint32_t npops, double fst, double pmin,
                   double pmax, double miss_rate, uint8_t* payload, double* p_anc, int nthreads);

/* Genotype likelihoods (host only), BEAGLE order: nsites x nsamples x 3
 * float32 (GL of 0, 1, 2 copies of the modelled allele).  Genotypes as in
 * pcaone_sim_bed, depth ~ Poisson(depth), per-read error err; each GL triple
 * is normalised to sum 1 (depth 0 gives 1/3 each). */
int pcaone_sim_gl(uint64_t seed, uint32_t nsamples, uint32_t nsites, uint32_t npops, double fst, double pmin,
                  double pmax, double depth, double err, float* out, int nthreads);
/* A BGEN file (host only) of Balding-Nichols genotypes as in pcaone_sim_bed,
 * written as genotype probabilities: each call certain with probability
 * 0.6, otherwise spread (the true genotype keeps a weight ~ U(0.5, 1), the
 * rest split at random), quantised to `bits` per probability so the stored
 * values sum to 2^bits - 1 (layout 1: three 16-bit values scaled by 32768).
 * Variant j's first allele is the modelled one for even j and the other for
 * odd j.  layout 1 or 2; compression 0 none, 1 zlib, 2 zstd; bits 1-32
 * (layout 2); missing samples (probability miss_rate) are flagged in the
 * ploidy byte with zero probabilities (layout 1: all zero); pmin_declared 1
 * declares a ploidy range 1-2 (every sample still diploid); sample_ids adds
 * the sample identifier block. */
int pcaone_sim_bgen(const char* path, uint64_t seed, uint32_t nsamples, uint32_t nsnps, uint32_t npops, double fst,
                    double pmin, double pmax, double miss_rate, int layout, int compression, int bits,
                    int pmin_declared, int sample_ids, int nthreads);
/* scRNA-like dense matrix (host only): ngenes rows of ncells float32 in
 * `out`.  Gene means base * LogNormal(0, sdlog), cell size factors
 * LogNormal(0, 0.3), counts ~ NegBinomial(mean, dispersion phi) drawn as
 * Gamma-Poisson, value = log1p(1e6 * count / library size). */
int pcaone_sim_counts(uint64_t seed, uint32_t ncells, uint32_t ngenes, double base, double sdlog, double phi,
                      float* out, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* PCAONE_GPU_H */
