// extern "C" boundary: every entry point catches, records the message on the
// handle and returns a negative status (the reference's cao.error throws
// std::runtime_error, src/Logger.hpp:96-105; the C++ wrapper on the other side
// of this ABI turns the status back into an exception).
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "../../include/pcaone_gpu.h"
#include "engine.hpp"

struct pcaone_gpu {
  pcaone::Engine* eng = nullptr;
  std::string err;
};

namespace {

template <class F>
int guarded(pcaone_gpu* h, F&& f) {
  if (!h || !h->eng) return PCAONE_EINVAL;
  try {
    f(*h->eng);
    h->err.clear();
    return PCAONE_OK;
  } catch (const pcaone::Error& e) {
    h->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    h->err = "host allocation failed";
    return PCAONE_EINVAL;
  } catch (const std::exception& e) {
    h->err = e.what();
    return PCAONE_EINVAL;
  }
}

}  // namespace

extern "C" {

int pcaone_gpu_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int pcaone_gpu_host_alloc(size_t bytes, void** out) {
  if (!out) return PCAONE_EINVAL;
  *out = nullptr;
  if (bytes == 0) return 0;
  return hipHostMalloc(out, bytes, hipHostMallocDefault) == hipSuccess ? 0 : -2;
}

int pcaone_gpu_host_free(void* p) {
  if (!p) return 0;
  return hipHostFree(p) == hipSuccess ? 0 : -2;
}

int pcaone_gpu_create(int device, pcaone_gpu** out) {
  if (!out) return PCAONE_EINVAL;
  *out = nullptr;
  auto* h = new (std::nothrow) pcaone_gpu();
  if (!h) return PCAONE_EINVAL;
  try {
    h->eng = new pcaone::Engine(device);
  } catch (const pcaone::Error& e) {
    const int c = e.code;
    delete h;
    return c;
  } catch (...) {
    delete h;
    return PCAONE_EHIP;
  }
  *out = h;
  return PCAONE_OK;
}

void pcaone_gpu_destroy(pcaone_gpu* h) {
  if (!h) return;
  delete h->eng;
  delete h;
}

const char* pcaone_gpu_last_error(const pcaone_gpu* h) { return h ? h->err.c_str() : "null handle"; }

int pcaone_gpu_load_bed(pcaone_gpu* h, const uint8_t* payload, uint32_t nsamples, uint32_t nsnps) {
  return guarded(h, [&](pcaone::Engine& e) { e.load_bed(payload, nsamples, nsnps); });
}

int pcaone_gpu_load_dense_f32(pcaone_gpu* h, const float* featmajor, uint32_t nsamples, uint32_t nfeatures) {
  return guarded(h, [&](pcaone::Engine& e) { e.load_dense(featmajor, nsamples, nfeatures); });
}

int pcaone_gpu_feature_means(pcaone_gpu* h, double* mu) {
  return guarded(h, [&](pcaone::Engine& e) {
    PCAONE_REQUIRE(mu != nullptr, PCAONE_EINVAL, "feature_means: null output");
    e.feature_means(mu);
  });
}

int pcaone_gpu_allele_freq(pcaone_gpu* h, double* F, int64_t* counts) {
  return guarded(h, [&](pcaone::Engine& e) { e.allele_freq(F, counts); });
}

int pcaone_gpu_load_bed_file(pcaone_gpu* h, const char* filebed, uint32_t nsamples, uint32_t nsnps,
                             int out_of_core) {
  return guarded(h, [&](pcaone::Engine& e) { e.load_bed_file(filebed, nsamples, nsnps, out_of_core != 0); });
}

int pcaone_gpu_rsvd_init(pcaone_gpu* h, const pcaone_rsvd_params* p, const double* omega, const int32_t* perm) {
  return guarded(h, [&](pcaone::Engine& e) {
    PCAONE_REQUIRE(p != nullptr, PCAONE_EINVAL, "rsvd_init: null params");
    pcaone::RsvdParams q;
    q.variant = p->variant;
    q.k = p->k;
    q.os = p->oversamples;
    q.bands = p->bands;
    q.precision = p->precision;
    q.seed = p->seed;
    q.rand = p->rand;
    q.perm = p->perm;
    q.standardize = p->standardize;
    q.ploidy = p->ploidy;
    q.blocksize = p->blocksize;
    q.band_factor = p->band_factor > 0 ? p->band_factor : 1;
    e.rsvd_init(q, omega, perm);
  });
}

int pcaone_gpu_rsvd_dims(const pcaone_gpu* h, int64_t* rows, int64_t* cols, int64_t* size) {
  return guarded(const_cast<pcaone_gpu*>(h), [&](pcaone::Engine& e) { e.dims(rows, cols, size); });
}

int pcaone_gpu_rsvd_perm(const pcaone_gpu* h, int32_t* perm) {
  return guarded(const_cast<pcaone_gpu*>(h), [&](pcaone::Engine& e) {
    PCAONE_REQUIRE(perm != nullptr, PCAONE_EINVAL, "rsvd_perm: null output");
    e.rsvd_perm(perm);
  });
}

int pcaone_gpu_rsvd_gandh(pcaone_gpu* h, int pi, double* G, double* H) {
  return guarded(h, [&](pcaone::Engine& e) { e.rsvd_gandh(pi, G, H); });
}

int pcaone_gpu_rsvd_usv(pcaone_gpu* h, int maxp, double tol, double* U, double* S, double* V, int* epochs,
                        double* diffs) {
  return guarded(h, [&](pcaone::Engine& e) { e.rsvd_usv(maxp, tol, U, S, V, epochs, diffs); });
}

int pcaone_gpu_rsvd_set_flags(pcaone_gpu* h, int update, int standardize) {
  return guarded(h, [&](pcaone::Engine& e) { e.set_flags(update != 0, standardize != 0); });
}

int pcaone_gpu_rsvd_em(pcaone_gpu* h, int maxp, double tol, int maxiter, double tolem, double* U, double* S,
                       double* V, int* iters, double* diffs) {
  return guarded(h, [&](pcaone::Engine& e) { e.rsvd_em(maxp, tol, maxiter, tolem, U, S, V, iters, diffs); });
}

int pcaone_gpu_load_gl(pcaone_gpu* h, const float* gl, uint32_t nsamples, uint32_t nsites, int per_site) {
  return guarded(h, [&](pcaone::Engine& e) { e.load_gl(gl, nsamples, nsites, per_site); });
}

int pcaone_gpu_gl_from_bed(pcaone_gpu* h) {
  return guarded(h, [&](pcaone::Engine& e) { e.gl_from_bed(); });
}

int pcaone_gpu_gl_em_maf(pcaone_gpu* h, int maxiter, double tol, double* F, int* iters) {
  return guarded(h, [&](pcaone::Engine& e) { e.gl_em_maf(maxiter, tol, F, iters); });
}

int pcaone_gpu_pcangsd_update(pcaone_gpu* h, const double* U, const double* S, const double* V, int k,
                              int standardize, float* E, double* Dc) {
  return guarded(h, [&](pcaone::Engine& e) { e.pcangsd_update(U, S, V, k, standardize, E, Dc); });
}

int pcaone_gpu_pcangsd(pcaone_gpu* h, int maxp, double tol, int maxiter, double tolem, double* U, double* S,
                       double* V, int* iters, double* diffs, double* Dc, double* C) {
  return guarded(h, [&](pcaone::Engine& e) { e.pcangsd(maxp, tol, maxiter, tolem, U, S, V, iters, diffs, Dc, C); });
}

int pcaone_gpu_pcangsd_iram(pcaone_gpu* h, int k, int ncv, int maxit, double tol, int maxiter, double tolem,
                            double* U, double* S, double* V, int* iters, double* diffs, double* Dc, double* C) {
  return guarded(h, [&](pcaone::Engine& e) {
    e.pcangsd_iram(k, ncv, maxit, tol, maxiter, tolem, U, S, V, iters, diffs, Dc, C);
  });
}

int pcaone_ld_windows(const int32_t* pos, uint32_t nsnps, const int32_t* chr_end, uint32_t nchr, int32_t ld_bp,
                      int32_t* ws, int32_t* we, uint32_t* nwin) {
  try {
    const uint32_t n = pcaone::ld_windows(pos, nsnps, chr_end, nchr, ld_bp, ws, we);
    if (nwin) *nwin = n;
    return PCAONE_OK;
  } catch (const pcaone::Error& e) {
    return e.code < 0 ? e.code : PCAONE_EINVAL;
  } catch (...) {
    return PCAONE_EINVAL;
  }
}

int pcaone_write_ld_gz(const char* fileout, const char* filebim, const int32_t* ws, const int32_t* we, uint32_t nwin,
                       const double* r2, uint64_t npairs, int level, int nthreads) {
  try {
    pcaone::write_ld_gz(fileout, filebim, ws, we, nwin, r2, npairs, level, nthreads);
    return PCAONE_OK;
  } catch (const pcaone::Error& e) {
    return e.code < 0 ? e.code : PCAONE_EINVAL;
  } catch (...) {
    return PCAONE_EINVAL;
  }
}

int pcaone_gpu_ld_usv(pcaone_gpu* h, const double* U, uint32_t k) {
  return guarded(h, [&](pcaone::Engine& e) { e.ld_usv(U, k); });
}

int pcaone_gpu_ld_r2(pcaone_gpu* h, const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2,
                     uint64_t capacity, uint64_t* npairs) {
  return guarded(h, [&](pcaone::Engine& e) {
    const uint64_t n = e.ld_r2(ws, we, nwin, r2, capacity);
    if (npairs) *npairs = n;
  });
}

int pcaone_gpu_ld_r2_geno(pcaone_gpu* h, const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2,
                          uint64_t capacity, uint64_t* npairs, int32_t* sums, int* tile_config) {
  return guarded(h, [&](pcaone::Engine& e) {
    const uint64_t n = e.ld_r2_geno(ws, we, nwin, r2, capacity, sums, tile_config);
    if (npairs) *npairs = n;
  });
}

int pcaone_gpu_ld_r2_geno_noalloc(pcaone_gpu* h, const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2,
                                  uint64_t capacity, uint64_t* npairs, int32_t* sums, int tile_config) {
  static_assert(PCAONE_LD_SCRATCH_BYTES == pcaone::kLdScratchBytes, "scratch size of the header and the library");
  return guarded(h, [&](pcaone::Engine& e) {
    const uint64_t n = e.ld_r2_geno_noalloc(ws, we, nwin, r2, capacity, sums, tile_config);
    if (npairs) *npairs = n;
  });
}

int pcaone_gpu_gram_init(pcaone_gpu* h, int precision, int standardize, int ploidy) {
  return guarded(h, [&](pcaone::Engine& e) { e.gram_init(precision, standardize != 0, ploidy); });
}

int pcaone_gpu_gram_matvec(pcaone_gpu* h, int side, const double* x, double* y) {
  return guarded(h, [&](pcaone::Engine& e) { e.gram_matvec(side, x, y); });
}

int pcaone_gpu_iram(pcaone_gpu* h, int k, int ncv, int maxit, double tol, double* U, double* S, double* V,
                    int* nconv, int* nops) {
  return guarded(h, [&](pcaone::Engine& e) { e.iram(k, ncv, maxit, tol, U, S, V, nconv, nops); });
}

int pcaone_fancy_plan(int bands, uint32_t nblocks, int band_factor, int pi, uint64_t* bandsize, int* acc,
                      int* update, int* zero, int* reset_both) {
  if (bands < 4 || bands % 2 || pi < 0 || !bandsize || !acc || !update || !zero) return PCAONE_EINVAL;
  if (nblocks > 0 && (nblocks < 2 || band_factor < 1)) return PCAONE_EINVAL;
  const pcaone::FancyPlan pl = pcaone::fancy_plan(bands, pi, *bandsize, nblocks, nblocks ? (uint32_t)band_factor : 1);
  for (size_t b = 0; b < pl.steps.size(); ++b) {
    acc[b] = pl.steps[b].acc;
    update[b] = pl.steps[b].update ? 1 : 0;
    zero[b] = pl.steps[b].zero;
  }
  if (reset_both) *reset_both = pl.reset_both ? 1 : 0;
  return PCAONE_OK;
}

int pcaone_ooc_plan(uint32_t nsamples, uint32_t nsnps, int svd, int k, int oversamples, int bands, double memory_gb,
                    uint32_t* blocksize, uint32_t* nblocks, uint32_t* band_factor) {
  if (!blocksize || !nblocks || !band_factor) return PCAONE_EINVAL;
  try {
    pcaone::ooc_plan(nsamples, nsnps, svd, k, oversamples, bands, memory_gb, *blocksize, *nblocks, *band_factor);
    return PCAONE_OK;
  } catch (const pcaone::Error& e) {
    return e.code < 0 ? e.code : PCAONE_EINVAL;
  } catch (...) {
    return PCAONE_EINVAL;
  }
}

int pcaone_plink_perm(uint32_t nsamples, uint32_t nsnps, uint32_t buffer_gb, uint32_t nbands, int32_t* perm) {
  if (!perm) return PCAONE_EINVAL;
  try {
    pcaone::plink_perm(nsamples, nsnps, buffer_gb, nbands, perm);
    return PCAONE_OK;
  } catch (const pcaone::Error& e) {
    return e.code < 0 ? e.code : PCAONE_EINVAL;
  } catch (...) {
    return PCAONE_EINVAL;
  }
}

int pcaone_write_eigs(const char* prefix, uint32_t nsamples, uint32_t nsnps, int k, const double* S,
                      const double* U, const double* V, int ploidy) {
  if (!prefix || !S || !U || k < 1 || nsamples == 0 || nsnps == 0) return PCAONE_EINVAL;
  const std::string p(prefix);
  std::ofstream outs(p + ".sigvals"), oute(p + ".eigvals"), outu(p + ".eigvecs");
  if (!outs || !oute || !outu) return PCAONE_EINVAL;
  for (auto* o : {&outs, &oute, &outu}) o->precision(6);
  outs << '#' << nsamples << ',' << nsnps << '\n';
  for (int c = 0; c < k; ++c) outs << S[c] << '\n';
  for (int c = 0; c < k; ++c) oute << S[c] * S[c] / (double)nsnps * ploidy << '\n';
  auto rows = [k](std::ofstream& o, const double* A, uint64_t n) {
    for (uint64_t r = 0; r < n; ++r)
      for (int c = 0; c < k; ++c) o << A[(uint64_t)c * n + r] << (c + 1 < k ? '\t' : '\n');
  };
  rows(outu, U, nsamples);
  if (V) {
    std::ofstream outv(p + ".loadings");
    if (!outv) return PCAONE_EINVAL;
    outv.precision(6);
    rows(outv, V, nsnps);
  }
  return PCAONE_OK;
}

int pcaone_gpu_load_csv(pcaone_gpu* h, const char* path, int scale, double scale_factor, int nthreads,
                        uint32_t* nsamples, uint32_t* nfeatures) {
  return guarded(h, [&](pcaone::Engine& e) {
    e.load_csv(path, scale, scale_factor, nthreads);
    e.loaded_dims(nsamples, nfeatures);
  });
}

int pcaone_gpu_load_bgen(pcaone_gpu* h, const char* path, int mode, double maf, int standardize, int ploidy,
                         int nthreads, uint32_t* nsamples, uint32_t* nsnps, double* F, uint8_t* minor) {
  return guarded(h, [&](pcaone::Engine& e) {
    e.load_bgen(path, mode, maf, standardize, ploidy, nthreads, F, minor);
    e.loaded_dims(nsamples, nsnps);
  });
}

int pcaone_gpu_project_gl(pcaone_gpu* h, const char* filebeagle, const char* filembim, const double* S,
                          const double* Vref, uint32_t k, uint32_t ref_nsnps, int standardize, int maxiter,
                          double tolem, double* U, double* VS, uint32_t* noverlap, int* iters, double* diffs) {
  return guarded(h, [&](pcaone::Engine& e) {
    const uint32_t nm =
        e.project_gl(filebeagle, filembim, S, Vref, k, ref_nsnps, standardize, maxiter, tolem, U, VS, iters, diffs);
    if (noverlap) *noverlap = nm;
  });
}

int pcaone_gpu_project_gl_matched(pcaone_gpu* h, const int32_t* ref_idx, const uint8_t* flip, const double* F,
                                  const double* S, const double* Vref, uint32_t k, uint32_t ref_nsnps,
                                  int standardize, int maxiter, double tolem, double* U, double* VS,
                                  uint32_t* noverlap, int* iters, double* diffs) {
  return guarded(h, [&](pcaone::Engine& e) {
    const uint32_t nm = e.project_gl_matched(ref_idx, flip, F, S, Vref, k, ref_nsnps, standardize, maxiter, tolem, U,
                                             VS, iters, diffs);
    if (noverlap) *noverlap = nm;
  });
}

int pcaone_gpu_dense_rows(pcaone_gpu* h, uint32_t r0, uint32_t nrows, float* out) {
  return guarded(h, [&](pcaone::Engine& e) { e.dense_rows(r0, nrows, out); });
}

int pcaone_bgen_info(const char* path, uint32_t* nsamples, uint32_t* nvariants, int* layout, int* compression) {
  try {
    pcaone::BgenIndex ix;
    pcaone::bgen_index(path, ix);
    if (nsamples) *nsamples = ix.nsamples;
    if (nvariants) *nvariants = ix.nvariants;
    if (layout) *layout = ix.layout;
    if (compression) *compression = ix.compression;
    return PCAONE_OK;
  } catch (const pcaone::Error& e) {
    return e.code < 0 ? e.code : PCAONE_EINVAL;
  } catch (...) {
    return PCAONE_EINVAL;
  }
}

int pcaone_read_csv(const char* path, int nthreads, float* values, uint64_t capacity, uint32_t* nsamples,
                    uint64_t* nfeatures, int64_t* libsize) {
  try {
    pcaone::read_csv_host(path, nthreads, values, capacity, nsamples, nfeatures, libsize);
    return PCAONE_OK;
  } catch (const pcaone::Error& e) {
    return e.code < 0 ? e.code : PCAONE_EINVAL;
  } catch (...) {
    return PCAONE_EINVAL;
  }
}

int pcaone_gpu_project(pcaone_gpu* h, const char* filebim, const char* filembim, const double* S,
                       const double* Vref, uint32_t k, uint32_t ref_nsnps, int method, int standardize, double* U,
                       double* W, uint32_t* noverlap) {
  return guarded(h, [&](pcaone::Engine& e) {
    const uint32_t nm = e.project(filebim, filembim, S, Vref, k, ref_nsnps, method, standardize, U, W);
    if (noverlap) *noverlap = nm;
  });
}

int pcaone_gpu_inbreeding(pcaone_gpu* h, const double* U, const double* S, const double* V, uint32_t k,
                          const double* F, int maxiter, double tolem, double* coef, double* lrt, double* hwe_p,
                          int* iters) {
  return guarded(h, [&](pcaone::Engine& e) {
    const int it = e.inbreeding(U, S, V, k, F, maxiter, tolem, coef, lrt, hwe_p);
    if (iters) *iters = it;
  });
}

int pcaone_write_hwe(const char* fileout, const char* filebim, const double* hwe_p, const double* lrt,
                     const double* coef, uint32_t nsnps) {
  try {
    pcaone::write_hwe(fileout, filebim, hwe_p, lrt, coef, nsnps);
    return PCAONE_OK;
  } catch (const pcaone::Error& e) {
    return e.code < 0 ? e.code : PCAONE_EINVAL;
  } catch (...) {
    return PCAONE_EINVAL;
  }
}

int pcaone_gpu_write_residuals(pcaone_gpu* h, const char* fileout, const double* U, const double* S,
                               const double* V, uint32_t k, int ld_stats) {
  return guarded(h, [&](pcaone::Engine& e) { e.write_residuals(fileout, U, S, V, k, ld_stats); });
}

int pcaone_write_mbim(const char* fileout, const char* filebim, const double* F, uint32_t nsnps) {
  try {
    pcaone::write_mbim(fileout, filebim, F, nsnps);
    return PCAONE_OK;
  } catch (const pcaone::Error& e) {
    return e.code < 0 ? e.code : PCAONE_EINVAL;
  } catch (...) {
    return PCAONE_EINVAL;
  }
}

int pcaone_gpu_ld_prune(pcaone_gpu* h, const char* filebim, int ld_bp, double r2_tol, const char* prefix) {
  return guarded(h, [&](pcaone::Engine& e) { e.ld_prune(filebim, ld_bp, r2_tol, prefix); });
}

int pcaone_gpu_ld_clump(pcaone_gpu* h, const char* fileassoc, const char* colnames, const char* filebim, int clump_bp,
                        double clump_r2, double clump_p1, double clump_p2, const char* fileout) {
  return guarded(h, [&](pcaone::Engine& e) {
    e.ld_clump(fileassoc, colnames, filebim, clump_bp, clump_r2, clump_p1, clump_p2, fileout);
  });
}

int pcaone_gpu_grm_eig(pcaone_gpu* h, double* evals, double* U2) {
  return guarded(h, [&](pcaone::Engine& e) { e.grm_eig(evals, U2); });
}

extern "C++" {
namespace {
// parse_beagle_samples (src/Utils.cpp:407-424): the header line's tokens (split
// at runs of tab, space, newline); after the first, every third one is a sample
std::vector<std::string> beagle_samples(const char* path) {
  gzFile fp = gzopen(path, "r");
  if (!fp) throw pcaone::Error(PCAONE_EINVAL, std::string("can not open ") + path);
  std::string line;
  char buf[1 << 16];
  while (gzgets(fp, buf, sizeof(buf))) {
    line += buf;
    if (!line.empty() && line.back() == '\n') break;
  }
  gzclose(fp);
  std::vector<std::string> tok;
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && (line[i] == '\t' || line[i] == ' ' || line[i] == '\n')) ++i;
    size_t j = i;
    while (j < line.size() && !(line[j] == '\t' || line[j] == ' ' || line[j] == '\n')) ++j;
    if (j > i) tok.emplace_back(line, i, j - i);
    i = j;
  }
  if (tok.empty() || tok.size() % 3)
    throw pcaone::Error(PCAONE_EINVAL, "Number of columns should be a multiple of 3.");
  std::vector<std::string> out;
  for (size_t c = 3; c < tok.size(); c += 3) out.push_back(tok[c]);
  return out;
}
}  // namespace
}  // extern "C++"

int pcaone_write_eigvecs2(const char* fileout, const char* filebeagle, uint32_t n, const double* U2) {
  try {
    if (!fileout || !filebeagle || !U2 || n == 0) return PCAONE_EINVAL;
    const std::vector<std::string> names = beagle_samples(filebeagle);
    if (names.size() != n) return PCAONE_EINVAL;
    std::ofstream out(fileout);
    if (!out) return PCAONE_EINVAL;
    std::string line = "#FID\tIID";
    for (uint32_t j = 0; j < n; ++j) line += "\tPC" + std::to_string(j + 1);
    out << line << '\n';
    // Eigen's default row print: "%g"-style 6 significant digits (the stream's
    // precision), every value right-aligned to the row's widest, " " between
    std::vector<std::string> v(n);
    char num[64];
    for (uint32_t i = 0; i < n; ++i) {
      size_t w = 0;
      for (uint32_t j = 0; j < n; ++j) {
        const int len = std::snprintf(num, sizeof(num), "%.6g", U2[(uint64_t)i * n + j]);
        v[j].assign(num, (size_t)len);
        w = std::max(w, v[j].size());
      }
      line = names[i] + "\t" + names[i] + "\t";
      for (uint32_t j = 0; j < n; ++j) {
        if (j) line += ' ';
        line.append(w - v[j].size(), ' ');
        line += v[j];
      }
      out << line << '\n';
    }
    return out ? PCAONE_OK : PCAONE_EINVAL;
  } catch (const pcaone::Error& e) {
    return e.code < 0 ? e.code : PCAONE_EINVAL;
  } catch (...) {
    return PCAONE_EINVAL;
  }
}

int pcaone_gpu_set_profiling(pcaone_gpu* h, int on) {
  return guarded(h, [&](pcaone::Engine& e) {
    e.sync();
    e.timer.resolve();
    e.timer.on = on != 0;
  });
}

int pcaone_gpu_reset_stats(pcaone_gpu* h) {
  return guarded(h, [&](pcaone::Engine& e) {
    e.sync();
    e.timer.clear();
  });
}

int pcaone_gpu_kernel_stats(pcaone_gpu* h, int idx, char* name, size_t name_len, int64_t* count,
                            double* total_ms) {
  int ret = 0;
  const int st = guarded(h, [&](pcaone::Engine& e) {
    e.sync();
    e.timer.resolve();
    if (idx < 0) {
      ret = (int)e.timer.agg.size();
      return;
    }
    PCAONE_REQUIRE((size_t)idx < e.timer.agg.size(), PCAONE_EINVAL, "kernel_stats: index out of range");
    auto it = e.timer.agg.begin();
    std::advance(it, idx);
    if (name && name_len) {
      std::strncpy(name, it->first.c_str(), name_len - 1);
      name[name_len - 1] = '\0';
    }
    if (count) *count = it->second.first;
    if (total_ms) *total_ms = it->second.second;
  });
  return st < 0 ? st : ret;
}

int pcaone_gpu_synchronize(pcaone_gpu* h) {
  return guarded(h, [&](pcaone::Engine& e) { e.sync(); });
}

}  // extern "C"
