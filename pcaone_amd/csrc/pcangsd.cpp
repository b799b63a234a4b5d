// PCAngsd on genotype likelihoods (BEAGLE input): the likelihoods stay in HBM
// as FP32, emMAF_with_GL runs on the device, and the E matrix the RSVD
// operator contracts is regenerated in place (laid-out rows + windowed
// transpose) at every EM iteration.  The operator runs on E exactly as on a
// dense FP32 matrix, without centring (setFlags(false, false)).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"

namespace pcaone {

// FileBeagle::read_all keeps the first two likelihoods of each sample
// (src/FileBeagle.cpp:99-100, 160-180); gl is [site][sample][per_site].
void Engine::load_gl(const float* gl, uint32_t nsamples, uint32_t nsites, int per_site) {
  PCAONE_REQUIRE(gl && nsamples > 0 && nsites > 0, -1, "load_gl: empty input");
  PCAONE_REQUIRE(per_site == 2 || per_site == 3, -1, "load_gl: per_site must be 2 or 3");
  PCAONE_HIP(hipSetDevice(device_));
  have_op_ = have_mv_ = false;
  N_ = nsamples;
  M_ = nsites;
  std::vector<float> two;
  const float* src = gl;
  if (per_site == 3) {
    two.resize((size_t)N_ * M_ * 2);
    for (size_t e = 0; e < (size_t)N_ * M_; ++e) {
      two[2 * e] = gl[3 * e];
      two[2 * e + 1] = gl[3 * e + 1];
    }
    src = two.data();
  }
  end_stream();
  glP_.reserve((uint64_t)N_ * M_ * 2);
  PCAONE_HIP(hipMemcpyAsync(glP_.p, src, (uint64_t)N_ * M_ * 2 * sizeof(float), hipMemcpyHostToDevice, stream_));
  sync();  // `two` goes out of scope
  // the operator sees E uncentred: mean 0, F 0 for the epilogues
  mu_orig_.assign(M_, 0.0);
  F_orig_.assign(M_, 0.0);
  counts_orig_.clear();
  have_bed_ = true;
  src_ = SRC_DENSE;
  gl_mode_ = true;
  have_glF_ = gl_bed_ = have_resid_ = have_ldg_ = have_grm_ = false;
}

// --pcangsd on in-core PLINK input (src/FilePlink.cpp:68, 89-114;
// src/Data.cpp:249-270): the likelihoods are the hard calls of the loaded
// codes, F stays the genotype-count frequency (the reference's emMAF_with_GL
// call is commented out there, src/FilePlink.cpp:127-130) and the initial E
// is the centred genotype.  The operator then runs on E as for BEAGLE input.
void Engine::gl_from_bed() {
  PCAONE_REQUIRE(have_bed_ && src_ == SRC_BED, -3, "gl_from_bed: load PLINK genotypes first");
  PCAONE_HIP(hipSetDevice(device_));
  have_op_ = have_mv_ = false;
  glP_.reserve((uint64_t)N_ * M_ * 2);
  timed("gl_from_bed", [&] { launch_bed_to_gl(resident_raw("gl_from_bed"), N_, M_, glP_.p, stream_); });
  glF_.reserve(M_);
  glF_host_ = F_orig_;
  PCAONE_HIP(hipMemcpyAsync(glF_.p, glF_host_.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
  sync();
  // the operator sees E uncentred: mean 0, F 0 for the epilogues
  mu_orig_.assign(M_, 0.0);
  F_orig_.assign(M_, 0.0);
  counts_orig_.clear();
  src_ = SRC_DENSE;
  gl_mode_ = have_glF_ = gl_bed_ = true;
  have_resid_ = have_ldg_ = have_grm_ = false;
}

// emMAF_with_GL (src/Utils.cpp:501-531) from F = 0.25 (src/FileBeagle.cpp:127):
// F_j <- 1/(2N) sum_i posterior(F_j) until sqrt(sum (dF)^2 / M) < tol.
void Engine::gl_em_maf(int maxiter, double tol, double* F, int* iters) {
  PCAONE_REQUIRE(gl_mode_, -3, "gl_em_maf: no genotype likelihoods loaded");
  PCAONE_REQUIRE(maxiter >= 1, -1, "gl_em_maf: maxiter must be >= 1");
  PCAONE_HIP(hipSetDevice(device_));
  have_op_ = false;
  DevBuf<double> fa, fb, part;
  fa.reserve(M_);
  fb.reserve(M_);
  const uint32_t nb = gl_em_maf_blocks(M_);
  part.reserve(nb);
  timed("gl_maf", [&] { launch_fill_f64(fa.p, M_, 0.25, stream_); });
  std::vector<double> hp(nb);
  int it = 0;
  for (it = 1; it <= maxiter; ++it) {
    timed("gl_maf", [&] { launch_gl_em_maf(glP_.p, N_, M_, fa.p, fb.p, part.p, stream_); });
    PCAONE_HIP(hipMemcpyAsync(hp.data(), part.p, (size_t)nb * 8, hipMemcpyDeviceToHost, stream_));
    sync();
    double d2 = 0.0;
    for (uint32_t b = 0; b < nb; ++b) d2 += hp[b];
    std::swap(fa.p, fb.p);
    if (std::sqrt(d2 / M_) < tol) break;
  }
  if (it > maxiter) it = maxiter;
  glF_.reserve(M_);
  PCAONE_HIP(hipMemcpyAsync(glF_.p, fa.p, (size_t)M_ * 8, hipMemcpyDeviceToDevice, stream_));
  glF_host_.resize(M_);
  PCAONE_HIP(hipMemcpyAsync(glF_host_.data(), fa.p, (size_t)M_ * 8, hipMemcpyDeviceToHost, stream_));
  sync();
  if (F) std::memcpy(F, glF_host_.data(), (size_t)M_ * 8);
  if (iters) *iters = it;
  have_glF_ = true;
}

// E into the laid-out rows (dns_) and the windowed transpose (dnt_): the
// initial E when U == null, else the PCAngsd update from U (N x k row-major),
// S (k) and V (M x k, laid-out rows), all device pointers.  With standardize
// the rows are divided by sqrt(2F(1-F)) and Dc (device, N) receives the GRM
// diagonal sums.
void Engine::gl_fill_E(const double* U, const double* S, const double* V, uint32_t k, bool standardize,
                       double* Dc) {
  GlE e{};
  e.P = glP_.p;
  e.nsamples = N_;
  e.nsites = M_;
  e.perm = d_perm_.p;
  e.F = glF_.p;
  e.U = U;
  e.S = S;
  e.V = V;
  e.k = k;
  e.standardize = standardize ? 1 : 0;
  e.bed_init = (gl_bed_ && !U) ? 1 : 0;
  e.out = dns_.p;
  e.ld = kg_;
  e.rows_per_block = gl_rows_per_block(M_);
  const uint32_t ny = (uint32_t)ceil_div(M_, e.rows_per_block);
  if (Dc) {
    gl_part_.reserve((uint64_t)ny * N_);
    e.dpart = gl_part_.p;
  }
  timed("gl_E", [&] { launch_gl_E(e, stream_); });
  if (Dc) timed("gl_E_dc", [&] { launch_sum_rows(gl_part_.p, ny, N_, Dc, stream_); });
  timed("dense_layout", [&] {
    launch_dense_transpose(dns_.p, (uint32_t)kg_, M_, bs_, nwin_, N_, dnt_.p, (uint32_t)kw_, stream_);
  });
}

// fit_with_pi, PCAngsd branch (src/Data.cpp:249-270), or with standardize
// pcangsd_standardize_E (src/Data.cpp:301-343), from host U (N x k), S (k),
// V (M x k, original site order), all column-major.  Replaces the operator's
// matrix like the reference's in-place update.  E (optional): M x N
// site-major FP32; Dc (optional): N.
void Engine::pcangsd_update(const double* U, const double* S, const double* V, int k, int standardize, float* E,
                            double* Dc) {
  PCAONE_REQUIRE(gl_mode_ && have_op_, -3, "pcangsd_update: load likelihoods, run gl_em_maf and rsvd_init first");
  PCAONE_REQUIRE(U && S && V && k >= 1 && k <= 64, -1, "pcangsd_update: bad U / S / V / k");
  PCAONE_HIP(hipSetDevice(device_));
  std::vector<double> ur((size_t)N_ * k), vr((size_t)M_ * k);
  for (uint32_t r = 0; r < N_; ++r)
    for (int c = 0; c < k; ++c) ur[(size_t)r * k + c] = U[(size_t)c * N_ + r];
  for (uint32_t r = 0; r < M_; ++r)
    for (int c = 0; c < k; ++c) vr[(size_t)r * k + c] = V[(size_t)c * M_ + perm_[r]];
  DevBuf<double> du, ds, dv;
  du.reserve(ur.size());
  ds.reserve(k);
  dv.reserve(vr.size());
  PCAONE_HIP(hipMemcpyAsync(du.p, ur.data(), ur.size() * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(ds.p, S, (size_t)k * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dv.p, vr.data(), vr.size() * 8, hipMemcpyHostToDevice, stream_));
  gl_Dc_.reserve(N_);
  gl_fill_E(du.p, ds.p, dv.p, (uint32_t)k, standardize != 0, (standardize && Dc) ? gl_Dc_.p : nullptr);
  std::vector<float> rows;
  if (E) {
    rows.resize((size_t)M_ * kg_);
    PCAONE_HIP(hipMemcpyAsync(rows.data(), dns_.p, rows.size() * 4, hipMemcpyDeviceToHost, stream_));
  }
  if (standardize && Dc) PCAONE_HIP(hipMemcpyAsync(Dc, gl_Dc_.p, (size_t)N_ * 8, hipMemcpyDeviceToHost, stream_));
  sync();
  if (E)
    for (uint32_t r = 0; r < M_; ++r)
      std::memcpy(E + (size_t)perm_[r] * N_, rows.data() + (size_t)r * kg_, (size_t)N_ * 4);
}

namespace {

// flip_UV with ubase (src/Utils.cpp:116-130): column i of U (n x k, column-major)
// and of V (m x k) change sign when U's entry of largest magnitude (the first
// such on ties) is negative.
void flip_uv(std::vector<double>& U, std::vector<double>& V, size_t n, size_t m, int k) {
  for (int i = 0; i < k; ++i) {
    double* u = U.data() + (size_t)i * n;
    size_t x = 0;
    for (size_t r = 1; r < n; ++r)
      if (std::fabs(u[r]) > std::fabs(u[x])) x = r;
    if (u[x] < 0) {
      for (size_t r = 0; r < n; ++r) u[r] = -u[r];
      double* v = V.data() + (size_t)i * m;
      for (size_t r = 0; r < m; ++r) v[r] = -v[r];
    }
  }
}

// rmse(X, Y) (src/Utils.cpp:153-167): Y's columns take the sign nearer X
// (flip_Y), then the root mean square of X - Y.
double rmse_flip(const std::vector<double>& X, const std::vector<double>& Y, size_t m, int k) {
  double sq = 0.0;
  for (int i = 0; i < k; ++i) {
    const double* x = X.data() + (size_t)i * m;
    const double* y = Y.data() + (size_t)i * m;
    double dm = 0.0, dp = 0.0;
    for (size_t r = 0; r < m; ++r) {
      dm += std::fabs(x[r] - y[r]);
      dp += std::fabs(x[r] + y[r]);
    }
    const double s = dm > 2.0 * dp ? -1.0 : 1.0;
    for (size_t r = 0; r < m; ++r) {
      const double d = x[r] - s * y[r];
      sq += d * d;
    }
  }
  return std::sqrt(sq / ((double)m * (double)k));
}

}  // namespace

// run_pca_with_arnoldi's EM branch in core for PCAngsd (src/Arnoldi.cpp:57-104):
// IRAM on the initial E, flip_UV, then E <- fit_with_pi(U, S, V) and a new IRAM
// until rmse(V_new, V) < tolem (at most maxiter times); then
// pcangsd_standardize_E gives Dc and, when C is given, the GRM C = E E^T / M with
// diagonal Dc / M.  The operator is the Gram mat-vec of gram_init on the
// likelihoods' E (identity site order).
//
// The solver object there keeps the eigenvectors it handed out first
// (PartialSVDSolver caches m_evecs on the first matrix_U / matrix_V call and
// never drops it, external/Spectra/contrib/PartialSVDSolver.h:177-206), so
// every later solve contributes only its singular values: on the smaller Gram
// side the singular vectors stay those of the first solve, and the other side is
// the current E times them over the new singular values.  The same is done here.
void Engine::pcangsd_iram(int k, int ncv, int maxit, double tol, int maxiter, double tolem, double* U, double* S,
                          double* V, int* iters, double* diffs, double* Dc, double* C) {
  PCAONE_REQUIRE(gl_mode_ && have_mv_, -3, "pcangsd_iram: load likelihoods, run gl_em_maf and gram_init first");
  PCAONE_REQUIRE(k >= 1 && maxiter >= 0, -1, "pcangsd_iram: k must be >= 1 and maxiter >= 0");
  PCAONE_HIP(hipSetDevice(device_));
  const size_t n = N_, m = M_;
  std::vector<double> u(n * k), s(k), v(m * k), v2(m * k), rowu(n * k), rowv(m * k);
  // fit_with_pi's operands on the device: U (n x k) and V (m x k) row-major
  Ucur_.reserve(n * k);
  V_.reserve(m * k);
  Sdev_.reserve(k);
  auto upload = [&]() {
    for (size_t r = 0; r < n; ++r)
      for (int c = 0; c < k; ++c) rowu[r * k + c] = u[(size_t)c * n + r];
    for (size_t r = 0; r < m; ++r)
      for (int c = 0; c < k; ++c) rowv[r * k + c] = v[(size_t)c * m + r];
    PCAONE_HIP(hipMemcpyAsync(Ucur_.p, rowu.data(), n * k * 8, hipMemcpyHostToDevice, stream_));
    PCAONE_HIP(hipMemcpyAsync(V_.p, rowv.data(), m * k * 8, hipMemcpyHostToDevice, stream_));
    PCAONE_HIP(hipMemcpyAsync(Sdev_.p, s.data(), (size_t)k * 8, hipMemcpyHostToDevice, stream_));
  };
  const int side = (N_ > M_) ? 1 : 0;
  const size_t ne = side ? m : n;
  std::vector<double> lam(k), ev0(ne * k), evn(ne * k);
  int nc = 0;
  eigs(side, k, ncv, maxit, tol, lam.data(), ev0.data(), &nc, nullptr);
  PCAONE_REQUIRE(nc == k, -4, "the nconv is not equal to k.");
  usv_from_eigvecs(k, lam.data(), ev0.data(), u.data(), s.data(), v.data());
  flip_uv(u, v, n, m, k);
  int it = 0;
  for (int i = 1; i <= maxiter; ++i) {
    upload();
    gl_fill_E(Ucur_.p, Sdev_.p, V_.p, (uint32_t)k, false, nullptr);
    eigs(side, k, ncv, maxit, tol, lam.data(), evn.data(), &nc, nullptr);
    PCAONE_REQUIRE(nc == k, -4, "pcangsd_iram: fewer than k eigenpairs converged");
    usv_from_eigvecs(k, lam.data(), ev0.data(), u.data(), s.data(), v2.data());
    flip_uv(u, v2, n, m, k);
    const double diff = rmse_flip(v2, v, m, k);
    v.swap(v2);
    it = i;
    if (diffs) diffs[i - 1] = diff;
    if (diff < tolem) break;
  }
  if (Dc || C) {
    upload();
    gl_Dc_.reserve(N_);
    gl_fill_E(Ucur_.p, Sdev_.p, V_.p, (uint32_t)k, true, gl_Dc_.p);
    if (Dc) PCAONE_HIP(hipMemcpyAsync(Dc, gl_Dc_.p, n * 8, hipMemcpyDeviceToHost, stream_));
    if (C) {  // kept on the device for grm_eig
      grm_.reserve((uint64_t)N_ * N_);
      timed("grm", [&] { launch_grm(dns_.p, kg_, N_, M_, gl_Dc_.p, grm_.p, stream_); });
      PCAONE_HIP(hipMemcpyAsync(C, grm_.p, n * n * 8, hipMemcpyDeviceToHost, stream_));
      have_grm_ = true;
    }
  }
  sync();  // the staging vectors go out of scope
  if (U) std::memcpy(U, u.data(), n * k * 8);
  if (S) std::memcpy(S, s.data(), (size_t)k * 8);
  if (V) std::memcpy(V, v.data(), m * k * 8);
  if (iters) *iters = it;
}

// run_pca_with_halko's EM branch for PCAngsd (src/Halko.cpp:294-331): one
// solve on the initial E, then E <- fit_with_pi(U, S, V) and a new solve until
// 1 - mev(V, Vpre) < tolem; then pcangsd_standardize_E gives Dc and, when C is
// given, the GRM C = E E^T / M with diagonal Dc / M (N x N, row-major).  U, S,
// V are those of the last solve; the operator's matrix is left standardised.
void Engine::pcangsd(int maxp, double tol, int maxiter, double tolem, double* U, double* S, double* V, int* iters,
                     double* diffs, double* Dc, double* C) {
  PCAONE_REQUIRE(gl_mode_ && have_op_, -3, "pcangsd: load likelihoods, run gl_em_maf and rsvd_init first");
  PCAONE_REQUIRE(maxiter >= 0, -1, "pcangsd: maxiter must be >= 0");
  PCAONE_HIP(hipSetDevice(device_));
  const uint32_t k = (uint32_t)p_.k;
  set_flags(false, false);
  rsvd_usv(maxp, tol, nullptr, nullptr, nullptr, nullptr, nullptr);
  int it = 0;
  for (int i = 0; i < maxiter; ++i) {
    PCAONE_HIP(hipMemcpyAsync(Vpre_.p, V_.p, (size_t)M_ * k * 8, hipMemcpyDeviceToDevice, stream_));
    PCAONE_HIP(hipMemcpyAsync(Sdev_.p, last_S_.data(), (size_t)k * 8, hipMemcpyHostToDevice, stream_));
    gl_fill_E(Ucur_.p, Sdev_.p, V_.p, k, false, nullptr);
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
  copy_usv(U, S, V);
  if (Dc || C) {
    gl_Dc_.reserve(N_);
    PCAONE_HIP(hipMemcpyAsync(Sdev_.p, last_S_.data(), (size_t)k * 8, hipMemcpyHostToDevice, stream_));
    gl_fill_E(Ucur_.p, Sdev_.p, V_.p, k, true, gl_Dc_.p);
    if (Dc) PCAONE_HIP(hipMemcpyAsync(Dc, gl_Dc_.p, (size_t)N_ * 8, hipMemcpyDeviceToHost, stream_));
    if (C) {  // kept on the device for grm_eig
      grm_.reserve((uint64_t)N_ * N_);
      timed("grm", [&] { launch_grm(dns_.p, kg_, N_, M_, gl_Dc_.p, grm_.p, stream_); });
      PCAONE_HIP(hipMemcpyAsync(C, grm_.p, (size_t)N_ * N_ * 8, hipMemcpyDeviceToHost, stream_));
      have_grm_ = true;
    }
    sync();
  }
  if (iters) *iters = it;
}

}  // namespace pcaone
