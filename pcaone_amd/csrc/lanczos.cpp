// FP64 restarted Lanczos driver around the GPU Gram mat-vec: the role Spectra's
// SymEigsSolver / PartialSVDSolver play for IRAM in the reference
// (src/Arnoldi.cpp:48-203; Lanczos three-term recurrence with
// re-orthogonalisation, external/Spectra/LinAlg/Lanczos.h:56-151).  The Krylov
// basis lives on the host (it is ncv x n doubles); every operator application
// runs on the GPU.  Restarts are thick restarts on the Ritz vectors, and a
// Ritz pair is converged when its residual is below tol * max(eps^(2/3), |theta|),
// Spectra's criterion (external/Spectra/SymEigsBase.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "engine.hpp"

namespace pcaone {

// Cyclic Jacobi on a symmetric n x n row-major matrix; eigenvalues descending
// in w, eigenvectors in the columns of Z (row-major n x n).
static void sym_eig_host(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& Z) {
  std::vector<double> V((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) (i == j ? dg : off) += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    if (off <= 1e-28 * dg || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2.0 * apq);
        double t = 1.0 / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        if (theta < 0) t = -t;
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double x = A[(size_t)p * n + k], y = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * x - s * y;
          A[(size_t)q * n + k] = s * x + c * y;
        }
        for (int k = 0; k < n; ++k) {  // columns p, q
          double x = A[(size_t)k * n + p], y = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * x - s * y;
          A[(size_t)k * n + q] = s * x + c * y;
          x = V[(size_t)k * n + p];
          y = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * x - s * y;
          V[(size_t)k * n + q] = s * x + c * y;
        }
      }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int a, int b) { return A[(size_t)a * n + a] > A[(size_t)b * n + b]; });
  w.resize(n);
  Z.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    w[j] = A[(size_t)ord[j] * n + ord[j]];
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + j] = V[(size_t)i * n + ord[j]];
  }
}

void Engine::eigs(int side, int k, int ncv, int maxit, double tol, double* evals, double* evecs, int* nconv_out,
                  int* nops_out) {
  PCAONE_REQUIRE(have_mv_, -3, "eigs: call gram_init first");
  const uint64_t n = side ? M_ : N_;
  PCAONE_REQUIRE(k >= 1 && (uint64_t)k < n, -1, "eigs: need 1 <= k < n");
  ncv = (int)std::min<uint64_t>((uint64_t)ncv, n);
  PCAONE_REQUIRE(ncv > k, -1, "eigs: ncv must exceed k");
  const int64_t ops0 = nops_;
  std::vector<double> V((size_t)(ncv + 1) * n, 0.0), T((size_t)ncv * ncv, 0.0), w(n), h(ncv), h2(ncv), r(n);
  auto vec = [&](int i) { return V.data() + (size_t)i * n; };
  // deterministic start vector
  {
    uint64_t x = 0x9E3779B97F4A7C15ull;
    double nrm = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      const double u = (double)(x >> 11) * 0x1.0p-53 - 0.5;
      vec(0)[i] = u;
      nrm += u * u;
    }
    nrm = std::sqrt(nrm);
    for (uint64_t i = 0; i < n; ++i) vec(0)[i] /= nrm;
  }
  const double eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);
  int m = 0, nconv = 0;
  double rnorm = 0.0;
  std::vector<double> theta, Y;
  for (int it = 0; it < maxit; ++it) {
    for (int j = m; j < ncv; ++j) {
      gram_matvec(side, vec(j), w.data());
      // classical Gram-Schmidt against the whole basis, twice
      for (int pass = 0; pass < 2; ++pass) {
        double* hp = pass ? h2.data() : h.data();
        for (int i = 0; i <= j; ++i) {
          const double* vi = vec(i);
          double s = 0.0;
          for (uint64_t q = 0; q < n; ++q) s += vi[q] * w[q];
          hp[i] = s;
        }
        for (int i = 0; i <= j; ++i) {
          const double* vi = vec(i);
          const double s = hp[i];
          for (uint64_t q = 0; q < n; ++q) w[q] -= s * vi[q];
        }
      }
      for (int i = 0; i <= j; ++i) {
        const double v = h[i] + h2[i];
        T[(size_t)i * ncv + j] = v;
        T[(size_t)j * ncv + i] = v;
      }
      double b = 0.0;
      for (uint64_t q = 0; q < n; ++q) b += w[q] * w[q];
      b = std::sqrt(b);
      if (j + 1 < ncv) {
        PCAONE_REQUIRE(b > 0.0, -4, "eigs: Krylov space exhausted (invariant subspace)");
        for (uint64_t q = 0; q < n; ++q) vec(j + 1)[q] = w[q] / b;
      } else {
        rnorm = b;
        r = w;
      }
    }
    sym_eig_host(ncv, T, theta, Y);
    nconv = 0;
    for (int i = 0; i < k; ++i) {
      const double res = rnorm * std::fabs(Y[(size_t)(ncv - 1) * ncv + i]);
      if (res <= tol * std::max(eps23, std::fabs(theta[i]))) ++nconv;
    }
    if (nconv >= k || it + 1 == maxit) break;
    // thick restart on the leading Ritz vectors
    const int keep = std::min(ncv - 1, k + (ncv - k) / 2);
    std::vector<double> U((size_t)keep * n, 0.0);
    for (int i = 0; i < keep; ++i)
      for (int p = 0; p < ncv; ++p) {
        const double c = Y[(size_t)p * ncv + i];
        const double* vp = vec(p);
        double* ui = U.data() + (size_t)i * n;
        for (uint64_t q = 0; q < n; ++q) ui[q] += c * vp[q];
      }
    std::copy(U.begin(), U.end(), V.begin());
    PCAONE_REQUIRE(rnorm > 0.0, -4, "eigs: Krylov space exhausted at restart (zero residual)");
    for (uint64_t q = 0; q < n; ++q) vec(keep)[q] = r[q] / rnorm;
    std::fill(T.begin(), T.end(), 0.0);
    for (int i = 0; i < keep; ++i) T[(size_t)i * ncv + i] = theta[i];
    m = keep;
  }
  for (int i = 0; i < k; ++i) evals[i] = theta[i];
  if (evecs) {
    std::fill(evecs, evecs + (size_t)n * k, 0.0);
    for (int i = 0; i < k; ++i)
      for (int p = 0; p < ncv; ++p) {
        const double c = Y[(size_t)p * ncv + i];
        const double* vp = vec(p);
        double* ei = evecs + (size_t)i * n;
        for (uint64_t q = 0; q < n; ++q) ei[q] += c * vp[q];
      }
  }
  if (nconv_out) *nconv_out = nconv;
  if (nops_out) *nops_out = (int)(nops_ - ops0);
}

// IRAM PCA (run_pca_with_arnoldi in-core, src/Arnoldi.cpp:57-68): the top-k
// singular triplets of X from the eigenpairs of X^T X (tall, N > M) or X X^T
// (wide), PartialSVDSolver's choice (external/Spectra/contrib/PartialSVDSolver.h:135-150).
void Engine::iram(int k, int ncv, int maxit, double tol, double* U, double* S, double* V, int* nconv, int* nops) {
  const int side = (N_ > M_) ? 1 : 0;
  const uint64_t n = side ? M_ : N_;
  std::vector<double> lam(k), vecs((size_t)n * k);
  int nc = 0;
  eigs(side, k, ncv, maxit, tol, lam.data(), vecs.data(), &nc, nops);
  if (nconv) *nconv = nc;
  // run_pca_with_arnoldi stops here (src/Arnoldi.cpp:64)
  PCAONE_REQUIRE(nc == k, -4, "the nconv is not equal to k.");
  usv_from_eigvecs(k, lam.data(), vecs.data(), U, S, V);
}

// PartialSVDSolver's matrix_U / matrix_V (external/Spectra/contrib/
// PartialSVDSolver.h:177-206): S = sqrt(lam); on the smaller Gram side the
// eigenvectors are the singular vectors, the other side is X V / s or X^T U / s
void Engine::usv_from_eigvecs(int k, const double* lam, const double* vecs, double* U, double* S, double* V) {
  const int side = (N_ > M_) ? 1 : 0;
  const uint64_t n = side ? M_ : N_;
  std::vector<double> s(k);
  for (int i = 0; i < k; ++i) {
    s[i] = std::sqrt(std::max(lam[i], 0.0));
    PCAONE_REQUIRE(s[i] > 0.0, -4, "iram: zero singular value among the top k (rank-deficient matrix)");
  }
  if (S) std::memcpy(S, s.data(), k * sizeof(double));
  std::vector<double> tmp(side ? N_ : M_);
  for (int i = 0; i < k; ++i) {
    const double* e = vecs + (size_t)i * n;
    if (side) {  // V = eigvecs, U = X V / s
      if (V) std::memcpy(V + (size_t)i * M_, e, M_ * sizeof(double));
      apply_X(e, tmp.data());
      if (U)
        for (uint32_t j = 0; j < N_; ++j) U[(size_t)i * N_ + j] = tmp[j] / s[i];
    } else {  // U = eigvecs, V = X^T U / s
      if (U) std::memcpy(U + (size_t)i * N_, e, N_ * sizeof(double));
      apply_Xt(e, tmp.data());
      if (V)
        for (uint32_t j = 0; j < M_; ++j) V[(size_t)i * M_ + j] = tmp[j] / s[i];
    }
  }
}

}  // namespace pcaone
