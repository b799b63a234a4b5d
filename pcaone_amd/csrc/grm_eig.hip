// Eigenvectors of the PCAngsd GRM: the reference decomposes C = E E^T / M
// (diagonal Dc / M) with Eigen::JacobiSVD and writes the left singular vectors
// to .eigvecs2 (src/Halko.cpp:317-331, src/Arnoldi.cpp:94-103).  C is
// symmetric, so its singular values are |lambda| of its eigenvalues and its
// singular vectors are its eigenvectors; here C stays on the device, rocSOLVER's
// FP64 divide-and-conquer eigensolver (dsyevd) decomposes a copy of it, and
// the eigenvectors are gathered into JacobiSVD's order (|lambda| descending,
// ties by index) as row-major rows per sample.  This step is an output of the
// EM driver, outside the RSVD hot path, so a library eigensolver serves it.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "engine.hpp"

namespace pcaone {
namespace {

// U2[i][j] = Q[i][order[j]]: Q column-major (column c = eigenvector c)
__global__ void k_gather_eigvecs(const double* __restrict__ Q, const int32_t* __restrict__ order, uint32_t n,
                                 double* __restrict__ U2) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (uint64_t)n * n) return;
  const uint32_t i = (uint32_t)(e / n), j = (uint32_t)(e % n);
  U2[e] = Q[(uint64_t)order[j] * n + i];
}

void check_blas(rocblas_status st, const char* what) {
  PCAONE_REQUIRE(st == rocblas_status_success, -2,
                 std::string("grm_eig: ") + what + " failed (rocBLAS status " + std::to_string((int)st) + ")");
}

}  // namespace

void destroy_blas(void* handle) {
  if (handle) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(handle));
}

void Engine::grm_eig(double* evals, double* U2) {
  PCAONE_REQUIRE(have_grm_, -3, "grm_eig: no GRM: run pcaone_gpu_pcangsd with C first");
  PCAONE_REQUIRE(evals || U2, -1, "grm_eig: nothing to return");
  PCAONE_HIP(hipSetDevice(device_));
  const uint32_t n = N_;
  if (!blas_) {
    rocblas_handle hd = nullptr;
    check_blas(rocblas_create_handle(&hd), "rocblas_create_handle");
    blas_ = hd;
  }
  rocblas_handle hd = static_cast<rocblas_handle>(blas_);
  check_blas(rocblas_set_stream(hd, stream_), "rocblas_set_stream");
  DevBuf<double> q, d, e, u2;
  DevBuf<int32_t> info, ord;
  q.reserve((uint64_t)n * n);
  d.reserve(n);
  e.reserve(n);
  info.reserve(1);
  PCAONE_HIP(hipMemcpyAsync(q.p, grm_.p, (size_t)n * n * 8, hipMemcpyDeviceToDevice, stream_));
  // row-major C read as column-major is C itself (symmetric)
  timed("grm_eig", [&] {
    check_blas(rocsolver_dsyevd(hd, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, q.p, (rocblas_int)n,
                                d.p, e.p, info.p),
               "rocsolver_dsyevd");
  });
  int32_t inf = 0;
  std::vector<double> lam(n);
  PCAONE_HIP(hipMemcpyAsync(&inf, info.p, 4, hipMemcpyDeviceToHost, stream_));
  PCAONE_HIP(hipMemcpyAsync(lam.data(), d.p, (size_t)n * 8, hipMemcpyDeviceToHost, stream_));
  sync();
  PCAONE_REQUIRE(inf == 0, -4, "grm_eig: the eigensolver did not converge (info " + std::to_string(inf) + ")");
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    const double x = std::fabs(lam[a]), y = std::fabs(lam[b]);
    return x > y;  // NaN never: info == 0
  });
  if (evals)
    for (uint32_t j = 0; j < n; ++j) evals[j] = std::fabs(lam[order[j]]);
  if (U2) {
    ord.reserve(n);
    u2.reserve((uint64_t)n * n);
    PCAONE_HIP(hipMemcpyAsync(ord.p, order.data(), (size_t)n * 4, hipMemcpyHostToDevice, stream_));
    const uint64_t tot = (uint64_t)n * n;
    timed("grm_eig", [&] {
      hipLaunchKernelGGL(k_gather_eigvecs, dim3((uint32_t)ceil_div(tot, 256)), dim3(256), 0, stream_, q.p, ord.p, n,
                         u2.p);
    });
    PCAONE_HIP(hipMemcpyAsync(U2, u2.p, (size_t)tot * 8, hipMemcpyDeviceToHost, stream_));
  }
  sync();  // the host order vector and the local device buffers
}

}  // namespace pcaone
