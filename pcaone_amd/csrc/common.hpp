// Shared host/device definitions for the PCAone MI355X hot path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace pcaone {

// Error raised by every host entry point; the C-ABI layer turns it into a
// negative status and keeps the message for pcaone_gpu_last_error().
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PCAONE_HIP(expr)                                                                  \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      throw ::pcaone::Error(-2, std::string("HIP error '") + hipGetErrorString(_e) +      \
                                    "' at " + __FILE__ + ":" + std::to_string(__LINE__) + \
                                    " in " #expr);                                        \
  } while (0)

#define PCAONE_REQUIRE(cond, code, msg)                    \
  do {                                                     \
    if (!(cond)) throw ::pcaone::Error((code), (msg));     \
  } while (0)

// Device buffer owned by a handle.  Never copied; moved only.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // (Re)allocate to hold at least count elements; contents undefined.
  void reserve(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) return;
    PCAONE_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  size_t bytes() const { return n * sizeof(T); }
};

__host__ __device__ inline uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }
__host__ __device__ inline uint64_t ceil_div(uint64_t x, uint64_t m) { return (x + m - 1) / m; }

// K order of the split-BF16 packed contractions: within every aligned group of
// 16 codes, operand position 4a + b holds code 4b + a (a 4 x 4 transpose, its
// own inverse).  decode16 produces the A operand in this order straight from
// one 32-bit word of packed codes, and the B operands are written in it.
__host__ __device__ inline uint64_t kperm16(uint64_t k) {
  return (k & ~(uint64_t)15) | ((k & 3) << 2) | ((k >> 2) & 3);
}

// Arithmetic arms of the packed-genotype contractions.
enum Precision : int {
  PREC_FP64 = 0,        // decode -> double, FP64 FMA (oracle arm)
  PREC_FP32 = 1,        // decode -> float, FP32 FMA (plain drop-in baseline)
  PREC_SPLIT_BF16 = 2,  // genotypes exact in BF16, operand split hi+lo, FP32 MFMA accumulate
};

}  // namespace pcaone
