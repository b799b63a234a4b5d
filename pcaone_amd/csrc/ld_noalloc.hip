// Allocation-free variant of the integer LD entry point, for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  Its device memory beyond the
// inputs (the handle's packed genotypes and per-SNP terms) is one fixed
// file-scope array, `scratch`, of kLdScratchBytes on every device the library
// runs on, held until the process ends.  Host code never uses the array's
// name as an address: its device address is taken once per device, with that
// device current, by hipGetSymbolAddress, and only that pointer is handed to
// copies and kernels.  Every call on a device shares that device's array and
// leaves the last call's contents behind; calls that overlap on one device
// race.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "engine.hpp"

namespace pcaone {

static __device__ __align__(256) unsigned char scratch[kLdScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void* ld_scratch(int device) {
  PCAONE_REQUIRE(device >= 0 && device < kMaxDevices, -1, "ld scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    PCAONE_HIP(hipGetDevice(&cur));
    PCAONE_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(cur);
    PCAONE_HIP(e);
    PCAONE_REQUIRE(p != nullptr, -2, "ld scratch: no device address for the scratch array");
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

// Scratch layout of one call, each segment on a 256-byte boundary: the lead
// map of the SNPs [ws[0], ws[nwin-1]], the window sizes and offsets, the tile
// list, the r^2 and (when asked) the integer sums.
uint64_t Engine::ld_r2_geno_noalloc(const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2,
                                    uint64_t capacity, int32_t* sums, int config) {
  PCAONE_REQUIRE(have_bed_ && src_ == SRC_BED && !gl_mode_, -3, "ld_r2_geno_noalloc: load PLINK genotypes first");
  PCAONE_REQUIRE(have_ldg_, -3,
                 "ld_r2_geno_noalloc: the LD layout is built by pcaone_gpu_ld_r2_geno: call it once after the load");
  PCAONE_REQUIRE(ws && we && nwin > 0, -1, "ld_r2_geno_noalloc: no windows");
  PCAONE_REQUIRE(config < ld_geno_nconfigs(), -1, "ld_r2_geno_noalloc: no such tile configuration");
  PCAONE_HIP(hipSetDevice(device_));
  std::vector<int64_t> off;
  std::vector<int32_t> win_of_row;
  const uint64_t total = window_offsets(ws, we, nwin, M_, off, win_of_row);  // checks the windows
  // the lead map covers the batch's leads only: SNPs [ws[0], ws[nwin - 1]]
  const uint32_t lo = (uint32_t)ws[0], rows = (uint32_t)ws[nwin - 1] - lo + 1;
  if (!r2) return total;
  PCAONE_REQUIRE(capacity >= total, -1, "ld_r2_geno_noalloc: output too small");
  if (total == 0) return 0;
  int cfg = config;
  if (cfg < 0) {  // the handle's tuned shape for this window-size class, else 64 x 64
    const int cls = (int)std::floor(std::log2(std::max(1.0, (double)total / nwin)));
    auto it = ldg_tuned_.find(cls);
    cfg = it != ldg_tuned_.end() ? it->second : 2;
  }
  int tm, tn;
  ld_geno_tile(cfg, &tm, &tn);
  const std::vector<int2> tiles = band_tiles(ws, we, nwin, M_, tm, tn);
  auto seg = [](uint64_t bytes) { return round_up(bytes, 256); };
  const uint64_t o_win = 0, o_we = o_win + seg((uint64_t)rows * 4), o_off = o_we + seg((uint64_t)nwin * 4);
  const uint64_t o_tiles = o_off + seg((uint64_t)nwin * 8), o_r2 = o_tiles + seg(tiles.size() * sizeof(int2));
  const uint64_t o_sums = o_r2 + seg(total * 8), end = o_sums + (sums ? total * 16 : 0);
  PCAONE_REQUIRE(end <= kLdScratchBytes, -1,
                 "ld_r2_geno_noalloc: the windows need " + std::to_string(end) + " bytes of scratch, more than the " +
                     std::to_string(kLdScratchBytes) + " it has: pass fewer windows per call");
  unsigned char* base = static_cast<unsigned char*>(ld_scratch(device_));
  PCAONE_HIP(hipMemcpyAsync(base + o_win, win_of_row.data() + lo, (size_t)rows * 4, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(base + o_we, we, (size_t)nwin * 4, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(base + o_off, off.data(), (size_t)nwin * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(base + o_tiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice, stream_));
  LdGeno g{};
  g.bed = ldg_bed_.p;
  g.row_bytes = ldg_row_bytes_;
  g.nsnps = M_;
  g.tiles = reinterpret_cast<const int2*>(base + o_tiles);
  g.ntiles = (uint32_t)tiles.size();
  g.F = ldg_F_.p;
  g.mun = ldg_mun_.p;
  g.ss = ldg_ss_.p;
  g.inv_n = 1.0 / (double)N_;
  g.win_of_row = reinterpret_cast<const int32_t*>(base + o_win);
  g.win_lo = lo;
  g.win_rows = rows;
  g.we = reinterpret_cast<const int32_t*>(base + o_we);
  g.off = reinterpret_cast<const int64_t*>(base + o_off);
  g.r2 = reinterpret_cast<double*>(base + o_r2);
  g.sums = sums ? reinterpret_cast<int32_t*>(base + o_sums) : nullptr;
  const uint32_t grid = ld_geno_grid(cfg, g.ntiles, (uint32_t)ncu_);
  timed("ld_geno", [&] { launch_ld_geno(g, cfg, grid, stream_); });
  PCAONE_HIP(hipMemcpyAsync(r2, base + o_r2, total * 8, hipMemcpyDeviceToHost, stream_));
  if (sums) PCAONE_HIP(hipMemcpyAsync(sums, base + o_sums, total * 16, hipMemcpyDeviceToHost, stream_));
  sync();  // the host staging vectors and the caller's outputs
  return total;
}

}  // namespace pcaone
