// LD r^2 (the --print-r2 path of run_ld_stuff on a .residuals matrix, or on
// PLINK input with --USV, src/LD.cpp:450-470): windows from positions on the
// host, the band of the Gram on the device.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"

namespace pcaone {

// divide_pos_by_window (src/LD.cpp:125-139), step for step: a SNP leads a
// window unless its position equals its chromosome's last position (which
// also moves on to the next chromosome); the window holds the SNPs j >= i of
// that chromosome with pos[j] - pos[i] <= ld_bp.
uint32_t ld_windows(const int32_t* pos, uint32_t nsnps, const int32_t* chr_end, uint32_t nchr, int32_t ld_bp,
                    int32_t* ws, int32_t* we) {
  PCAONE_REQUIRE(pos && chr_end && ws && we && nsnps > 0 && nchr > 0, -1, "ld_windows: empty input");
  for (uint32_t c = 0; c < nchr; ++c)
    PCAONE_REQUIRE(chr_end[c] >= 0 && (uint32_t)chr_end[c] < nsnps && (c == 0 || chr_end[c] > chr_end[c - 1]), -1,
                   "ld_windows: chromosome ends must increase within [0, nsnps)");
  PCAONE_REQUIRE((uint32_t)chr_end[nchr - 1] == nsnps - 1, -1, "ld_windows: the last chromosome must end at nsnps-1");
  uint32_t nwin = 0, c = 0;
  for (uint32_t i = 0; i < nsnps; ++i) {
    PCAONE_REQUIRE(c < nchr, -1, "ld_windows: positions run past the last chromosome");
    if (pos[i] == pos[chr_end[c]]) {
      ++c;
      continue;
    }
    uint32_t j = i;
    for (; j <= (uint32_t)chr_end[c]; ++j)
      if ((int64_t)pos[j] - pos[i] > ld_bp) break;
    ws[nwin] = (int32_t)i;
    we[nwin] = (int32_t)(j - i);
    ++nwin;
  }
  return nwin;
}

// run_ld_stuff's --USV branch for PLINK input (src/Main.cpp:64-75,
// src/LD.cpp:460-466): FileBed::read_all's centred genotypes G = BED2GENO - F
// (missing 0, no MAF filter at the default --maf 0) become (I - U U^T) G, U
// the N x k left singular vectors as read from .eigvecs (column-major here).
// The result stays in HBM in FP64, as the reference's data->G, and is not
// re-centred; ld_r2 then runs on it.
void Engine::ld_usv(const double* U, uint32_t k) {
  PCAONE_REQUIRE(have_bed_ && src_ == SRC_BED && !gl_mode_, -3, "ld_usv: load PLINK genotypes first");
  PCAONE_REQUIRE(U && k >= 1 && k <= 64, -1, "ld_usv: U must be nsamples x k with 1 <= k <= 64");
  PCAONE_HIP(hipSetDevice(device_));
  std::vector<double> ur((size_t)N_ * k);
  for (uint32_t j = 0; j < N_; ++j)
    for (uint32_t c = 0; c < k; ++c) ur[(size_t)j * k + c] = U[(size_t)c * N_ + j];
  DevBuf<double> du, df, dw;
  du.reserve(ur.size());
  df.reserve(M_);
  dw.reserve((uint64_t)M_ * k);
  resid_.reserve((uint64_t)M_ * N_);
  PCAONE_HIP(hipMemcpyAsync(du.p, ur.data(), ur.size() * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(df.p, F_orig_.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
  timed("ld_usv", [&] { launch_usv_residual(resident_raw("ld_usv"), N_, M_, df.p, du.p, k, dw.p, resid_.p, stream_); });
  sync();  // the staging buffers are freed on return
  have_resid_ = true;
}

// ld_r2_big (src/LD.cpp:419-442) on FileBin::read_all's centred matrix (or the
// --USV-adjusted genotypes after ld_usv): the r^2 of pairs (ws[w], ws[w] + j),
// j = 1 .. we[w] - 1, window after window.  Returns the number of pairs
// (written when r2 != null and capacity allows).
uint64_t Engine::ld_r2(const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2, uint64_t capacity) {
  const bool resid = have_resid_;
  PCAONE_REQUIRE(have_bed_ && ((src_ == SRC_DENSE && !gl_mode_) || resid), -3,
                 "ld_r2: load a dense (.residuals) matrix, or PLINK genotypes and ld_usv, first");
  PCAONE_REQUIRE(N_ >= 2, -1, "ld_r2: needs at least two samples");
  PCAONE_REQUIRE(ws && we, -1, "ld_r2: null windows");
  PCAONE_HIP(hipSetDevice(device_));
  std::vector<int64_t> off(nwin);
  std::vector<int32_t> win_of_row(M_, -1);
  uint64_t total = 0;
  for (uint32_t w = 0; w < nwin; ++w) {
    PCAONE_REQUIRE(ws[w] >= 0 && (uint32_t)ws[w] < M_ && we[w] >= 1 && (uint64_t)ws[w] + we[w] <= M_, -1,
                   "ld_r2: window out of range");
    PCAONE_REQUIRE(w == 0 || ws[w] > ws[w - 1], -1, "ld_r2: lead SNPs must increase");
    off[w] = (int64_t)total;
    total += (uint64_t)(we[w] - 1);
    win_of_row[ws[w]] = (int32_t)w;
  }
  if (!r2) return total;
  PCAONE_REQUIRE(capacity >= total, -1, "ld_r2: output too small");
  if (total == 0) return 0;
  // tiles of the band: row block I, column blocks I .. (last column of I's windows) / 64
  const uint32_t nb = (uint32_t)ceil_div(M_, 64);
  std::vector<int32_t> last(nb, -1);
  for (uint32_t w = 0; w < nwin; ++w) {
    if (we[w] <= 1) continue;
    const uint32_t I = (uint32_t)ws[w] / 64;
    last[I] = std::max<int32_t>(last[I], ws[w] + we[w] - 1);
  }
  std::vector<int2> tiles;
  for (uint32_t I = 0; I < nb; ++I)
    if (last[I] >= 0)
      for (uint32_t J = I; J <= (uint32_t)last[I] / 64; ++J) tiles.push_back(int2{(int)I, (int)J});
  DevBuf<double> mu, sds, out;
  DevBuf<int2> dt;
  DevBuf<int32_t> dwin, dwe;
  DevBuf<int64_t> doff;
  mu.reserve(M_);
  sds.reserve(M_);
  out.reserve(total);
  dt.reserve(tiles.size());
  dwin.reserve(M_);
  dwe.reserve(nwin);
  doff.reserve(nwin);
  if (resid)  // the adjusted genotypes are used as they are
    timed("ld_sds", [&] { launch_fill_f64(mu.p, M_, 0.0, stream_); });
  else
    PCAONE_HIP(hipMemcpyAsync(mu.p, mu_orig_.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dt.p, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dwin.p, win_of_row.data(), (size_t)M_ * 4, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dwe.p, we, (size_t)nwin * 4, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(doff.p, off.data(), (size_t)nwin * 8, hipMemcpyHostToDevice, stream_));
  const void* X = resid ? static_cast<const void*>(resid_.p) : static_cast<const void*>(rawf_.p);
  timed("ld_sds", [&] { launch_ld_sds(X, resid ? 1 : 0, N_, M_, mu.p, sds.p, stream_); });
  LdTiles t{};
  t.X = X;
  t.x_f64 = resid ? 1 : 0;
  t.mu = mu.p;
  t.sds = sds.p;
  t.n = N_;
  t.m = M_;
  t.tiles = dt.p;
  t.win_of_row = dwin.p;
  t.we = dwe.p;
  t.off = doff.p;
  t.r2 = out.p;
  timed("ld_tile", [&] { launch_ld_tiles(t, (uint32_t)tiles.size(), stream_); });
  PCAONE_HIP(hipMemcpyAsync(r2, out.p, total * 8, hipMemcpyDeviceToHost, stream_));
  sync();
  return total;
}

// Window pairs -> output offsets and the lead-SNP map; returns the pair count.
uint64_t window_offsets(const int32_t* ws, const int32_t* we, uint32_t nwin, uint32_t m,
                               std::vector<int64_t>& off, std::vector<int32_t>& win_of_row) {
  off.assign(nwin, 0);
  win_of_row.assign(m, -1);
  uint64_t total = 0;
  for (uint32_t w = 0; w < nwin; ++w) {
    PCAONE_REQUIRE(ws[w] >= 0 && (uint32_t)ws[w] < m && we[w] >= 1 && (uint64_t)ws[w] + we[w] <= m, -1,
                   "ld_r2: window out of range");
    PCAONE_REQUIRE(w == 0 || ws[w] > ws[w - 1], -1, "ld_r2: lead SNPs must increase");
    off[w] = (int64_t)total;
    total += (uint64_t)(we[w] - 1);
    win_of_row[ws[w]] = (int32_t)w;
  }
  return total;
}

// Band tiles of the windows [0, nwin): row block I (tm SNPs) against the
// column blocks (tn SNPs) from its own first SNP to the last SNP its windows
// reach.
std::vector<int2> band_tiles(const int32_t* ws, const int32_t* we, uint32_t nwin, uint32_t m, int tm,
                                    int tn) {
  const uint32_t nb = (uint32_t)ceil_div(m, (uint64_t)tm);
  std::vector<int64_t> last(nb, -1);
  for (uint32_t w = 0; w < nwin; ++w)
    if (we[w] > 1) {
      const uint32_t I = (uint32_t)ws[w] / tm;
      last[I] = std::max<int64_t>(last[I], (int64_t)ws[w] + we[w] - 1);
    }
  std::vector<int2> tiles;
  for (uint32_t I = 0; I < nb; ++I)
    if (last[I] >= 0)
      for (int64_t J = (int64_t)I * tm / tn; J <= last[I] / tn; ++J) tiles.push_back(int2{(int)I, (int)J});
  return tiles;
}

// ld_r2_big (src/LD.cpp:419-442) on the matrix Data::write_residuals writes
// for PLINK input with --ld --ld-stats 1 (src/Data.cpp:196-221): G = BED2GENO
// - F (missing 0, not standardised), each SNP re-centred by its mean.  The r^2
// of each window pair is computed from the exact integer sums of the packed
// genotypes (int8 MFMA) and the per-SNP terms below, in FP64; the reference
// computes the same quantity on a float32 copy of G.
uint64_t Engine::ld_r2_geno(const int32_t* ws, const int32_t* we, uint32_t nwin, double* r2, uint64_t capacity,
                            int32_t* sums, int* config) {
  PCAONE_REQUIRE(have_bed_ && src_ == SRC_BED && !gl_mode_, -3, "ld_r2_geno: load PLINK genotypes first");
  PCAONE_REQUIRE(N_ >= 2, -1, "ld_r2_geno: needs at least two samples");
  PCAONE_REQUIRE(ws && we, -1, "ld_r2_geno: null windows");
  PCAONE_HIP(hipSetDevice(device_));
  std::vector<int64_t> off;
  std::vector<int32_t> win_of_row;
  const uint64_t total = window_offsets(ws, we, nwin, M_, off, win_of_row);
  if (!r2) return total;
  PCAONE_REQUIRE(capacity >= total, -1, "ld_r2_geno: output too small");
  if (total == 0) return 0;
  if (!have_ldg_) {
    // identity SNP-major layout (64-byte rows, padding codes missing) and the
    // per-SNP terms from the exact code counts of load_bed
    ldg_row_bytes_ = (uint32_t)round_up((N_ + 3) >> 2, 64);
    ldg_bed_.reserve((uint64_t)ldg_row_bytes_ * M_);
    timed("ld_geno_layout",
          [&] { launch_bed_to_snpmajor(resident_raw("ld_r2_geno"), N_, M_, nullptr, ldg_bed_.p, ldg_row_bytes_, stream_); });
    std::vector<double> mun(M_), ss(M_);
    for (uint32_t i = 0; i < M_; ++i) {
      const int64_t n00 = counts_orig_[4 * (size_t)i], n10 = counts_orig_[4 * (size_t)i + 2],
                    n11 = counts_orig_[4 * (size_t)i + 3];
      const double F = F_orig_[i], sy = (double)(2 * n00 + n10), syy = (double)(4 * n00 + n10),
                   n = (double)(n00 + n10 + n11);
      mun[i] = 0.5 * sy - F * n;
      ss[i] = 0.25 * syy - F * sy + F * F * n - mun[i] * mun[i] / (double)N_;
    }
    ldg_F_.reserve(M_);
    ldg_mun_.reserve(M_);
    ldg_ss_.reserve(M_);
    PCAONE_HIP(hipMemcpyAsync(ldg_F_.p, F_orig_.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
    PCAONE_HIP(hipMemcpyAsync(ldg_mun_.p, mun.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
    PCAONE_HIP(hipMemcpyAsync(ldg_ss_.p, ss.data(), (size_t)M_ * 8, hipMemcpyHostToDevice, stream_));
    sync();
    have_ldg_ = true;
  }
  DevBuf<int32_t> dwin, dwe;
  DevBuf<int64_t> doff;
  DevBuf<double> out;
  DevBuf<int32_t> dsums;
  dwin.reserve(M_);
  dwe.reserve(nwin);
  doff.reserve(nwin);
  out.reserve(total);
  if (sums) dsums.reserve(total * 4);
  PCAONE_HIP(hipMemcpyAsync(dwin.p, win_of_row.data(), (size_t)M_ * 4, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(dwe.p, we, (size_t)nwin * 4, hipMemcpyHostToDevice, stream_));
  PCAONE_HIP(hipMemcpyAsync(doff.p, off.data(), (size_t)nwin * 8, hipMemcpyHostToDevice, stream_));
  LdGeno g{};
  g.bed = ldg_bed_.p;
  g.row_bytes = ldg_row_bytes_;
  g.nsnps = M_;
  g.F = ldg_F_.p;
  g.mun = ldg_mun_.p;
  g.ss = ldg_ss_.p;
  g.inv_n = 1.0 / (double)N_;
  g.win_of_row = dwin.p;
  g.win_lo = 0;
  g.win_rows = M_;
  g.we = dwe.p;
  g.off = doff.p;
  g.r2 = out.p;
  g.sums = sums ? dsums.p : nullptr;
  DevBuf<int2> dt;
  uint32_t grid = 0;
  // tile list of a configuration over the windows [0, nw), uploaded
  auto prepare = [&](int cfg, uint32_t nw) {
    int tm, tn;
    ld_geno_tile(cfg, &tm, &tn);
    const std::vector<int2> tiles = band_tiles(ws, we, nw, M_, tm, tn);
    dt.reserve(std::max<size_t>(tiles.size(), 1));
    PCAONE_HIP(hipMemcpyAsync(dt.p, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice, stream_));
    g.tiles = dt.p;
    g.ntiles = (uint32_t)tiles.size();
    grid = ld_geno_grid(cfg, g.ntiles, (uint32_t)ncu_);
    sync();  // `tiles` goes out of scope
  };
  int cfg = config ? *config : -1;
  PCAONE_REQUIRE(cfg < ld_geno_nconfigs(), -1, "ld_r2_geno: no such tile configuration");
  if (cfg < 0) {
    // tile shape by window-size class (log2 of the mean pairs per window):
    // every candidate timed on a prefix of the windows holding about 4M pairs
    const int cls = (int)std::floor(std::log2(std::max(1.0, (double)total / nwin)));
    auto it = ldg_tuned_.find(cls);
    if (it != ldg_tuned_.end()) {
      cfg = it->second;
    } else {
      uint32_t nw = nwin;
      for (uint32_t w = 0; w < nwin; ++w)
        if ((uint64_t)off[w] >= (4ull << 20)) {
          nw = w;
          break;
        }
      float best = 0.f;
      hipEvent_t e0, e1;
      PCAONE_HIP(hipEventCreate(&e0));
      PCAONE_HIP(hipEventCreate(&e1));
      for (int c = 0; c < ld_geno_nconfigs(); ++c) {
        prepare(c, nw);
        launch_ld_geno(g, c, grid, stream_);  // code load and caches
        PCAONE_HIP(hipEventRecord(e0, stream_));
        launch_ld_geno(g, c, grid, stream_);
        PCAONE_HIP(hipGetLastError());
        PCAONE_HIP(hipEventRecord(e1, stream_));
        PCAONE_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PCAONE_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (cfg < 0 || ms < best) {
          best = ms;
          cfg = c;
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      ldg_tuned_[cls] = cfg;
    }
  }
  if (config) *config = cfg;
  prepare(cfg, nwin);
  timed("ld_geno", [&] { launch_ld_geno(g, cfg, grid, stream_); });
  PCAONE_HIP(hipMemcpyAsync(r2, out.p, total * 8, hipMemcpyDeviceToHost, stream_));
  if (sums) PCAONE_HIP(hipMemcpyAsync(sums, dsums.p, total * 16, hipMemcpyDeviceToHost, stream_));
  sync();
  return total;
}

}  // namespace pcaone
