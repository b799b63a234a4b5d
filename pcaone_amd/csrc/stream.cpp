// Out-of-core genotypes (-m, src/Data.cpp:27-71; FileBed::read_block_initial,
// src/FilePlink.cpp:133-224).  The .bed stays in host memory (a read-only
// mapping of the file, paged in from disk on demand); each use of a window
// gathers its rows, in operator order, into a pinned staging buffer on host
// threads, copies it on a copy stream into one slot of a two-slot device
// ring, and lays it out there (SNP-major rows, sample-major block) for the
// same contractions the resident path runs.  While window w computes on the
// main stream, window w + 1 is gathered and copied into the other slot.  Only
// F, d, the missing lists and the operator's N x l / M x l matrices stay
// resident: a dataset larger than HBM runs in about two windows of it.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace pcaone {
namespace {

// pinned host memory owned by a scope
struct Pinned {
  uint8_t* p = nullptr;
  explicit Pinned(size_t n) { PCAONE_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1), 0)); }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  Pinned(const Pinned&) = delete;
  Pinned& operator=(const Pinned&) = delete;
};

}  // namespace

void Engine::load_bed_file(const char* filebed, uint32_t nsamples, uint32_t nsnps, bool out_of_core) {
  PCAONE_REQUIRE(filebed && nsamples > 0 && nsnps > 0, -1, "load_bed_file: empty input");
  PCAONE_HIP(hipSetDevice(device_));
  const int fd = ::open(filebed, O_RDONLY);
  PCAONE_REQUIRE(fd >= 0, -1, std::string("Cannot open bed file: ") + filebed);
  struct stat st {};
  const bool stat_ok = fstat(fd, &st) == 0;
  const uint64_t nb = (nsamples + 3) >> 2;
  const uint64_t need = 3 + nb * nsnps;
  if (!stat_ok || (uint64_t)st.st_size < need) {
    ::close(fd);
    throw Error(-1, std::string("load_bed_file: ") + filebed + " holds fewer than 3 + nsnps x ceil(nsamples / 4) bytes");
  }
  void* base = mmap(nullptr, need, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  PCAONE_REQUIRE(base != MAP_FAILED, -2, std::string("load_bed_file: cannot map ") + filebed);
  const uint8_t* b = static_cast<const uint8_t*>(base);
  if (b[0] != 0x6c || b[1] != 0x1b || b[2] != 0x01) {  // src/FilePlink.cpp:20-26
    munmap(base, need);
    throw Error(-1, "Incorrect magic number in plink bed file.");
  }
  if (!out_of_core) {
    try {
      load_bed(b + 3, nsamples, nsnps);
    } catch (...) {
      munmap(base, need);
      throw;
    }
    munmap(base, need);
    return;
  }
  have_op_ = have_mv_ = false;
  end_stream();
  raw_.release();  // nothing of a previous payload stays in HBM
  bed_snp_.release();
  bed_smp_.release();
  N_ = nsamples;
  M_ = nsnps;
  map_base_ = base;
  map_len_ = need;
  host_bed_ = b + 3;
  streamed_ = true;
  if (!copy_stream_) PCAONE_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
  stream_counts();
}

void Engine::end_stream() {
  if (stream_) PCAONE_HIP(hipStreamSynchronize(stream_));
  if (copy_stream_) PCAONE_HIP(hipStreamSynchronize(copy_stream_));
  for (auto& sl : slots_) {
    sl.raw.release();
    sl.snp.release();
    sl.smp.release();
    if (sl.stage) (void)hipHostFree(sl.stage);
    sl.stage = nullptr;
    sl.stage_bytes = 0;
    if (sl.copied) (void)hipEventDestroy(sl.copied);
    if (sl.used) (void)hipEventDestroy(sl.used);
    sl.copied = sl.used = nullptr;
    sl.copy_rec = sl.used_rec = sl.laid = false;
    sl.win = -1;
  }
  last_slot_ = -1;
  if (map_base_) munmap(map_base_, map_len_);
  map_base_ = nullptr;
  map_len_ = 0;
  host_bed_ = nullptr;
  streamed_ = false;
}

// rows [r0, r0 + nr) of the operator (permuted: through perm_, else the file's
// own rows) into dst, ceil(N / 4) bytes each, on up to 16 host threads
void Engine::gather_rows(uint8_t* dst, uint32_t r0, uint32_t nr, bool permuted) const {
  const uint64_t nb = (N_ + 3) >> 2;
  auto work = [&](uint32_t a, uint32_t e) {
    if (!permuted) {
      std::memcpy(dst + (uint64_t)a * nb, host_bed_ + (uint64_t)(r0 + a) * nb, (uint64_t)(e - a) * nb);
      return;
    }
    for (uint32_t i = a; i < e; ++i)
      std::memcpy(dst + (uint64_t)i * nb, host_bed_ + (uint64_t)perm_[r0 + i] * nb, nb);
  };
  const uint64_t bytes = (uint64_t)nr * nb;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const uint32_t nt = (uint32_t)std::min<uint64_t>({16u, hw, std::max<uint64_t>(1, bytes >> 22), nr});
  if (nt <= 1) {
    work(0, nr);
    return;
  }
  std::vector<std::thread> th;
  const uint32_t per = (nr + nt - 1) / nt;
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t a = t * per, e = std::min(nr, a + per);
    if (a < e) th.emplace_back(work, a, e);
  }
  for (auto& x : th) x.join();
}

// code counts of every SNP (F, src/FilePlink.cpp:36-62) by one pass over the
// mapped rows in 64 MB pieces
void Engine::stream_counts() {
  const uint64_t nb = (N_ + 3) >> 2;
  const uint32_t rb = (uint32_t)round_up(nb, 64);
  const uint32_t R = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(M_, (64ull << 20) / nb));
  Pinned stage((size_t)R * nb);
  DevBuf<uint8_t> raw, snp;
  DevBuf<int32_t> cnt;
  raw.reserve((uint64_t)R * nb);
  snp.reserve((uint64_t)R * rb);
  cnt.reserve((uint64_t)R * 4);
  std::vector<int32_t> h((size_t)M_ * 4);
  for (uint32_t r0 = 0; r0 < M_; r0 += R) {
    const uint32_t nr = std::min(R, M_ - r0);
    gather_rows(stage.p, r0, nr, false);
    PCAONE_HIP(hipMemcpyAsync(raw.p, stage.p, (size_t)nr * nb, hipMemcpyHostToDevice, stream_));
    timed("bed_layout", [&] { launch_bed_to_snpmajor(raw.p, N_, nr, nullptr, snp.p, rb, stream_); });
    timed("count_codes", [&] { launch_count_codes(snp.p, nr, rb, N_, cnt.p, stream_); });
    PCAONE_HIP(hipMemcpyAsync(h.data() + (size_t)r0 * 4, cnt.p, (size_t)nr * 16, hipMemcpyDeviceToHost, stream_));
    sync();  // the staging is refilled next
  }
  set_code_counts(h);
}

// the ring for the current windows (bs_ rows each)
void Engine::stream_setup() {
  sync();
  if (copy_stream_) PCAONE_HIP(hipStreamSynchronize(copy_stream_));
  const uint64_t nb = (N_ + 3) >> 2;
  for (auto& sl : slots_) {
    sl.raw.reserve((uint64_t)bs_ * nb);
    sl.snp.reserve((uint64_t)bs_ * row_bytes_);
    sl.smp.reserve((uint64_t)npad_ * win_bytes_);
    if (sl.stage_bytes < (size_t)bs_ * nb) {
      if (sl.stage) (void)hipHostFree(sl.stage);
      sl.stage = nullptr;
      sl.stage_bytes = 0;
      PCAONE_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.stage), (size_t)bs_ * nb, 0));
      sl.stage_bytes = (size_t)bs_ * nb;
    }
    if (!sl.copied) PCAONE_HIP(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
    if (!sl.used) PCAONE_HIP(hipEventCreateWithFlags(&sl.used, hipEventDisableTiming));
    sl.copy_rec = sl.used_rec = sl.laid = false;
    sl.win = -1;
  }
  last_slot_ = -1;
}

// window w's rows into slot s: host gather into the staging once its last
// copy is done, then the copy, after the last work that read the slot
void Engine::stream_fill(int s, uint32_t w) {
  StreamSlot& sl = slots_[s];
  const uint64_t nb = (N_ + 3) >> 2;
  const uint32_t nr = win_rows(w);
  if (sl.copy_rec) PCAONE_HIP(hipEventSynchronize(sl.copied));
  gather_rows(sl.stage, w * bs_, nr, true);
  if (sl.used_rec) PCAONE_HIP(hipStreamWaitEvent(copy_stream_, sl.used, 0));
  PCAONE_HIP(hipMemcpyAsync(sl.raw.p, sl.stage, (size_t)nr * nb, hipMemcpyHostToDevice, copy_stream_));
  PCAONE_HIP(hipEventRecord(sl.copied, copy_stream_));
  sl.copy_rec = true;
  sl.win = w;
  sl.laid = false;
}

int Engine::stream_acquire(uint32_t w) {
  PCAONE_REQUIRE(w < nwin_, -1, "stream: window out of range");
  int s = -1;
  for (int t = 0; t < 2; ++t)
    if (slots_[t].win == (int64_t)w) s = t;
  if (s < 0) {
    s = last_slot_ == 0 ? 1 : 0;
    stream_fill(s, w);
  }
  StreamSlot& sl = slots_[s];
  if (!sl.laid) {
    const uint32_t nr = win_rows(w);
    PCAONE_HIP(hipStreamWaitEvent(stream_, sl.copied, 0));
    timed("stream_layout", [&] { launch_bed_to_snpmajor(sl.raw.p, N_, nr, nullptr, sl.snp.p, row_bytes_, stream_); });
    // code 01 (missing, contributes nothing) past the window's last SNP
    PCAONE_HIP(hipMemsetAsync(sl.smp.p, 0x55, sl.smp.bytes(), stream_));
    timed("stream_layout", [&] {
      launch_transpose_windows(sl.snp.p, row_bytes_, nr, bs_, 1, sl.smp.p, npad_, win_bytes_, stream_);
    });
    sl.laid = true;
  }
  last_slot_ = s;
  return s;
}

void Engine::stream_release(int s) {
  PCAONE_HIP(hipEventRecord(slots_[s].used, stream_));
  slots_[s].used_rec = true;
}

void Engine::stream_prefetch(uint32_t w) {
  for (int t = 0; t < 2; ++t)
    if (slots_[t].win == (int64_t)w) return;
  stream_fill(last_slot_ == 0 ? 1 : 0, w);
}

// Both missing lists (SNP rows; window x sample rows) by two passes over the
// windows: counts, then positions, exactly as the resident layout builds them.
void Engine::stream_missing_lists() {
  const uint64_t rows_win = (uint64_t)nwin_ * npad_;
  DevBuf<int32_t> cs, cw;
  cs.reserve(M_);
  cw.reserve(rows_win);
  PCAONE_HIP(hipMemsetAsync(cw.p, 0, rows_win * 4, stream_));
  auto rows_of = [&](int s, uint32_t w, PackedRows& ps, PackedRows& pw) {
    const uint32_t nr = win_rows(w);
    ps = PackedRows{slots_[s].snp.p, row_bytes_, nr, nr, nr, N_, N_, 1, 0, 0};
    pw = PackedRows{slots_[s].smp.p, win_bytes_, npad_, npad_, N_, nr, nr, 1, 0, w * bs_};
  };
  for (uint32_t w = 0; w < nwin_; ++w) {
    const int s = stream_acquire(w);
    PackedRows ps, pw;
    rows_of(s, w, ps, pw);
    timed("missing_lists", [&] { launch_count_missing(ps, cs.p + (uint64_t)w * bs_, stream_); });
    timed("missing_lists", [&] { launch_count_missing(pw, cw.p + (uint64_t)w * npad_, stream_); });
    stream_release(s);
    stream_prefetch((w + 1) % nwin_);
  }
  auto prefix = [&](const DevBuf<int32_t>& c, uint64_t n, DevBuf<int64_t>& ptr) -> int64_t {
    std::vector<int32_t> hc(n);
    PCAONE_HIP(hipMemcpyAsync(hc.data(), c.p, n * 4, hipMemcpyDeviceToHost, stream_));
    sync();
    std::vector<int64_t> hp(n + 1, 0);
    for (uint64_t r = 0; r < n; ++r) hp[r + 1] = hp[r] + hc[r];
    ptr.reserve(n + 1);
    PCAONE_HIP(hipMemcpyAsync(ptr.p, hp.data(), (n + 1) * 8, hipMemcpyHostToDevice, stream_));
    sync();  // hp must outlive the copy
    return hp.back();
  };
  nnz_snp_ = prefix(cs, M_, miss_snp_ptr_);
  const int64_t nnz_win = prefix(cw, rows_win, miss_win_ptr_);
  PCAONE_REQUIRE(nnz_win == nnz_snp_, -2, "missing lists disagree between the two layouts");
  miss_snp_idx_.reserve(std::max<int64_t>(nnz_snp_, 1));
  miss_win_idx_.reserve(std::max<int64_t>(nnz_win, 1));
  for (uint32_t w = 0; w < nwin_; ++w) {
    const int s = stream_acquire(w);
    PackedRows ps, pw;
    rows_of(s, w, ps, pw);
    timed("missing_lists", [&] {
      launch_fill_missing(ps, miss_snp_ptr_.p + (uint64_t)w * bs_, miss_snp_idx_.p, stream_);
    });
    timed("missing_lists", [&] {
      launch_fill_missing(pw, miss_win_ptr_.p + (uint64_t)w * npad_, miss_win_idx_.p, stream_);
    });
    stream_release(s);
    stream_prefetch((w + 1) % nwin_);
  }
  miss_win_pos_.reserve(std::max<int64_t>(nnz_win, 1));
  timed("missing_lists", [&] {
    launch_map_missing(miss_win_ptr_.p, miss_win_idx_.p, nwin_, npad_, N_, miss_snp_ptr_.p, miss_snp_idx_.p,
                       miss_win_pos_.p, stream_);
  });
  sync();
}

}  // namespace pcaone
