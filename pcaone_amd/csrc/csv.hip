// The CSV reader of single-cell counts (FileCsv::read_all and parse_csvzstd,
// src/FileCsv.cpp:11-62, 82-142): one feature per line, one sample per
// comma-separated field, the file zstd-compressed (as the reference reads it)
// or plain text.  The host decompresses and parses the lines on worker
// threads into an FP32 feature-major matrix; the library sizes (per sample,
// the sum over features of each field read as an integer, as std::stoi reads
// it) come from the same pass.  On the device one block per feature applies
// the --scale transform in FP64 and standardises the feature over samples
// (src/Utils.cpp:34-43), and the result enters the dense path of load_dense.
//   scale 2 (cpmed): log10(count * median(libsize) / libsize + 1)
//   scale 3:         log1p(count / libsize * scale_factor)
//   scale 4:         count / libsize * scale_factor
//   scale 0, 1:      count (1: standardised, 0: as is)
// zstd is the system's libzstd.so.1, opened at run time.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "zstd_dl.hpp"

namespace pcaone {
namespace {

// std::stoi's reading of a field: leading blanks, sign, digits (none: an error)
bool field_int(const char* b, const char* e, int64_t& v) {
  while (b < e && (*b == ' ' || *b == '\t')) ++b;
  bool neg = false;
  if (b < e && (*b == '+' || *b == '-')) neg = *b++ == '-';
  if (b == e || *b < '0' || *b > '9') return false;
  int64_t x = 0;
  while (b < e && *b >= '0' && *b <= '9') x = x * 10 + (*b++ - '0');
  v = neg ? -x : x;
  return true;
}

// A batch of complete lines [0, n) starting at `text + starts[i]` (ends exclusive of '\n').
struct LineBatch {
  std::vector<size_t> start, end;
};

}  // namespace

struct CsvParser {
  uint32_t ncol = 0;
  uint64_t nrow = 0;
  std::vector<float> values;     // feature-major
  std::vector<int64_t> libsize;  // per sample
  bool non_int = false;  // a field std::stoi would refuse (only --scale 2 reads them as integers)
  int nthreads = 1;

  // parses the complete lines of buf; returns the bytes consumed
  size_t feed(const std::string& buf) {
    LineBatch lb;
    size_t p = 0;
    for (;;) {
      const size_t e = buf.find('\n', p);
      if (e == std::string::npos) break;
      lb.start.push_back(p);
      lb.end.push_back(e);
      p = e + 1;
    }
    const size_t n = lb.start.size();
    if (!n) return 0;
    if (ncol == 0) {  // the column count from the first line
      uint32_t c = 1;
      for (size_t i = lb.start[0]; i < lb.end[0]; ++i) c += buf[i] == ',';
      ncol = c;
      libsize.assign(ncol, 0);
    }
    const uint64_t r0 = nrow;
    values.resize((size_t)(nrow + n) * ncol);
    const int nt = std::max(1, std::min<int>(nthreads, (int)std::max<size_t>(1, n / 64)));
    std::vector<std::vector<int64_t>> lib(nt, std::vector<int64_t>(ncol, 0));
    std::vector<std::string> errs(nt);
    std::vector<uint8_t> bad_int(nt, 0);
    auto work = [&](int t) {
      for (size_t l = (size_t)t; l < n; l += (size_t)nt) {
        const char* s = buf.data() + lb.start[l];
        const char* e = buf.data() + lb.end[l];
        float* row = values.data() + (size_t)(r0 + l) * ncol;
        uint32_t c = 0;
        const char* f = s;
        for (const char* q = s;; ++q) {
          if (q == e || *q == ',') {
            if (c >= ncol) {
              errs[t] = "load_csv: the csv file has unaligned columns (line " + std::to_string(r0 + l + 1) + ")";
              return;
            }
            std::string field(f, q);
            char* endp = nullptr;
            errno = 0;
            const float v = std::strtof(field.c_str(), &endp);
            int64_t iv = 0;
            if (endp == field.c_str()) {
              errs[t] = "load_csv: not a number at line " + std::to_string(r0 + l + 1) + ": '" + field + "'";
              return;
            }
            if (!field_int(f, q, iv)) bad_int[t] = 1;
            row[c] = v;
            lib[t][c] += iv;
            ++c;
            f = q + 1;
            if (q == e) break;
          }
        }
        if (c != ncol) {
          errs[t] = "load_csv: the csv file has unaligned columns (line " + std::to_string(r0 + l + 1) + ")";
          return;
        }
      }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    for (auto& e : errs) PCAONE_REQUIRE(e.empty(), -1, e);
    for (int t = 0; t < nt; ++t) non_int = non_int || bad_int[t];
    for (int t = 0; t < nt; ++t)
      for (uint32_t c = 0; c < ncol; ++c) libsize[c] += lib[t][c];
    nrow += n;
    return p;
  }
};

void read_csv(const char* path, int nthreads, CsvParser& out) {
  PCAONE_REQUIRE(path, -1, "load_csv: no file name");
  FILE* fp = std::fopen(path, "rb");
  PCAONE_REQUIRE(fp != nullptr, -1, std::string("load_csv: can not open ") + path);
  out.nthreads = std::max(1, nthreads);
  std::string pending;
  std::vector<char> in(1 << 22);
  try {
    size_t got = std::fread(in.data(), 1, in.size(), fp);
    const bool zst = got >= 4 && (unsigned char)in[0] == 0x28 && (unsigned char)in[1] == 0xB5 &&
                     (unsigned char)in[2] == 0x2F && (unsigned char)in[3] == 0xFD;
    auto consume = [&]() {
      const size_t used = out.feed(pending);
      pending.erase(0, used);
    };
    if (zst) {
      Zstd z("load_csv");
      void* dctx = z.create();
      PCAONE_REQUIRE(dctx != nullptr, -1, "load_csv: ZSTD_createDCtx failed");
      std::vector<char> ob(1 << 22);
      size_t last = 1;
      try {
        while (got > 0) {
          ZInBuf ib{in.data(), got, 0};
          // until the input is used up and the output buffer was not filled
          // (a full one may leave decoded bytes inside the decoder)
          for (;;) {
            ZOutBuf o{ob.data(), ob.size(), 0};
            last = z.stream(dctx, &o, &ib);
            PCAONE_REQUIRE(!z.is_error(last), -1, std::string("load_csv: ZSTD decompression failed: ") + z.error_name(last));
            pending.append(ob.data(), o.pos);
            if (ib.pos == ib.size && o.pos < o.size) break;
          }
          if (pending.size() >= (16u << 20)) consume();
          got = std::fread(in.data(), 1, in.size(), fp);
        }
      } catch (...) {
        z.release(dctx);
        throw;
      }
      z.release(dctx);
      PCAONE_REQUIRE(last == 0, -1, "load_csv: EOF before end of ZSTD_decompressStream");
    } else {
      while (got > 0) {
        pending.append(in.data(), got);
        if (pending.size() >= (16u << 20)) consume();
        got = std::fread(in.data(), 1, in.size(), fp);
      }
    }
    consume();  // a last line without '\n' is left out, as the reference counts lines
  } catch (...) {
    std::fclose(fp);
    throw;
  }
  std::fclose(fp);
  PCAONE_REQUIRE(out.nrow > 0 && out.ncol > 0, -1, "load_csv: input file is empty.");
}

namespace {

// one block per feature: transform in FP64, then centre and scale by the
// sample standard deviation (N - 1) when it exceeds 1e-10
__global__ __launch_bounds__(256) void k_csv_normalize(float* __restrict__ X, uint32_t n, const double* __restrict__ lib,
                                                       int scale, double median, double factor) {
  __shared__ double red[4];
  __shared__ double stat[2];
  float* row = X + (uint64_t)blockIdx.x * n;
  const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
  auto tr = [&](uint32_t j) {
    const double c = (double)row[j];
    if (scale == 2) return log10(c * median / lib[j] + 1.0);
    if (scale == 3) return log1p(c / lib[j] * factor);
    if (scale == 4) return c / lib[j] * factor;
    return c;
  };
  auto block_sum = [&](double x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o);
    __syncthreads();  // red is reused
    if (lane == 0) red[wv] = x;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
  };
  double s = 0.0;
  for (uint32_t j = t; j < n; j += blockDim.x) s += tr(j);
  const double mean = block_sum(s) / (double)n;
  double q = 0.0;
  for (uint32_t j = t; j < n; j += blockDim.x) {
    const double d = tr(j) - mean;
    q += d * d;
  }
  const double sd = sqrt(block_sum(q)) / sqrt((double)n - 1.0);
  if (t == 0) {
    stat[0] = mean;
    stat[1] = sd;
  }
  __syncthreads();
  const double m = stat[0], d = stat[1];
  const bool div = scale > 0 && d > 1e-10;
  for (uint32_t j = t; j < n; j += blockDim.x) {
    const double v = scale > 0 ? tr(j) - m : tr(j);
    row[j] = (float)(div ? v / d : v);
  }
}

}  // namespace

void read_csv_host(const char* path, int nthreads, float* values, uint64_t capacity, uint32_t* nsamples,
                   uint64_t* nfeatures, int64_t* libsize) {
  CsvParser p;
  read_csv(path, nthreads, p);
  if (nsamples) *nsamples = p.ncol;
  if (nfeatures) *nfeatures = p.nrow;
  if (!values) return;
  PCAONE_REQUIRE(capacity >= (uint64_t)p.ncol * p.nrow, -1, "read_csv: values too small");
  std::memcpy(values, p.values.data(), (size_t)p.ncol * p.nrow * sizeof(float));
  if (libsize) std::memcpy(libsize, p.libsize.data(), (size_t)p.ncol * sizeof(int64_t));
}

void Engine::load_csv(const char* path, int scale, double scale_factor, int nthreads) {
  PCAONE_REQUIRE(scale >= 0 && scale <= 4, -1, "load_csv: --scale must be 0..4");
  CsvParser p;
  read_csv(path, nthreads, p);
  PCAONE_REQUIRE(p.nrow <= 0xffffffffull, -1, "load_csv: too many features");
  PCAONE_REQUIRE(p.ncol >= 2, -1, "load_csv: needs at least two samples");
  PCAONE_REQUIRE(scale != 2 || !p.non_int, -1, "load_csv: --scale 2 reads every count as an integer (std::stoi)");
  double median = 0.0;
  if (scale == 2) {  // get_median of the int library sizes (integer mean of the middle pair)
    std::vector<int64_t> v = p.libsize;
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    median = (double)(n % 2 == 0 ? (v[n / 2 - 1] + v[n / 2]) / 2 : v[n / 2]);
  }
  PCAONE_HIP(hipSetDevice(device_));
  have_op_ = false;
  N_ = p.ncol;
  M_ = (uint32_t)p.nrow;
  rawf_.reserve((uint64_t)N_ * M_);
  PCAONE_HIP(hipMemcpyAsync(rawf_.p, p.values.data(), (size_t)N_ * M_ * 4, hipMemcpyHostToDevice, stream_));
  if (scale > 0) {
    std::vector<double> lib(p.libsize.begin(), p.libsize.end());
    DevBuf<double> dlib;
    dlib.reserve(N_);
    PCAONE_HIP(hipMemcpyAsync(dlib.p, lib.data(), (size_t)N_ * 8, hipMemcpyHostToDevice, stream_));
    timed("csv_normalize", [&] {
      hipLaunchKernelGGL(k_csv_normalize, dim3(M_), dim3(256), 0, stream_, rawf_.p, N_, dlib.p, scale, median,
                         scale_factor);
    });
    sync();  // lib and the host values
  }
  // the reference's CSV reader centres only through the standardisation
  // (scale > 0); the standardised features have mean 0, the raw ones keep theirs
  load_dense_finish(false);
}

}  // namespace pcaone
