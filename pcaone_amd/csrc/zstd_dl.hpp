// zstd from the system's libzstd.so.1, opened at run time (the library is not
// linked, so a host without it still loads everything that does not read
// zstd data).  The streaming decoder serves the CSV reader, the one-shot
// calls the BGEN genotype blocks and the BGEN writer of the simulator.
#pragma once

#include <dlfcn.h>

#include <cstddef>
#include <string>

#include "common.hpp"

namespace pcaone {

// the zstd streaming buffers (zstd.h, stable since v1.3)
struct ZInBuf {
  const void* src;
  size_t size;
  size_t pos;
};
struct ZOutBuf {
  void* dst;
  size_t size;
  size_t pos;
};

struct Zstd {
  void* so = nullptr;
  void* (*create)() = nullptr;
  size_t (*release)(void*) = nullptr;
  size_t (*stream)(void*, ZOutBuf*, ZInBuf*) = nullptr;
  size_t (*decompress)(void*, size_t, const void*, size_t) = nullptr;
  size_t (*compress)(void*, size_t, const void*, size_t, int) = nullptr;
  size_t (*bound)(size_t) = nullptr;
  unsigned (*is_error)(size_t) = nullptr;
  const char* (*error_name)(size_t) = nullptr;
  explicit Zstd(const char* who) {
    so = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
    PCAONE_REQUIRE(so != nullptr, -1, std::string(who) + ": the data is zstd-compressed and libzstd.so.1 is not available");
    create = reinterpret_cast<void* (*)()>(dlsym(so, "ZSTD_createDCtx"));
    release = reinterpret_cast<size_t (*)(void*)>(dlsym(so, "ZSTD_freeDCtx"));
    stream = reinterpret_cast<size_t (*)(void*, ZOutBuf*, ZInBuf*)>(dlsym(so, "ZSTD_decompressStream"));
    decompress = reinterpret_cast<size_t (*)(void*, size_t, const void*, size_t)>(dlsym(so, "ZSTD_decompress"));
    compress = reinterpret_cast<size_t (*)(void*, size_t, const void*, size_t, int)>(dlsym(so, "ZSTD_compress"));
    bound = reinterpret_cast<size_t (*)(size_t)>(dlsym(so, "ZSTD_compressBound"));
    is_error = reinterpret_cast<unsigned (*)(size_t)>(dlsym(so, "ZSTD_isError"));
    error_name = reinterpret_cast<const char* (*)(size_t)>(dlsym(so, "ZSTD_getErrorName"));
    PCAONE_REQUIRE(create && release && stream && decompress && compress && bound && is_error && error_name, -1,
                   std::string(who) + ": libzstd.so.1 lacks the API");
  }
  ~Zstd() {
    if (so) dlclose(so);
  }
  Zstd(const Zstd&) = delete;
  Zstd& operator=(const Zstd&) = delete;
};

}  // namespace pcaone
