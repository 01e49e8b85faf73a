// Shared host/device helpers for libbrianmi (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "devcommon.hpp"
#include <stdexcept>
#include <string>

namespace bm {

// Error carried to the C-ABI boundary as (code, message).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define BM_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      throw ::bm::Error(e_ == hipErrorOutOfMemory ? 3 : 2,                        \
                        std::string(#call) + ": " + hipGetErrorString(e_));       \
    }                                                                             \
  } while (0)

#define BM_REQUIRE(cond, msg)                                                     \
  do {                                                                            \
    if (!(cond)) throw ::bm::Error(1, std::string(msg));                          \
  } while (0)

// Run a C-ABI body: translate exceptions into codes + last error text.
template <typename F>
int guard(F &&f) {
  try {
    f();
    return 0;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return 2;
  } catch (...) {
    set_last_error("unknown error");
    return 2;
  }
}

// Device buffer owned by the library.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n) {
    release();
    if (n == 0) return;
    BM_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace bm
