// Shared host/device helpers for libbrianmi (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
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

// splitmix64 finaliser; the Poisson sources use two rounds keyed by
// (seed, stream) and the step counter (same integer arithmetic on host tests).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint32_t counter_u32(uint64_t seed, uint64_t stream, uint64_t counter) {
  const uint64_t G = 0x9E3779B97F4A7C15ull;
  uint64_t key = mix64(seed * G + stream);
  return (uint32_t)(mix64(key + counter * G) >> 32);
}

// Cheaper per-step stream for inputs drawn every step (batched STDP): a
// 32-bit key per stream, then MurmurHash3's finaliser of a Weyl sequence
// key + counter * 0x9E3779B9 (a bijection of the counter for each key).
__host__ __device__ inline uint32_t stream_key32(uint64_t seed, uint64_t stream) {
  const uint64_t G = 0x9E3779B97F4A7C15ull;
  return (uint32_t)(mix64(seed * G + stream) >> 32);
}
__host__ __device__ inline uint32_t weyl_u32(uint32_t key, uint32_t counter) {
  uint32_t h = key + counter * 0x9E3779B9u;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

}  // namespace bm
