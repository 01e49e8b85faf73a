// Integer hash helpers shared by the host code, the compiled kernels and the
// kernels specialised at run time (hipRTC): device-safe, no host headers.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <cstdint>
#else
// hipRTC declares the fixed-width integer types in its own namespace
using __hip_internal::int8_t;
using __hip_internal::int16_t;
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint8_t;
using __hip_internal::uint16_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#endif

namespace bm {

// splitmix64 finaliser; the Poisson sources use two rounds keyed by
// (seed, stream) and the step counter (same integer arithmetic on host tests).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t counter_key(uint64_t seed, uint64_t stream) {
  return mix64(seed * 0x9E3779B97F4A7C15ull + stream);
}
__host__ __device__ inline uint64_t counter_u64(uint64_t key, uint64_t counter) {
  return mix64(key + counter * 0x9E3779B97F4A7C15ull);
}
__host__ __device__ inline uint32_t counter_u32(uint64_t seed, uint64_t stream, uint64_t counter) {
  return (uint32_t)(counter_u64(counter_key(seed, stream), counter) >> 32);
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
