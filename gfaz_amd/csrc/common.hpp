// Shared device/host helpers for the gfx950 grammar codec.
#pragma once

#include <hip/hip_runtime.h>

#include "error.hpp"

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace gfz {

constexpr int kWave = 64;
constexpr uint32_t kBlock = 256;  // 4 waves

#define GFZ_HIP(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::gfz::Error(e_ == hipErrorOutOfMemory ? ::gfz::kErrOutOfMemory              \
                                                   : ::gfz::kErrHip,                      \
                         std::string("HIP error at " __FILE__ ":") +                     \
                             std::to_string(__LINE__) + " - " + hipGetErrorString(e_));  \
  } while (0)

#define GFZ_CHECK_LAUNCH() GFZ_HIP(hipGetLastError())

// ---- packed 2-mer (include/grammar/packed_2mer.hpp:15-48) -----------------
__host__ __device__ inline int64_t pack2(int32_t a, int32_t b) {
  return (int64_t)(((uint64_t)(uint32_t)a << 32) | (uint64_t)(uint32_t)b);
}
__host__ __device__ inline int32_t key_first(int64_t k) { return (int32_t)(k >> 32); }
__host__ __device__ inline int32_t key_second(int64_t k) { return (int32_t)k; }
__host__ __device__ inline int64_t reverse2(int64_t k) {
  return pack2(-key_second(k), -key_first(k));
}
__host__ __device__ inline int64_t canon2(int64_t k) {
  const int64_t r = reverse2(k);
  return k < r ? k : r;
}
__host__ __device__ inline uint32_t uabs(int32_t v) {
  return v < 0 ? (uint32_t)0 - (uint32_t)v : (uint32_t)v;
}

// Open-addressing hash of a canonical key; EMPTY = pack(INT32_MIN, 0), which no
// traversal or rule symbol can form (|symbol| < 2^31).
constexpr uint64_t kEmptyKey = 0x8000000000000000ull;
__host__ __device__ inline uint64_t key_hash(uint64_t k) {
  k ^= k >> 31;
  k *= 0x9E3779B97F4A7C15ull;
  k ^= k >> 29;
  return k;
}

// ---- wave helpers ---------------------------------------------------------
__device__ inline uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

inline uint64_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v)
    p <<= 1;
  return p;
}

}  // namespace gfz
