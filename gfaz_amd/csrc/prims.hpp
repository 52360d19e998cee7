// Device-wide primitives written for gfx950 (no hipCUB/rocPRIM on this path):
// exclusive scans, a sample sort of distinct 64-bit keys, fills, head bitmaps.
#pragma once

#include "ctx.hpp"

#include <initializer_list>

namespace gfz {

// Block-wide exclusive scan for NT threads; `lds` needs NT/kWave + 1 slots.
// Returns this thread's exclusive prefix and the block total.
template <uint32_t NT, typename T>
__device__ inline T block_exclusive_scan(T v, T *lds, T &total) {
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  T x = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T y = __shfl_up(x, d);
    if (lane >= (uint32_t)d)
      x += y;
  }
  if (lane == kWave - 1)
    lds[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (uint32_t w = 0; w < NT / kWave; ++w) {
      T t = lds[w];
      lds[w] = run;
      run += t;
    }
    lds[NT / kWave] = run;
  }
  __syncthreads();
  T res = lds[wid] + x - v;
  total = lds[NT / kWave];
  __syncthreads();
  return res;
}

template <typename T>
__device__ inline T wave_reduce_max(T v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    T o = __shfl_xor(v, d);
    v = o > v ? o : v;
  }
  return v;
}

// out[0..n) = exclusive prefix of in, out[n] = total (out holds n + 1).
void exclusive_scan(Ctx &ctx, const uint32_t *in, uint64_t *out, uint64_t n);
void exclusive_scan(Ctx &ctx, const uint64_t *in, uint64_t *out, uint64_t n);

// Ascending sort of n distinct 64-bit keys below 2^key_bits (tmp holds n): a
// sample sort up to 512 K keys, an LSD radix sort beyond (over the bytes of
// key_bits). local_cap (0 = the default 4096) bounds the sample sort's buckets
// sorted in registers; larger ones take its in-memory radix path (a smaller
// cap lets tests reach it). err (device) is set to 4 if a radix look-back wait
// hit its bound.
void sort_unique_u64(Ctx &ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, uint32_t *err,
                     uint32_t local_cap = 0, uint32_t key_bits = 64);

// Positions where a[i] != b[i] over n int32 (synchronises).
uint64_t count_mismatch_i32(Ctx &ctx, const int32_t *a, const int32_t *b, uint64_t n);

void fill_u32(Ctx &ctx, uint32_t *p, uint32_t v, uint64_t n);
void fill_u64(Ctx &ctx, uint64_t *p, uint64_t v, uint64_t n);
void fill_u8(Ctx &ctx, uint8_t *p, uint8_t v, uint64_t n);
// Several fills in one launch: n elements of `width` bytes (1, 4 or 8) at ptr
// set to the low bytes of value; at most kMaxFillRuns runs.
struct FillRun {
  void *ptr;
  uint64_t n;
  uint64_t value;
  uint32_t width;
};
constexpr uint32_t kMaxFillRuns = 8;
struct FillRuns {
  FillRun run[kMaxFillRuns];
  uint32_t count;
};
void fill_runs(Ctx &ctx, std::initializer_list<FillRun> runs);

// Set bit off[t] of `heads` (pre-zeroed, ceil(n_pos/32) words) for every
// t < n_off with off[t] < n_pos.
void set_head_bits(Ctx &ctx, const uint64_t *off, uint64_t n_off, uint32_t *heads,
                   uint64_t n_pos);

// Positions of the set bits of `bits` (n_bits bits) in ascending order into
// `pos` (at most pos_cap written); returns the full count (synchronises).
uint64_t bitmap_positions(Ctx &ctx, const uint32_t *bits, uint64_t n_bits, uint64_t *pos,
                          uint64_t pos_cap);

}  // namespace gfz
