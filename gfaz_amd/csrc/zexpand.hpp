// Device side of expansion through prefix-summed rule bodies (Z), shared by the
// whole-batch decoder and the resident (extract / serving) decoder. See the
// comment at the top of expand.hip for the arithmetic.
#pragma once

#include "codec.hpp"

namespace gfz {
namespace zx {

__device__ inline bool is_rule(int32_t v, uint32_t min_id, uint64_t R) {
  const uint32_t a = uabs(v);
  return a >= min_id && (uint64_t)(a - min_id) < R;
}

// ---- block order ------------------------------------------------------------------
// Which tile a block takes. The tiles [0, rows * cols) are read as `rows` rows of
// `cols` (row r = tiles [r * cols, (r + 1) * cols)); blocks take them column by
// column, and whole columns are dealt to the XCDs in turn (the dispatcher hands
// consecutive blocks to the eight XCDs round-robin), so one XCD's L2 holds what
// the same column of every row gathers. With one traversal per row that column
// is the same stretch of every haplotype, which gathers the same rule bodies.
// Returns UINT64_MAX for the padding blocks of a short last column group;
// rows <= 1 keeps the launch order.
constexpr uint32_t kXcds = 8;

__device__ inline uint64_t block_tile(uint32_t rows, uint64_t cols) {
  const uint64_t b = blockIdx.x;
  if (rows <= 1)
    return b;
  const uint64_t i = b / kXcds;
  const uint64_t c = (i / rows) * kXcds + b % kXcds;
  return c < cols ? (i % rows) * cols + c : UINT64_MAX;
}

// Grid of a launch ordered by block_tile over n_tiles tiles (cols set for it).
inline uint64_t row_grid(uint64_t n_tiles, uint32_t rows, uint64_t &cols) {
  if (rows <= 1) {
    cols = n_tiles;
    return n_tiles;
  }
  cols = (n_tiles + rows - 1) / rows;
  return (uint64_t)kXcds * rows * ((cols + kXcds - 1) / kXcds);
}

// ---- symbols of an expansion ------------------------------------------------------
// What one symbol writes at its local position q: base + Z[zidx0 + q], or
// base + Z[zidx0 - q] when bit 31 of zz (reverse orientation) is set, with
// zidx0 = zz & 0x7fffffff. A node ID (or the leading zero of a Z being built)
// points at the two reserved zero entries at the end of Z, so every symbol
// takes the same path. With kDiff (no inverse delta wanted) the value is
// Z[j] - Z[j -/+ 1] instead, a body element.
constexpr uint32_t kRevBit = 0x80000000u;

// szo[r] = (expanded size, sum, Z offset) of rule r.
__device__ inline SymInfo rule_sym(int32_t v, uint32_t pre, bool delta, const uint4 *szo,
                                   uint32_t min_id) {
  const uint4 w = szo[uabs(v) - min_id];
  SymInfo s;
  if (v > 0) {
    s.base = pre;
    s.zz = w.z + 1;
  } else {
    s.base = delta ? pre - w.y : pre;
    s.zz = (w.z + w.x - 1) | kRevBit;
  }
  return s;
}

// Bitmap of symbol starts in a tile, with per-64-bit-word ranks, so the symbol
// covering a position is rank + popcount.
template <uint32_t kTile>
struct StartMap {
  unsigned long long bits[kTile / 64];
  uint32_t rank[kTile / 64];

  __device__ void clear() {
    for (uint32_t w = threadIdx.x; w < kTile / 64; w += kBlock)
      bits[w] = 0;
  }
  __device__ void mark(uint32_t rel) { atomicOr(&bits[rel >> 6], 1ull << (rel & 63)); }
  // Call after a barrier that follows the last mark(); ends with a barrier.
  __device__ void rank_words() {
    static_assert(kTile / 64 <= kBlock, "one thread per bitmap word");
    const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    constexpr uint32_t kWords = kTile / 64;
    __shared__ uint32_t wsum[kBlock / kWave];
    uint32_t c = threadIdx.x < kWords ? (uint32_t)__popcll(bits[threadIdx.x]) : 0u;
    uint32_t x = c;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t g = __shfl_up(x, d);
      if (lane >= (uint32_t)d)
        x += g;
    }
    if (lane == kWave - 1)
      wsum[wid] = x;
    __syncthreads();
    uint32_t carry = 0;
    for (uint32_t w = 0; w < wid; ++w)
      carry += wsum[w];
    if (threadIdx.x < kWords)
      rank[threadIdx.x] = carry + x - c;
    __syncthreads();
  }
  __device__ uint32_t symbol(uint32_t p) const {
    const uint32_t w = p >> 6, b = p & 63;
    const unsigned long long upto = b == 63 ? ~0ull : (2ull << b) - 1ull;
    return rank[w] + (uint32_t)__popcll(bits[w] & upto) - 1u;
  }
};

// Shared memory of one expansion tile.
template <uint32_t kItems>
struct ZTileSmem {
  static constexpr uint32_t kTile = kBlock * kItems;
  StartMap<kTile> map;
  uint32_t sbase[kTile];
  uint32_t sstart[kTile];
  uint32_t srev[kTile / 32];
};

// The three steps of a tile: clear, stage every covering symbol (index s from
// 0, its start rel relative to the tile's first position: the first symbol may
// start before it), then expand. Block-wide except ztile_stage.
template <uint32_t kItems>
__device__ inline void ztile_begin(ZTileSmem<kItems> &sm) {
  constexpr uint32_t kTile = kBlock * kItems;
  sm.map.clear();
  for (uint32_t w = threadIdx.x; w < kTile / 32; w += kBlock)
    sm.srev[w] = 0;
  __syncthreads();
}
template <uint32_t kItems>
__device__ inline void ztile_stage(ZTileSmem<kItems> &sm, uint32_t s, uint32_t rel,
                                   const SymInfo &si) {
  const bool rev = si.zz & kRevBit;
  const uint32_t zi = si.zz & ~kRevBit;
  sm.sbase[s] = si.base;
  sm.sstart[s] = rev ? zi + rel : zi - rel;
  if (rev)
    atomicOr(&sm.srev[s >> 5], 1u << (s & 31));
  sm.map.mark(s == 0 ? 0u : rel);
}
// Each thread issues all of its Z gathers before its stores; ends with no
// barrier pending (LDS free for the caller's next tile).
template <uint32_t kItems, bool kDiff>
__device__ inline void ztile_finish(ZTileSmem<kItems> &sm, uint32_t n_here, const int32_t *Z,
                                    uint32_t zraw, int32_t *out_tile) {
  __syncthreads();
  sm.map.rank_words();
  uint32_t jz[kItems], vb[kItems];
  uint32_t revm = 0;
#pragma unroll
  for (uint32_t j = 0; j < kItems; ++j) {
    const uint32_t p = j * kBlock + threadIdx.x;
    const uint32_t s = sm.map.symbol(p < n_here ? p : n_here - 1);
    const uint32_t rv = (sm.srev[s >> 5] >> (s & 31)) & 1u;
    revm |= rv << j;
    jz[j] = p < n_here ? (rv ? sm.sstart[s] - p : sm.sstart[s] + p) : zraw;
    vb[j] = sm.sbase[s];
  }
  uint32_t z[kItems];
#pragma unroll
  for (uint32_t j = 0; j < kItems; ++j)
    z[j] = (uint32_t)Z[jz[j]];
  if (kDiff) {
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j)
      z[j] -= (uint32_t)Z[((revm >> j) & 1u) ? jz[j] + 1 : jz[j] - 1];
  }
#pragma unroll
  for (uint32_t j = 0; j < kItems; ++j) {
    const uint32_t p = j * kBlock + threadIdx.x;
    if (p < n_here)
      out_tile[p] = (int32_t)(vb[j] + z[j]);
  }
  __syncthreads();
}

// Expand the n_here (<= tile) output positions starting at decoded position
// `origin` into out_tile[0..n_here): symbols k0 .. k0 + nsym - 1 cover them,
// symbol k starting at sym_off[k] (the first may start before origin). Each
// symbol's base and the Z index of tile position 0 are staged in LDS. Block-wide:
// every thread of the block must call it.
template <uint32_t kItems, bool kDiff>
__device__ inline void expand_tile(ZTileSmem<kItems> &sm, const SymInfo *__restrict__ syms,
                                   const uint64_t *__restrict__ sym_off, uint64_t k0,
                                   uint32_t nsym, uint64_t origin, uint32_t n_here,
                                   const int32_t *Z, uint32_t zraw, int32_t *out_tile) {
  ztile_begin(sm);
  for (uint32_t s = threadIdx.x; s < nsym; s += kBlock)
    ztile_stage(sm, s, (uint32_t)(sym_off[k0 + s] - origin), syms[k0 + s]);
  ztile_finish<kItems, kDiff>(sm, n_here, Z, zraw, out_tile);
}

}  // namespace zx
}  // namespace gfz
