// Rule expansion and inverse delta on gfx950 (decompress / extract).
//
// Restates src/workflows/decompression_workflow.cpp:70-174 (recursive
// expand_rule, then delta_round inverse prefix sums per traversal):
//   expand(+r) = expand(first[r]) ++ expand(second[r])
//   expand(-r) = expand(-second[r]) ++ expand(-first[r])
//   a symbol is a rule iff min_id <= |s| < min_id + R.
// Instead of recursing per encoded symbol, the rulebook is pre-expanded once
// (bodies, built level by level on the device) and every output position finds
// its encoded symbol (a rank over a start bitmap staged in LDS) and reads one
// body element, so one lane writes one output element and a wave's stores are
// consecutive. A rulebook whose bodies do not fit the budget is expanded by
// descending the rule tree by expanded sizes instead.
#include "codec.hpp"
#include "prims.hpp"

#include <algorithm>
#include <cstdlib>

namespace gfz {

namespace {

constexpr uint32_t kExItems = 16;
constexpr uint32_t kExTile = kBlock * kExItems;  // output elements per block
constexpr uint32_t kSsItems = 16;
constexpr uint32_t kSsTile = kBlock * kSsItems;
constexpr uint32_t kSsScanThreads = 1024;

__device__ inline bool is_rule(int32_t v, uint32_t min_id, uint64_t R) {
  const uint32_t a = uabs(v);
  return a >= min_id && (uint64_t)(a - min_id) < R;
}

__global__ __launch_bounds__(kBlock) void k_rule_init(const int32_t *__restrict__ f,
                                                      const int32_t *__restrict__ s, uint64_t R,
                                                      uint32_t min_id, RuleRec *__restrict__ rec,
                                                      uint32_t *__restrict__ size,
                                                      uint32_t *__restrict__ depth) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * kBlock) {
    const int32_t a = f[r], b = s[r];
    RuleRec q;
    q.a = a;
    q.b = b;
    q.sa = is_rule(a, min_id, R) ? 0u : 1u;
    q.sb = is_rule(b, min_id, R) ? 0u : 1u;
    rec[r] = q;
    size[r] = (q.sa && q.sb) ? 2u : 0u;
    depth[r] = (q.sa && q.sb) ? 1u : 0u;
  }
}

// One bottom-up pass. In pass k a rule resolves when each child is a node ID or a
// rule resolved before this pass (depth <= k + 1); it then gets depth k + 2 and
// its size. Children resolved during the same pass are not used, so every depth
// is exact whatever the thread order. `changed` is raised once per wave.
__global__ __launch_bounds__(kBlock) void k_rule_pass(RuleRec *__restrict__ rec,
                                                      uint32_t *__restrict__ size,
                                                      uint32_t *__restrict__ depth, uint64_t R,
                                                      uint32_t min_id, uint32_t pass,
                                                      uint32_t *changed, uint32_t *err) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  for (uint64_t r0 = (uint64_t)blockIdx.x * kBlock; r0 < R; r0 += (uint64_t)gridDim.x * kBlock) {
    const uint64_t r = r0 + threadIdx.x;
    bool resolved = false;
    if (r < R && !depth[r]) {
      RuleRec q = rec[r];
      bool ok = true;
      uint32_t sa = q.sa, sb = q.sb;
      if (!sa) {
        const uint64_t c = uabs(q.a) - min_id;
        const uint32_t dc = depth[c];
        ok = ok && dc != 0 && dc <= pass + 1;
        if (ok)
          sa = size[c];
      }
      if (!sb) {
        const uint64_t c = uabs(q.b) - min_id;
        const uint32_t dc = depth[c];
        ok = ok && dc != 0 && dc <= pass + 1;
        if (ok)
          sb = size[c];
      }
      if (ok) {
        const uint32_t sum = sa + sb;
        if (sum < sa) {
          atomicExch(err, 1u);
        } else {
          q.sa = sa;
          q.sb = sb;
          rec[r] = q;
          size[r] = sum;
          depth[r] = pass + 2;
          resolved = true;
        }
      }
    }
    const uint64_t m = __ballot(resolved);
    if (m && lane == (uint32_t)__ffsll((unsigned long long)m) - 1)
      *changed = 1u;
  }
}

__global__ __launch_bounds__(kBlock) void k_rule_unresolved(const uint32_t *__restrict__ depth,
                                                            uint64_t R, uint32_t *err) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * kBlock)
    if (!depth[r])
      atomicExch(err, 2u);
}

__global__ __launch_bounds__(kBlock) void k_sym_sizes(const int32_t *__restrict__ enc, uint64_t n,
                                                      const uint32_t *__restrict__ size, uint64_t R,
                                                      uint32_t min_id, uint32_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    const int32_t v = enc[i];
    out[i] = is_rule(v, min_id, R) ? size[uabs(v) - min_id] : 1u;
  }
}

__global__ void k_gather_u64(const uint64_t *__restrict__ src, const uint64_t *__restrict__ idx,
                             uint64_t n, uint64_t *__restrict__ dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}

// Per output tile: the symbols covering its first and last positions and where
// the first one starts (relative to base). Symbols are disjoint, non-empty
// intervals, so every field has exactly one writer.
struct TileDesc {
  uint64_t k0;
  uint64_t first_off;
  uint64_t k_last;
  uint64_t pad;
};

__global__ void k_tile_desc(const uint64_t *__restrict__ sym_off, uint64_t n_sym, uint64_t base,
                            uint64_t n_out, uint64_t n_tiles, TileDesc *__restrict__ desc) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_sym;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = sym_off[k] - base, hi = sym_off[k + 1] - base;
    for (uint64_t t = lo / kExTile; t < n_tiles && t * kExTile < hi; ++t) {
      if (t * kExTile >= lo) {
        desc[t].k0 = k;
        desc[t].first_off = lo;
      }
      const uint64_t end = ((t + 1) * kExTile < n_out ? (t + 1) * kExTile : n_out) - 1;
      if (end >= lo && end < hi)
        desc[t].k_last = k;
    }
  }
}

// Output of one block: kExTile consecutive positions [base + o0, base + o_end).
// The symbols overlapping the tile are staged in LDS with a bitmap of their start
// positions; a position's symbol is the rank of the set bits at or before it.
// With kBodies, a rule symbol reads its pre-expanded body (reversed and negated
// for a reverse-oriented symbol); otherwise it descends the rule tree by sizes.
template <bool kBodies>
__global__ __launch_bounds__(kBlock) void k_expand(
    const int32_t *__restrict__ syms, const uint64_t *__restrict__ sym_off, uint64_t n_sym,
    uint64_t base, uint64_t n_out, const TileDesc *__restrict__ tdesc,
    const RuleRec *__restrict__ rec, const uint32_t *__restrict__ rsize,
    const uint64_t *__restrict__ body_off, const int32_t *__restrict__ bodies, uint64_t R,
    uint32_t min_id, int32_t *__restrict__ out) {
  __shared__ uint32_t srel[kExTile];
  __shared__ int32_t ssym[kExTile];
  __shared__ unsigned long long sbits[kExTile / 64];
  __shared__ uint32_t srank[kExTile / 64];
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1);
  const uint64_t o0 = (uint64_t)blockIdx.x * kExTile;
  const uint64_t o_end = o0 + kExTile < n_out ? o0 + kExTile : n_out;
  if (tid < kExTile / 64)
    sbits[tid] = 0;
  const TileDesc td = tdesc[blockIdx.x];
  const uint64_t k0 = td.k0;
  const uint32_t nsym = (uint32_t)(td.k_last - k0 + 1);
  (void)n_sym;
  __syncthreads();
  for (uint32_t s = tid; s < nsym; s += kBlock) {
    const uint32_t rel = s == 0 ? 0u : (uint32_t)(sym_off[k0 + s] - base - o0);
    srel[s] = rel;
    ssym[s] = syms[k0 + s];
    atomicOr(&sbits[rel >> 6], 1ull << (rel & 63));
  }
  __syncthreads();
  if (tid < kWave) {
    const uint32_t c = (uint32_t)__popcll(sbits[lane]);
    uint32_t x = c;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t g = __shfl_up(x, d);
      if (lane >= (uint32_t)d)
        x += g;
    }
    srank[lane] = x - c;
  }
  __syncthreads();
  const uint64_t first_off = td.first_off;
#pragma unroll 4
  for (uint32_t j = 0; j < kExItems; ++j) {
    const uint32_t p = j * kBlock + tid;
    const uint64_t o = o0 + p;
    if (o >= o_end)
      break;
    const uint32_t w = p >> 6, b = p & 63;
    const uint64_t upto = (2ull << b) - 1ull;
    const uint32_t s = srank[w] + (uint32_t)__popcll(sbits[w] & upto) - 1u;
    uint64_t q = s == 0 ? o - first_off : (uint64_t)(p - srel[s]);
    int32_t v = ssym[s];
    if (is_rule(v, min_id, R)) {
      if (kBodies) {
        const uint32_t r = uabs(v) - min_id;
        const uint64_t src = v > 0 ? body_off[r] + q : body_off[r] + (rsize[r] - 1 - q);
        const int32_t x = bodies[src];
        v = v > 0 ? x : -x;
      } else {
        do {
          const RuleRec c = rec[uabs(v) - min_id];
          int32_t left, right;
          uint32_t ls;
          if (v > 0) {
            left = c.a;
            right = c.b;
            ls = c.sa;
          } else {
            left = -c.b;
            right = -c.a;
            ls = c.sb;
          }
          if (q < ls) {
            v = left;
          } else {
            q -= ls;
            v = right;
          }
        } while (is_rule(v, min_id, R));
      }
    }
    out[base + o] = v;
  }
}

// Pair stream of the rulebook: symbols (first[r], second[r]) at body offsets
// (body_off[r], body_off[r] + |first[r]|), so that expanding it one level over
// already-built bodies writes the bodies of the next depth.
__global__ void k_pair_stream(const RuleRec *__restrict__ rec, const uint64_t *__restrict__ body_off,
                              uint64_t R, int32_t *__restrict__ ps, uint64_t *__restrict__ ps_off) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const RuleRec c = rec[r];
    ps[2 * r] = c.a;
    ps[2 * r + 1] = c.b;
    ps_off[2 * r] = body_off[r];
    ps_off[2 * r + 1] = body_off[r] + c.sa;
    if (r == R - 1)
      ps_off[2 * R] = body_off[R];
  }
}

// Level boundaries when depth is non-decreasing in rule ID (true for every
// rulebook the reference CPU compressor writes: a rule's depth is the round that
// created it, and rounds hand out increasing IDs). bound[d] = first rule of
// depth d; *bad = 1 if depth ever decreases.
__global__ void k_depth_bounds(const uint32_t *__restrict__ depth, uint64_t R, uint32_t max_depth,
                               uint64_t *__restrict__ bound, uint32_t *bad) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = depth[r];
    if (r == 0) {
      for (uint32_t e = 1; e <= d && e <= max_depth; ++e)
        bound[e] = 0;
    } else {
      const uint32_t p = depth[r - 1];
      if (d < p)
        atomicExch(bad, 1u);
      for (uint32_t e = p + 1; e <= d && e <= max_depth; ++e)
        bound[e] = r;
    }
  }
}

// ---- segmented inclusive prefix sum -----------------------------------------
// Value (sum, flag): flag = a segment head was seen. (a) then (b) =
// (b.flag ? b.sum : a.sum + b.sum, a.flag | b.flag).
struct Seg {
  uint32_t sum;
  uint32_t flag;
};
__device__ inline Seg seg_then(Seg a, Seg b) {
  return Seg{b.flag ? b.sum : a.sum + b.sum, a.flag | b.flag};
}

template <uint32_t NT>
__device__ inline Seg block_seg_exclusive(Seg v, Seg *lds, Seg &total) {
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  Seg x = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    Seg g;
    g.sum = __shfl_up(x.sum, d);
    g.flag = __shfl_up(x.flag, d);
    if (lane >= (uint32_t)d)
      x = seg_then(g, x);
  }
  if (lane == kWave - 1)
    lds[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    Seg run{0, 0};
    for (uint32_t w = 0; w < NT / kWave; ++w) {
      const Seg t = lds[w];
      lds[w] = run;
      run = seg_then(run, t);
    }
    lds[NT / kWave] = run;
  }
  __syncthreads();
  Seg ex;
  ex.sum = __shfl_up(x.sum, 1);
  ex.flag = __shfl_up(x.flag, 1);
  if (lane == 0)
    ex = Seg{0, 0};
  const Seg res = seg_then(lds[wid], ex);
  total = lds[NT / kWave];
  __syncthreads();
  return res;
}

__device__ inline bool bit_at(const uint32_t *bits, uint64_t i) {
  return (bits[i >> 5] >> (i & 31)) & 1u;
}

// ---- expansion fused with the first inverse delta ----------------------------
// One pass over the output: each block expands a tile (positions staged as in
// k_expand), runs the segmented prefix sum over the tile in LDS, and chains the
// carry across tiles with a decoupled look-back (tiles taken in order from an
// atomic ticket, so every tile a block waits on belongs to a block already
// running). A status word per tile packs (state, head flag, 32-bit sum).
constexpr uint64_t kStNotReady = 0, kStAggregate = 1ull << 62, kStInclusive = 2ull << 62;
constexpr uint32_t kPad = kExTile + kExTile / 16;  // LDS transpose with a 17-word stride

__device__ inline uint32_t pad_idx(uint32_t p) { return p + (p >> 4); }

__device__ inline uint64_t st_pack(uint64_t state, Seg v) {
  return state | ((uint64_t)(v.flag & 1) << 32) | v.sum;
}

template <bool kBodies>
__global__ __launch_bounds__(kBlock) void k_expand_delta(
    const int32_t *__restrict__ syms, const uint64_t *__restrict__ sym_off, uint64_t n_sym,
    uint64_t n_out, const TileDesc *__restrict__ tdesc, const RuleRec *__restrict__ rec,
    const uint32_t *__restrict__ rsize, const uint64_t *__restrict__ body_off,
    const int32_t *__restrict__ bodies, uint64_t R, uint32_t min_id,
    const uint32_t *__restrict__ heads, unsigned long long *status, uint32_t *ticket,
    uint32_t *err, int32_t *__restrict__ out) {
  // srel/ssym stage the tile's symbols; once every lane holds its expanded
  // values in registers the same LDS is reused as the transpose buffer sval.
  __shared__ uint32_t sbuf[2 * kExTile];
  static_assert(kPad <= 2 * kExTile, "transpose buffer fits the staging area");
  uint32_t *const srel = sbuf;
  int32_t *const ssym = reinterpret_cast<int32_t *>(sbuf + kExTile);
  uint32_t *const sval = sbuf;
  __shared__ unsigned long long sbits[kExTile / 64];
  __shared__ uint32_t srank[kExTile / 64];
  __shared__ Seg slds[kBlock / kWave + 1];
  __shared__ uint32_t stile;
  __shared__ Seg scarry;
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1);
  if (tid == 0)
    stile = atomicAdd(ticket, 1u);
  if (tid < kExTile / 64)
    sbits[tid] = 0;
  __syncthreads();
  const uint32_t tile = stile;
  const uint64_t o0 = (uint64_t)tile * kExTile;
  const uint64_t o_end = o0 + kExTile < n_out ? o0 + kExTile : n_out;
  const TileDesc td = tdesc[tile];
  const uint64_t k0 = td.k0;
  const uint32_t nsym = (uint32_t)(td.k_last - k0 + 1);
  (void)n_sym;
  for (uint32_t s = tid; s < nsym; s += kBlock) {
    const uint32_t rel = s == 0 ? 0u : (uint32_t)(sym_off[k0 + s] - o0);
    srel[s] = rel;
    ssym[s] = syms[k0 + s];
    atomicOr(&sbits[rel >> 6], 1ull << (rel & 63));
  }
  __syncthreads();
  if (tid < kWave) {
    const uint32_t c = (uint32_t)__popcll(sbits[lane]);
    uint32_t x = c;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t g = __shfl_up(x, d);
      if (lane >= (uint32_t)d)
        x += g;
    }
    srank[lane] = x - c;
  }
  __syncthreads();
  const uint64_t first_off = td.first_off;
  // Expand in the striped layout (consecutive lanes, consecutive positions).
  int32_t ev[kExItems];
#pragma unroll
  for (uint32_t j = 0; j < kExItems; ++j) {
    const uint32_t p = j * kBlock + tid;
    const uint64_t o = o0 + p;
    int32_t v = 0;
    if (o < o_end) {
      const uint32_t w = p >> 6, b = p & 63;
      const uint64_t upto = (2ull << b) - 1ull;
      const uint32_t s = srank[w] + (uint32_t)__popcll(sbits[w] & upto) - 1u;
      uint64_t q = s == 0 ? o - first_off : (uint64_t)(p - srel[s]);
      v = ssym[s];
      if (is_rule(v, min_id, R)) {
        if (kBodies) {
          const uint32_t r = uabs(v) - min_id;
          const uint64_t src = v > 0 ? body_off[r] + q : body_off[r] + (rsize[r] - 1 - q);
          const int32_t x = bodies[src];
          v = v > 0 ? x : -x;
        } else {
          do {
            const RuleRec c = rec[uabs(v) - min_id];
            int32_t left, right;
            uint32_t ls;
            if (v > 0) {
              left = c.a;
              right = c.b;
              ls = c.sa;
            } else {
              left = -c.b;
              right = -c.a;
              ls = c.sb;
            }
            if (q < ls) {
              v = left;
            } else {
              q -= ls;
              v = right;
            }
          } while (is_rule(v, min_id, R));
        }
      }
    }
    ev[j] = v;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kExItems; ++j)
    sval[pad_idx(j * kBlock + tid)] = (uint32_t)ev[j];
  __syncthreads();
  // Blocked layout for the scan: thread t owns positions [16t, 16t + 16).
  const uint32_t pb = tid * kExItems;
  uint32_t hbits = 0;
  if (o0 + pb < o_end) {
    const uint64_t g = o0 + pb;  // 16-aligned, so the 16 bits sit in one word
    hbits = (heads[g >> 5] >> (g & 31)) & 0xFFFFu;
  }
  uint32_t vals[kExItems];
  Seg a{0, 0};
#pragma unroll
  for (uint32_t i = 0; i < kExItems; ++i) {
    vals[i] = sval[pad_idx(pb + i)];
    a = seg_then(a, Seg{vals[i], (hbits >> i) & 1u});
  }
  Seg total;
  const Seg pre = block_seg_exclusive<kBlock>(a, slds, total);
  // Look-back by wave 0 over windows of 64 predecessors: lane l reads the status
  // of tile (base - l). The nearest predecessor that is inclusive, or whose own
  // aggregate already contains a traversal start, ends the walk; everything
  // nearer must have published its aggregate, else the window is re-read.
  if (tid < kWave) {
    if (tile == 0) {
      if (lane == 0) {
        __hip_atomic_store(&status[0], st_pack(kStInclusive, total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        scarry = Seg{0, 0};
      }
    } else {
      if (lane == 0)
        __hip_atomic_store(&status[tile],
                           st_pack(total.flag ? kStInclusive : kStAggregate, total),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      Seg carry{0, 0};
      int64_t wbase = (int64_t)tile - 1;
      uint32_t spins = 0;
      for (;;) {
        const int64_t pt = wbase - (int64_t)lane;
        uint64_t st = kStInclusive;  // before tile 0: the identity, inclusive
        if (pt >= 0)
          st = __hip_atomic_load(&status[pt], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t state = st & (3ull << 62);
        const uint32_t fl = (uint32_t)(st >> 32) & 1u;
        const bool stop = state == kStInclusive || (state == kStAggregate && fl);
        const uint64_t stop_mask = __ballot(stop);
        const uint64_t nr_mask = __ballot(state == kStNotReady);
        const uint32_t L = stop_mask ? (uint32_t)__builtin_ctzll(stop_mask) : (uint32_t)kWave;
        const uint64_t need = L >= (uint32_t)kWave ? ~0ull : ((1ull << L) - 1ull);
        if (nr_mask & need) {
          if (++spins > (1u << 24)) {
            if (lane == 0)
              atomicExch(err, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        const uint32_t top = L >= (uint32_t)kWave ? (uint32_t)kWave - 1 : L;
        Seg acc{0, 0};
        for (int l = (int)top; l >= 0; --l) {
          const uint32_t ps = __shfl((uint32_t)st, l);
          const uint32_t pf = __shfl(fl, l);
          acc = seg_then(acc, Seg{ps, pf});
        }
        carry = seg_then(acc, carry);
        if (L < (uint32_t)kWave)
          break;
        wbase -= kWave;
      }
      if (lane == 0) {
        if (!total.flag)
          __hip_atomic_store(&status[tile], st_pack(kStInclusive, seg_then(carry, total)),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        scarry = carry;
      }
    }
  }
  __syncthreads();
  Seg run = seg_then(scarry, pre);
#pragma unroll
  for (uint32_t i = 0; i < kExItems; ++i) {
    run = seg_then(run, Seg{vals[i], (hbits >> i) & 1u});
    sval[pad_idx(pb + i)] = run.sum;
  }
  __syncthreads();
#pragma unroll 4
  for (uint32_t j = 0; j < kExItems; ++j) {
    const uint32_t p = j * kBlock + tid;
    if (o0 + p < o_end)
      out[o0 + p] = (int32_t)sval[pad_idx(p)];
  }
}

__global__ __launch_bounds__(kBlock) void k_seg_reduce(const int32_t *__restrict__ x, uint64_t n,
                                                       const uint32_t *__restrict__ heads,
                                                       Seg *__restrict__ agg) {
  __shared__ Seg lds[kBlock / kWave + 1];
  const uint64_t base = (uint64_t)blockIdx.x * kSsTile + (uint64_t)threadIdx.x * kSsItems;
  Seg a{0, 0};
#pragma unroll
  for (uint32_t j = 0; j < kSsItems; ++j) {
    const uint64_t i = base + j;
    if (i < n)
      a = seg_then(a, Seg{(uint32_t)x[i], bit_at(heads, i) ? 1u : 0u});
  }
  Seg total;
  block_seg_exclusive<kBlock>(a, lds, total);
  if (threadIdx.x == 0)
    agg[blockIdx.x] = total;
}

__global__ __launch_bounds__(kSsScanThreads) void k_seg_tiles(Seg *agg, uint64_t nb) {
  __shared__ Seg lds[kSsScanThreads / kWave + 1];
  const uint64_t chunk = (nb + kSsScanThreads - 1) / kSsScanThreads;
  const uint64_t begin = (uint64_t)threadIdx.x * chunk;
  const uint64_t end = begin + chunk < nb ? begin + chunk : nb;
  Seg s{0, 0};
  for (uint64_t i = begin; i < end; ++i)
    s = seg_then(s, agg[i]);
  Seg total;
  Seg run = block_seg_exclusive<kSsScanThreads>(s, lds, total);
  for (uint64_t i = begin; i < end; ++i) {
    const Seg t = agg[i];
    agg[i] = run;
    run = seg_then(run, t);
  }
}

__global__ __launch_bounds__(kBlock) void k_seg_apply(int32_t *__restrict__ x, uint64_t n,
                                                      const uint32_t *__restrict__ heads,
                                                      const Seg *__restrict__ carry) {
  __shared__ Seg lds[kBlock / kWave + 1];
  const uint64_t base = (uint64_t)blockIdx.x * kSsTile + (uint64_t)threadIdx.x * kSsItems;
  uint32_t v[kSsItems];
  uint32_t hb = 0;
  Seg a{0, 0};
#pragma unroll
  for (uint32_t j = 0; j < kSsItems; ++j) {
    const uint64_t i = base + j;
    v[j] = i < n ? (uint32_t)x[i] : 0u;
    const uint32_t h = (i < n && bit_at(heads, i)) ? 1u : 0u;
    hb |= h << j;
    a = seg_then(a, Seg{v[j], h});
  }
  Seg total;
  Seg run = seg_then(carry[blockIdx.x], block_seg_exclusive<kBlock>(a, lds, total));
#pragma unroll
  for (uint32_t j = 0; j < kSsItems; ++j) {
    const uint64_t i = base + j;
    run = seg_then(run, Seg{v[j], (hb >> j) & 1u});
    if (i < n)
      x[i] = (int32_t)run.sum;
  }
}

uint32_t stride_grid(Ctx &ctx, uint64_t n) {
  const uint64_t want = ceil_div(n, kBlock * 4);
  const uint64_t cap = (uint64_t)ctx.num_cus() * 8;
  return (uint32_t)std::max<uint64_t>(1, std::min(want, cap));
}

}  // namespace

void rulebook_build(Ctx &ctx, const int32_t *d_first, const int32_t *d_second, uint64_t R,
                    uint32_t min_id, Rulebook &rb) {
  rb = Rulebook();
  rb.n_rules = R;
  rb.min_id = min_id;
  if ((uint64_t)min_id + R > (uint64_t)INT32_MAX + 1)
    throw Error(kErrCorrupt, "rule-ID window exceeds the int32 symbol range");
  rb.rec = DBuf<RuleRec>(ctx.pool(), std::max<uint64_t>(R, 1));
  rb.size = DBuf<uint32_t>(ctx.pool(), std::max<uint64_t>(R, 1));
  if (!R)
    return;
  DBuf<uint32_t> depth(ctx.pool(), R);
  DBuf<uint32_t> flags(ctx.pool(), 2);  // changed, err
  hipLaunchKernelGGL(k_rule_init, dim3(stride_grid(ctx, R)), dim3(kBlock), 0, ctx.stream(),
                     d_first, d_second, R, min_id, rb.rec.get(), rb.size.get(), depth.get());
  GFZ_CHECK_LAUNCH();
  uint32_t max_depth = 1;
  for (uint64_t pass = 0; pass <= R; ++pass) {
    fill_u32(ctx, flags.get(), 0, 2);
    ctx.timed("rule_sizes", [&] {
      hipLaunchKernelGGL(k_rule_pass, dim3(stride_grid(ctx, R)), dim3(kBlock), 0, ctx.stream(),
                         rb.rec.get(), rb.size.get(), depth.get(), R, min_id, (uint32_t)pass,
                         flags.get(), flags.get() + 1);
    });
    GFZ_CHECK_LAUNCH();
    uint32_t h[2];
    GFZ_HIP(hipMemcpyAsync(h, flags.get(), sizeof(h), hipMemcpyDeviceToHost, ctx.stream()));
    ctx.sync();
    if (h[1])
      throw Error(kErrCorrupt, "rule expansion size exceeds 2^32 symbols");
    if (!h[0])
      break;
    max_depth = (uint32_t)pass + 2;
  }
  hipLaunchKernelGGL(k_rule_unresolved, dim3(stride_grid(ctx, R)), dim3(kBlock), 0, ctx.stream(),
                     depth.get(), R, flags.get() + 1);
  GFZ_CHECK_LAUNCH();
  uint32_t e = 0;
  GFZ_HIP(hipMemcpyAsync(&e, flags.get() + 1, 4, hipMemcpyDeviceToHost, ctx.stream()));
  ctx.sync();
  if (e)
    throw Error(kErrCorrupt, "rulebook is cyclic: some rule never bottoms out in node IDs");
  rb.depth = max_depth;

  // Level ranges (contiguous when depth is non-decreasing in rule ID).
  std::vector<uint64_t> bound(max_depth + 2, R);
  {
    DBuf<uint64_t> d_bound(ctx.pool(), max_depth + 2);
    fill_u64(ctx, d_bound.get(), R, max_depth + 2);
    fill_u32(ctx, flags.get(), 0, 1);
    hipLaunchKernelGGL(k_depth_bounds, dim3(stride_grid(ctx, R)), dim3(kBlock), 0, ctx.stream(),
                       depth.get(), R, max_depth, d_bound.get(), flags.get());
    GFZ_CHECK_LAUNCH();
    GFZ_HIP(hipMemcpyAsync(bound.data(), d_bound.get(), bound.size() * 8, hipMemcpyDeviceToHost,
                           ctx.stream()));
    GFZ_HIP(hipMemcpyAsync(&e, flags.get(), 4, hipMemcpyDeviceToHost, ctx.stream()));
    ctx.sync();
  }
  if (e)
    return;  // depth not monotone in rule ID: expansion descends the tree instead
  depth.reset();

  // Bodies, one depth level at a time (children are always shallower).
  rb.body_off = DBuf<uint64_t>(ctx.pool(), R + 1);
  exclusive_scan(ctx, rb.size.get(), rb.body_off.get(), R);
  std::vector<uint64_t> lvl_off(max_depth + 2);
  {
    DBuf<uint64_t> d_lvl(ctx.pool(), max_depth + 2), d_idx(ctx.pool(), max_depth + 2);
    GFZ_HIP(hipMemcpyAsync(d_idx.get(), bound.data(), bound.size() * 8, hipMemcpyHostToDevice,
                           ctx.stream()));
    hipLaunchKernelGGL(k_gather_u64, dim3(1), dim3(kBlock), 0, ctx.stream(), rb.body_off.get(),
                       d_idx.get(), max_depth + 2, d_lvl.get());
    GFZ_CHECK_LAUNCH();
    GFZ_HIP(hipMemcpyAsync(lvl_off.data(), d_lvl.get(), lvl_off.size() * 8, hipMemcpyDeviceToHost,
                           ctx.stream()));
    ctx.sync();
  }
  const uint64_t total = lvl_off[max_depth + 1];
  uint64_t budget = uint64_t(1) << 30;  // elements (4 GiB)
  if (const char *env = std::getenv("GFAZ_BODY_BUDGET_MB"))
    budget = std::strtoull(env, nullptr, 10) * (uint64_t(1) << 20) / 4;
  if (total > budget)
    return;
  rb.bodies = DBuf<int32_t>(ctx.pool(), std::max<uint64_t>(total, 1));
  rb.body_elems = total;
  DBuf<int32_t> ps(ctx.pool(), 2 * R);
  DBuf<uint64_t> ps_off(ctx.pool(), 2 * R + 1);
  hipLaunchKernelGGL(k_pair_stream, dim3(stride_grid(ctx, R)), dim3(kBlock), 0, ctx.stream(),
                     rb.rec.get(), rb.body_off.get(), R, ps.get(), ps_off.get());
  GFZ_CHECK_LAUNCH();
  for (uint32_t d = 1; d <= max_depth; ++d) {
    const uint64_t r0 = bound[d], r1 = bound[d + 1];
    if (r0 >= r1)
      continue;
    const uint64_t lo = lvl_off[d], hi = lvl_off[d + 1];
    const uint64_t n_tiles = ceil_div(hi - lo, kExTile);
    DBuf<TileDesc> tf(ctx.pool(), n_tiles);
    ctx.timed("build_bodies", [&] {
      hipLaunchKernelGGL(k_tile_desc, dim3(stride_grid(ctx, 2 * (r1 - r0))), dim3(kBlock), 0,
                         ctx.stream(), ps_off.get() + 2 * r0, 2 * (r1 - r0), lo, hi - lo, n_tiles,
                         tf.get());
      GFZ_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_expand<true>, dim3((uint32_t)n_tiles), dim3(kBlock), 0, ctx.stream(),
                         ps.get() + 2 * r0, ps_off.get() + 2 * r0, 2 * (r1 - r0), lo, hi - lo,
                         tf.get(), rb.rec.get(), rb.size.get(), rb.body_off.get(), rb.bodies.get(),
                         R, min_id, rb.bodies.get());
      GFZ_CHECK_LAUNCH();
    });
  }
  ctx.sync();
  rb.has_bodies = true;
}

void decode_plan(Ctx &ctx, const Rulebook &rb, const int32_t *d_enc, uint64_t n_enc,
                 const uint32_t *enc_lengths, const uint32_t *orig_lengths, uint64_t n_trav,
                 DecodePlan &plan, DBuf<uint64_t> &sym_off) {
  plan = DecodePlan();
  uint64_t total_enc = 0;
  std::vector<uint64_t> enc_start(n_trav);
  for (uint64_t t = 0; t < n_trav; ++t) {
    enc_start[t] = total_enc;
    total_enc += enc_lengths[t];
  }
  if (total_enc != n_enc)
    throw Error(kErrCorrupt, "flattened sequence block is truncated (lengths sum " +
                                 std::to_string(total_enc) + ", block holds " +
                                 std::to_string(n_enc) + ")");
  sym_off = DBuf<uint64_t>(ctx.pool(), n_enc + 1);
  {
    DBuf<uint32_t> sz(ctx.pool(), std::max<uint64_t>(n_enc, 1));
    if (n_enc) {
      hipLaunchKernelGGL(k_sym_sizes, dim3(stride_grid(ctx, n_enc)), dim3(kBlock), 0,
                         ctx.stream(), d_enc, n_enc, rb.size.get(), rb.n_rules, rb.min_id,
                         sz.get());
      GFZ_CHECK_LAUNCH();
    }
    exclusive_scan(ctx, sz.get(), sym_off.get(), n_enc);
  }
  // Decoded offset of each traversal = symbol offset at its first encoded symbol.
  plan.out_offsets.assign(n_trav + 1, 0);
  if (n_trav) {
    DBuf<uint64_t> idx(ctx.pool(), n_trav + 1), val(ctx.pool(), n_trav + 1);
    std::vector<uint64_t> h_idx(enc_start);
    h_idx.push_back(n_enc);
    GFZ_HIP(hipMemcpyAsync(idx.get(), h_idx.data(), (n_trav + 1) * 8, hipMemcpyHostToDevice,
                           ctx.stream()));
    hipLaunchKernelGGL(k_gather_u64, dim3(stride_grid(ctx, n_trav + 1)), dim3(kBlock), 0,
                       ctx.stream(), sym_off.get(), idx.get(), n_trav + 1, val.get());
    GFZ_CHECK_LAUNCH();
    GFZ_HIP(hipMemcpyAsync(plan.out_offsets.data(), val.get(), (n_trav + 1) * 8,
                           hipMemcpyDeviceToHost, ctx.stream()));
    ctx.sync();
  } else {
    GFZ_HIP(hipMemcpyAsync(plan.out_offsets.data(), sym_off.get() + n_enc, 8,
                           hipMemcpyDeviceToHost, ctx.stream()));
    ctx.sync();
  }
  plan.n_out = plan.out_offsets[n_trav];
  if (orig_lengths)
    for (uint64_t t = 0; t < n_trav; ++t)
      if (plan.out_offsets[t + 1] - plan.out_offsets[t] != orig_lengths[t])
        plan.length_mismatch = true;
}

void decode_run(Ctx &ctx, const Rulebook &rb, const int32_t *d_enc, uint64_t n_enc,
                const uint32_t *enc_lengths, uint64_t n_trav, int delta_round,
                const DecodePlan &plan, const uint64_t *d_sym_off, int32_t *d_out) {
  const uint64_t n_out = plan.n_out;
  if (!n_out)
    return;
  const uint64_t n_tiles = ceil_div(n_out, kExTile);
  if (n_tiles > UINT32_MAX)
    throw Error(kErrInvalid, "decoded batch too large for one launch");
  DBuf<TileDesc> tf(ctx.pool(), n_tiles);
  hipLaunchKernelGGL(k_tile_desc, dim3(stride_grid(ctx, n_enc)), dim3(kBlock), 0, ctx.stream(),
                     d_sym_off, n_enc, 0, n_out, n_tiles, tf.get());
  GFZ_CHECK_LAUNCH();
  const uint64_t *bo = rb.has_bodies ? rb.body_off.get() : nullptr;
  const int32_t *bd = rb.has_bodies ? rb.bodies.get() : nullptr;
  if (delta_round <= 0) {
    ctx.timed("expand", [&] {
      if (rb.has_bodies)
        hipLaunchKernelGGL(k_expand<true>, dim3((uint32_t)n_tiles), dim3(kBlock), 0,
                           ctx.stream(), d_enc, d_sym_off, n_enc, 0, n_out, tf.get(),
                           rb.rec.get(), rb.size.get(), bo, bd, rb.n_rules, rb.min_id, d_out);
      else
        hipLaunchKernelGGL(k_expand<false>, dim3((uint32_t)n_tiles), dim3(kBlock), 0,
                           ctx.stream(), d_enc, d_sym_off, n_enc, 0, n_out, tf.get(),
                           rb.rec.get(), rb.size.get(), bo, bd, rb.n_rules, rb.min_id, d_out);
    });
    GFZ_CHECK_LAUNCH();
    ctx.add_bytes("expand", 4.0 * n_enc + 4.0 * n_out + 8.0 * rb.n_rules);
    return;
  }
  // Heads at the decoded start of every non-empty traversal.
  std::vector<uint64_t> starts;
  for (uint64_t t = 0; t < n_trav; ++t)
    if (enc_lengths[t] && plan.out_offsets[t + 1] > plan.out_offsets[t])
      starts.push_back(plan.out_offsets[t]);
  const uint64_t words = ceil_div(n_out, 32);
  DBuf<uint32_t> heads(ctx.pool(), words);
  fill_u32(ctx, heads.get(), 0, words);
  DBuf<uint64_t> d_starts;
  if (!starts.empty()) {
    d_starts = DBuf<uint64_t>(ctx.pool(), starts.size());
    GFZ_HIP(hipMemcpyAsync(d_starts.get(), starts.data(), starts.size() * 8,
                           hipMemcpyHostToDevice, ctx.stream()));
    set_head_bits(ctx, d_starts.get(), starts.size(), heads.get(), n_out);
  }
  // First inverse delta fused into the expansion; any further rounds separately.
  DBuf<unsigned long long> status(ctx.pool(), n_tiles);
  DBuf<uint32_t> ctl(ctx.pool(), 2);  // ticket, err
  fill_u64(ctx, reinterpret_cast<uint64_t *>(status.get()), 0, n_tiles);
  fill_u32(ctx, ctl.get(), 0, 2);
  ctx.timed("expand", [&] {
    if (rb.has_bodies)
      hipLaunchKernelGGL(k_expand_delta<true>, dim3((uint32_t)n_tiles), dim3(kBlock), 0,
                         ctx.stream(), d_enc, d_sym_off, n_enc, n_out, tf.get(), rb.rec.get(),
                         rb.size.get(), bo, bd, rb.n_rules, rb.min_id, heads.get(), status.get(),
                         ctl.get(), ctl.get() + 1, d_out);
    else
      hipLaunchKernelGGL(k_expand_delta<false>, dim3((uint32_t)n_tiles), dim3(kBlock), 0,
                         ctx.stream(), d_enc, d_sym_off, n_enc, n_out, tf.get(), rb.rec.get(),
                         rb.size.get(), bo, bd, rb.n_rules, rb.min_id, heads.get(), status.get(),
                         ctl.get(), ctl.get() + 1, d_out);
  });
  GFZ_CHECK_LAUNCH();
  ctx.add_bytes("expand", 4.0 * n_enc + 4.0 * n_out + 8.0 * rb.n_rules);
  for (int r = 1; r < delta_round; ++r)
    segmented_prefix_sum(ctx, d_out, n_out, heads.get());
  uint32_t e = 0;
  GFZ_HIP(hipMemcpyAsync(&e, ctl.get() + 1, 4, hipMemcpyDeviceToHost, ctx.stream()));
  ctx.sync();
  if (e)
    throw Error(kErrHip, "expansion look-back did not complete (internal error)");
}

void segmented_prefix_sum(Ctx &ctx, int32_t *x, uint64_t n, const uint32_t *heads) {
  if (!n)
    return;
  const uint64_t nb = ceil_div(n, kSsTile);
  DBuf<Seg> agg(ctx.pool(), nb);
  ctx.timed("inverse_delta", [&] {
    hipLaunchKernelGGL(k_seg_reduce, dim3((uint32_t)nb), dim3(kBlock), 0, ctx.stream(), x, n,
                       heads, agg.get());
    GFZ_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_seg_tiles, dim3(1), dim3(kSsScanThreads), 0, ctx.stream(), agg.get(), nb);
    GFZ_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_seg_apply, dim3((uint32_t)nb), dim3(kBlock), 0, ctx.stream(), x, n, heads,
                       agg.get());
    GFZ_CHECK_LAUNCH();
  });
}

}  // namespace gfz
