// Rule expansion and inverse delta on gfx950 (decompress / extract).
//
// Restates src/workflows/decompression_workflow.cpp:70-174 (recursive
// expand_rule, then delta_round inverse prefix sums per traversal):
//   expand(+r) = expand(first[r]) ++ expand(second[r])
//   expand(-r) = expand(-second[r]) ++ expand(-first[r])
//   a symbol is a rule iff min_id <= |s| < min_id + R.
//
// Prefix-summed bodies (Z). For every rule the rulebook keeps, in HBM, the
// running sums of its expansion with a leading zero: Z_r[0] = 0,
// Z_r[u + 1] = body_r[0] + ... + body_r[u]. With sum_r = Z_r[size_r]:
//   prefix of expand(+r) up to q  =          Z_r[q + 1]
//   prefix of expand(-r) up to q  = -sum_r + Z_r[size_r - 1 - q]
// so after one inverse delta the value at position q of encoded symbol s is
//   (sum of the traversal's earlier symbols) + that prefix:
// one gather per output element and no carry chain between output tiles (the
// symbol-level sums are scanned over the encoded stream, which is much
// shorter). Without an inverse delta a body element is a difference of two
// neighbouring Z entries. Every output position finds its encoded symbol by a
// rank over a bitmap of symbol starts staged in LDS, so one lane writes one
// output element and a wave's stores are consecutive.
//
// Z is built level by level on the device: a rule's Z is its first child's
// prefixes followed by its second child's shifted by the first child's sum,
// which is the same gather over a three-symbol stream per rule (0, first,
// second). Rulebooks whose depth is not non-decreasing in rule ID, or whose Z
// exceeds the budget, are expanded by descending the rule tree by expanded
// sizes, then inverse-delta'd by segmented scans.
#include "codec.hpp"
#include "prims.hpp"
#include "zexpand.hpp"

#include <algorithm>
#include <cstring>
#include <cstdlib>

namespace gfz {

namespace {

using namespace zx;

constexpr uint32_t kExItems = 16;  // default output elements per thread
constexpr uint32_t kSsItems = 16;
constexpr uint32_t kSsTile = kBlock * kSsItems;
constexpr uint32_t kSsScanThreads = 1024;

// ---- rule sizes, sums and depths --------------------------------------------
__global__ __launch_bounds__(kBlock) void k_rule_init(const int32_t *__restrict__ f,
                                                      const int32_t *__restrict__ s, uint64_t R,
                                                      uint32_t min_id, RuleRec *__restrict__ rec,
                                                      uint2 *__restrict__ rss,
                                                      uint32_t *__restrict__ depth,
                                                      uint32_t *lo_out) {
  // The first non-leaf rule goes to *lo_out (one atomic per wave at most).
  const uint32_t lane = threadIdx.x & (kWave - 1);
  for (uint64_t r0 = (uint64_t)blockIdx.x * kBlock; r0 < R; r0 += (uint64_t)gridDim.x * kBlock) {
    const uint64_t r = r0 + threadIdx.x;
    if (r >= R)
      continue;  // a wave's lane 0 holds its first rule, so lane 0 stays active below
    const int32_t a = f[r], b = s[r];
    RuleRec q;
    q.a = a;
    q.b = b;
    q.sa = is_rule(a, min_id, R) ? 0u : 1u;
    q.sb = is_rule(b, min_id, R) ? 0u : 1u;
    rec[r] = q;
    const bool leaf = q.sa && q.sb;
    rss[r] = leaf ? make_uint2(2u, (uint32_t)a + (uint32_t)b) : make_uint2(0u, 0u);
    depth[r] = leaf ? 1u : 0u;
    const uint64_t inner = __ballot(!leaf);
    if (inner && lane == 0) {
      const uint32_t first = (uint32_t)(r + (uint32_t)__ffsll((unsigned long long)inner) - 1);
      if (first < __atomic_load_n(lo_out, __ATOMIC_RELAXED))
        atomicMin(lo_out, first);
    }
  }
}

// Size, sum and Z offset of every rule in one record (the Z offsets are below
// 2^31 whenever bodies are built).
__global__ __launch_bounds__(kBlock) void k_rule_szo(const uint2 *__restrict__ rss,
                                                     const uint64_t *__restrict__ zoff,
                                                     uint64_t R, uint4 *__restrict__ szo) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * kBlock) {
    const uint2 w = rss[r];
    szo[r] = make_uint4(w.x, w.y, (uint32_t)zoff[r], 0u);
  }
}

// One bottom-up pass. In pass k a rule resolves when each child is a node ID or a
// rule resolved before this pass (depth <= k + 1); it then gets depth k + 2, its
// size and its sum. Children resolved during the same pass are not used, so
// every depth is exact whatever the thread order. `changed` is raised once per
// wave. The pass starts at *lo_in, the first rule left unresolved by the pass
// before (every rule below it has its depth), and leaves in *lo_out the first
// rule it left unresolved: rule IDs grow with depth, so the passes scan the
// shrinking suffix of deeper rules instead of the whole rulebook.
__global__ __launch_bounds__(kBlock) void k_rule_pass(RuleRec *__restrict__ rec,
                                                      uint2 *__restrict__ rss,
                                                      uint32_t *__restrict__ depth, uint64_t R,
                                                      uint32_t min_id, uint32_t pass,
                                                      uint32_t *changed, uint32_t *err,
                                                      const uint32_t *__restrict__ lo_in,
                                                      uint32_t *lo_out) {
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const uint64_t start = (uint64_t)(*lo_in & ~(kBlock - 1));
  for (uint64_t r0 = start + (uint64_t)blockIdx.x * kBlock; r0 < R;
       r0 += (uint64_t)gridDim.x * kBlock) {
    const uint64_t r = r0 + threadIdx.x;
    bool resolved = false;
    const bool pending = r < R && !depth[r];
    if (pending) {
      RuleRec q = rec[r];
      const bool ra = q.sa != 0, rb = q.sb != 0;  // node-ID children
      bool ok = true;
      uint32_t sa = q.sa, sb = q.sb, suma = (uint32_t)q.a, sumb = (uint32_t)q.b;
      if (!ra) {
        const uint32_t c = uabs(q.a) - min_id;
        const uint32_t dc = depth[c];
        ok = ok && dc != 0 && dc <= pass + 1;
        if (ok) {
          const uint2 w = rss[c];
          sa = w.x;
          suma = q.a > 0 ? w.y : 0u - w.y;
        }
      }
      if (!rb) {
        const uint32_t c = uabs(q.b) - min_id;
        const uint32_t dc = depth[c];
        ok = ok && dc != 0 && dc <= pass + 1;
        if (ok) {
          const uint2 w = rss[c];
          sb = w.x;
          sumb = q.b > 0 ? w.y : 0u - w.y;
        }
      }
      if (ok) {
        const uint32_t total = sa + sb;
        if (total < sa) {
          atomicExch(err, 1u);
        } else {
          rss[r] = make_uint2(total, suma + sumb);
          q.sa = sa;
          q.sb = sb;
          rec[r] = q;
          depth[r] = pass + 2;
          resolved = true;
        }
      }
    }
    const uint64_t m = __ballot(resolved);
    if (m && lane == (uint32_t)__ffsll((unsigned long long)m) - 1)
      *changed = 1u;
    const uint64_t left = __ballot(pending && !resolved);
    if (left && lane == 0) {
      const uint32_t first = (uint32_t)(r0 + wid * kWave + (uint32_t)__ffsll((unsigned long long)left) - 1);
      if (first < __atomic_load_n(lo_out, __ATOMIC_RELAXED))
        atomicMin(lo_out, first);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_rule_unresolved(const uint32_t *__restrict__ depth,
                                                            uint64_t R, uint32_t *err) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * kBlock)
    if (!depth[r])
      atomicExch(err, 2u);
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

__global__ void k_gather_u64(const uint64_t *__restrict__ src, const uint64_t *__restrict__ idx,
                             uint64_t n, uint64_t *__restrict__ dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}

// Expanded size + 1 (the rule's Z entries) of every rule.
__global__ void k_plus_one(const uint2 *__restrict__ in, uint64_t n, uint32_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[i].x + 1u;
}

// ---- decode planning: symbol offsets and records in three launches ---------
// Per encoded symbol: its expanded size (1 for a node ID), the wrapping sum of
// its expansion, and whether a traversal starts there. A reduce pass gives
// every tile of kPlanTile symbols its aggregate, one block scans the tiles, and
// an apply pass re-derives the tile's symbols and writes each one's decoded
// offset (exclusive sum of sizes) and record (SymInfo: the traversal's sum of
// earlier expansions, i.e. a scan of the sums restarted at traversal starts,
// plus the Z position). No block waits for another.
constexpr uint32_t kPlanItems = 8;                  // symbols per lane, 64 apart
constexpr uint32_t kPlanTile = kBlock * kPlanItems;  // 2048
constexpr uint32_t kPlanScanThreads = 1024;

// Aggregate of a run of symbols: total size; whether a traversal starts in it;
// sum of the expansions after its last traversal start (all of them if none).
struct PlanAgg {
  uint64_t size;
  uint32_t head;
  uint32_t sum;
};
__device__ inline PlanAgg plan_combine(const PlanAgg &a, const PlanAgg &b) {
  return {a.size + b.size, a.head | b.head, b.head ? b.sum : a.sum + b.sum};
}
__device__ inline PlanAgg plan_shfl_up(const PlanAgg &x, int d) {
  return {__shfl_up(x.size, d), __shfl_up(x.head, d), __shfl_up(x.sum, d)};
}

// Ordered exclusive scan of one PlanAgg per thread over a block of NT threads;
// `lds` holds NT / kWave + 1 entries. Returns the exclusive prefix, sets total.
template <uint32_t NT>
__device__ inline PlanAgg plan_block_scan(PlanAgg x, PlanAgg *lds, PlanAgg &total) {
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  PlanAgg inc = x;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const PlanAgg y = plan_shfl_up(inc, d);
    if (lane >= (uint32_t)d)
      inc = plan_combine(y, inc);
  }
  if (lane == kWave - 1)
    lds[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    PlanAgg run{0, 0, 0};
    for (uint32_t w = 0; w < NT / kWave; ++w) {
      const PlanAgg t = lds[w];
      lds[w] = run;
      run = plan_combine(run, t);
    }
    lds[NT / kWave] = run;
  }
  __syncthreads();
  PlanAgg ex = plan_shfl_up(inc, 1);
  if (lane == 0)
    ex = PlanAgg{0, 0, 0};
  const PlanAgg res = plan_combine(lds[wid], ex);
  total = lds[NT / kWave];
  __syncthreads();
  return res;
}

// Entries of a[0, n) below v (a non-decreasing); all 64 lanes of one wave call
// it and get the count. Each step probes 64 evenly spaced entries of the range
// the count can still be in, which shrinks it to one gap between probes.
__device__ inline uint64_t wave_count_below(const uint64_t *__restrict__ a, uint64_t n,
                                            uint64_t v) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  uint64_t lo = 0, hi = n;  // the count lies in [lo, hi]
  while (lo < hi) {
    const uint64_t step = (hi - lo + kWave - 1) / kWave;
    const uint64_t k = lo + (uint64_t)lane * step;
    const uint32_t c = (uint32_t)__popcll(__ballot(k < hi && a[k] < v));
    if (c == 0)
      break;  // a[lo] >= v: the count is lo
    const uint64_t nlo = lo + (uint64_t)(c - 1) * step + 1;
    const uint64_t nhi = lo + (uint64_t)c * step;
    lo = nlo;
    hi = nhi < hi ? nhi : hi;
  }
  return lo;
}

// The tile's traversal starts as a bitmap in LDS (starts: encoded start of
// every traversal, non-decreasing).
__device__ inline void plan_heads(const uint64_t *__restrict__ starts, uint64_t n_starts,
                                  uint64_t s0, uint64_t s1, uint32_t *hbits, uint64_t *range) {
  for (uint32_t w = threadIdx.x; w < kPlanTile / 32; w += kBlock)
    hbits[w] = 0;
  if (threadIdx.x < 2 * kWave) {  // wave 0 finds the first start, wave 1 the end
    const uint32_t w = threadIdx.x / kWave;
    const uint64_t c = wave_count_below(starts, n_starts, w ? s1 : s0);
    if ((threadIdx.x & (kWave - 1)) == 0)
      range[w] = c;
  }
  __syncthreads();
  for (uint64_t t = range[0] + threadIdx.x; t < range[1]; t += kBlock) {
    const uint32_t q = (uint32_t)(starts[t] - s0);
    atomicOr(&hbits[q >> 5], 1u << (q & 31));
  }
  __syncthreads();
}

// A wave's kPlanItems x 64 consecutive symbols, item j of lane l at
// chunk + j * 64 + l (every load and store is one coalesced run per item);
// indexes are clamped into the stream (n_enc >= 1 whenever a tile exists) so
// every load is issued before any is used.
struct PlanSyms {
  int32_t v[kPlanItems];
  uint32_t sz[kPlanItems], ss[kPlanItems];
  uint32_t zo[kPlanItems];  // a rule symbol's Z offset (plans that decode through Z)
  uint32_t head;   // bit j: a traversal starts at this lane's item j
  uint32_t valid;  // bit j: item j is inside the stream
};
__device__ inline uint64_t plan_chunk(uint64_t s0) {
  return s0 + (uint64_t)(threadIdx.x / kWave) * (kPlanItems * kWave);
}
// Symbols and their rules' sizes and sums (and Z offsets, from szo when it is
// given, else sizes and sums from rss); issued before the tile's traversal
// starts are looked up, so the two latencies overlap.
__device__ inline void plan_syms(const int32_t *__restrict__ enc, uint64_t n_enc, uint64_t s0,
                                 const uint2 *__restrict__ rss, const uint4 *__restrict__ szo,
                                 uint64_t R, uint32_t min_id, PlanSyms &p) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t c0 = plan_chunk(s0);
  p.head = p.valid = 0;
#pragma unroll
  for (uint32_t j = 0; j < kPlanItems; ++j) {
    const uint64_t i = c0 + (uint64_t)j * kWave + lane;
    p.v[j] = enc[i < n_enc ? i : n_enc - 1];
    p.valid |= i < n_enc ? 1u << j : 0u;
  }
  if (szo) {
#pragma unroll
    for (uint32_t j = 0; j < kPlanItems; ++j) {
      const bool rule = is_rule(p.v[j], min_id, R);
      const uint4 w = rule ? szo[uabs(p.v[j]) - min_id] : make_uint4(1u, 0u, 0u, 0u);
      p.sz[j] = w.x;
      p.ss[j] = w.y;
      p.zo[j] = w.z;
    }
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kPlanItems; ++j) {
      const bool rule = is_rule(p.v[j], min_id, R);
      const uint2 w = rule ? rss[uabs(p.v[j]) - min_id] : make_uint2(1u, 0u);
      p.sz[j] = w.x;
      p.ss[j] = w.y;
      p.zo[j] = 0u;
    }
  }
}
// Then the start flags (from the LDS bitmap) and the sizes and signed sums.
__device__ inline void plan_finish(uint64_t s0, const uint32_t *hbits, uint64_t R,
                                   uint32_t min_id, PlanSyms &p) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t c0 = plan_chunk(s0);
#pragma unroll
  for (uint32_t j = 0; j < kPlanItems; ++j) {
    const uint32_t q = (uint32_t)(c0 + (uint64_t)j * kWave + lane - s0);
    p.head |= ((hbits[q >> 5] >> (q & 31)) & 1) << j;
    const bool ok = (p.valid >> j) & 1;
    const uint32_t rs = p.ss[j];
    p.ss[j] = !ok ? 0u
                  : is_rule(p.v[j], min_id, R) ? (p.v[j] > 0 ? rs : 0u - rs) : (uint32_t)p.v[j];
    p.sz[j] = ok ? p.sz[j] : 0u;
  }
  p.head &= p.valid;
}

// The apply pass's symbols: values, and sizes, signed sums and Z offsets as the
// reduce pass stored them (no rule lookups here).
__device__ inline void plan_syms_stored(const int32_t *__restrict__ enc, uint64_t n_enc,
                                        uint64_t s0, const uint2 *__restrict__ szss,
                                        const uint32_t *__restrict__ zos, PlanSyms &p) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t c0 = plan_chunk(s0);
  p.head = p.valid = 0;
#pragma unroll
  for (uint32_t j = 0; j < kPlanItems; ++j) {
    const uint64_t i = c0 + (uint64_t)j * kWave + lane;
    const uint64_t ic = i < n_enc ? i : n_enc - 1;
    p.v[j] = enc[ic];
    const uint2 w = szss[ic];
    p.zo[j] = zos ? zos[ic] : 0u;
    const bool ok = i < n_enc;
    p.valid |= ok ? 1u << j : 0u;
    p.sz[j] = ok ? w.x : 0u;
    p.ss[j] = ok ? w.y : 0u;
  }
}
__device__ inline void plan_heads_only(uint64_t s0, const uint32_t *hbits, PlanSyms &p) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t c0 = plan_chunk(s0);
#pragma unroll
  for (uint32_t j = 0; j < kPlanItems; ++j) {
    const uint32_t q = (uint32_t)(c0 + (uint64_t)j * kWave + lane - s0);
    p.head |= ((hbits[q >> 5] >> (q & 31)) & 1) << j;
  }
  p.head &= p.valid;
}

// Inclusive ordered scan of one PlanAgg per lane across a wave.
__device__ inline PlanAgg plan_wave_incl(PlanAgg x) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const PlanAgg y = plan_shfl_up(x, d);
    if (lane >= (uint32_t)d)
      x = plan_combine(y, x);
  }
  return x;
}
__device__ inline PlanAgg plan_lane(const PlanAgg &x, int l) {
  return {__shfl(x.size, l), __shfl(x.head, l), __shfl(x.sum, l)};
}

// The wave's aggregate over its chunk, in position order (every lane gets it).
__device__ inline PlanAgg plan_wave_total(const PlanSyms &p) {
  PlanAgg t{0, 0, 0};
#pragma unroll
  for (uint32_t j = 0; j < kPlanItems; ++j) {
    const PlanAgg inc = plan_wave_incl(PlanAgg{p.sz[j], (p.head >> j) & 1, p.ss[j]});
    t = plan_combine(t, plan_lane(inc, kWave - 1));
  }
  return t;
}

// Exclusive prefixes of the block's waves (in order) and the block total.
__device__ inline PlanAgg plan_waves_prefix(const PlanAgg &wave_total, PlanAgg *lds,
                                            PlanAgg &total) {
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0)
    lds[wid] = wave_total;
  __syncthreads();
  PlanAgg run{0, 0, 0}, mine{0, 0, 0};
  for (uint32_t w = 0; w < kBlock / kWave; ++w) {
    if (w == wid)
      mine = run;
    run = plan_combine(run, lds[w]);
  }
  total = run;
  return mine;
}

__global__ __launch_bounds__(kBlock) void k_plan_reduce(
    const int32_t *__restrict__ enc, uint64_t n_enc, const uint64_t *__restrict__ starts,
    uint64_t n_starts, const uint2 *__restrict__ rss, const uint4 *__restrict__ szo,
    uint64_t R, uint32_t min_id, uint64_t n_tiles, uint32_t rows, uint64_t cols,
    PlanAgg *__restrict__ tiles, uint2 *__restrict__ szss, uint32_t *__restrict__ zos) {
  __shared__ uint32_t hbits[kPlanTile / 32];
  __shared__ uint64_t range[2];
  __shared__ PlanAgg lds[kBlock / kWave];
  const uint64_t t = block_tile(rows, cols);
  if (t >= n_tiles)
    return;  // uniform per block, before any barrier
  const uint64_t s0 = t * kPlanTile;
  const uint64_t s1 = s0 + kPlanTile < n_enc ? s0 + kPlanTile : n_enc;
  PlanSyms p;
  plan_syms(enc, n_enc, s0, rss, szo, R, min_id, p);
  plan_heads(starts, n_starts, s0, s1, hbits, range);
  plan_finish(s0, hbits, R, min_id, p);
  {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t c0 = plan_chunk(s0);
#pragma unroll
    for (uint32_t j = 0; j < kPlanItems; ++j)
      if ((p.valid >> j) & 1) {
        szss[c0 + (uint64_t)j * kWave + lane] = make_uint2(p.sz[j], p.ss[j]);
        if (zos)
          zos[c0 + (uint64_t)j * kWave + lane] = p.zo[j];
      }
  }
  PlanAgg total;
  plan_waves_prefix(plan_wave_total(p), lds, total);
  if (threadIdx.x == 0)
    tiles[t] = total;
}

// One block: tile aggregates -> exclusive tile prefixes (in place); the grand
// total size goes to *end.
__global__ __launch_bounds__(kPlanScanThreads) void k_plan_scan(PlanAgg *__restrict__ tiles,
                                                               uint64_t n_tiles,
                                                               uint64_t *__restrict__ end) {
  __shared__ PlanAgg lds[kPlanScanThreads / kWave + 1];
  const uint64_t chunk = (n_tiles + kPlanScanThreads - 1) / kPlanScanThreads;
  const uint64_t b = (uint64_t)threadIdx.x * chunk;
  const uint64_t e = b + chunk < n_tiles ? b + chunk : n_tiles;
  PlanAgg mine{0, 0, 0};
  for (uint64_t t = b; t < e; ++t)
    mine = plan_combine(mine, tiles[t]);
  PlanAgg total;
  PlanAgg run = plan_block_scan<kPlanScanThreads>(mine, lds, total);
  for (uint64_t t = b; t < e; ++t) {
    const PlanAgg x = tiles[t];
    tiles[t] = run;
    run = plan_combine(run, x);
  }
  if (threadIdx.x == 0)
    *end = total.size;
}

// delta < 0: offsets only; otherwise records too (delta > 0: with the first
// inverse delta folded in).
__global__ __launch_bounds__(kBlock) void k_plan_apply(
    const int32_t *__restrict__ enc, uint64_t n_enc, const uint64_t *__restrict__ starts,
    uint64_t n_starts, const uint2 *__restrict__ szss, const uint32_t *__restrict__ zos,
    uint64_t R, uint32_t min_id, uint32_t zraw, int delta,
    const PlanAgg *__restrict__ tile_prefix, uint64_t n_tiles, uint32_t rows, uint64_t cols,
    uint64_t *__restrict__ sym_off, SymInfo *__restrict__ syms) {
  __shared__ uint32_t hbits[kPlanTile / 32];
  __shared__ uint64_t range[2];
  __shared__ PlanAgg lds[kBlock / kWave];
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t t = block_tile(rows, cols);
  if (t >= n_tiles)
    return;  // uniform per block, before any barrier
  const uint64_t s0 = t * kPlanTile;
  const uint64_t s1 = s0 + kPlanTile < n_enc ? s0 + kPlanTile : n_enc;
  PlanSyms p;
  plan_syms_stored(enc, n_enc, s0, szss, delta >= 0 ? zos : nullptr, p);
  plan_heads(starts, n_starts, s0, s1, hbits, range);
  plan_heads_only(s0, hbits, p);
  PlanAgg total;
  PlanAgg carry =
      plan_combine(tile_prefix[t], plan_waves_prefix(plan_wave_total(p), lds, total));
  const uint64_t c0 = plan_chunk(s0);
#pragma unroll
  for (uint32_t j = 0; j < kPlanItems; ++j) {
    const bool head = (p.head >> j) & 1;
    const PlanAgg x{p.sz[j], head ? 1u : 0u, p.ss[j]};
    const PlanAgg inc = plan_wave_incl(x);
    PlanAgg ex = plan_shfl_up(inc, 1);
    if (lane == 0)
      ex = PlanAgg{0, 0, 0};
    const PlanAgg at = plan_combine(carry, ex);
    carry = plan_combine(carry, plan_lane(inc, kWave - 1));
    const uint64_t i = c0 + (uint64_t)j * kWave + lane;
    if ((p.valid >> j) & 1) {
      sym_off[i] = at.size;
      if (delta >= 0) {
        const uint32_t pre = delta > 0 && !head ? at.sum : 0u;
        SymInfo s;
        const int32_t v = p.v[j];
        if (is_rule(v, min_id, R)) {  // as rule_sym: ss = -sum for a reversed rule
          if (v > 0) {
            s.base = pre;
            s.zz = p.zo[j] + 1;
          } else {
            s.base = delta > 0 ? pre + p.ss[j] : pre;
            s.zz = (p.zo[j] + p.sz[j] - 1) | kRevBit;
          }
        } else {
          s.base = pre + (uint32_t)v;
          s.zz = zraw;
        }
        syms[i] = s;
      }
    }
  }
}

// ---- tiles ------------------------------------------------------------------------
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
                            uint64_t n_out, uint32_t tile, uint64_t n_tiles,
                            TileDesc *__restrict__ desc) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_sym;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = sym_off[k] - base, hi = sym_off[k + 1] - base;
    for (uint64_t t = lo / tile; t < n_tiles && t * tile < hi; ++t) {
      if (t * tile >= lo) {
        desc[t].k0 = k;
        desc[t].first_off = lo;
      }
      const uint64_t end = ((t + 1) * tile < n_out ? (t + 1) * tile : n_out) - 1;
      if (end >= lo && end < hi)
        desc[t].k_last = k;
    }
  }
}

// Expansion through Z (see the file comment): out[base + o] for o in [0, n_out),
// one tile per block, symbols found through the tile descriptors. Z and out may
// be one buffer when the ranges read and written are disjoint. Blocks take tiles
// in block_tile order.
template <uint32_t kItems, bool kDiff>
__global__ __launch_bounds__(kBlock) void k_expand_z(const SymInfo *__restrict__ syms,
                                                     const uint64_t *__restrict__ sym_off,
                                                     uint64_t base, uint64_t n_out,
                                                     const TileDesc *__restrict__ tdesc,
                                                     uint64_t n_tiles, uint32_t rows,
                                                     uint64_t cols, const int32_t *Z,
                                                     uint32_t zraw, int32_t *out) {
  constexpr uint32_t kTile = kBlock * kItems;
  __shared__ ZTileSmem<kItems> sm;
  const uint64_t t = block_tile(rows, cols);
  if (t >= n_tiles)
    return;  // uniform per block, before any barrier
  const uint64_t o0 = t * kTile;
  const uint32_t n_here = (uint32_t)(o0 + kTile < n_out ? kTile : n_out - o0);
  const TileDesc td = tdesc[t];
  expand_tile<kItems, kDiff>(sm, syms, sym_off, td.k0, (uint32_t)(td.k_last - td.k0 + 1),
                             base + o0, n_here, Z, zraw, out + base + o0);
}

// Largest k in [lo, hi) with a[k] <= v, given a non-decreasing and a[lo] <= v.
// All 64 lanes of one wave call it; every lane gets the result.
__device__ inline uint64_t wave_last_le_u64(const uint64_t *__restrict__ a, uint64_t lo,
                                            uint64_t hi, uint64_t v) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  while (hi - lo > 1) {
    const uint64_t step = (hi - lo + kWave - 1) / kWave;
    const uint64_t k = lo + (uint64_t)lane * step;
    const uint64_t m = __ballot(k < hi && a[k] <= v);  // lane 0 always holds
    const uint32_t top = 63u - (uint32_t)__clzll((long long)m);
    lo = lo + (uint64_t)top * step;
    hi = hi < lo + step ? hi : lo + step;
  }
  return lo;
}

// Z bodies of one level, built in place: tile b covers Z[lo + b * kTile, ...),
// lo = zoff[r0]. The level's rules [r0, r1) have increasing zoff, and a rule's
// body is three symbols: the leading zero at zoff[r], child a at zoff[r] + 1,
// child b shifted by a's sum at zoff[r] + 1 + |a|. Each block finds the rules
// overlapping its tile (wave-wide searches of zoff) and stages their symbols
// straight into LDS; children are at shallower, already-built levels.
template <uint32_t kItems>
__global__ __launch_bounds__(kBlock) void k_expand_level(
    const RuleRec *__restrict__ rec, const uint4 *__restrict__ szo,
    const uint64_t *__restrict__ zoff, uint64_t r0,
    uint64_t r1, uint32_t min_id, uint64_t R, uint64_t lo, uint64_t n_out, int32_t *Z,
    uint32_t zraw) {
  constexpr uint32_t kTile = kBlock * kItems;
  __shared__ ZTileSmem<kItems> sm;
  __shared__ uint64_t s_rf, s_rl;
  __shared__ uint32_t s_q0, s_n;
  const uint64_t o0 = (uint64_t)blockIdx.x * kTile;
  const uint32_t n_here = (uint32_t)(o0 + kTile < n_out ? kTile : n_out - o0);
  const uint64_t origin = lo + o0, last = origin + n_here - 1;
  ztile_begin(sm);
  if (threadIdx.x < kWave) {
    const uint64_t rf = wave_last_le_u64(zoff, r0, r1, origin);
    const uint64_t rl = wave_last_le_u64(zoff, rf, r1, last);
    if (threadIdx.x == 0) {
      // The first symbol is the last of rule rf's three that starts at or before
      // origin; the last symbol likewise for rule rl and the tile's last position.
      const uint64_t zf = zoff[rf], zl = zoff[rl];
      const uint32_t saf = rec[rf].sa, sal = rec[rl].sa;
      const uint32_t q0 = origin >= zf + 1 + saf ? 2u : (origin >= zf + 1 ? 1u : 0u);
      const uint32_t q1 = last >= zl + 1 + sal ? 2u : (last >= zl + 1 ? 1u : 0u);
      s_rf = rf;
      s_rl = rl;
      s_q0 = q0;
      s_n = (uint32_t)(3 * (rl - rf) + q1 - q0 + 1);
    }
  }
  __syncthreads();
  const uint64_t rf = s_rf, rl = s_rl;
  const uint32_t q0 = s_q0, nsym = s_n;
  for (uint64_t r = rf + threadIdx.x; r <= rl; r += kBlock) {
    const RuleRec c = rec[r];
    const uint64_t z0 = zoff[r];
    const bool ra = !is_rule(c.a, min_id, R), rb = !is_rule(c.b, min_id, R);
    // A rule child's record gives its body position and, for child a, the sum
    // that shifts child b (one gather per rule child).
    const uint4 wa = ra ? make_uint4(0u, 0u, 0u, 0u) : szo[uabs(c.a) - min_id];
    const uint32_t suma = ra ? (uint32_t)c.a : (c.a > 0 ? wa.y : 0u - wa.y);
    const SymInfo sa = ra ? SymInfo{(uint32_t)c.a, zraw}
                          : (c.a > 0 ? SymInfo{0u, wa.z + 1}
                                     : SymInfo{0u - wa.y, (wa.z + wa.x - 1) | kRevBit});
    const SymInfo sb = rb ? SymInfo{suma + (uint32_t)c.b, zraw}
                          : rule_sym(c.b, suma, true, szo, min_id);
    const uint64_t g0 = 3 * (r - rf);  // symbol q of rule r is g0 + q - q0
#pragma unroll
    for (uint32_t q = 0; q < 3; ++q) {
      if (g0 + q < q0 || g0 + q - q0 >= nsym)
        continue;
      const uint64_t st = q == 0 ? z0 : (q == 1 ? z0 + 1 : z0 + 1 + c.sa);
      const SymInfo si = q == 0 ? SymInfo{0u, zraw} : (q == 1 ? sa : sb);
      ztile_stage(sm, (uint32_t)(g0 + q - q0), (uint32_t)(st - origin), si);
    }
  }
  ztile_finish<kItems, false>(sm, n_here, Z, zraw, Z + origin);
}

// Expansion by descending the rule tree (no Z): node IDs, no inverse delta.
constexpr uint32_t kDescItems = 16;
constexpr uint32_t kDescTile = kBlock * kDescItems;

__global__ __launch_bounds__(kBlock) void k_expand_descent(
    const int32_t *__restrict__ enc, const uint64_t *__restrict__ sym_off, uint64_t n_out,
    const TileDesc *__restrict__ tdesc, const RuleRec *__restrict__ rec, uint64_t R,
    uint32_t min_id, int32_t *__restrict__ out) {
  __shared__ StartMap<kDescTile> map;
  __shared__ uint16_t srel[kDescTile];
  const uint64_t o0 = (uint64_t)blockIdx.x * kDescTile;
  const uint64_t o_end = o0 + kDescTile < n_out ? o0 + kDescTile : n_out;
  const TileDesc td = tdesc[blockIdx.x];
  const uint64_t k0 = td.k0;
  const uint32_t nsym = (uint32_t)(td.k_last - k0 + 1);
  map.clear();
  __syncthreads();
  for (uint32_t s = threadIdx.x; s < nsym; s += kBlock) {
    const uint32_t rel = s == 0 ? 0u : (uint32_t)(sym_off[k0 + s] - o0);
    srel[s] = (uint16_t)rel;
    map.mark(rel);
  }
  __syncthreads();
  map.rank_words();
  const uint64_t first_off = td.first_off;
  for (uint32_t j = 0; j < kDescItems; ++j) {
    const uint32_t p = j * kBlock + threadIdx.x;
    const uint64_t o = o0 + p;
    if (o >= o_end)
      break;
    const uint32_t s = map.symbol(p);
    uint64_t q = s == 0 ? o - first_off : (uint64_t)(p - srel[s]);
    int32_t v = enc[k0 + s];
    while (is_rule(v, min_id, R)) {
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
    }
    out[o] = v;
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

// Tile-size candidates of the Z expansion (output elements per thread).
constexpr uint32_t kZItems[] = {8, 16, 32};

template <uint32_t kItems>
void launch_z(Ctx &ctx, bool diff, const SymInfo *syms, const uint64_t *off, uint64_t base,
              uint64_t n_out, const TileDesc *td, uint64_t n_tiles, uint32_t rows,
              const int32_t *Z, uint32_t zraw, int32_t *out) {
  uint64_t cols = 0;
  const uint64_t grid = row_grid(n_tiles, rows, cols);
  if (grid > UINT32_MAX)
    throw Error(kErrInvalid, "expansion too large for one launch");
  if (diff)
    hipLaunchKernelGGL((k_expand_z<kItems, true>), dim3((uint32_t)grid), dim3(kBlock), 0,
                       ctx.stream(), syms, off, base, n_out, td, n_tiles, rows, cols, Z, zraw,
                       out);
  else
    hipLaunchKernelGGL((k_expand_z<kItems, false>), dim3((uint32_t)grid), dim3(kBlock), 0,
                       ctx.stream(), syms, off, base, n_out, td, n_tiles, rows, cols, Z, zraw,
                       out);
  GFZ_CHECK_LAUNCH();
}

// Rows of the cross-traversal tile order for a batch of n_trav traversals over
// n_tiles tiles: one per traversal while every row keeps a few columns, else the
// plain order. The tuned key "expand_rows" = 0 pins the plain order.
uint32_t expansion_rows(Ctx &ctx, uint64_t n_trav, uint64_t n_tiles) {
  if (ctx.tuned("expand_rows") == 0 || n_trav < 2 || n_trav > (1u << 20) ||
      n_tiles < 4 * kXcds * n_trav)
    return 1;
  return (uint32_t)n_trav;
}

// out[base + o] for o in [0, n_out) from the symbols syms[0..n_sym) whose
// starts (absolute, offset by base) are off[0..n_sym] - through Z. The output
// holds n_trav traversals (used for the tile order only).
void expand_through_z(Ctx &ctx, const char *group, bool diff, const SymInfo *syms,
                      const uint64_t *off, uint64_t n_sym, uint64_t base, uint64_t n_out,
                      uint64_t n_trav, const int32_t *Z, uint32_t zraw, int32_t *out) {
  if (!n_out)
    return;
  auto run = [&](uint32_t items, bool timed) {
    const uint64_t tile = (uint64_t)kBlock * items;
    const uint64_t n_tiles = ceil_div(n_out, tile);
    if (n_tiles > UINT32_MAX)
      throw Error(kErrInvalid, "expansion too large for one launch");
    const uint32_t rows = expansion_rows(ctx, n_trav, n_tiles);
    DBuf<TileDesc> td(ctx.pool(), n_tiles);
    hipLaunchKernelGGL(k_tile_desc, dim3(stride_grid(ctx, n_sym)), dim3(kBlock), 0, ctx.stream(),
                       off, n_sym, base, n_out, (uint32_t)tile, n_tiles, td.get());
    GFZ_CHECK_LAUNCH();
    auto go = [&] {
      switch (items) {
        case 8: launch_z<8>(ctx, diff, syms, off, base, n_out, td.get(), n_tiles, rows, Z, zraw, out); break;
        case 32: launch_z<32>(ctx, diff, syms, off, base, n_out, td.get(), n_tiles, rows, Z, zraw, out); break;
        default: launch_z<16>(ctx, diff, syms, off, base, n_out, td.get(), n_tiles, rows, Z, zraw, out); break;
      }
    };
    if (timed)
      return ctx.time_ms(go);
    ctx.timed(group, go);
    return 0.0f;
  };
  // Size classes of 4x; small expansions keep the default.
  uint32_t cls = 0;
  for (uint64_t n = n_out; n >= 4; n >>= 2)
    ++cls;
  const std::string key = std::string("expand_items/") + std::to_string(cls);
  int items = ctx.tuned(key);
  if (items < 0 && ctx.autotune() && n_out >= (uint64_t(1) << 20)) {
    float best = 0;
    for (uint32_t c : kZItems) {
      run(c, true);  // warm
      const float ms = run(c, true);
      if (items < 0 || ms < best) {
        best = ms;
        items = (int)c;
      }
    }
    ctx.set_tuned(key, items);
  }
  run(items < 0 ? kExItems : (uint32_t)items, false);
}

}  // namespace

template <uint32_t kItems>
void launch_level(Ctx &ctx, const Rulebook &rb, uint64_t r0, uint64_t r1, uint32_t min_id,
                  uint64_t R, uint64_t lo, uint64_t n_out) {
  const uint64_t n_tiles = ceil_div(n_out, (uint64_t)kBlock * kItems);
  if (n_tiles > UINT32_MAX)
    throw Error(kErrInvalid, "rule bodies of one level too large for one launch");
  hipLaunchKernelGGL(k_expand_level<kItems>, dim3((uint32_t)n_tiles), dim3(kBlock), 0,
                     ctx.stream(), rb.rec.get(), rb.szo.get(), rb.zoff.get(), r0,
                     r1, min_id, R, lo, n_out, rb.z.get(), rb.z_raw);
  GFZ_CHECK_LAUNCH();
}

// Z[lo, lo + n_out) for the rules [r0, r1) of one level, with the tile size
// autotuned per size class like the decode expansion (every candidate writes
// the same bodies).
void build_level(Ctx &ctx, const Rulebook &rb, uint64_t r0, uint64_t r1, uint32_t min_id,
                 uint64_t R, uint64_t lo, uint64_t n_out) {
  if (!n_out)
    return;
  auto go = [&](uint32_t items) {
    switch (items) {
      case 8: launch_level<8>(ctx, rb, r0, r1, min_id, R, lo, n_out); break;
      case 32: launch_level<32>(ctx, rb, r0, r1, min_id, R, lo, n_out); break;
      default: launch_level<16>(ctx, rb, r0, r1, min_id, R, lo, n_out); break;
    }
  };
  uint32_t cls = 0;
  for (uint64_t n = n_out; n >= 4; n >>= 2)
    ++cls;
  const std::string key = std::string("level_items/") + std::to_string(cls);
  int items = ctx.tuned(key);
  if (items < 0 && ctx.autotune() && n_out >= (uint64_t(1) << 20)) {
    float best = 0;
    for (uint32_t c : kZItems) {
      ctx.time_ms([&] { go(c); });  // warm
      const float ms = ctx.time_ms([&] { go(c); });
      if (items < 0 || ms < best) {
        best = ms;
        items = (int)c;
      }
    }
    ctx.set_tuned(key, items);
  }
  const uint32_t use = items < 0 ? kExItems : (uint32_t)items;
  ctx.timed("build_bodies", [&] { go(use); });
}

void rulebook_build(Ctx &ctx, const int32_t *d_first, const int32_t *d_second, uint64_t R,
                    uint32_t min_id, Rulebook &rb) {
  rb = Rulebook();
  rb.n_rules = R;
  rb.min_id = min_id;
  if ((uint64_t)min_id + R > (uint64_t)INT32_MAX + 1)
    throw Error(kErrCorrupt, "rule-ID window exceeds the int32 symbol range");
  rb.rec = DBuf<RuleRec>(ctx.pool(), std::max<uint64_t>(R, 1));
  rb.ss = DBuf<uint2>(ctx.pool(), std::max<uint64_t>(R, 1));
  if (!R) {  // no rules: Z is just the two zeros node IDs point at
    rb.zoff = DBuf<uint64_t>(ctx.pool(), 1);
    fill_u64(ctx, rb.zoff.get(), 0, 1);
    rb.z = DBuf<int32_t>(ctx.pool(), 2);
    fill_u32(ctx, reinterpret_cast<uint32_t *>(rb.z.get()), 0, 2);
    rb.z_raw = 1;
    rb.szo = DBuf<uint4>(ctx.pool(), 1);
    rb.has_bodies = true;
    ctx.sync();
    return;
  }
  DBuf<uint32_t> depth(ctx.pool(), R);
  // Passes run in batches without a host round trip: a pass that resolves
  // nothing leaves nothing for the later passes of its batch either.
  constexpr uint32_t kPassBatch = 8;
  DBuf<uint32_t> flags(ctx.pool(), kPassBatch + 4);  // changed per pass, err, bad, pad
  uint32_t *err = flags.get() + kPassBatch, *bad = flags.get() + kPassBatch + 1;
  // First unresolved rule before each pass of a batch (lows[0] carries the last
  // batch's lows[kPassBatch]); the first pass starts at the first non-leaf rule.
  DBuf<uint32_t> lows(ctx.pool(), kPassBatch + 1);
  fill_u32(ctx, lows.get(), ~0u, 1);
  hipLaunchKernelGGL(k_rule_init, dim3(stride_grid(ctx, R)), dim3(kBlock), 0, ctx.stream(),
                     d_first, d_second, R, min_id, rb.rec.get(), rb.ss.get(), depth.get(),
                     lows.get());
  GFZ_CHECK_LAUNCH();
  uint32_t max_depth = 1;
  bool converged = false;
  for (uint64_t pass0 = 0; !converged && pass0 <= R; pass0 += kPassBatch) {
    fill_u32(ctx, flags.get(), 0, kPassBatch + 4);
    fill_u32(ctx, lows.get() + 1, ~0u, kPassBatch);
    ctx.timed("rule_sizes", [&] {
      for (uint32_t k = 0; k < kPassBatch; ++k)
        hipLaunchKernelGGL(k_rule_pass, dim3(stride_grid(ctx, R)), dim3(kBlock), 0,
                           ctx.stream(), rb.rec.get(), rb.ss.get(), depth.get(),
                           R, min_id, (uint32_t)(pass0 + k), flags.get() + k, err,
                           lows.get() + k, lows.get() + k + 1);
    });
    GFZ_CHECK_LAUNCH();
    GFZ_HIP(hipMemcpyAsync(lows.get(), lows.get() + kPassBatch, 4, hipMemcpyDeviceToDevice,
                           ctx.stream()));
    const uint32_t *h =
        reinterpret_cast<const uint32_t *>(ctx.fetch({{flags.get(), (kPassBatch + 4) / 2, 0}}));
    if (h[kPassBatch])
      throw Error(kErrCorrupt, "rule expansion size exceeds 2^32 symbols");
    for (uint32_t k = 0; k < kPassBatch; ++k) {
      if (!h[k]) {
        converged = true;
        break;
      }
      max_depth = (uint32_t)(pass0 + k) + 2;
    }
  }
  rb.depth = max_depth;

  // Unresolved rules (a cycle), level boundaries, Z offsets (size + 1 entries
  // per rule) and the Z offset of each level, read back together.
  std::vector<uint64_t> bound(max_depth + 2, R), lvl_off(max_depth + 2);
  rb.zoff = DBuf<uint64_t>(ctx.pool(), R + 1);
  {
    fill_u32(ctx, flags.get(), 0, kPassBatch + 4);
    hipLaunchKernelGGL(k_rule_unresolved, dim3(stride_grid(ctx, R)), dim3(kBlock), 0,
                       ctx.stream(), depth.get(), R, err);
    GFZ_CHECK_LAUNCH();
    DBuf<uint64_t> d_bound(ctx.pool(), max_depth + 2), d_lvl(ctx.pool(), max_depth + 2);
    fill_u64(ctx, d_bound.get(), R, max_depth + 2);
    hipLaunchKernelGGL(k_depth_bounds, dim3(stride_grid(ctx, R)), dim3(kBlock), 0, ctx.stream(),
                       depth.get(), R, max_depth, d_bound.get(), bad);
    GFZ_CHECK_LAUNCH();
    DBuf<uint32_t> szp1(ctx.pool(), R);
    hipLaunchKernelGGL(k_plus_one, dim3(stride_grid(ctx, R)), dim3(kBlock), 0, ctx.stream(),
                       rb.ss.get(), R, szp1.get());
    GFZ_CHECK_LAUNCH();
    exclusive_scan(ctx, szp1.get(), rb.zoff.get(), R);
    hipLaunchKernelGGL(k_gather_u64, dim3(1), dim3(kBlock), 0, ctx.stream(), rb.zoff.get(),
                       d_bound.get(), max_depth + 2, d_lvl.get());
    GFZ_CHECK_LAUNCH();
    const size_t nb = max_depth + 2;
    constexpr uint32_t kFlagWords = (kPassBatch + 4) / 2;
    uint32_t hf[kPassBatch + 4];
    if (2 * nb + kFlagWords <= 512) {
      const uint64_t *m = ctx.fetch({{flags.get(), kFlagWords, 0},
                                     {d_bound.get(), (uint32_t)nb, kFlagWords},
                                     {d_lvl.get(), (uint32_t)nb, (uint32_t)(kFlagWords + nb)}});
      std::memcpy(hf, m, sizeof(hf));
      std::copy(m + kFlagWords, m + kFlagWords + nb, bound.begin());
      std::copy(m + kFlagWords + nb, m + kFlagWords + 2 * nb, lvl_off.begin());
    } else {  // deep rulebook: pageable copies
      GFZ_HIP(hipMemcpyAsync(bound.data(), d_bound.get(), nb * 8, hipMemcpyDeviceToHost,
                             ctx.stream()));
      GFZ_HIP(hipMemcpyAsync(lvl_off.data(), d_lvl.get(), nb * 8, hipMemcpyDeviceToHost,
                             ctx.stream()));
      std::memcpy(hf, ctx.fetch({{flags.get(), kFlagWords, 0}}), sizeof(hf));
    }
    if (hf[kPassBatch])
      throw Error(kErrCorrupt, "rulebook is cyclic: some rule never bottoms out in node IDs");
    if (hf[kPassBatch + 1]) {
      rb.zoff.reset();
      return;  // depth not monotone in rule ID: expansion descends the tree instead
    }
  }
  depth.reset();
  const uint64_t total = lvl_off[max_depth + 1];
  // Default budget: a quarter of the free device memory, at most 2^31 - 2
  // elements (31-bit Z indices; 8 GiB).
  uint64_t budget = uint64_t(1) << 30;
  {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
      budget = std::max<uint64_t>(budget, (uint64_t)free_b / 4 / 4);
    else
      (void)hipGetLastError();
  }
  if (const char *env = std::getenv("GFAZ_BODY_BUDGET_MB"))
    budget = std::strtoull(env, nullptr, 10) * (uint64_t(1) << 20) / 4;
  budget = std::min<uint64_t>(budget, (uint64_t(1) << 31) - 2);  // 31-bit Z indices
  if (total > budget) {
    rb.zoff.reset();
    return;
  }
  rb.z = DBuf<int32_t>(ctx.pool(), total + 2);
  rb.z_elems = total;
  rb.z_raw = (uint32_t)(total + 1);
  fill_u32(ctx, reinterpret_cast<uint32_t *>(rb.z.get()) + total, 0, 2);
  rb.szo = DBuf<uint4>(ctx.pool(), R);
  hipLaunchKernelGGL(k_rule_szo, dim3(stride_grid(ctx, R)), dim3(kBlock), 0, ctx.stream(),
                     rb.ss.get(), rb.zoff.get(), R, rb.szo.get());
  GFZ_CHECK_LAUNCH();
  // One level at a time: children are always at shallower, already-built levels.
  for (uint32_t d = 1; d <= max_depth; ++d) {
    const uint64_t r0 = bound[d], r1 = bound[d + 1];
    if (r0 >= r1)
      continue;
    const uint64_t lo = lvl_off[d], hi = lvl_off[d + 1];
    build_level(ctx, rb, r0, r1, min_id, R, lo, hi - lo);
  }
  // No synchronisation: the bodies are built on the context stream, which every
  // use of the rulebook runs on, and the rule columns were last read before the
  // readbacks above returned.
  rb.has_bodies = true;
}

void decode_plan(Ctx &ctx, const Rulebook &rb, const int32_t *d_enc, uint64_t n_enc,
                 const uint32_t *enc_lengths, const uint32_t *orig_lengths, uint64_t n_trav,
                 int delta_round, DecodePlan &plan, DBuf<uint64_t> &sym_off) {
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
  // Encoded start of every traversal, then n_enc: the traversal starts for the
  // planning kernels and the gather positions of the decoded offsets. Staged in
  // pinned memory (an asynchronous copy).
  uint64_t *h_idx = static_cast<uint64_t *>(ctx.pinned((n_trav + 1) * 8));
  std::copy(enc_start.begin(), enc_start.end(), h_idx);
  h_idx[n_trav] = n_enc;
  DBuf<uint64_t> idx(ctx.pool(), n_trav + 1), val(ctx.pool(), n_trav + 1);
  GFZ_HIP(hipMemcpyAsync(idx.get(), h_idx, (n_trav + 1) * 8, hipMemcpyHostToDevice,
                         ctx.stream()));
  // Records only when the expansion goes through rule bodies and a decode follows.
  const int rec = (delta_round >= 0 && rb.has_bodies) ? (delta_round > 0 ? 1 : 0) : -1;
  if (rec >= 0)
    plan.syms = DBuf<SymInfo>(ctx.pool(), std::max<uint64_t>(n_enc, 1));
  if (n_enc) {
    const uint64_t n_tiles = ceil_div(n_enc, kPlanTile);
    DBuf<PlanAgg> tiles(ctx.pool(), n_tiles);
    DBuf<uint2> szss(ctx.pool(), n_enc);  // each symbol's size and signed sum
    DBuf<uint32_t> zos;                   // and its rule's Z offset (decodes through Z)
    if (rec >= 0)
      zos = DBuf<uint32_t>(ctx.pool(), n_enc);
    // Same stretch of every traversal together (block_tile): its symbols gather
    // the same rules' sizes, sums and Z offsets.
    const uint32_t rows = expansion_rows(ctx, n_trav, n_tiles);
    uint64_t cols = 0;
    const uint64_t grid = row_grid(n_tiles, rows, cols);
    if (grid > UINT32_MAX)
      throw Error(kErrInvalid, "encoded block too large for one launch");
    ctx.timed("decode_plan", [&] {
      hipLaunchKernelGGL(k_plan_reduce, dim3((uint32_t)grid), dim3(kBlock), 0, ctx.stream(),
                         d_enc, n_enc, idx.get(), n_trav + 1, rb.ss.get(),
                         rec >= 0 ? rb.szo.get() : nullptr, rb.n_rules, rb.min_id, n_tiles,
                         rows, cols, tiles.get(), szss.get(), zos.get());
      GFZ_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_plan_scan, dim3(1), dim3(kPlanScanThreads), 0, ctx.stream(),
                         tiles.get(), n_tiles, sym_off.get() + n_enc);
      GFZ_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_plan_apply, dim3((uint32_t)grid), dim3(kBlock), 0, ctx.stream(),
                         d_enc, n_enc, idx.get(), n_trav + 1, szss.get(), zos.get(),
                         rb.n_rules, rb.min_id, rb.z_raw, rec, tiles.get(), n_tiles, rows, cols,
                         sym_off.get(), plan.syms.get());
      GFZ_CHECK_LAUNCH();
    });
  } else {
    fill_u64(ctx, sym_off.get(), 0, 1);
  }
  // Decoded offset of each traversal = symbol offset at its first encoded symbol.
  plan.out_offsets.assign(n_trav + 1, 0);
  hipLaunchKernelGGL(k_gather_u64, dim3(stride_grid(ctx, n_trav + 1)), dim3(kBlock), 0,
                     ctx.stream(), sym_off.get(), idx.get(), n_trav + 1, val.get());
  GFZ_CHECK_LAUNCH();
  if (n_trav + 1 <= 512) {
    const uint64_t *m = ctx.fetch({{val.get(), (uint32_t)(n_trav + 1), 0}});
    std::copy(m, m + n_trav + 1, plan.out_offsets.begin());
  } else {
    GFZ_HIP(hipMemcpyAsync(plan.out_offsets.data(), val.get(), (n_trav + 1) * 8,
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
  const bool delta = delta_round > 0;
  int scans_left = delta_round;
  if (rb.has_bodies) {
    // Symbol records from the plan (with an inverse delta to undo, the first one
    // is folded in through the traversal-level sums of the encoded symbols).
    if (!plan.syms.get())
      throw Error(kErrInvalid, "decode plan was made without symbol records");
    const SymInfo *syms = plan.syms.get();
    if (delta)
      scans_left -= 1;
    expand_through_z(ctx, "expand", !delta, syms, d_sym_off, n_enc, 0, n_out, n_trav,
                     rb.z.get(), rb.z_raw, d_out);
    ctx.add_bytes("expand", 4.0 * n_enc + 4.0 * n_out + 8.0 * rb.n_rules);
  } else {
    const uint64_t n_tiles = ceil_div(n_out, kDescTile);
    if (n_tiles > UINT32_MAX)
      throw Error(kErrInvalid, "decoded batch too large for one launch");
    DBuf<TileDesc> tdesc(ctx.pool(), n_tiles);
    hipLaunchKernelGGL(k_tile_desc, dim3(stride_grid(ctx, n_enc)), dim3(kBlock), 0, ctx.stream(),
                       d_sym_off, n_enc, 0, n_out, kDescTile, n_tiles, tdesc.get());
    GFZ_CHECK_LAUNCH();
    ctx.timed("expand", [&] {
      hipLaunchKernelGGL(k_expand_descent, dim3((uint32_t)n_tiles), dim3(kBlock), 0, ctx.stream(),
                         d_enc, d_sym_off, n_out, tdesc.get(), rb.rec.get(), rb.n_rules,
                         rb.min_id, d_out);
    });
    GFZ_CHECK_LAUNCH();
    ctx.add_bytes("expand", 4.0 * n_enc + 4.0 * n_out + 8.0 * rb.n_rules);
  }
  if (scans_left <= 0) {
    ctx.sync();
    return;
  }
  // Remaining inverse deltas: segmented scans over the decoded stream.
  std::vector<uint64_t> starts;
  for (uint64_t t = 0; t < n_trav; ++t)
    if (plan.out_offsets[t + 1] > plan.out_offsets[t])
      starts.push_back(plan.out_offsets[t]);
  const uint64_t words = ceil_div(n_out, 32);
  DBuf<uint32_t> heads(ctx.pool(), words);
  fill_u32(ctx, heads.get(), 0, words);
  DBuf<uint64_t> d_starts(ctx.pool(), std::max<size_t>(starts.size(), 1));
  if (!starts.empty()) {
    GFZ_HIP(hipMemcpyAsync(d_starts.get(), starts.data(), starts.size() * 8,
                           hipMemcpyHostToDevice, ctx.stream()));
    set_head_bits(ctx, d_starts.get(), starts.size(), heads.get(), n_out);
  }
  for (int r = 0; r < scans_left; ++r)
    segmented_prefix_sum(ctx, d_out, n_out, heads.get());
  ctx.sync();
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
