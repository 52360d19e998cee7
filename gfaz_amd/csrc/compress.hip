// Grammar compression of GFA traversals on gfx950.
//
// Restates, with the same output bit for bit, the reference CPU pipeline:
//   delta transform + max |symbol|  (src/workflows/compression_workflow.cpp:235-265,
//                                    src/codec/codec.cpp:45-76)
//   per round (compression_workflow.cpp:567-690):
//     repeated canonical pairs      (src/grammar/rule_generator.cpp:148-260)
//     greedy non-overlapping replace (src/grammar/path_encoder.cpp:29-83)
//     used rules sorted by signed key -> IDs next+rank, remap
//                                   (compression_workflow.cpp:601-680,
//                                    src/workflows/compression_utils.cpp:201-221)
//
// Layout in HBM: one flat int32 stream of all traversals (paths, then walks), a
// head bitmap marking each non-empty traversal's first element, a 64-bit
// element index everywhere. Work is split into fixed tiles of the flat stream,
// independent of traversal lengths; the greedy pick recurrence crosses tiles
// through a one-bit carry (bits.hpp).
#include "bits.hpp"
#include "codec.hpp"
#include "prims.hpp"


#include <algorithm>
#include <chrono>

namespace gfz {

namespace {

constexpr uint32_t kInsItems = 8;                   // pair-insert tile: 2048 elements
constexpr uint32_t kInsTile = kBlock * kInsItems;
constexpr uint32_t kRtItems = 16;                   // cand/emit tile: 4096 elements
constexpr uint32_t kRtTile = kBlock * kRtItems;     // = 64 words of 64 candidates
constexpr uint32_t kRtWords = kRtTile / 64;
static_assert(kRtWords == kWave, "one candidate word per lane of one wave");
constexpr uint32_t kTileScanThreads = 1024;
constexpr uint64_t kTileScanSingle = 32768;  // tiles one block scans alone

__device__ inline bool head_at(const uint32_t *heads, uint64_t i) {
  return (heads[i >> 5] >> (i & 31)) & 1u;
}

// ---- cross-traversal tile order ------------------------------------------------
// Which kRtTile tile of the flat stream a block takes. Traversal h owns the tiles
// whose first element lies in [start[h], start[h + 1]) (start[n] = L), so every
// tile has exactly one owner. Blocks take column c (the c-th tile of every
// traversal) for all traversals together, and the columns are dealt to the XCDs
// in turn (the dispatcher hands consecutive blocks to the eight XCDs
// round-robin): one XCD then marks the same stretch of every haplotype at once,
// so the rule-table lines of the pairs those haplotypes share can still be in its
// L2 when the next haplotype looks them up. Most keys of the later rounds occur
// two or three times, so the gain is small (on plain_10m about 1 ms of the
// 29 ms of candidate marking per compress, in the first two rounds).
// start == nullptr keeps the launch order.
constexpr uint32_t kXcdCount = 8;
constexpr uint64_t kNoTile = ~0ull;
struct TileOrder {
  const uint64_t *start;  // first element of each non-empty traversal
  uint64_t L;
  uint32_t n;             // traversals
  uint32_t cols;          // most tiles any traversal owns
};

__device__ inline uint64_t order_tile(const TileOrder &o) {
  const uint64_t b = blockIdx.x;
  if (!o.start)
    return b;
  const uint64_t i = b / kXcdCount;
  const uint64_t c = (i / o.n) * kXcdCount + b % kXcdCount;
  if (c >= o.cols)
    return kNoTile;
  const uint32_t h = (uint32_t)(i % o.n);
  const uint64_t e = h + 1 < o.n ? o.start[h + 1] : o.L;
  const uint64_t t = (o.start[h] + kRtTile - 1) / kRtTile + c;
  return t * kRtTile < e ? t : kNoTile;
}

// Blocks of a launch in that order over traversals of at most max_len elements.
inline uint64_t order_grid(uint32_t n, uint64_t max_len, uint32_t &cols) {
  cols = (uint32_t)((max_len + kRtTile - 1) / kRtTile);
  return (uint64_t)kXcdCount * n * ((cols + kXcdCount - 1) / kXcdCount);
}

// *longest = the longest traversal of the next stream (length *L_next).
__global__ void k_trav_longest(const uint64_t *__restrict__ start, uint32_t n,
                               const uint64_t *__restrict__ L_next,
                               unsigned long long *__restrict__ longest) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h < n) {
    const uint64_t e = h + 1 < n ? start[h + 1] : *L_next;
    atomicMax(longest, (unsigned long long)(e - start[h]));
  }
}

// Keys are compacted before they reach any table: every symbol of a round has
// |v| < next = span / 2, so ck = (a + next) * span + (b >= 0 ? b : b + span)
// keeps the signed packed-key order of (a, b) (a signed, then b as unsigned 32
// bits, include/grammar/packed_2mer.hpp:15-17 and the int64 sort of
// src/workflows/compression_workflow.cpp:648-655) in fewer than 62 bits. That
// leaves the top bit free for the "repeated" mark of the counting table, and
// all-ones free as the empty slot.
constexpr uint64_t kFree = ~0ull;
constexpr uint64_t kRep = 1ull << 63;
// Slot words of the candidate pass (cold candidates only): a rule slot (below
// 2^30: the rule table has a power-of-two capacity with next + capacity < 2^31),
// bit 31 for the reverse orientation.
constexpr uint32_t kSlotRev = 1u << 31;

// The quotient by span (< 2^31) from a double estimate, which is within one of
// it: c < span^2 is exact in a double below 2^53, and above that its rounding
// moves c / span by under 2^-21.
__device__ inline int64_t expand_key(uint64_t c, uint64_t span) {
  const uint32_t sp = (uint32_t)span, half = sp / 2;
  uint32_t q = (uint32_t)((double)c / (double)sp);
  int64_t r = (int64_t)c - (int64_t)((uint64_t)q * sp);
  if (r < 0) {
    --q;
    r += sp;
  } else if (r >= (int64_t)sp) {
    ++q;
    r -= sp;
  }
  const int32_t a = (int32_t)(q - half);
  const int32_t b = (uint32_t)r < half ? (int32_t)r : (int32_t)((uint32_t)r - sp);
  return pack2(a, b);
}
// compact_key(canon2(pack2(a, b))) in 32-bit arithmetic: the canonical pair is
// (a, b) or its reverse (-b, -a), whichever packs smaller (a signed, then b as
// an unsigned 32-bit word), and both factors of the product fit 31 bits.
__device__ inline uint64_t pair_key(int32_t a, int32_t b, uint64_t span) {
  const int32_t ra = (int32_t)(0u - (uint32_t)b), rb = (int32_t)(0u - (uint32_t)a);
  const bool rev = ra < a || (ra == a && (uint32_t)rb < (uint32_t)b);
  const int32_t ca = rev ? ra : a, cb = rev ? rb : b;
  const uint32_t sp = (uint32_t)span, half = sp / 2;
  return (uint64_t)((uint32_t)ca + half) * sp + (uint32_t)(cb >= 0 ? cb : (int32_t)((uint32_t)cb + sp));
}

// Fresh pairs. The reference encoder is a left-to-right stack greedy
// (src/grammar/path_encoder.cpp:29-83): two symbols that stay adjacent and
// unreplaced through a round did not form a rule of that round, so their
// canonical key occurred at most once in it (rule_generator.cpp:148-260), and
// it occurs at most once in the next stream too (each of its occurrences there
// is one of the round before). So from the second round on, only pairs with a
// symbol of the last round's rules (|v| >= fresh, the floor of those IDs) can
// repeat or be rules; the others are skipped by counting and candidate marking
// alike, which leaves the grammar unchanged. (Equivalently: a rule made in
// round r > 1 has depth r, so one of its children is a rule of round r - 1.)
// fresh = 0 keeps every pair.
__device__ inline bool fresh_pair(int32_t a, int32_t b, uint32_t fresh) {
  const uint32_t ua = uabs(a), ub = uabs(b);
  return (ua > ub ? ua : ub) >= fresh;
}

// The pairs starting at elements first + j * stride (j < N, one thread's pass):
// a[j], b[j] and a bit set of the valid ones (i + 1 < L, no traversal starts
// at i + 1, and the pair is fresh). Indexes are clamped into [0, L - 2] instead
// of guarded (callers run with L >= 2), so every load of the pass is in flight
// before any is used.
template <uint32_t N>
__device__ inline uint32_t load_pairs(const int32_t *__restrict__ x,
                                      const uint32_t *__restrict__ heads, uint64_t L,
                                      uint64_t first, uint64_t stride, int32_t (&a)[N],
                                      int32_t (&b)[N], uint32_t fresh) {
  uint32_t hw[N];
#pragma unroll
  for (uint32_t j = 0; j < N; ++j) {
    const uint64_t i = first + (uint64_t)j * stride;
    const uint64_t ia = i + 1 < L ? i : L - 2;
    a[j] = x[ia];
    b[j] = x[ia + 1];
    hw[j] = heads[(ia + 1) >> 5];
  }
  uint32_t valid = 0;
#pragma unroll
  for (uint32_t j = 0; j < N; ++j) {
    const uint64_t i = first + (uint64_t)j * stride;
    valid |= (i + 1 < L && !((hw[j] >> ((i + 1) & 31)) & 1) && fresh_pair(a[j], b[j], fresh))
                 ? 1u << j
                 : 0u;
  }
  return valid;
}

// Pair keys of the same pass (kFree where no pair starts); returns the valid set.
template <uint32_t N>
__device__ inline uint32_t load_pair_keys(const int32_t *__restrict__ x,
                                          const uint32_t *__restrict__ heads, uint64_t L,
                                          uint64_t span, uint64_t first, uint64_t stride,
                                          uint64_t (&kv)[N], uint32_t fresh) {
  int32_t a[N], b[N];
  const uint32_t valid = load_pairs<N>(x, heads, L, first, stride, a, b, fresh);
#pragma unroll
  for (uint32_t j = 0; j < N; ++j)
    kv[j] = (valid >> j) & 1 ? pair_key(a[j], b[j], span) : kFree;
  return valid;
}

// Counting table (one 8-byte word per slot): key | kRep once the key was seen
// twice. Its only job is to find the repeated keys of a round.
// Inserts give up after kMaxProbe slots and raise `overflow` (the host then
// retries with a larger table), so no block needs a shared running count and
// every probe loop is short. Per-block key counts go to `block_counts` with
// plain stores and are summed afterwards.
constexpr uint32_t kMaxProbe = 256;

struct CountTable {
  unsigned long long *slots;
  uint64_t mask;
  uint32_t *overflow;
  uint2 *block_counts;  // (new keys, newly repeated keys) per block
  // Keys that became repeated, kStageCap per block (block b at b * kStageCap);
  // a block with more raises *stage_overflow and the table is scanned instead.
  unsigned long long *stage;
  uint32_t *stage_overflow;
};
constexpr uint32_t kStageCap = 512;

// Rule table: the repeated keys of a round, kLine to a 64-byte line. A key sits
// in the first line, from its hash on, that had a free slot when it was
// inserted; keys never leave, so a lookup stops at the first line holding the
// key or a free slot. At load <= 1/4 nearly every lookup reads one line (a
// probe chain of single slots at load 1/2 reads 15-20 in the worst of a
// wave's 512 lookups). The slot index is the key's provisional rule ID within
// the round; vals[slot] becomes its final ID.
// The lines form regions of rlines lines, one per counting bucket (nbmask + 1
// of them; one region when the round was counted in the HBM table): a key's
// first line is in its bucket's region (picked by a multiply-shift of its hash,
// so a region may hold any number of lines), and its probe wraps within it, so
// one block can build a bucket's region alone.
constexpr uint32_t kLine = 8;
struct RuleTable {
  unsigned long long *keys;  // kFree = empty; 64-byte aligned
  uint32_t *vals;
  uint32_t rlines;           // lines per region
  uint32_t nbmask;           // regions - 1 (a power of two minus one)
  // Rounds where most pairs are no rule: one bit per hashed key of the table
  // (2^fbits bits, set as keys go in). A clear bit proves a key absent without
  // reading its line; nullptr in rounds where most pairs are rules.
  uint32_t *filter;
  uint32_t fbits;
};

__device__ inline uint32_t filter_bit(const RuleTable &rt, uint64_t k) {
  return (uint32_t)(key_hash(k) >> 13) & ((1u << rt.fbits) - 1);
}
__device__ inline void filter_add(const RuleTable &rt, uint64_t k) {
  if (rt.filter) {
    const uint32_t b = filter_bit(rt, k);
    atomicOr(&rt.filter[b >> 5], 1u << (b & 31));
  }
}
__device__ inline bool filter_has(const RuleTable &rt, uint64_t k) {
  const uint32_t b = filter_bit(rt, k);
  return (rt.filter[b >> 5] >> (b & 31)) & 1u;
}

// Bucket of a key among mask + 1 (hash bits above those that pick its line).
__device__ inline uint32_t bucket_of(uint64_t k, uint32_t mask) {
  return (uint32_t)(key_hash(k) >> 40) & mask;
}

// A key's first line within its region (of `lines` lines).
__device__ inline uint32_t region_line(uint64_t k, uint32_t lines) {
  return (uint32_t)(((key_hash(k) & 0xffffffffull) * lines) >> 32);
}

// A probe position: the region's first line and the line within the region.
struct RuleProbe {
  uint64_t base;
  uint32_t off;
};

__device__ inline RuleProbe rule_line(const RuleTable &rt, uint64_t k) {
  return RuleProbe{(uint64_t)bucket_of(k, rt.nbmask) * rt.rlines, region_line(k, rt.rlines)};
}

__device__ inline void rule_next_line(const RuleTable &rt, RuleProbe &p) {
  p.off = p.off + 1 == rt.rlines ? 0u : p.off + 1;
}

__device__ inline void load_line(const unsigned long long *keys, uint64_t line,
                                 uint64_t (&w)[kLine]) {
  const uint4 *p = reinterpret_cast<const uint4 *>(keys + line * kLine);
#pragma unroll
  for (uint32_t q = 0; q < kLine / 2; ++q) {
    const uint4 v = p[q];
    w[2 * q] = (uint64_t)v.x | ((uint64_t)v.y << 32);
    w[2 * q + 1] = (uint64_t)v.z | ((uint64_t)v.w << 32);
  }
}

// Looks up N keys at once: every round of the loop issues the line loads of all
// pending keys before looking at any of them. slot[j] = -1 when absent or not
// valid.
template <int N>
__device__ inline void rule_find_batch(const RuleTable &rt, const uint64_t (&k)[N],
                                       const bool (&valid)[N], int64_t (&slot)[N]) {
  RuleProbe l[N];
  uint32_t pending = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    l[j] = rule_line(rt, k[j]);
    slot[j] = -1;
    if (valid[j])
      pending |= 1u << j;
  }
  for (uint32_t p = 0; pending && p < rt.rlines; ++p) {
    uint64_t w[N][kLine];
#pragma unroll
    for (int j = 0; j < N; ++j)
      if ((pending >> j) & 1)
        load_line(rt.keys, l[j].base + l[j].off, w[j]);
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (!((pending >> j) & 1))
        continue;
      int hit = -1;
      bool has_free = false;
#pragma unroll
      for (uint32_t q = 0; q < kLine; ++q) {
        hit = w[j][q] == k[j] ? (int)q : hit;
        has_free |= w[j][q] == kFree;
      }
      if (hit >= 0) {
        slot[j] = (int64_t)((l[j].base + l[j].off) * kLine + (uint32_t)hit);
        pending &= ~(1u << j);
      } else if (has_free) {
        pending &= ~(1u << j);
      } else {
        rule_next_line(rt, l[j]);
      }
    }
  }
}

__device__ inline int64_t rule_find(const RuleTable &rt, uint64_t k) {
  const uint64_t kk[1] = {k};
  const bool v[1] = {true};
  int64_t s[1];
  rule_find_batch<1>(rt, kk, v, s);
  return s[0];
}

// Inserts a key that is not in the table yet; false when kMaxProbe lines were
// full. Which slots are taken is decided by the CAS results alone: a plain read
// of the line may show a slot free that another block has just filled (a line
// cached before that block's CAS; the CAS then fails and the next slot is
// tried), never the reverse, so starting at the first slot seen free keeps
// every key in the first line that had room.
__device__ inline bool rule_insert(const RuleTable &rt, uint64_t k) {
  RuleProbe l = rule_line(rt, k);
  for (uint32_t p = 0; p < kMaxProbe; ++p) {
    unsigned long long *line = rt.keys + (l.base + l.off) * kLine;
    uint64_t w[kLine];
    load_line(rt.keys, l.base + l.off, w);
    uint32_t j = 0;
    while (j < kLine && w[j] != kFree)
      ++j;
    for (; j < kLine; ++j) {
      const unsigned long long old =
          atomicCAS(line + j, (unsigned long long)kFree, (unsigned long long)k);
      if (old == kFree || old == k)
        return true;
    }
    rule_next_line(rt, l);
  }
  return false;
}

// Count up to N keys (bit j of `have`), key j seen cnt[j] times in the caller's
// tile. A plain probe runs first: a key never leaves its slot, so finding it is
// final, and the hot keys every tile shares (the backbone's (+1,+1) and its
// bubble neighbours) already carry kRep and cost no atomic at all. Missing keys
// claim an empty slot with a CAS. Every pending key issues its memory operation
// in the same round. fresh/reps count keys new to the table and keys this call
// turned repeated (each key is turned repeated exactly once, so the block sums
// give exact totals).
template <int N>
__device__ inline void count_insert_batch(const CountTable &t, const uint64_t (&k)[N],
                                          const uint32_t (&cnt)[N], uint32_t have,
                                          uint32_t &fresh, uint32_t &reps, uint32_t &newrep) {
  uint64_t s[N];
#pragma unroll
  for (int j = 0; j < N; ++j)
    s[j] = key_hash(k[j]) & t.mask;
  uint32_t probe = have, cas = 0, mark = 0;
  for (uint32_t p = 0; probe && p < kMaxProbe; ++p) {
    uint64_t c[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
      c[j] = (probe >> j) & 1 ? t.slots[s[j]] : 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (!((probe >> j) & 1))
        continue;
      if ((c[j] & ~kRep) == k[j]) {
        if (!(c[j] & kRep))
          mark |= 1u << j;
        probe &= ~(1u << j);
      } else if (c[j] == kFree) {
        probe &= ~(1u << j);
        cas |= 1u << j;
      } else {
        s[j] = (s[j] + 1) & t.mask;
      }
    }
  }
  for (uint32_t p = 0; cas && p < kMaxProbe; ++p) {
    unsigned long long old[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
      if ((cas >> j) & 1)
        old[j] = atomicCAS(&t.slots[s[j]], (unsigned long long)kFree,
                           (unsigned long long)(cnt[j] >= 2 ? (k[j] | kRep) : k[j]));
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (!((cas >> j) & 1))
        continue;
      if (old[j] == kFree) {
        ++fresh;
        if (cnt[j] >= 2) {
          ++reps;
          newrep |= 1u << j;
        }
        cas &= ~(1u << j);
      } else if ((old[j] & ~kRep) == k[j]) {
        if (!(old[j] & kRep))
          mark |= 1u << j;
        cas &= ~(1u << j);
      } else {
        s[j] = (s[j] + 1) & t.mask;
      }
    }
  }
  if (mark) {
    unsigned long long old[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
      if ((mark >> j) & 1)
        old[j] = atomicOr(&t.slots[s[j]], (unsigned long long)kRep);
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (((mark >> j) & 1) && !(old[j] & kRep)) {
        ++reps;
        newrep |= 1u << j;
      }
  }
  if (probe | cas)
    atomicExch(t.overflow, 1u);
}

// Appends the keys of the items in `mask` to the block's staging list.
template <int N>
__device__ inline void stage_keys(const CountTable &t, uint32_t *s_n, const uint64_t (&k)[N],
                                  uint32_t mask) {
  while (mask) {
    const int j = __builtin_ctz(mask);
    mask &= mask - 1;
    const uint32_t pos = atomicAdd(s_n, 1u);
    if (pos < kStageCap)
      t.stage[(uint64_t)blockIdx.x * kStageCap + pos] = k[j];
    else
      atomicExch(t.stage_overflow, 1u);
  }
}

// ---- delta (codec.cpp:45-76) ------------------------------------------------
// x[i] -= x[i-1] inside each traversal (heads keep their value), and the max of
// |value| over the result. Each thread takes 8 consecutive elements per step
// (two 16-byte loads and stores); one atomic per block for the max.
__device__ inline uint32_t block_max(uint32_t m) {
  __shared__ uint32_t wmax[kBlock / kWave];
  m = wave_reduce_max(m);
  if ((threadIdx.x & (kWave - 1)) == 0)
    wmax[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    for (uint32_t w = 1; w < kBlock / kWave; ++w)
      m = wmax[w] > m ? wmax[w] : m;
  return m;
}

__global__ __launch_bounds__(kBlock) void k_delta(const int32_t *__restrict__ in,
                                                  int32_t *__restrict__ out,
                                                  const uint32_t *__restrict__ heads, uint64_t n,
                                                  int transform, uint32_t *max_abs) {
  uint32_t m = 0;
  const uint64_t n8 = n / 8;
  // kDeltaChunks chunks of 8 per thread and step, every load of the step
  // (chunks, head words, the element before each chunk) issued before any use.
  constexpr uint32_t kDeltaChunks = 4;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t c0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; c0 < n8;
       c0 += stride * kDeltaChunks) {
    int4 a[kDeltaChunks], b[kDeltaChunks];
    uint32_t hw[kDeltaChunks];
    int32_t pv[kDeltaChunks];
#pragma unroll
    for (uint32_t u = 0; u < kDeltaChunks; ++u) {
      const uint64_t c = c0 + u * stride;
      const uint64_t i0 = (c < n8 ? c : n8 - 1) * 8;
      a[u] = *reinterpret_cast<const int4 *>(in + i0);
      b[u] = *reinterpret_cast<const int4 *>(in + i0 + 4);
      if (transform) {
        hw[u] = heads[i0 >> 5];
        pv[u] = in[i0 ? i0 - 1 : 0];  // unused at a head (the stream starts with one)
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < kDeltaChunks; ++u) {
      const uint64_t c = c0 + u * stride;
      if (c >= n8)
        break;
      const uint64_t i0 = c * 8;
      int32_t v[8] = {a[u].x, a[u].y, a[u].z, a[u].w, b[u].x, b[u].y, b[u].z, b[u].w};
      if (transform) {
        const uint32_t hb = (hw[u] >> (i0 & 31)) & 0xFFu;  // i0 % 8 == 0
        int32_t d[8];
        d[0] = (hb & 1u) ? v[0] : (int32_t)((uint32_t)v[0] - (uint32_t)pv[u]);
#pragma unroll
        for (int j = 1; j < 8; ++j)
          d[j] = ((hb >> j) & 1u) ? v[j] : (int32_t)((uint32_t)v[j] - (uint32_t)v[j - 1]);
#pragma unroll
        for (int j = 0; j < 8; ++j)
          v[j] = d[j];
        *reinterpret_cast<int4 *>(out + i0) = make_int4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<int4 *>(out + i0 + 4) = make_int4(v[4], v[5], v[6], v[7]);
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t w = uabs(v[j]);
        m = w > m ? w : m;
      }
    }
  }
  // Tail (n % 8 elements), one block.
  if (blockIdx.x == 0 && threadIdx.x < n - n8 * 8) {
    const uint64_t i = n8 * 8 + threadIdx.x;
    int32_t v = in[i];
    if (transform) {
      if (!head_at(heads, i))
        v = (int32_t)((uint32_t)v - (uint32_t)in[i - 1]);
      out[i] = v;
    }
    const uint32_t u = uabs(v);
    m = u > m ? u : m;
  }
  m = block_max(m);
  if (threadIdx.x == 0 && m)
    atomicMax(max_abs, m);
}

// ---- wide symbols -----------------------------------------------------------------
// Compact pair keys need every |symbol| below 2^30. Node symbols (IDs or deltas)
// can reach 2^31, and rule IDs start above the largest. The grammar reads a
// symbol only through equality, sign and the signed order of packed pairs, and
// all three survive a renumbering that keeps magnitudes below kWideFloor, maps
// the distinct magnitudes from kWideFloor up to kWideFloor + their rank, and
// commutes with negation. So a stream whose rule-ID floor is kWideTrigger or more
// is renumbered that way, compressed with the floor kWideFloor + (distinct large
// magnitudes), and mapped back: node magnitudes through the table, rule IDs by
// the shift between the two floors.
constexpr uint32_t kWideFloor = 1u << 28;
constexpr uint32_t kWideTrigger = 1u << 29;
constexpr uint64_t kWideCap = uint64_t(1) << 27;  // steps with a large magnitude

__global__ __launch_bounds__(kBlock) void k_wide_gather(const int32_t *__restrict__ x,
                                                        uint64_t n, uint32_t *__restrict__ vals,
                                                        uint64_t cap,
                                                        unsigned long long *__restrict__ count) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < n;
       base += (uint64_t)gridDim.x * kBlock) {
    const uint64_t i = base + threadIdx.x;
    const uint32_t a = i < n ? uabs(x[i]) : 0u;
    const bool big = a >= kWideFloor;
    const uint64_t mask = __ballot(big);
    if (!mask)
      continue;
    const uint32_t leader = (uint32_t)__builtin_ctzll(mask);
    unsigned long long at = 0;
    if (lane == leader)
      at = atomicAdd(count, (unsigned long long)__popcll(mask));
    at = __shfl(at, (int)leader);
    if (big) {
      const uint64_t k = at + __popcll(mask & ((1ull << lane) - 1));
      if (k < cap)
        vals[k] = a;
    }
  }
}

// Renumbers (in -> out) with the sorted distinct large magnitudes u[0..n_u).
__global__ __launch_bounds__(kBlock) void k_wide_map(const int32_t *__restrict__ in,
                                                     int32_t *__restrict__ out, uint64_t n,
                                                     const uint32_t *__restrict__ u, uint32_t n_u) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    int32_t v = in[i];
    const uint32_t a = uabs(v);
    if (a >= kWideFloor) {
      uint32_t lo = 0, hi = n_u;  // lower_bound; a is in the table
      while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (u[mid] < a)
          lo = mid + 1;
        else
          hi = mid;
      }
      const uint32_t r = kWideFloor + lo;
      v = v < 0 ? -(int32_t)r : (int32_t)r;
    }
    out[i] = v;
  }
}

// Maps renumbered symbols back in place: node magnitudes through the table, rule
// IDs (from next_wide up) to next_orig + their offset.
__global__ __launch_bounds__(kBlock) void k_wide_unmap(int32_t *__restrict__ x, uint64_t n,
                                                       const uint32_t *__restrict__ u,
                                                       uint32_t next_wide, uint32_t next_orig) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    const int32_t v = x[i];
    const uint32_t a = uabs(v);
    if (a >= kWideFloor) {
      const uint32_t o = a < next_wide ? u[a - kWideFloor] : a - next_wide + next_orig;
      x[i] = v < 0 ? (int32_t)(0u - o) : (int32_t)o;
    }
  }
}

// ---- K1: repeated-pair set ----------------------------------------------------
// Three levels of aggregation before HBM: (1) within a wave, the few most common
// keys (the backbone's (+1,+1) dominates early rounds) are matched across all
// lanes with ballots, and the rest collapse into runs of equal consecutive keys;
// (2) each block folds its tile into an LDS table with counts; (3) every distinct
// key of the tile is counted once in the HBM table, and the block adds its new
// and newly repeated keys to the global counters with one atomic each.
constexpr int kDict = 4;

// LDS slot = key | kRep once the tile has seen the key twice (kRep is free:
// compact keys are below 2^62), so the table is 8 bytes per slot.
template <uint32_t kSlots>
__device__ inline void lds_add(unsigned long long *skeys, uint64_t k, uint32_t n) {
  uint32_t s = (uint32_t)key_hash(k) & (kSlots - 1);
  const unsigned long long v = n >= 2 ? (k | kRep) : k;
  for (uint32_t p = 0; p < kSlots; ++p) {
    const unsigned long long old = atomicCAS(&skeys[s], (unsigned long long)kFree, v);
    if (old == kFree)
      return;
    if ((old & ~kRep) == k) {
      if (!(old & kRep))
        atomicOr(&skeys[s], (unsigned long long)kRep);
      return;
    }
    s = (s + 1) & (kSlots - 1);
  }
}

// One key per lane (v: the lane holds one) into the wave's dictionary and the
// block's LDS table: (1a) keys already in the dictionary, a free entry taking
// the first other key; (1b) the rest as runs of equal consecutive keys, one
// table add per run.
template <uint32_t kSlots>
__device__ inline void dict_step(unsigned long long *skeys, uint64_t (&dkey)[kDict],
                                 uint32_t (&dcnt)[kDict], uint64_t k, bool v) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  uint64_t rem = __ballot(v);
#pragma unroll
  for (int d = 0; d < kDict; ++d) {
    if (!rem)
      break;
    if (dkey[d] == kFree)
      dkey[d] = __shfl(k, (int)__builtin_ctzll(rem));
    const uint64_t same = __ballot(k == dkey[d]) & rem;
    dcnt[d] += (uint32_t)__popcll(same);
    rem &= ~same;
  }
  const bool mine = (rem >> lane) & 1;
  const uint64_t prev = __shfl_up(k, 1);
  const bool prev_mine = lane > 0 && ((rem >> (lane - 1)) & 1);
  const bool run_head = mine && !(prev_mine && prev == k);
  const uint64_t hmask = __ballot(run_head);
  if (run_head) {
    const uint64_t ends = hmask | ~rem;
    const uint64_t after = lane == kWave - 1 ? 0ull : ends >> (lane + 1);
    const uint32_t len = after ? (uint32_t)__builtin_ctzll(after) + 1u : (uint32_t)(kWave - lane);
    lds_add<kSlots>(skeys, k, len);
  }
}

// A tile holds at most kInsTile pairs, so with kSlots >= kInsTile every key
// finds a slot (a new key arrives while fewer than kSlots are stored, and a
// probe visits every slot). Fewer slots keep more blocks resident; more slots
// keep probes short when most of a tile's keys are distinct.
template <uint32_t kSlots>
__global__ __launch_bounds__(kBlock) void k_pair_aggregate(const int32_t *__restrict__ x,
                                                           const uint32_t *__restrict__ heads,
                                                           uint64_t L, uint64_t span,
                                                           uint32_t fresh,
                                                           unsigned long long *__restrict__ out,
                                                           uint32_t *__restrict__ out_n) {
  __shared__ unsigned long long skeys[kSlots];
  __shared__ uint32_t lds[kBlock / kWave + 1];
  const uint32_t lane = threadIdx.x & (kWave - 1);
  for (uint32_t s = threadIdx.x; s < kSlots; s += kBlock)
    skeys[s] = kFree;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kInsTile;
  uint64_t kv[kInsItems];
  const uint32_t valid =
      load_pair_keys<kInsItems>(x, heads, L, span, base + threadIdx.x, kBlock, kv, fresh);
  // A wave-wide dictionary of up to kDict keys with counts (wave-uniform values)
  // absorbs the hot keys without LDS atomics; it is flushed once at the end.
  uint64_t dkey[kDict];
  uint32_t dcnt[kDict];
#pragma unroll
  for (int d = 0; d < kDict; ++d) {
    dkey[d] = kFree;
    dcnt[d] = 0;
  }
  for (uint32_t j = 0; j < kInsItems; ++j)
    dict_step<kSlots>(skeys, dkey, dcnt, kv[j], (valid >> j) & 1);
  if (lane < kDict) {
    uint64_t fk = kFree;
    uint32_t fc = 0;
#pragma unroll
    for (int d = 0; d < kDict; ++d)
      if ((int)lane == d) {
        fk = dkey[d];
        fc = dcnt[d];
      }
    if (fc)
      lds_add<kSlots>(skeys, fk, fc);
  }
  __syncthreads();
  // The tile's distinct keys (kRep: seen twice in the tile), compacted into the
  // block's region of the staging buffer.
  constexpr int kPer = kSlots / kBlock;
  unsigned long long v[kPer];
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    v[i] = skeys[threadIdx.x * kPer + i];
    c += v[i] != kFree;
  }
  uint32_t total;
  uint32_t off = block_exclusive_scan<kBlock>(c, lds, total);
  unsigned long long *dst = out + (uint64_t)blockIdx.x * kInsTile;
#pragma unroll
  for (int i = 0; i < kPer; ++i)
    if (v[i] != kFree)
      dst[off++] = v[i];
  if (threadIdx.x == 0)
    out_n[blockIdx.x] = total;
}

// Global inserts of the aggregated tiles: all of a thread's probe chains in flight.
__global__ __launch_bounds__(kBlock) void k_pair_insert_staged(
    const unsigned long long *__restrict__ in, const uint32_t *__restrict__ in_n, CountTable t) {
  __shared__ uint32_t sfresh[kBlock / kWave], sreps[kBlock / kWave];
  __shared__ uint32_t s_stage_n;
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (threadIdx.x == 0)
    s_stage_n = 0;
  __syncthreads();
  const uint32_t n = in_n[blockIdx.x];
  const unsigned long long *src = in + (uint64_t)blockIdx.x * kInsTile;
  uint64_t kv[kInsItems];
  uint32_t cnt[kInsItems];
  uint32_t have = 0;
#pragma unroll
  for (uint32_t j = 0; j < kInsItems; ++j) {
    const uint32_t e = j * kBlock + threadIdx.x;
    const uint64_t v = e < n ? src[e] : kFree;
    kv[j] = v & ~kRep;
    cnt[j] = (v & kRep) ? 2u : 1u;
    if (e < n)
      have |= 1u << j;
  }
  uint32_t fresh = 0, reps = 0;
  if (have) {
    uint32_t newrep = 0;
    count_insert_batch<(int)kInsItems>(t, kv, cnt, have, fresh, reps, newrep);
    stage_keys<(int)kInsItems>(t, &s_stage_n, kv, newrep);
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    fresh += __shfl_xor(fresh, d);
    reps += __shfl_xor(reps, d);
  }
  if (lane == 0) {
    sfresh[wid] = fresh;
    sreps[wid] = reps;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t nf = 0, r = 0;
    for (uint32_t w = 0; w < kBlock / kWave; ++w) {
      nf += sfresh[w];
      r += sreps[w];
    }
    t.block_counts[blockIdx.x] = make_uint2(nf, r);
  }
}

// Rounds where nearly every pair is distinct gain nothing from the tile's LDS
// table: each thread inserts its pairs directly, all probe chains in flight.
__global__ __launch_bounds__(kBlock) void k_pair_insert_direct(const int32_t *__restrict__ x,
                                                               const uint32_t *__restrict__ heads,
                                                               uint64_t L, uint64_t span,
                                                               uint32_t fresh, CountTable t) {
  __shared__ uint32_t sfresh[kBlock / kWave], sreps[kBlock / kWave];
  __shared__ uint32_t s_stage_n;
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (threadIdx.x == 0)
    s_stage_n = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kInsTile;
  uint64_t kv[kInsItems];
  uint32_t cnt[kInsItems];
  const uint32_t have =
      load_pair_keys<kInsItems>(x, heads, L, span, base + threadIdx.x, kBlock, kv, fresh);
#pragma unroll
  for (uint32_t j = 0; j < kInsItems; ++j)
    cnt[j] = 1;
  uint32_t newk = 0, reps = 0;
  if (have) {
    uint32_t newrep = 0;
    count_insert_batch<(int)kInsItems>(t, kv, cnt, have, newk, reps, newrep);
    stage_keys<(int)kInsItems>(t, &s_stage_n, kv, newrep);
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    newk += __shfl_xor(newk, d);
    reps += __shfl_xor(reps, d);
  }
  if (lane == 0) {
    sfresh[wid] = newk;
    sreps[wid] = reps;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t n = 0, r = 0;
    for (uint32_t w = 0; w < kBlock / kWave; ++w) {
      n += sfresh[w];
      r += sreps[w];
    }
    t.block_counts[blockIdx.x] = make_uint2(n, r);
  }
}

// ---- K1 for rounds of mostly distinct pairs: partitioned counting -----------
// (1) A partition block takes `per` consecutive input tiles and writes their
// pair keys into a region of its own, grouped by hash bucket (a counting sort
// in LDS), with one row of bucket starts; no global atomics. (2) One block per
// bucket gathers that bucket's segment from every region, counts its keys in a
// 128 KB LDS table and writes the repeated ones to the bucket's list. A table
// or list overflow raises a flag; the host then counts the round with the
// global table instead.
constexpr uint32_t kBktSlots = 16384;   // LDS table of one bucket
constexpr uint32_t kBktProbe = 512;
constexpr uint32_t kBktThreads = 1024;
constexpr uint32_t kMaxBuckets1 = 4096;  // buckets of a one-level partition
constexpr uint32_t kFineBuckets = 128;   // second-level buckets per coarse bucket
constexpr uint32_t kFineBits = 7;
constexpr uint32_t kMaxBuckets = 32768;  // buckets of a two-level partition
constexpr uint32_t kMaxParts = 2048;    // partition blocks (segments per bucket)
constexpr uint32_t kMaxParts2 = 16384;  // partition blocks of a two-level round
constexpr uint32_t kL2Parts = 64;        // second-level blocks per coarse bucket
constexpr uint32_t kRepStride = kBktSlots / 2;  // repeated keys listed per bucket
constexpr uint32_t kDedupSlots = 2048;   // per-block set of keys already known repeated
static_assert((uint64_t)kMaxBuckets * kBktSlots < (1ull << 32),
              "a round's bucket totals fit half of a 64-bit word");
static_assert(kFineBuckets == 1u << kFineBits, "fine buckets are a power of two");

// Partitioned keys: regions, each with a row of nb + 1 bucket starts. A key's
// row bucket is (bucket_of(k, tmask) >> pshift) & (nb - 1). One-level rounds
// have one set of regions whose rows split all tmask + 1 buckets; two-level
// rounds first split by coarse bucket (the high bits), then each coarse bucket
// c is split by fine bucket into regions c * count .. c * count + count - 1.
// Counting bucket b then reads row bucket b & (nb - 1) of the regions of coarse
// bucket b / nb.
struct Parts {
  unsigned long long *entries;  // region g at base[g] if base, else at g * region
  uint64_t region;
  const uint64_t *base;
  uint32_t *starts;             // row g at g * (nb + 1): bucket starts, then the total
  uint32_t nb;                  // buckets per row (a power of two)
  uint32_t count;               // regions per coarse bucket (all of them when one-level)
  uint32_t tmask;               // the round's buckets - 1
  uint32_t pshift;
};

__device__ inline uint32_t row_bucket(uint64_t k, const Parts &pt) {
  return (bucket_of(k, pt.tmask) >> pt.pshift) & (pt.nb - 1);
}


// Exclusive scan of the block's bucket histogram in place; writes the block's
// row of starts (nb + 1 words).
__device__ inline void part_starts(uint32_t *hist, const Parts &pt) {
  __shared__ uint32_t lds[kBlock / kWave + 1];
  const uint32_t nb = pt.nb;
  const uint32_t per_t = (nb + kBlock - 1) / kBlock;
  const uint32_t b0 = threadIdx.x * per_t;
  uint32_t sum = 0;
  for (uint32_t j = 0; j < per_t && b0 + j < nb; ++j)
    sum += hist[b0 + j];
  uint32_t total;
  uint32_t run = block_exclusive_scan<kBlock>(sum, lds, total);
  for (uint32_t j = 0; j < per_t && b0 + j < nb; ++j) {
    const uint32_t c = hist[b0 + j];
    hist[b0 + j] = run;
    run += c;
  }
  __syncthreads();
  uint32_t *row = pt.starts + (uint64_t)blockIdx.x * (nb + 1);
  for (uint32_t b = threadIdx.x; b < nb; b += kBlock)
    row[b] = hist[b];
  if (threadIdx.x == 0)
    row[nb] = total;
  __syncthreads();  // the starts are read before the placing pass moves the cursors
}

// Keys of one pair-insert tile straight from the stream. The wave's most common
// key (its first lane's) is written once; if the wave holds it twice it is
// already known repeated (kRep set in the entry).
__device__ inline void tile_keys(const int32_t *__restrict__ x, const uint32_t *__restrict__ heads,
                                 uint64_t L, uint64_t span, uint32_t fresh, uint64_t t,
                                 uint64_t (&kv)[kInsItems]) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t have = load_pair_keys<kInsItems>(x, heads, L, span, t * kInsTile + threadIdx.x,
                                                  kBlock, kv, fresh);
#pragma unroll
  for (uint32_t j = 0; j < kInsItems; ++j) {
    const bool v = (have >> j) & 1;
    uint64_t k = kv[j];
    const uint64_t valid = __ballot(v);
    if (valid) {
      const uint32_t leader = (uint32_t)__builtin_ctzll(valid);
      const uint64_t lk = __shfl(k, (int)leader);
      const uint64_t same = __ballot(v && k == lk);
      if (__popcll(same) >= 2) {
        if (lane == leader)
          k = lk | kRep;
        else if ((same >> lane) & 1)
          k = kFree;
      }
    }
    kv[j] = k;
  }
}

__global__ __launch_bounds__(kBlock) void k_bucket_part(const int32_t *__restrict__ x,
                                                        const uint32_t *__restrict__ heads,
                                                        uint64_t L, uint64_t span, uint32_t fresh,
                                                        uint64_t n_tiles, uint32_t per, Parts pt) {
  extern __shared__ uint32_t hist[];  // one counter per bucket (dynamic LDS)
  for (uint32_t i = threadIdx.x; i < pt.nb; i += kBlock)
    hist[i] = 0;
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * per;
  const uint64_t t1 = t0 + per < n_tiles ? t0 + per : n_tiles;
  uint64_t kv[kInsItems];
  for (uint64_t t = t0; t < t1; ++t) {
    tile_keys(x, heads, L, span, fresh, t, kv);
#pragma unroll
    for (uint32_t j = 0; j < kInsItems; ++j)
      if (kv[j] != kFree)
        atomicAdd(&hist[row_bucket(kv[j] & ~kRep, pt)], 1u);
  }
  __syncthreads();
  part_starts(hist, pt);
  // Second pass (the same keys, re-read from cache): LDS cursors place them.
  unsigned long long *dst = pt.entries + (uint64_t)blockIdx.x * pt.region;
  for (uint64_t t = t0; t < t1; ++t) {
    tile_keys(x, heads, L, span, fresh, t, kv);
#pragma unroll
    for (uint32_t j = 0; j < kInsItems; ++j)
      if (kv[j] != kFree)
        dst[atomicAdd(&hist[row_bucket(kv[j] & ~kRep, pt)], 1u)] = kv[j];
  }
}

// A block's entries spread over segments (aggregated tiles, or one coarse
// bucket's segment of several regions), visited as one flat sequence: the
// segments' lengths are scanned in LDS, kFlatSegs at a time, and every thread
// takes entries kBlock apart, kFlatItems loads in flight, walking the prefix
// forward to find each entry's segment (entries of a wave are consecutive, so
// loads coalesce; segments are tens to hundreds of entries, so the walk is a
// step or two). fn(entry, segment index, offset in the segment).
constexpr uint32_t kFlatSegs = 1024;
constexpr uint32_t kFlatItems = 8;
struct FlatLds {
  uint32_t pre[kFlatSegs + 1];
  const unsigned long long *src[kFlatSegs];
  uint32_t scan[kBlock / kWave + 1];
};

template <class Seg, class Fn>
__device__ inline void for_flat_entries(FlatLds &fl, uint32_t nseg, Seg seg, Fn fn) {
  constexpr uint32_t kPerThread = kFlatSegs / kBlock;
  for (uint32_t c0 = 0; c0 < nseg; c0 += kFlatSegs) {
    const uint32_t nt = nseg - c0 < kFlatSegs ? nseg - c0 : kFlatSegs;
    uint32_t len[kPerThread], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPerThread; ++k) {
      const uint32_t q = threadIdx.x * kPerThread + k;
      len[k] = 0;
      if (q < nt)
        seg(c0 + q, fl.src[q], len[k]);
      sum += len[k];
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan<kBlock>(sum, fl.scan, total);
#pragma unroll
    for (uint32_t k = 0; k < kPerThread; ++k) {
      fl.pre[threadIdx.x * kPerThread + k] = off;
      off += len[k];
    }
    if (threadIdx.x == 0)
      fl.pre[kFlatSegs] = total;
    __syncthreads();
    uint32_t cur = 0;  // this thread's segment, non-decreasing
    for (uint32_t i0 = threadIdx.x; i0 < total; i0 += kBlock * kFlatItems) {
      unsigned long long v[kFlatItems];
      uint32_t sg[kFlatItems], at[kFlatItems];
#pragma unroll
      for (uint32_t j = 0; j < kFlatItems; ++j) {
        const uint32_t i = i0 + j * kBlock < total ? i0 + j * kBlock : total - 1;
        while (fl.pre[cur + 1] <= i)
          ++cur;
        sg[j] = cur;
        at[j] = i - fl.pre[cur];
        v[j] = fl.src[cur][at[j]];
      }
#pragma unroll
      for (uint32_t j = 0; j < kFlatItems; ++j)
        if (i0 + j * kBlock < total)
          fn(v[j], c0 + sg[j], at[j]);
    }
    __syncthreads();  // the prefix is rewritten for the next segments
  }
}

// The same partition of aggregated tiles (entries already carry kRep). A key
// already known repeated adds nothing after its first copy, so each block keeps
// one copy of it: the first pass claims it in an LDS set (the claiming element's
// index recorded), the second places only the claimant's copy. The hottest
// keys, repeated in every tile, then reach their bucket once per block instead
// of once per tile. A key the full set cannot take is kept in every copy; both
// passes see the same set, so their counts agree.
__device__ inline uint32_t dedup_slot(uint64_t k) {
  return (uint32_t)(key_hash(k) >> 13) & (kDedupSlots - 1);
}

__global__ __launch_bounds__(kBlock) void k_bucket_part_tiles(
    const unsigned long long *__restrict__ tiles, const uint32_t *__restrict__ tiles_n,
    uint64_t n_tiles, uint32_t per, Parts pt) {
  extern __shared__ uint32_t hist[];  // one counter per bucket (dynamic LDS)
  __shared__ unsigned long long dkey[kDedupSlots];
  __shared__ uint32_t didx[kDedupSlots];
  __shared__ FlatLds fl;
  for (uint32_t i = threadIdx.x; i < pt.nb; i += kBlock)
    hist[i] = 0;
  for (uint32_t i = threadIdx.x; i < kDedupSlots; i += kBlock)
    dkey[i] = kFree;
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * per;
  const uint64_t t1 = t0 + per < n_tiles ? t0 + per : n_tiles;
  const uint32_t nseg = (uint32_t)(t1 - t0);
  auto seg = [&](uint32_t q, const unsigned long long *&src, uint32_t &len) {
    src = tiles + (t0 + q) * kInsTile;
    len = tiles_n[t0 + q];
  };
  for_flat_entries(fl, nseg, seg, [&](unsigned long long k, uint32_t q, uint32_t at) {
    if (k & kRep) {
      const uint32_t me = q * kInsTile + at;
      uint32_t sl = dedup_slot(k);
      for (uint32_t p = 0; p < 16; ++p) {
        const unsigned long long old = atomicCAS(&dkey[sl], (unsigned long long)kFree, k);
        if (old == kFree) {
          didx[sl] = me;
          break;
        }
        if (old == k)
          return;  // a later copy of a known-repeated key
        sl = (sl + 1) & (kDedupSlots - 1);
      }
    }
    atomicAdd(&hist[row_bucket(k & ~kRep, pt)], 1u);
  });
  part_starts(hist, pt);
  unsigned long long *dst = pt.entries + (uint64_t)blockIdx.x * pt.region;
  for_flat_entries(fl, nseg, seg, [&](unsigned long long k, uint32_t q, uint32_t at) {
    if (k & kRep) {
      const uint32_t me = q * kInsTile + at;
      uint32_t sl = dedup_slot(k);
      for (uint32_t p = 0; p < 16; ++p) {
        const unsigned long long cur = dkey[sl];
        if (cur == k) {
          if (didx[sl] != me)
            return;
          break;
        }
        if (cur == kFree)
          break;
        sl = (sl + 1) & (kDedupSlots - 1);
      }
    }
    dst[atomicAdd(&hist[row_bucket(k & ~kRep, pt)], 1u)] = k;
  });
}

// Entries of a coarse bucket's level-one segments, per second-level block: block
// g = c * p2count + j takes coarse bucket c of level-one regions
// [j * P / p2count, (j + 1) * P / p2count).
__global__ void k_part2_sizes(Parts p1, uint32_t p2count, uint32_t n_coarse,
                              uint32_t *__restrict__ sizes) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (uint64_t)n_coarse * p2count)
    return;
  const uint32_t c = (uint32_t)(g / p2count), j = (uint32_t)(g % p2count);
  const uint32_t P = p1.count;
  const uint32_t r0 = (uint32_t)((uint64_t)j * P / p2count);
  const uint32_t r1 = (uint32_t)((uint64_t)(j + 1) * P / p2count);
  uint32_t n = 0;
  for (uint32_t r = r0; r < r1; ++r) {
    const uint32_t *row = p1.starts + (uint64_t)r * (p1.nb + 1);
    n += row[c + 1] - row[c];
  }
  sizes[g] = n;
}

// Second level of a two-level partition: block c * p2.count + j splits coarse
// bucket c of its level-one regions by fine bucket into its own region (at
// p2.base[block]), with a row of fine-bucket starts.
__global__ __launch_bounds__(kBlock) void k_bucket_part2(Parts p1, Parts p2) {
  extern __shared__ uint32_t hist[];  // one counter per fine bucket (dynamic LDS)
  __shared__ FlatLds fl;
  for (uint32_t i = threadIdx.x; i < p2.nb; i += kBlock)
    hist[i] = 0;
  __syncthreads();
  const uint32_t c = blockIdx.x / p2.count, j = blockIdx.x % p2.count;
  const uint32_t P = p1.count;
  const uint32_t r0 = (uint32_t)((uint64_t)j * P / p2.count);
  const uint32_t r1 = (uint32_t)((uint64_t)(j + 1) * P / p2.count);
  auto seg = [&](uint32_t q, const unsigned long long *&src, uint32_t &len) {
    const uint64_t r = r0 + q;
    const uint32_t *row = p1.starts + r * (p1.nb + 1);
    const uint32_t s0 = row[c];
    len = row[c + 1] - s0;
    src = p1.entries + r * p1.region + s0;
  };
  for_flat_entries(fl, r1 - r0, seg, [&](unsigned long long k, uint32_t, uint32_t) {
    atomicAdd(&hist[row_bucket(k & ~kRep, p2)], 1u);
  });
  part_starts(hist, p2);
  unsigned long long *dst = p2.entries + p2.base[blockIdx.x];
  for_flat_entries(fl, r1 - r0, seg, [&](unsigned long long k, uint32_t, uint32_t) {
    dst[atomicAdd(&hist[row_bucket(k & ~kRep, p2)], 1u)] = k;
  });
}

// One block per bucket: its segment of every partition region, counted in an
// LDS table; the repeated keys go to rep[b * kRepStride ...].
__global__ __launch_bounds__(kBktThreads) void k_bucket_count(Parts pt,
                                                              unsigned long long *__restrict__ rep,
                                                              uint2 *__restrict__ block_counts,
                                                              unsigned long long *__restrict__ sums,
                                                              uint32_t *__restrict__ overflow,
                                                              uint32_t rep_cap) {
  __shared__ unsigned long long tab[kBktSlots];
  __shared__ uint32_t lds[kBktThreads / kWave + 1];
  constexpr uint32_t kPer = kBktSlots / kBktThreads;
  for (uint32_t i = threadIdx.x; i < kBktSlots; i += kBktThreads)
    tab[i] = kFree;
  __syncthreads();
  const uint32_t b = blockIdx.x;
  bool fail = false;
  auto insert = [&](uint64_t ev) {  // ev: key, with kRep if already known repeated
    const uint64_t k = ev & ~kRep;
    uint32_t sl = (uint32_t)key_hash(k) & (kBktSlots - 1);
    for (uint32_t p = 0; p < kBktProbe; ++p) {
      const unsigned long long old = atomicCAS(&tab[sl], (unsigned long long)kFree,
                                               (unsigned long long)ev);
      if (old == kFree)
        return;
      if ((old & ~kRep) == k) {
        if (!(old & kRep))
          atomicOr(&tab[sl], (unsigned long long)kRep);
        return;
      }
      sl = (sl + 1) & (kBktSlots - 1);
    }
    fail = true;
  };
  // Segment g of the bucket: [starts[r][lb], starts[r][lb + 1]) of region
  // r = coarse * count + g.
  const uint32_t lb = b & (pt.nb - 1);
  const uint64_t r0 = (uint64_t)(b / pt.nb) * pt.count;
  constexpr uint32_t kBatch = 4;
  if (pt.count <= kWave) {
    // Few long segments (two-level rounds): every thread takes entries of the
    // bucket's concatenated segments, found by a search of their prefix in LDS.
    __shared__ uint32_t pre[kWave + 1];
    __shared__ const unsigned long long *src_of[kWave];
    if (threadIdx.x < kWave) {
      const uint32_t g = threadIdx.x;
      uint32_t len = 0;
      if (g < pt.count) {
        const uint64_t r = r0 + g;
        const uint32_t *row = pt.starts + r * (pt.nb + 1);
        const uint32_t s0 = row[lb];
        len = row[lb + 1] - s0;
        src_of[g] = pt.entries + (pt.base ? pt.base[r] : r * pt.region) + s0;
      }
      uint32_t x = len;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (g >= (uint32_t)d)
          x += y;
      }
      pre[g + 1] = x;
      if (g == 0)
        pre[0] = 0;
    }
    __syncthreads();
    const uint32_t total = pre[pt.count];
    for (uint32_t i0 = threadIdx.x; i0 < total; i0 += kBktThreads * kBatch) {
      uint64_t ev[kBatch];
#pragma unroll
      for (uint32_t j = 0; j < kBatch; ++j) {
        const uint32_t i = i0 + j * kBktThreads < total ? i0 + j * kBktThreads : total - 1;
        uint32_t lo = 0, hi = pt.count;  // last g with pre[g] <= i
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) / 2;
          if (pre[mid] <= i)
            lo = mid;
          else
            hi = mid;
        }
        ev[j] = src_of[lo][i - pre[lo]];
      }
#pragma unroll
      for (uint32_t j = 0; j < kBatch; ++j)
        if (i0 + j * kBktThreads < total)
          insert(ev[j]);
    }
  } else {
    // Many short segments (one-level rounds): a thread per segment, four entries
    // loaded together (clamped into the segment), then inserted.
    for (uint32_t g = threadIdx.x; g < pt.count; g += kBktThreads) {
      const uint64_t r = r0 + g;
      const uint32_t *row = pt.starts + r * (pt.nb + 1);
      const uint32_t s0 = row[lb], s1 = row[lb + 1];
      const unsigned long long *src = pt.entries + (pt.base ? pt.base[r] : r * pt.region);
      for (uint32_t i = s0; i < s1; i += kBatch) {
        uint64_t ev[kBatch];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j)
          ev[j] = src[i + j < s1 ? i + j : s1 - 1];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j)
          if (i + j < s1)
            insert(ev[j]);
      }
    }
  }
  // Counts, and the repeated keys compacted into the bucket's list.
  uint32_t d = 0, r = 0;
  unsigned long long v[kPer];
  __syncthreads();
#pragma unroll
  for (uint32_t i = 0; i < kPer; ++i) {
    v[i] = tab[threadIdx.x * kPer + i];
    d += v[i] != kFree;
    r += v[i] != kFree && (v[i] & kRep);
  }
  uint32_t total_r;
  uint32_t off = block_exclusive_scan<kBktThreads>(r, lds, total_r);
  unsigned long long *out = rep + (uint64_t)b * kRepStride;
#pragma unroll
  for (uint32_t i = 0; i < kPer; ++i)
    if (v[i] != kFree && (v[i] & kRep)) {
      if (off < kRepStride)
        out[off] = v[i] & ~kRep;
      ++off;
    }
  uint32_t total_d;
  block_exclusive_scan<kBktThreads>(d, lds, total_d);
  if (__syncthreads_or(fail) && threadIdx.x == 0)
    atomicExch(overflow, 1u);
  if (threadIdx.x == 0) {
    if (total_r > rep_cap)
      atomicExch(overflow, 1u);
    block_counts[b] = make_uint2(total_d, total_r);
    // Distinct keys (high half) and repeated keys (low half) in one add: each
    // total is below 2^32 (at most kMaxBuckets * kBktSlots keys). The longest
    // list, on a line of its own, sizes the rule-table regions.
    atomicAdd(&sums[0], ((unsigned long long)total_d << 32) | total_r);
    atomicMax(&sums[8], (unsigned long long)total_r);
  }
}

// Sum of the per-block counts: out[0] = distinct keys, out[1] = repeated keys.
__global__ __launch_bounds__(1024) void k_sum_counts(const uint2 *__restrict__ c, uint64_t n,
                                                     unsigned long long *__restrict__ out) {
  __shared__ unsigned long long s0[1024 / kWave], s1[1024 / kWave];
  unsigned long long a = 0, b = 0;
  for (uint64_t i = threadIdx.x; i < n; i += 1024) {
    const uint2 v = c[i];
    a += v.x;
    b += v.y;
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    a += __shfl_xor(a, d);
    b += __shfl_xor(b, d);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s0[threadIdx.x / kWave] = a;
    s1[threadIdx.x / kWave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long x = 0, y = 0;
    for (uint32_t w = 0; w < 1024 / kWave; ++w) {
      x += s0[w];
      y += s1[w];
    }
    out[0] = x;
    out[1] = y;
  }
}

// ---- K1b: the repeated keys, moved into the rule table -------------------------
// From the staging lists (one per counting block or bucket): one block per
// list, so a bucket's few hundred keys take one or two inserts per thread.
__global__ __launch_bounds__(kBlock) void k_extract_staged(
    const unsigned long long *__restrict__ stage, uint64_t stride,
    const uint2 *__restrict__ block_counts, RuleTable rt, uint32_t *__restrict__ err) {
  const uint64_t b = blockIdx.x;
  const uint64_t n = block_counts[b].y < stride ? block_counts[b].y : stride;
  for (uint64_t j = threadIdx.x; j < n; j += kBlock) {
    const unsigned long long k = stage[b * stride + j];
    if (!rule_insert(rt, k))
      atomicExch(err, 3u);
    filter_add(rt, k);
  }
}

// From the counting table itself (when some block staged too many keys).
constexpr uint32_t kExtractItems = 16;

// Bucket-counted rounds: bucket b's region of the rule table built by one block
// in LDS (rt.rlines lines, at most kBktSlots slots) from the bucket's list,
// then written out whole, free slots included. A key takes the first line of
// its probe within the region that has a free slot, as rule_insert would place
// it; the host sizes regions above the longest list, so every key finds one.
__global__ __launch_bounds__(kBktThreads) void k_region_build(
    const unsigned long long *__restrict__ stage, uint64_t stride,
    const uint2 *__restrict__ block_counts, RuleTable rt, uint32_t *__restrict__ err) {
  __shared__ unsigned long long reg[kBktSlots];
  const uint32_t lines = rt.rlines, slots = lines * kLine;
  for (uint32_t i = threadIdx.x; i < slots; i += kBktThreads)
    reg[i] = kFree;
  __syncthreads();
  const uint64_t b = blockIdx.x;
  const uint64_t n = block_counts[b].y < stride ? block_counts[b].y : stride;
  for (uint64_t j = threadIdx.x; j < n; j += kBktThreads) {
    const unsigned long long k = stage[b * stride + j];
    uint32_t l = region_line(k, lines);
    bool placed = false;
    for (uint32_t p = 0; p < lines && !placed; ++p) {
      for (uint32_t q = 0; q < kLine && !placed; ++q)
        placed = atomicCAS(&reg[l * kLine + q], (unsigned long long)kFree, k) == kFree;
      l = l + 1 == lines ? 0u : l + 1;
    }
    if (!placed)  // a region smaller than its list: never sized so by the host
      atomicExch(err, 3u);
    filter_add(rt, k);
  }
  __syncthreads();
  unsigned long long *dst = rt.keys + b * rt.rlines * kLine;
  for (uint32_t i = threadIdx.x; i < slots; i += kBktThreads)
    dst[i] = reg[i];
}

__global__ __launch_bounds__(kBlock) void k_extract_repeated(
    const unsigned long long *__restrict__ count_slots, uint64_t cap, RuleTable rt,
    uint32_t *__restrict__ err) {
  const uint64_t s0 = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * kExtractItems;
  for (uint32_t i = 0; i < kExtractItems && s0 + i < cap; ++i) {
    const uint64_t v = count_slots[s0 + i];
    if (v != kFree && (v & kRep)) {
      if (!rule_insert(rt, v & ~kRep))
        atomicExch(err, 3u);
      filter_add(rt, v & ~kRep);
    }
  }
}

// Inclusive scan of transfer functions across the 64 lanes of a wave.
__device__ inline uint32_t wave_fn_inclusive(uint32_t f) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t g = __shfl_up(f, d);
    if (lane >= (uint32_t)d)
      f = fn_then(g, f);
  }
  return f;
}

// ---- K1c: the round's hottest keys, cached in LDS by K2 and K3 ---------------
// A few keys can make up most pairs of early rounds (the backbone's own pairs:
// one key is three quarters of all pairs of the first round of a 64-haplotype
// pangenome). Every wave then reads the same rule-table lines, and those reads
// queue on one cache channel per die. One block samples 64 windows of 64
// consecutive pairs spread over the stream, takes the keys seen at least
// kHotMin times (at most 64), resolves their slots and publishes them as a
// small open-addressing table that K2 and K3 copy into LDS. A key missing from
// it is simply looked up in the rule table, so the sample only decides speed.
constexpr uint32_t kHotWindows = 64, kHotWin = 64;   // 4096 sampled pairs
constexpr uint32_t kHotSampleSlots = 8192;
constexpr uint32_t kHotMin = 64;                     // 1.6% of the sample
constexpr uint32_t kHotSlots = 128;                  // <= 64 keys: load <= 1/2
constexpr uint32_t kHotThreads = 1024;

struct alignas(16) HotEntry {
  unsigned long long key;  // kFree = empty
  uint32_t slot;
  uint32_t pad;
};

// The kHotTop most frequent sampled keys as canonical pairs, which K2 and K3
// compare raw pairs against before any key arithmetic.
constexpr uint32_t kHotTop = 4;
struct alignas(16) HotPair {
  int32_t a, b;  // canonical pair
  uint32_t slot;
  uint32_t set;  // 0 = empty entry
};

__global__ __launch_bounds__(kHotThreads) void k_hot_sample(const int32_t *__restrict__ x,
                                                          const uint32_t *__restrict__ heads,
                                                          uint64_t L, uint64_t span,
                                                          uint32_t fresh, RuleTable rt,
                                                          HotEntry *__restrict__ hot,
                                                          HotPair *__restrict__ top) {
  __shared__ unsigned long long skey[kHotSampleSlots];
  __shared__ uint32_t scnt[kHotSampleSlots];
  __shared__ unsigned long long hkey[kHotSlots];
  __shared__ uint32_t hslot[kHotSlots];
  // The keys seen kHotMin times (at most kHotWindows * kHotWin / kHotMin = 64).
  __shared__ unsigned long long lkey[kWave];
  __shared__ uint32_t lcnt[kWave];
  __shared__ uint32_t ln;
  __shared__ HotPair stop[kHotTop];
  static_assert(kHotWindows * kHotWin / kHotMin <= kWave, "one wave ranks the hot keys");
  for (uint32_t s = threadIdx.x; s < kHotSampleSlots; s += kHotThreads) {
    skey[s] = kFree;
    scnt[s] = 0;
  }
  for (uint32_t s = threadIdx.x; s < kHotSlots; s += kHotThreads) {
    hkey[s] = kFree;
    hslot[s] = 0;
  }
  if (threadIdx.x < kHotTop)
    stop[threadIdx.x] = HotPair{0, 0, 0, 0};
  if (threadIdx.x == 0)
    ln = 0;
  __syncthreads();
  // Sampled positions are distinct: windows are spread only when they cannot
  // overlap (window starts at least kHotWin apart); a shorter stream is
  // sampled whole.
  const uint64_t room = L > kHotWin + 1 ? L - kHotWin - 1 : 0;
  const bool spread = room >= (uint64_t)kHotWindows * kHotWin;
  for (uint32_t q = threadIdx.x; q < kHotWindows * kHotWin; q += kHotThreads) {
    const uint64_t i =
        spread ? (uint64_t)(q / kHotWin) * room / kHotWindows + q % kHotWin : (uint64_t)q;
    if (i + 1 >= L || head_at(heads, i + 1) || !fresh_pair(x[i], x[i + 1], fresh))
      continue;
    const unsigned long long k = pair_key(x[i], x[i + 1], span);
    uint32_t sl = (uint32_t)key_hash(k) & (kHotSampleSlots - 1);
    for (uint32_t p = 0; p < kHotSampleSlots; ++p) {
      const unsigned long long old = atomicCAS(&skey[sl], (unsigned long long)kFree, k);
      if (old == kFree || old == k) {
        atomicAdd(&scnt[sl], 1u);
        break;
      }
      sl = (sl + 1) & (kHotSampleSlots - 1);
    }
  }
  __syncthreads();
  for (uint32_t s = threadIdx.x; s < kHotSampleSlots; s += kHotThreads)
    if (scnt[s] >= kHotMin) {
      const uint32_t at = atomicAdd(&ln, 1u);
      lkey[at] = skey[s];
      lcnt[at] = scnt[s];
    }
  __syncthreads();
  // One wave: each lane one listed key; its slot, its rank by count (then key),
  // the top kHotTop as pairs, all of them into the small table.
  if (threadIdx.x < kWave) {
    const uint32_t lane = threadIdx.x, n = ln;
    const bool have = lane < n;
    const unsigned long long k = have ? lkey[lane] : kFree;
    const uint32_t c = have ? lcnt[lane] : 0;
    const int64_t slot = have ? rule_find(rt, k) : -1;
    uint32_t rank = 0;
    for (uint32_t o = 0; o < n; ++o) {
      const unsigned long long ko = __shfl(k, (int)o);
      const uint32_t co = __shfl(c, (int)o);
      rank += co > c || (co == c && ko < k);
    }
    if (slot >= 0) {
      if (rank < kHotTop) {
        const int64_t cp = expand_key(k, span);
        stop[rank] = HotPair{key_first(cp), key_second(cp), (uint32_t)slot, 1};
      }
      uint32_t h = (uint32_t)key_hash(k) & (kHotSlots - 1);
      for (uint32_t p = 0; p < kHotSlots; ++p) {
        if (atomicCAS(&hkey[h], (unsigned long long)kFree, k) == kFree) {
          hslot[h] = (uint32_t)slot;
          break;
        }
        h = (h + 1) & (kHotSlots - 1);
      }
    }
  }
  __syncthreads();
  for (uint32_t s = threadIdx.x; s < kHotSlots; s += kHotThreads) {
    HotEntry e;
    e.key = hkey[s];
    e.slot = hslot[s];
    e.pad = 0;
    hot[s] = e;
  }
  // Unset entries repeat the first set one, so a reader compares every entry
  // without testing `set` (entry 0 is set iff any is).
  if (threadIdx.x < kHotTop) {
    HotPair e = stop[threadIdx.x];
    for (uint32_t t = 0; t < kHotTop && !e.set; ++t)
      if (stop[t].set)
        e = stop[t];
    top[threadIdx.x] = e;
  }
}

struct HotLds {
  unsigned long long key[kHotSlots];
  uint32_t slot[kHotSlots];
};

__device__ inline void hot_load(HotLds &h, const HotEntry *__restrict__ hot) {
  for (uint32_t s = threadIdx.x; s < kHotSlots; s += blockDim.x) {
    const HotEntry e = hot[s];
    h.key[s] = e.key;
    h.slot[s] = e.slot;
  }
}

__device__ inline int64_t hot_find(const HotLds &h, uint64_t k) {
  uint32_t s = (uint32_t)key_hash(k) & (kHotSlots - 1);
  for (uint32_t p = 0; p < kHotSlots; ++p) {
    const unsigned long long c = h.key[s];
    if (c == k)
      return (int64_t)h.slot[s];
    if (c == kFree)
      return -1;
    s = (s + 1) & (kHotSlots - 1);
  }
  return -1;
}

// ---- K2: candidate bits and each tile's transfer function --------------------
// A pair is a candidate iff its key is in the rule table (the repeated keys).
// Per pass of kPassItems items a lane reads its pairs and compares them, raw,
// with the kHotTop hottest canonical pairs (scalar registers); every other
// pair goes to a per-wave LDS list. The wave then resolves its list densely:
// key arithmetic, the LDS table of the other hot keys, and rule-table lookups
// kColdBatch per lane at a time, each result written over its entry. The key
// arithmetic and line reads thus run over only the pairs that need them (a
// wave's lanes would otherwise all wait for the few that do), and the line
// reads' registers are never live beside the stream's. One lookup per lane at a
// time: two in flight held two 64-byte lines per lane, 93 VGPRs and spilled
// SGPRs (five waves per SIMD); one keeps the kernel under 64 VGPRs (eight
// waves): on plain_10m candidate marking 31.5 -> 28.9 ms per compress.
constexpr uint32_t kPassItems = 8;
constexpr uint32_t kColdCap = kPassItems * kWave;   // every item of a pass listed
constexpr int kColdBatch = 1;

// Resolves the wave's n listed raw pairs: each entry becomes its key's rule
// slot, or kFree when the key is not in the table.
__device__ inline void resolve_list(const RuleTable &rt, const HotLds &sh, uint64_t span,
                                    unsigned long long *list, uint32_t n) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  for (uint32_t e0 = 0; e0 < n; e0 += kWave * kColdBatch) {
    uint64_t kk[kColdBatch];
    bool cold[kColdBatch];
    int64_t hs[kColdBatch];
#pragma unroll
    for (int q = 0; q < kColdBatch; ++q) {
      const uint32_t e = e0 + (uint32_t)q * kWave + lane;
      kk[q] = 0;
      hs[q] = -1;
      if (e < n) {
        const int64_t raw = (int64_t)list[e];
        kk[q] = pair_key(key_first(raw), key_second(raw), span);
        hs[q] = hot_find(sh, kk[q]);
      }
      cold[q] = e < n && hs[q] < 0 && (!rt.filter || filter_has(rt, kk[q]));
    }
    int64_t sl[kColdBatch];
    rule_find_batch<kColdBatch>(rt, kk, cold, sl);
#pragma unroll
    for (int q = 0; q < kColdBatch; ++q) {
      const uint32_t e = e0 + (uint32_t)q * kWave + lane;
      if (e < n) {
        const int64_t r = hs[q] >= 0 ? hs[q] : sl[q];
        list[e] = r >= 0 ? (unsigned long long)r : kFree;
      }
    }
  }
}

// The hot canonical pairs as packed 64-bit words (x in the low half) in both
// orientations: a raw pair is tested with two 64-bit compares per entry.
struct HotKeys {
  uint64_t fwd[kHotTop], bwd[kHotTop];
  uint32_t slot[kHotTop];
  bool any;
};

__device__ inline uint64_t pack_raw(int32_t x, int32_t y) {
  return (uint64_t)(uint32_t)x | ((uint64_t)(uint32_t)y << 32);
}

__device__ inline void hot_keys_load(HotKeys &h, const HotPair *__restrict__ top) {
#pragma unroll
  for (uint32_t t = 0; t < kHotTop; ++t) {
    const HotPair e = top[t];
    h.fwd[t] = pack_raw(e.a, e.b);
    h.bwd[t] = pack_raw((int32_t)(0u - (uint32_t)e.b), (int32_t)(0u - (uint32_t)e.a));
    h.slot[t] = e.slot;
  }
  h.any = top[0].set != 0;
}

// True when the raw pair (x, y) is a hot pair, with its slot and whether it is
// the reverse of the canonical key (a pair equal to its own reverse is not).
__device__ inline bool hot_match(const HotKeys &h, int32_t x, int32_t y, bool &rev,
                                 uint32_t &slot) {
  const uint64_t p = pack_raw(x, y);
  bool f = false, r = false;
  uint32_t s = 0;
#pragma unroll
  for (uint32_t t = 0; t < kHotTop; ++t) {
    const bool ft = p == h.fwd[t], rt = p == h.bwd[t];
    s = (ft || rt) ? h.slot[t] : s;
    f |= ft;
    r |= rt;
  }
  rev = r && !f;
  slot = s;
  return h.any && (f || r);
}

// Pairs of one pass of a tile (item j of this lane at tile offset
// first + j * stride) and the valid set. Interior tiles (every pair's second
// element in the stream) are addressed by 32-bit offsets from the tile's base
// with no clamping; the last tile takes load_pairs.
template <uint32_t N, bool kInterior>
__device__ inline uint32_t load_tile_pairs(const int32_t *__restrict__ x,
                                           const uint32_t *__restrict__ heads, uint64_t L,
                                           uint64_t base, uint32_t first, uint32_t stride,
                                           int32_t (&a)[N], int32_t (&b)[N], uint32_t fresh) {
  if constexpr (!kInterior) {
    return load_pairs<N>(x, heads, L, base + first, stride, a, b, fresh);
  } else {
    const int32_t *__restrict__ xt = x + base;
    const uint32_t *__restrict__ ht = heads + (base >> 5);  // base is a multiple of 32
    uint32_t hw[N];
#pragma unroll
    for (uint32_t j = 0; j < N; ++j) {
      const uint32_t r = first + j * stride;
      a[j] = xt[r];
      b[j] = xt[r + 1];
      hw[j] = ht[(r + 1) >> 5];
    }
    uint32_t valid = 0;
#pragma unroll
    for (uint32_t j = 0; j < N; ++j)
      valid |= (((hw[j] >> ((first + j * stride + 1) & 31)) & 1) || !fresh_pair(a[j], b[j], fresh))
                   ? 0u
                   : 1u << j;
    return valid;
  }
}


// ---- K1 for hot rounds: known repeated pairs dropped before aggregation ------
// In early rounds a few keys make up most pairs (the backbone's own pair is
// three quarters of the first round's). A key seen twice among distinct sampled pairs
// repeats, so before counting one block samples the stream as k_hot_sample does
// and publishes the kHotTop most frequent keys seen at least kHotMin times, as
// canonical raw pairs. Aggregation then drops their pairs with two raw 64-bit
// compares each and queues the wave's other pairs raw in LDS, aggregating them
// densely a wave-full at a time, so the key, dictionary and table work runs over
// those only. The known keys enter the counting as one extra tile of entries
// already marked repeated (extra[0 .. *extra_n)).
__global__ __launch_bounds__(kHotThreads) void k_count_hot(const int32_t *__restrict__ x,
                                                         const uint32_t *__restrict__ heads,
                                                         uint64_t L, uint64_t span, uint32_t fresh,
                                                         HotPair *__restrict__ known,
                                                         unsigned long long *__restrict__ extra,
                                                         uint32_t *__restrict__ extra_n) {
  __shared__ unsigned long long skey[kHotSampleSlots];
  __shared__ uint32_t scnt[kHotSampleSlots];
  __shared__ unsigned long long lkey[kWave];
  __shared__ uint32_t lcnt[kWave];
  __shared__ uint32_t ln;
  __shared__ HotPair stop[kHotTop];
  for (uint32_t s = threadIdx.x; s < kHotSampleSlots; s += kHotThreads) {
    skey[s] = kFree;
    scnt[s] = 0;
  }
  if (threadIdx.x < kHotTop)
    stop[threadIdx.x] = HotPair{0, 0, 0, 0};
  if (threadIdx.x == 0)
    ln = 0;
  __syncthreads();
  // Distinct sampled positions, as in k_hot_sample.
  const uint64_t room = L > kHotWin + 1 ? L - kHotWin - 1 : 0;
  const bool spread = room >= (uint64_t)kHotWindows * kHotWin;
  for (uint32_t q = threadIdx.x; q < kHotWindows * kHotWin; q += kHotThreads) {
    const uint64_t i =
        spread ? (uint64_t)(q / kHotWin) * room / kHotWindows + q % kHotWin : (uint64_t)q;
    if (i + 1 >= L || head_at(heads, i + 1) || !fresh_pair(x[i], x[i + 1], fresh))
      continue;
    const unsigned long long k = pair_key(x[i], x[i + 1], span);
    uint32_t sl = (uint32_t)key_hash(k) & (kHotSampleSlots - 1);
    for (uint32_t p = 0; p < kHotSampleSlots; ++p) {
      const unsigned long long old = atomicCAS(&skey[sl], (unsigned long long)kFree, k);
      if (old == kFree || old == k) {
        atomicAdd(&scnt[sl], 1u);
        break;
      }
      sl = (sl + 1) & (kHotSampleSlots - 1);
    }
  }
  __syncthreads();
  for (uint32_t s = threadIdx.x; s < kHotSampleSlots; s += kHotThreads)
    if (scnt[s] >= kHotMin) {
      const uint32_t at = atomicAdd(&ln, 1u);
      lkey[at] = skey[s];
      lcnt[at] = scnt[s];
    }
  __syncthreads();
  if (threadIdx.x < kWave) {
    const uint32_t lane = threadIdx.x, n = ln;
    const bool have = lane < n;
    const unsigned long long k = have ? lkey[lane] : kFree;
    const uint32_t c = have ? lcnt[lane] : 0;
    uint32_t rank = 0;
    for (uint32_t o = 0; o < n; ++o) {
      const unsigned long long ko = __shfl(k, (int)o);
      const uint32_t co = __shfl(c, (int)o);
      rank += co > c || (co == c && ko < k);
    }
    if (have && rank < kHotTop) {
      const int64_t cp = expand_key(k, span);
      stop[rank] = HotPair{key_first(cp), key_second(cp), 0, 1};
      extra[rank] = k | kRep;
    }
    if (lane == 0)
      *extra_n = n < kHotTop ? n : kHotTop;
  }
  __syncthreads();
  if (threadIdx.x < kHotTop) {
    HotPair e = stop[threadIdx.x];
    for (uint32_t t = 0; t < kHotTop && !e.set; ++t)
      if (stop[t].set)
        e = stop[t];
    known[threadIdx.x] = e;
  }
}

// Tile aggregation with the known repeated pairs dropped (k_count_hot): each
// wave queues its other pairs raw in LDS (stream order within a pass) and
// aggregates the queue a wave-full at a time as k_pair_aggregate aggregates a
// tile, so the queue holds at most two wave-fulls.
constexpr uint32_t kColdQueue = 2 * kWave;
template <uint32_t kSlots>
__global__ __launch_bounds__(kBlock) void k_pair_aggregate_cold(
    const int32_t *__restrict__ x, const uint32_t *__restrict__ heads, uint64_t L, uint64_t span,
    uint32_t fresh, const HotPair *__restrict__ known, unsigned long long *__restrict__ out,
    uint32_t *__restrict__ out_n) {
  __shared__ unsigned long long skeys[kSlots];
  __shared__ unsigned long long queue[kBlock / kWave][kColdQueue];
  __shared__ uint32_t lds[kBlock / kWave + 1];
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  for (uint32_t s = threadIdx.x; s < kSlots; s += kBlock)
    skeys[s] = kFree;
  HotKeys hk;
  hot_keys_load(hk, known);
  const uint64_t base = (uint64_t)blockIdx.x * kInsTile;
  int32_t a[kInsItems], b[kInsItems];
  uint32_t valid;
  if (base + kInsTile < L)  // interior tile: every pair's second element in the stream
    valid = load_tile_pairs<kInsItems, true>(x, heads, L, base, threadIdx.x, kBlock, a, b, fresh);
  else
    valid = load_tile_pairs<kInsItems, false>(x, heads, L, base, threadIdx.x, kBlock, a, b, fresh);
  const uint64_t below = lanemask_lt();
  __syncthreads();  // the table is clear
  uint64_t dkey[kDict];
  uint32_t dcnt[kDict];
#pragma unroll
  for (int d = 0; d < kDict; ++d) {
    dkey[d] = kFree;
    dcnt[d] = 0;
  }
  auto drain = [&](uint32_t e0, uint32_t cnt) {  // queue entries [e0, e0 + cnt), cnt <= kWave
    const bool v = lane < cnt;
    uint64_t k = kFree;
    if (v) {
      const unsigned long long r = queue[wid][e0 + lane];
      k = pair_key((int32_t)(uint32_t)r, (int32_t)(uint32_t)(r >> 32), span);
    }
    dict_step<kSlots>(skeys, dkey, dcnt, k, v);
  };
  uint32_t n = 0;  // queued entries (< kWave between items)
#pragma unroll
  for (uint32_t j = 0; j < kInsItems; ++j) {
    bool rev = false;
    uint32_t sl = 0;
    const bool c = ((valid >> j) & 1) && !hot_match(hk, a[j], b[j], rev, sl);
    const uint64_t m = __ballot(c);
    if (c)
      queue[wid][n + (uint32_t)__popcll(m & below)] = pack_raw(a[j], b[j]);
    n += (uint32_t)__popcll(m);
    if (n >= kWave) {  // a wave-full: aggregate it, keep the rest at the front
      drain(0, kWave);
      const unsigned long long rest = lane < n - kWave ? queue[wid][kWave + lane] : 0ull;
      if (lane < n - kWave)
        queue[wid][lane] = rest;
      n -= kWave;
    }
  }
  if (n)
    drain(0, n);
  if (lane < kDict) {
    uint64_t fk = kFree;
    uint32_t fc = 0;
#pragma unroll
    for (int d = 0; d < kDict; ++d)
      if ((int)lane == d) {
        fk = dkey[d];
        fc = dcnt[d];
      }
    if (fc)
      lds_add<kSlots>(skeys, fk, fc);
  }
  __syncthreads();
  constexpr int kPer = kSlots / kBlock;
  unsigned long long v[kPer];
  uint32_t cnt = 0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    v[i] = skeys[threadIdx.x * kPer + i];
    cnt += v[i] != kFree;
  }
  uint32_t total;
  uint32_t off = block_exclusive_scan<kBlock>(cnt, lds, total);
  unsigned long long *dst = out + (uint64_t)blockIdx.x * kInsTile;
#pragma unroll
  for (int i = 0; i < kPer; ++i)
    if (v[i] != kFree)
      dst[off++] = v[i];
  if (threadIdx.x == 0)
    out_n[blockIdx.x] = total;
}

// Candidate bits, each tile's pick function, and the slot words of the cold
// candidates (the rule slot of the pair starting there, kSlotRev when the pair
// is the reverse of the rule's canonical key), with a bit per cold candidate.
// Each wave's slot words are packed, in position order, at the front of its
// kWavePos entries of the tile's kRtTile (a wave takes candidate words w with
// w % kWaves == wave), so the k-th cold candidate of a wave (counted on its
// cold bits) has word k there: emit then reads the few words of its cold picks
// from a short dense run instead of one line each, and no wave waits for
// another to place its words. Hot pairs need no word: emit tests its other
// picks against the hot pairs again, from the raw pair.
constexpr uint32_t kWaves = kBlock / kWave;
constexpr uint32_t kWavePos = kRtItems * kWave;  // tile positions of one wave
template <bool kInterior>
__device__ inline void cand_tile(const int32_t *__restrict__ x, const uint32_t *__restrict__ heads,
                                 uint64_t L, uint64_t span, uint32_t fresh, const RuleTable &rt,
                                 const HotKeys &hk, HotLds &sh,
                                 unsigned long long (*list)[kColdCap], uint64_t *wbits,
                                 uint64_t *__restrict__ cand, uint64_t *__restrict__ coldw,
                                 uint64_t n_words, uint32_t *__restrict__ slots, uint64_t tile) {
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const uint64_t base = tile * kRtTile;
  uint32_t *__restrict__ st = slots + base + wid * kWavePos;  // this wave's slot words
  const uint64_t below = lanemask_lt();
  uint32_t cold_run = 0;  // this wave's cold candidates so far
  for (uint32_t j0 = 0; j0 < kRtItems; j0 += kPassItems) {
    // Validity is a per-lane bit set (bool arrays live as scalar lane masks).
    int32_t xa[kPassItems], xb[kPassItems];
    const uint32_t valid = load_tile_pairs<kPassItems, kInterior>(
        x, heads, L, base, j0 * kBlock + threadIdx.x, kBlock, xa, xb, fresh);
    // Listed items keep their list position per lane (no ballot masks held
    // across the lookups: they would spill scalar registers).
    // Listed (cold) pairs store their slot word after the wave's lookups, and
    // only when they are rules.
    uint32_t hit = 0, listed = 0, neg = 0, n = 0;
    uint32_t at[kPassItems];
#pragma unroll
    for (uint32_t b = 0; b < kPassItems; ++b) {
      bool rev = false;
      uint32_t hs = 0;
      const bool v = (valid >> b) & 1;
      const bool h = v && hot_match(hk, xa[b], xb[b], rev, hs);
      const bool c = v && !h;
      hit |= h ? 1u << b : 0u;
      const uint64_t m = __ballot(c);
      at[b] = n + (uint32_t)__popcll(m & below);
      listed |= c ? 1u << b : 0u;
      if (c) {
        const int64_t raw = pack2(xa[b], xb[b]);
        list[wid][at[b]] = (unsigned long long)raw;
        neg |= raw != canon2(raw) ? 1u << b : 0u;
      }
      n += (uint32_t)__popcll(m);
    }
    __syncthreads();
    resolve_list(rt, sh, span, list[wid], n);
    __syncthreads();
#pragma unroll
    for (uint32_t b = 0; b < kPassItems; ++b) {
      bool f = (hit >> b) & 1, c = false;
      unsigned long long sl = kFree;
      if ((listed >> b) & 1) {
        sl = list[wid][at[b]];
        f = c = sl != kFree;
      }
      const uint64_t word = __ballot(f);
      const uint64_t cword = __ballot(c);
      if (c)
        st[cold_run + (uint32_t)__popcll(cword & below)] =
            (uint32_t)sl | ((neg >> b) & 1 ? kSlotRev : 0u);
      cold_run += (uint32_t)__popcll(cword);
      if (lane == 0) {
        const uint32_t w = (j0 + b) * kWaves + wid;
        wbits[w] = word;
        const uint64_t gw = (base >> 6) + w;
        if (gw < n_words) {
          cand[gw] = word;
          coldw[gw] = cword;
        }
      }
    }
    __syncthreads();  // the list is refilled by the next pass
  }
}

__global__ __launch_bounds__(kBlock) void k_cand(const int32_t *__restrict__ x,
                                                 const uint32_t *__restrict__ heads, uint64_t L,
                                                 uint64_t span, uint32_t fresh, RuleTable rt,
                                                 const HotEntry *__restrict__ hot,
                                                 const HotPair *__restrict__ top,
                                                 uint64_t *__restrict__ cand,
                                                 uint64_t *__restrict__ coldw,
                                                 uint64_t n_words, uint32_t *__restrict__ tile_fn,
                                                 uint32_t *__restrict__ slots, TileOrder order) {
  const uint64_t tile = order_tile(order);
  if (tile == kNoTile)
    return;
  __shared__ uint64_t wbits[kRtWords];
  __shared__ HotLds sh;
  __shared__ unsigned long long list[kBlock / kWave][kColdCap];
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  HotKeys hk;
  hot_keys_load(hk, top);
  hot_load(sh, hot);
  __syncthreads();
  // Every tile but the last is interior (its last pair ends inside the stream).
  if ((tile + 1) * kRtTile < L)
    cand_tile<true>(x, heads, L, span, fresh, rt, hk, sh, list, wbits, cand, coldw, n_words,
                    slots, tile);
  else
    cand_tile<false>(x, heads, L, span, fresh, rt, hk, sh, list, wbits, cand, coldw, n_words,
                     slots, tile);
  if (wid == 0) {
    const uint32_t f = wave_fn_inclusive(word_fn(wbits[lane]));
    if (lane == kWave - 1)
      tile_fn[tile] = f;
  }
}

// ---- K2b: carry bit and output offset of every tile --------------------------
// tile_info[t] = (survivors before tile t) << 1 | pick[-1] of tile t;
// *total = survivors overall (the next round's length). A scan of the tiles'
// transfer functions in three launches: each block of kTileScanThreads threads
// composes its tiles (one per thread while there are at most kTileScanThreads
// blocks), one block scans the blocks' compositions, and each block then
// replays its tiles from its carry-in.
__device__ inline FnWide tile_widen(const uint32_t *__restrict__ tile_fn, uint64_t t, uint64_t L) {
  const uint32_t f = tile_fn[t];
  const uint64_t len = (t + 1) * kRtTile <= L ? kRtTile : L - t * kRtTile;
  FnWide w;
  w.o0 = fn_out(f, 0);
  w.o1 = fn_out(f, 1);
  w.s0 = len - fn_cnt(f, 0) + w.o0;
  w.s1 = len - 1 - fn_cnt(f, 1) + w.o1;
  return w;
}

// Exclusive scan of one FnWide per thread across the block (kTileScanThreads);
// returns this thread's prefix, *block_total the block's composition.
__device__ inline FnWide fnw_block_exclusive(FnWide mine, FnWide *wsum, FnWide *carry_in,
                                             FnWide *block_total) {
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  FnWide inc = mine;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    FnWide g;
    g.s0 = __shfl_up(inc.s0, d);
    g.s1 = __shfl_up(inc.s1, d);
    g.o0 = __shfl_up(inc.o0, d);
    g.o1 = __shfl_up(inc.o1, d);
    if (lane >= (uint32_t)d)
      inc = fnw_then(g, inc);
  }
  if (lane == kWave - 1)
    wsum[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    FnWide run = fnw_identity();
    for (uint32_t w = 0; w < kTileScanThreads / kWave; ++w) {
      carry_in[w] = run;
      run = fnw_then(run, wsum[w]);
    }
    *block_total = run;
  }
  __syncthreads();
  FnWide ex;
  ex.s0 = __shfl_up(inc.s0, 1);
  ex.s1 = __shfl_up(inc.s1, 1);
  ex.o0 = __shfl_up(inc.o0, 1);
  ex.o1 = __shfl_up(inc.o1, 1);
  if (lane == 0)
    ex = fnw_identity();
  return fnw_then(carry_in[wid], ex);
}

// Tiles [b * per_block, (b + 1) * per_block) of block b, chunk per thread.
struct TileRange {
  uint64_t begin, end;
};
__device__ inline TileRange tile_range(uint64_t n_tiles, uint64_t per_block) {
  const uint64_t b0 = (uint64_t)blockIdx.x * per_block;
  const uint64_t b1 = b0 + per_block < n_tiles ? b0 + per_block : n_tiles;
  const uint64_t chunk = (per_block + kTileScanThreads - 1) / kTileScanThreads;
  const uint64_t begin = b0 + (uint64_t)threadIdx.x * chunk;
  TileRange r;
  r.begin = begin < b1 ? begin : b1;
  r.end = begin + chunk < b1 ? begin + chunk : b1;
  return r;
}

__global__ __launch_bounds__(kTileScanThreads) void k_tile_reduce(
    const uint32_t *__restrict__ tile_fn, uint64_t n_tiles, uint64_t per_block, uint64_t L,
    FnWide *__restrict__ block_fn) {
  __shared__ FnWide wsum[kTileScanThreads / kWave];
  __shared__ FnWide carry_in[kTileScanThreads / kWave];
  __shared__ FnWide tot;
  const TileRange r = tile_range(n_tiles, per_block);
  FnWide mine = fnw_identity();
  for (uint64_t t = r.begin; t < r.end; ++t)
    mine = fnw_then(mine, tile_widen(tile_fn, t, L));
  fnw_block_exclusive(mine, wsum, carry_in, &tot);
  if (threadIdx.x == 0)
    block_fn[blockIdx.x] = tot;
}

// One block: exclusive prefixes of the blocks' compositions (in place; at most
// kTileScanThreads of them) and the survivors overall.
__global__ __launch_bounds__(kTileScanThreads) void k_tile_blocks(FnWide *__restrict__ block_fn,
                                                                  uint64_t n_blocks,
                                                                  uint64_t *__restrict__ total) {
  __shared__ FnWide wsum[kTileScanThreads / kWave];
  __shared__ FnWide carry_in[kTileScanThreads / kWave];
  __shared__ FnWide tot;
  const FnWide mine = threadIdx.x < n_blocks ? block_fn[threadIdx.x] : fnw_identity();
  const FnWide pre = fnw_block_exclusive(mine, wsum, carry_in, &tot);
  if (threadIdx.x < n_blocks)
    block_fn[threadIdx.x] = pre;
  if (threadIdx.x == 0)
    *total = tot.s0;
}

// block_pre null: a single block with no carry-in, which also writes *total.
__global__ __launch_bounds__(kTileScanThreads) void k_tile_apply(
    const uint32_t *__restrict__ tile_fn, uint64_t n_tiles, uint64_t per_block, uint64_t L,
    const FnWide *__restrict__ block_pre, uint64_t *__restrict__ tile_info,
    uint64_t *__restrict__ total) {
  __shared__ FnWide wsum[kTileScanThreads / kWave];
  __shared__ FnWide carry_in[kTileScanThreads / kWave];
  __shared__ FnWide tot;
  const TileRange r = tile_range(n_tiles, per_block);
  FnWide mine = fnw_identity();
  for (uint64_t t = r.begin; t < r.end; ++t)
    mine = fnw_then(mine, tile_widen(tile_fn, t, L));
  const FnWide in = block_pre ? block_pre[blockIdx.x] : fnw_identity();
  const FnWide pre = fnw_then(in, fnw_block_exclusive(mine, wsum, carry_in, &tot));
  if (!block_pre && threadIdx.x == 0)
    *total = tot.s0;
  uint32_t p = pre.o0;
  uint64_t off = pre.s0;
  for (uint64_t t = r.begin; t < r.end; ++t) {
    tile_info[t] = (off << 1) | p;
    const FnWide w = tile_widen(tile_fn, t, L);
    off += p ? w.s1 : w.s0;
    p = p ? w.o1 : w.o0;
  }
}

// ---- K3: emit the next round's stream ---------------------------------------
// One wave, lane l: the picks and survivors of candidate word l of tile t
// (p_tile = the pick carried into the tile). Survivors are the valid elements
// whose predecessor was not picked.
__device__ inline void word_survivors(const uint64_t *__restrict__ cand, uint64_t n_words,
                                      uint64_t L, uint64_t t, uint32_t p_tile, uint64_t &picks,
                                      uint64_t &surv) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t gw = t * kRtWords + lane;
  const uint64_t c = gw < n_words ? cand[gw] : 0;
  const uint32_t inc = wave_fn_inclusive(word_fn(c));
  uint32_t ex = __shfl_up(inc, 1);
  if (lane == 0)
    ex = fn_identity();
  const uint32_t p_in = fn_out(ex, p_tile);
  picks = word_picks(c, p_in);
  const uint64_t lo = t * kRtTile + (uint64_t)lane * 64;
  const uint64_t nvalid = lo >= L ? 0 : (L - lo >= 64 ? 64 : L - lo);
  const uint64_t vmask = nvalid == 64 ? ~0ull : ((1ull << nvalid) - 1);
  surv = ~((picks << 1) | (uint64_t)p_in) & vmask;
}

// Each traversal's first element in the next stream: one wave per traversal
// counts the survivors of its start's tile before it (heads always survive).
__global__ __launch_bounds__(kBlock) void k_starts_next(
    const uint64_t *__restrict__ start, uint32_t n, const uint64_t *__restrict__ cand,
    uint64_t n_words, const uint64_t *__restrict__ tile_info, uint64_t L,
    uint64_t *__restrict__ start_next) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t h = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (h >= n)
    return;
  const uint64_t p = start[h], t = p / kRtTile;
  const uint64_t info = tile_info[t];
  uint64_t picks, surv;
  word_survivors(cand, n_words, L, t, (uint32_t)(info & 1), picks, surv);
  const uint64_t o = p - t * kRtTile, lo = (uint64_t)lane * 64;
  const uint64_t before = o <= lo ? 0 : (o - lo >= 64 ? ~0ull : ((1ull << (o - lo)) - 1));
  uint32_t cnt = (uint32_t)__popcll(surv & before);
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1)
    cnt += __shfl_xor(cnt, d);
  if (lane == 0)
    start_next[h] = (info >> 1) + cnt;
}

// Survivors are elements whose predecessor was not picked; a picked element
// emits +/-(next + slot) (provisional, remapped after the sort). A pick's slot
// and orientation come from the hot pairs (tested raw, as in K2) or, for a cold
// pick, from its slot word (its wave's k-th for the wave's k-th cold
// candidate). Heads always survive; their bits in the next stream are set from
// the next traversal starts (k_starts_next).
constexpr uint32_t kEmitPass = 8;  // items per pass of emit (their loads in flight together)
template <bool kInterior>
__device__ inline void emit_tile(const int32_t *__restrict__ x, uint64_t L,
                                 const uint32_t *__restrict__ slots, uint64_t rcap,
                                 const HotKeys &hk, const uint64_t *spicks,
                                 const uint64_t *ssurv, const uint32_t *soff,
                                 const uint64_t *scold, uint64_t off_tile,
                                 uint8_t *__restrict__ used, uint32_t next,
                                 int32_t *__restrict__ y, uint32_t *__restrict__ err,
                                 uint32_t &hot_used) {
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const uint64_t below = lanemask_lt();
  const uint64_t base = (uint64_t)blockIdx.x * kRtTile;
  const int32_t *__restrict__ xt = x + base;
  const uint32_t *__restrict__ st = slots + base;
  // This wave's cold candidates so far, from its slot words' region on: the
  // next cold candidate's slot word.
  uint32_t cold_run = wid * kWavePos;
  for (uint32_t j0 = 0; j0 < kRtItems; j0 += kEmitPass) {
    // Every load of the pass first, the slot words of its cold picks included
    // (interior tiles: every pair inside the stream; the last tile clamps its
    // indexes into it): one wait per pass, not one per cold pick.
    int32_t xa[kEmitPass], xb[kEmitPass];
    uint32_t sw[kEmitPass];
    uint32_t coldpick = 0;
#pragma unroll
    for (uint32_t b = 0; b < kEmitPass; ++b) {
      const uint32_t r = (j0 + b) * kBlock + threadIdx.x;
      if constexpr (kInterior) {
        xa[b] = xt[r];
        xb[b] = xt[r + 1];
      } else {
        const uint64_t i = base + r;
        const uint64_t ia = i < L ? i : L - 1;
        const uint64_t ib = ia + 1 < L ? ia + 1 : L - 1;
        xa[b] = x[ia];
        xb[b] = x[ib];
      }
      const uint32_t w = (j0 + b) * kWaves + wid;
      const uint64_t cw = scold[w];
      const bool cp = ((spicks[w] & cw) >> lane) & 1;
      sw[b] = cp ? st[cold_run + (uint32_t)__popcll(cw & below)] : 0u;
      coldpick |= cp ? 1u << b : 0u;
      cold_run += (uint32_t)__popcll(cw);
    }
#pragma unroll
    for (uint32_t b = 0; b < kEmitPass; ++b) {
      const uint32_t w = (j0 + b) * kWaves + wid;
      const uint32_t r = w * 64 + lane;
      const uint64_t svm = ssurv[w];
      if ((kInterior || base + r < L) && ((svm >> lane) & 1)) {
        const uint64_t pos = off_tile + soff[w] + (uint64_t)__popcll(svm & below);
        int32_t o = xa[b];
        if ((spicks[w] >> lane) & 1) {
          bool rev = false;
          uint32_t sl = 0;
          if ((coldpick >> b) & 1) {
            sl = sw[b] & ~kSlotRev;
            rev = (sw[b] & kSlotRev) != 0;
            if (sl < rcap) {
              used[sl] = 1;
            } else {
              atomicExch(err, 1u);  // a slot word always holds a table slot
              sl = 0;
            }
          } else if (hot_match(hk, xa[b], xb[b], rev, sl)) {
#pragma unroll
            for (uint32_t q = 0; q < kHotTop; ++q)
              if (hk.slot[q] == sl)
                hot_used |= 1u << q;
          } else {
            atomicExch(err, 1u);  // a pick is a hot or a cold candidate
            sl = 0;
          }
          o = (int32_t)(next + sl);
          if (rev)
            o = -o;
        }
        y[pos] = o;
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_emit(
    const int32_t *__restrict__ x, uint64_t L,
    const uint64_t *__restrict__ cand, const uint64_t *__restrict__ coldw, uint64_t n_words,
    const uint64_t *__restrict__ tile_info,
    const uint32_t *__restrict__ slots, uint64_t rcap, const HotPair *__restrict__ top,
    uint8_t *__restrict__ used, uint32_t next, int32_t *__restrict__ y,
    uint32_t *__restrict__ err) {
  __shared__ uint64_t spicks[kRtWords];
  __shared__ uint64_t ssurv[kRtWords];
  __shared__ uint32_t soff[kRtWords];
  __shared__ uint64_t scold[kRtWords];
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const uint64_t base = (uint64_t)blockIdx.x * kRtTile;
  const uint64_t info = tile_info[blockIdx.x];
  const uint32_t p_tile = (uint32_t)(info & 1);
  const uint64_t off_tile = info >> 1;
  HotKeys hk;
  hot_keys_load(hk, top);
  if (wid == 0) {
    const uint64_t gw = (base >> 6) + lane;
    uint64_t picks, surv;
    word_survivors(cand, n_words, L, blockIdx.x, p_tile, picks, surv);
    const uint32_t sc = (uint32_t)__popcll(surv);
    uint32_t sx = sc;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t g = __shfl_up(sx, d);
      if (lane >= (uint32_t)d)
        sx += g;
    }
    spicks[lane] = picks;
    ssurv[lane] = surv;
    soff[lane] = sx - sc;
    scold[lane] = gw < n_words ? coldw[gw] : 0;
  }
  __syncthreads();
  uint32_t hot_used = 0;  // hot entries this lane picked (their used flags, once per wave)
  if ((uint64_t)(blockIdx.x + 1) * kRtTile < L)
    emit_tile<true>(x, L, slots, rcap, hk, spicks, ssurv, soff, scold, off_tile, used, next,
                    y, err, hot_used);
  else
    emit_tile<false>(x, L, slots, rcap, hk, spicks, ssurv, soff, scold, off_tile, used, next,
                     y, err, hot_used);
#pragma unroll
  for (uint32_t t = 0; t < kHotTop; ++t)
    if (__ballot((hot_used >> t) & 1) && lane == 0)
      used[hk.slot[t]] = 1;
}

// ---- K4: keys of the used rules (order irrelevant: they are sorted next) -----
// Each block owns kCollectTile consecutive rule-table slots: a block-wide scan of
// the used flags gives every key its place, and one atomic per block claims the
// range. Keys are already compacted, so their unsigned order is the signed
// packed-key order the reference sorts by.
constexpr uint32_t kCollectItems = 16;
constexpr uint32_t kCollectTile = kBlock * kCollectItems;

__global__ __launch_bounds__(kBlock) void k_collect_used(const unsigned long long *__restrict__ rkeys,
                                                         const uint8_t *__restrict__ used,
                                                         uint64_t cap, uint64_t *__restrict__ out,
                                                         unsigned long long *scal) {
  // scal[1] = keys collected.
  __shared__ uint32_t lds[kBlock / kWave + 1];
  __shared__ unsigned long long sbase;
  const uint64_t s0 = (uint64_t)blockIdx.x * kCollectTile + (uint64_t)threadIdx.x * kCollectItems;
  uint32_t flags = 0;
  if (s0 + kCollectItems <= cap) {
    const uint4 v = *reinterpret_cast<const uint4 *>(used + s0);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (uint32_t q = 0; q < 16; ++q)
      flags |= ((w[q >> 2] >> (8 * (q & 3))) & 0xFFu) ? (1u << q) : 0u;
  } else {
    for (uint32_t q = 0; q < kCollectItems; ++q)
      if (s0 + q < cap && used[s0 + q])
        flags |= 1u << q;
  }
  uint32_t total;
  uint32_t off = block_exclusive_scan<kBlock>((uint32_t)__popc(flags), lds, total);
  if (threadIdx.x == 0)
    sbase = total ? atomicAdd(&scal[1], (unsigned long long)total) : 0ull;
  __syncthreads();
  const uint64_t b = sbase + off;
  uint32_t k = 0;
  while (flags) {
    const uint32_t q = (uint32_t)__ffs(flags) - 1;
    out[b + k++] = rkeys[s0 + q];
    flags &= flags - 1;
  }
}

// ---- K6: final IDs next + rank, rule columns ----------------------------------
// Each thread looks up kAssignBatch sorted keys at once (their key loads, then
// their line reads, in flight together).
constexpr int kAssignBatch = 4;
__global__ __launch_bounds__(kBlock) void k_assign(const uint64_t *__restrict__ sorted, uint64_t R,
                                                   uint64_t span, RuleTable rt, uint32_t next,
                                                   int32_t *__restrict__ rf,
                                                   int32_t *__restrict__ rs,
                                                   uint32_t *__restrict__ err) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t r0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r0 < R;
       r0 += stride * kAssignBatch) {
    uint64_t ck[kAssignBatch];
    bool v[kAssignBatch];
    int64_t s[kAssignBatch];
#pragma unroll
    for (int q = 0; q < kAssignBatch; ++q) {
      const uint64_t r = r0 + (uint64_t)q * stride;
      v[q] = r < R;
      ck[q] = v[q] ? sorted[r] : 0;
    }
    rule_find_batch<kAssignBatch>(rt, ck, v, s);
#pragma unroll
    for (int q = 0; q < kAssignBatch; ++q) {
      const uint64_t r = r0 + (uint64_t)q * stride;
      if (!v[q])
        continue;
      if (s[q] < 0) {
        atomicExch(err, 2u);
        continue;
      }
      rt.vals[s[q]] = next + (uint32_t)r;
      const int64_t k = expand_key(ck[q], span);
      rf[r] = key_first(k);
      rs[r] = key_second(k);
    }
  }
}

// ---- K7: provisional -> final IDs (compression_utils.cpp:201-221) -------------
// Provisional IDs (next + slot) to final ones (vals[slot]), in place, one
// kRtTile tile per block, four symbols per 16-byte access with every load of the
// tile issued first. The round's hot rules (most picks of the early rounds) take
// their final IDs from registers instead of a gather each. In the
// cross-traversal tile order (TileOrder over the next stream's traversals) the
// same stretch of every haplotype is remapped together on one XCD, so the
// final-ID lines of the rules they share are gathered into its L2 once.
constexpr uint32_t kRemapQuads = kRtTile / 4 / kBlock;
__global__ __launch_bounds__(kBlock) void k_remap(int32_t *__restrict__ y, uint64_t n,
                                                  const uint32_t *__restrict__ vals, uint32_t next,
                                                  const HotPair *__restrict__ top,
                                                  TileOrder order) {
  const uint64_t tile = order_tile(order);
  if (tile == kNoTile)
    return;
  uint32_t hs[kHotTop], hf[kHotTop];
#pragma unroll
  for (uint32_t t = 0; t < kHotTop; ++t) {
    hs[t] = top[t].slot;
    hf[t] = vals[hs[t]];  // a slot of the table (unset entries copy a set one or hold 0)
  }
  const bool any = top[0].set != 0;
  auto map = [&](int32_t v) -> int32_t {
    const uint32_t a = uabs(v);
    if (a < next)
      return v;
    const uint32_t sl = a - next;
    bool hit = false;
    uint32_t f = 0;
#pragma unroll
    for (uint32_t t = 0; t < kHotTop; ++t) {
      const bool m = any && sl == hs[t];
      f = m ? hf[t] : f;
      hit |= m;
    }
    if (!hit)
      f = vals[sl];
    return v > 0 ? (int32_t)f : -(int32_t)f;
  };
  const uint64_t q0 = tile * (kRtTile / 4);  // first 16-byte quad of the tile
  const uint64_t n4 = n / 4;
  int4 *__restrict__ y4 = reinterpret_cast<int4 *>(y);  // y: a pool buffer, 256-byte aligned
  int4 v[kRemapQuads];
#pragma unroll
  for (uint32_t q = 0; q < kRemapQuads; ++q) {
    const uint64_t c = q0 + (uint64_t)q * kBlock + threadIdx.x;
    v[q] = c < n4 ? y4[c] : make_int4(0, 0, 0, 0);
  }
#pragma unroll
  for (uint32_t q = 0; q < kRemapQuads; ++q) {
    const uint64_t c = q0 + (uint64_t)q * kBlock + threadIdx.x;
    if (c < n4) {
      int4 w = v[q];
      w.x = map(w.x);
      w.y = map(w.y);
      w.z = map(w.z);
      w.w = map(w.w);
      y4[c] = w;
    }
  }
  // The stream's last n % 4 symbols (inside the last tile).
  if ((tile + 1) * kRtTile >= n && threadIdx.x < n - n4 * 4) {
    const uint64_t i = n4 * 4 + threadIdx.x;
    y[i] = map(y[i]);
  }
}

uint32_t stride_grid(Ctx &ctx, uint64_t n) {
  const uint64_t want = ceil_div(n, kBlock * 4);
  const uint64_t cap = (uint64_t)ctx.num_cus() * 8;
  return (uint32_t)std::max<uint64_t>(1, std::min(want, cap));
}

}  // namespace

void compress_device(Ctx &ctx, const int32_t *d_nodes, uint64_t N, const uint32_t *lengths,
                     uint64_t T, const CompressParams &p, CompressResult &out) {
  if (p.rounds < 0)
    throw Error(kErrInvalid, "rounds must be >= 0");
  const int delta_round = p.delta_round < 0 ? 0 : p.delta_round;  // clamp, as the reference
  out = CompressResult();
  out.encoded_lengths.assign(T, 0);
  std::vector<uint64_t> starts;
  std::vector<uint64_t> ne_index;
  uint64_t total = 0;
  for (uint64_t t = 0; t < T; ++t) {
    if (lengths[t]) {
      starts.push_back(total);
      ne_index.push_back(t);
    }
    total += lengths[t];
  }
  if (total != N)
    throw Error(kErrInvalid, "sum of traversal lengths (" + std::to_string(total) +
                                 ") != number of nodes (" + std::to_string(N) + ")");
  Pool &pool = ctx.pool();
  const uint64_t words32 = ceil_div(std::max<uint64_t>(N, 1), 32);
  DBuf<uint32_t> heads(pool, words32), heads_next(pool, words32);
  fill_u32(ctx, heads.get(), 0, words32);
  // Each non-empty traversal's first element in the current stream (kept by
  // emit from round to round), and the longest traversal's length.
  if (starts.size() >= (1ull << 31))
    throw Error(kErrInvalid, "too many traversals (" + std::to_string(starts.size()) + ")");
  const uint32_t n_start = (uint32_t)starts.size();
  DBuf<uint64_t> tstart(pool, std::max<uint64_t>(n_start, 1)),
      tstart_next(pool, std::max<uint64_t>(n_start, 1));
  uint64_t longest = 0;
  for (uint64_t t = 0; t < T; ++t)
    longest = std::max<uint64_t>(longest, lengths[t]);
  if (n_start) {
    GFZ_HIP(hipMemcpyAsync(tstart.get(), starts.data(), starts.size() * 8,
                           hipMemcpyHostToDevice, ctx.stream()));
    set_head_bits(ctx, tstart.get(), n_start, heads.get(), N);
    ctx.sync();  // starts is host memory of this call
  }

  // Delta transform and the rule-ID floor.
  DBuf<int32_t> xa(pool, std::max<uint64_t>(N, 1)), xb(pool, std::max<uint64_t>(N, 1));
  DBuf<uint32_t> d_max(pool, 2);  // [0] max |delta|; read back as one 64-bit word
  fill_u32(ctx, d_max.get(), 0, 2);
  const int32_t *cur = d_nodes;
  int32_t *cur_buf = nullptr;
  if (N) {
    // The kernel reads 16-byte vectors: a caller's pointer that is not 16-byte
    // aligned is first copied into an aligned buffer.
    if (reinterpret_cast<uintptr_t>(d_nodes) % 16) {
      GFZ_HIP(hipMemcpyAsync(xb.get(), d_nodes, N * 4, hipMemcpyDeviceToDevice, ctx.stream()));
      cur = xb.get();
    }
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(ceil_div(N, (uint64_t)kBlock * 8), (uint64_t)ctx.num_cus() * 8));
    if (delta_round == 0) {
      ctx.timed("delta", [&] {
        hipLaunchKernelGGL(k_delta, dim3(grid), dim3(kBlock), 0, ctx.stream(), cur,
                           (int32_t *)nullptr, heads.get(), N, 0, d_max.get());
      });
      GFZ_CHECK_LAUNCH();
    }
    for (int r = 0; r < delta_round; ++r) {
      int32_t *dst = (cur == xa.get()) ? xb.get() : xa.get();
      ctx.timed("delta", [&] {
        hipLaunchKernelGGL(k_delta, dim3(grid), dim3(kBlock), 0, ctx.stream(), cur, dst,
                           heads.get(), N, 1, d_max.get());
      });
      GFZ_CHECK_LAUNCH();
      cur = dst;
      cur_buf = dst;
    }
  }
  const uint32_t max_abs = (uint32_t)ctx.fetch({{d_max.get(), 1, 0}})[0];
  if (max_abs == UINT32_MAX)
    throw Error(kErrOverflow, "delta values too large for rule ID assignment");
  uint32_t next = max_abs + 1;
  out.start_id = next;

  // Wide symbols: renumber large magnitudes (see k_wide_map).
  const uint32_t next_orig = next;
  uint32_t next_wide = 0;  // nonzero: the stream is renumbered
  DBuf<uint32_t> d_wide;
  if (N && next >= kWideTrigger) {
    const uint64_t cap = std::min<uint64_t>(N, kWideCap);
    DBuf<uint32_t> vals(pool, cap);
    DBuf<unsigned long long> cnt(pool, 1);
    fill_u32(ctx, reinterpret_cast<uint32_t *>(cnt.get()), 0, 2);
    hipLaunchKernelGGL(k_wide_gather, dim3(stride_grid(ctx, N)), dim3(kBlock), 0, ctx.stream(),
                       cur, N, vals.get(), cap, cnt.get());
    GFZ_CHECK_LAUNCH();
    const uint64_t n_big = ctx.fetch({{cnt.get(), 1, 0}})[0];
    if (n_big > cap)
      throw Error(kErrOverflow, std::to_string(n_big) + " steps carry symbols of magnitude >= " +
                                    std::to_string(kWideFloor) + " (at most " +
                                    std::to_string(kWideCap) + " supported)");
    std::vector<uint32_t> u(n_big);
    if (n_big)
      GFZ_HIP(hipMemcpyAsync(u.data(), vals.get(), n_big * 4, hipMemcpyDeviceToHost,
                             ctx.stream()));
    ctx.sync();
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    next_wide = kWideFloor + (uint32_t)u.size();
    d_wide = DBuf<uint32_t>(pool, std::max<size_t>(u.size(), 1));
    if (!u.empty())
      GFZ_HIP(hipMemcpyAsync(d_wide.get(), u.data(), u.size() * 4, hipMemcpyHostToDevice,
                             ctx.stream()));
    int32_t *dst = (cur == xa.get()) ? xb.get() : xa.get();
    hipLaunchKernelGGL(k_wide_map, dim3(stride_grid(ctx, N)), dim3(kBlock), 0, ctx.stream(), cur,
                       dst, N, d_wide.get(), (uint32_t)u.size());
    GFZ_CHECK_LAUNCH();
    ctx.sync();  // u is released at scope end
    cur = dst;
    cur_buf = dst;
    next = next_wide;
  }

  uint64_t L = N;
  uint64_t rules_cap = 0;
  DBuf<int32_t> rf, rs;
  DBuf<uint32_t> d_err(pool, 2);
  fill_u32(ctx, d_err.get(), 0, 2);
  uint64_t prev_distinct = 0, prev_len = 0, prev_elig = 0;
  // The floor of the last round's rule IDs: only pairs with a symbol at or
  // above it can repeat (fresh_pair); 0 in the first round.
  uint32_t fresh = 0;

  // Graph mode (tuned "round_graphs" = 1): each round's three stretches of
  // launches between host steps (counting; rule table to used keys; sort to
  // remap) run as HIP graphs, kept per round and updated in place on later
  // calls (Ctx::segment).
  struct GraphMode {
    Ctx &c;
    bool prev;
    GraphMode(Ctx &ctx_, bool on) : c(ctx_), prev(ctx_.graphs()) { c.set_graphs(on); }
    ~GraphMode() { c.set_graphs(prev); }
  } graph_mode(ctx, ctx.tuned("round_graphs") > 0);
  for (int round = 0; round < p.rounds && L >= 2; ++round) {
    RoundInfo info;
    info.length_in = L;
    auto seg_key = [&](const char *part) {
      return std::string("compress/") + part + "/" + std::to_string(round);
    };
    // Pairs that can be fresh: each of the last round's picks left one new
    // symbol, and a symbol starts at most two pairs (all pairs in the first round).
    const uint64_t elig = round == 0 ? L : std::min<uint64_t>(L, 2 * (prev_len - L));
    int32_t *dst = (cur_buf == xa.get()) ? xb.get() : xa.get();
    const uint64_t n_tiles = ceil_div(L, kRtTile);
    const uint64_t n_words = ceil_div(L, 64);
    // Compacted keys need every |symbol| < next < 2^30 (keys below 2^62).
    if (next >= (1u << 30))
      throw Error(kErrOverflow, "symbol space too large for 62-bit pair keys (next id " +
                                    std::to_string(next) + ")");
    const uint64_t span = 2 * (uint64_t)next;

    // K1 with capacity retry: the counting table holds at most cap/2 distinct
    // keys. Distinct pairs grow slowly from round to round (each round's keys are
    // pairs of the previous round's symbols); 4x the last count, at most 2L
    // (there are fewer than L pairs), leaves the table at most half full.
    uint64_t cap = next_pow2(std::max<uint64_t>(
        uint64_t(1) << 22,
        round == 0 ? L / 8 : std::min<uint64_t>(2 * L, 4 * prev_distinct)));
    DBuf<unsigned long long> ctab;
    const uint64_t n_ins = ceil_div(L, kInsTile);
    DBuf<uint2> bcounts(pool, std::max<uint64_t>(n_ins + 1, kMaxBuckets));
    DBuf<unsigned long long> stage;
    // distinct, repeated, overflow, stage overflow; bucket-counted rounds: distinct
    // and repeated packed in [0], the longest bucket list in [8]
    DBuf<unsigned long long> sums(pool, 9);
    uint64_t distinct = 0, repeated = 0;
    // Tile aggregation pays while keys recur within a tile. Once half of the last
    // round's pairs were distinct, tiles hold few repeats (a key's copies lie in
    // different traversals): the partition then reads the pairs straight from the
    // stream. "count_direct_share" (tuned or fixed) replaces the 2 (on plain_10m
    // and plain_500k, 2, 4 and 8 measured the same within 1 %).
    const int tuned_share = ctx.tuned("count_direct_share");
    const uint64_t direct_share = tuned_share > 0 ? (uint64_t)tuned_share : 2;
    // Rounds where most pairs are not fresh aggregate too: the tiles then hold
    // few keys, and the stream is read once instead of twice.
    const bool mostly_distinct =
        round > 0 && direct_share * prev_distinct > prev_elig && 2 * elig > L;
    DBuf<unsigned long long> tiles;  // aggregated tiles (rounds with hot keys)
    DBuf<uint32_t> tiles_n;
    // Tiles of few distinct keys (under a quarter of the last round's pairs were
    // distinct) aggregate in the smaller LDS table.
    // The aggregation table (2048 or 4096 slots) follows the last round's share of
    // distinct pairs unless the tuned "count_agg_slots/<size class>" fixes it.
    const int tuned_agg = ctx.tuned("count_agg_slots/" + std::to_string(Ctx::size_class(N)));
    const bool few_keys = tuned_agg > 0 ? tuned_agg <= (int)kInsTile
                                        : round == 0 || 4 * prev_distinct < prev_elig;
    // Rounds of few keys drop the sampled known repeated pairs before aggregating
    // (k_count_hot); their keys follow as one extra aggregated tile. The knob
    // "count_hot_drop" = n limits it to the first n rounds, 0 turns it off (the
    // grammar is the same either way). On plain_10m the first two rounds take it
    // (aggregation 8.2 -> 5.6 and 4.6 -> 3.4 ms).
    const int hot_knob = ctx.tuned("count_hot_drop");
    const bool hot_drop = few_keys && (hot_knob < 0 || round < hot_knob);
    const uint64_t n_agg = n_ins + (hot_drop ? 1 : 0);  // aggregated tiles
    DBuf<HotPair> known;
    auto aggregate = [&] {
      tiles = DBuf<unsigned long long>(pool, n_agg * kInsTile);
      tiles_n = DBuf<uint32_t>(pool, n_agg);
      if (hot_drop) {
        known = DBuf<HotPair>(pool, kHotTop);
        hipLaunchKernelGGL(k_count_hot, dim3(1), dim3(kHotThreads), 0, ctx.stream(), cur,
                           heads.get(), L, span, fresh, known.get(),
                           tiles.get() + n_ins * kInsTile, tiles_n.get() + n_ins);
        GFZ_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_pair_aggregate_cold<kInsTile>, dim3((uint32_t)n_ins), dim3(kBlock),
                           0, ctx.stream(), cur, heads.get(), L, span, fresh, known.get(),
                           tiles.get(), tiles_n.get());
      } else if (few_keys)
        hipLaunchKernelGGL(k_pair_aggregate<kInsTile>, dim3((uint32_t)n_ins), dim3(kBlock), 0,
                           ctx.stream(), cur, heads.get(), L, span, fresh, tiles.get(),
                           tiles_n.get());
      else
        hipLaunchKernelGGL(k_pair_aggregate<2 * kInsTile>, dim3((uint32_t)n_ins), dim3(kBlock),
                           0, ctx.stream(), cur, heads.get(), L, span, fresh, tiles.get(),
                           tiles_n.get());
      GFZ_CHECK_LAUNCH();
    };
    const bool direct = round > 0 && 2 * prev_distinct > prev_elig;  // table fallback mode
    bool stage_overflow = false;
    uint64_t stage_stride = kStageCap, stage_blocks = direct ? n_ins : n_agg;
    bool counted = false;
    uint64_t regions = 0, region_max = 0;  // bucket-counted: rule table by bucket regions
    // Rounds count in buckets while their distinct keys (the first round's
    // estimated at L/32) fit kMaxBuckets buckets of kBktSlots / 2 (two-level
    // partitions beyond kMaxBuckets1); larger rounds, and any round whose buckets
    // overflow, count in the HBM table.
    const uint64_t est =
        round == 0 ? L / 32 : std::min<uint64_t>(elig, prev_distinct + prev_distinct / 4);
    // Tuning and test knobs (ctx tuned values; every setting gives the same
    // grammar): "count_buckets" fixes the bucket count (a power of two, 64 ..
    // kMaxBuckets), "count_rep_cap" lowers the repeated keys a bucket may list
    // (a bucket above it sends the round to the HBM table).
    const int forced_buckets = ctx.tuned("count_buckets");
    // "count_bucket_scale/<size class>": 1, 2 or 4 halves, keeps or doubles the
    // estimated bucket count (the autotuner's knob).
    const int bucket_scale = ctx.tuned("count_bucket_scale/" + std::to_string(Ctx::size_class(N)));
    const int forced_rep_cap = ctx.tuned("count_rep_cap");
    const uint32_t rep_cap =
        forced_rep_cap > 0 ? std::min<uint32_t>((uint32_t)forced_rep_cap, kRepStride) : kRepStride;
    const bool use_buckets =
        forced_buckets > 0 || est <= (uint64_t)kMaxBuckets * (kBktSlots / 2);
    if (use_buckets) {
      const uint64_t nbk =
          forced_buckets > 0
              ? std::min<uint64_t>(kMaxBuckets, next_pow2(std::max<uint64_t>(64, forced_buckets)))
              : std::min<uint64_t>(
                    kMaxBuckets,
                    std::max<uint64_t>(
                        64, next_pow2(std::max<uint64_t>(64, ceil_div(est, kBktSlots / 2))) *
                                (bucket_scale == 1 || bucket_scale == 4 ? bucket_scale : 2) /
                                2));
      // At most kMaxParts partition blocks, so a bucket gathers at most that many
      // segments. Beyond kMaxBuckets1 buckets the keys are split twice: by coarse
      // bucket (the bucket's high bits) into the partition blocks' regions, then
      // each coarse bucket by fine bucket into kL2Parts regions, so a bucket
      // gathers kL2Parts long segments instead of kMaxParts short ones.
      // (the partition of aggregated tiles takes the known keys' extra tile too)
      const uint64_t n_src = mostly_distinct ? n_ins : n_agg;
      const bool two_level = nbk > kMaxBuckets1;
      // Two-level rounds gather second-level regions, not the partition blocks'
      // segments, so they use more partition blocks (kMaxParts2, or the knob
      // "count_parts2"): on plain_10m 2048 -> 16384 blocks with 16 -> 64 regions
      // per coarse bucket took the counting from 50.1 to 45.9 ms per compress.
      const int tuned_parts = ctx.tuned("count_parts2");
      const uint64_t max_parts =
          !two_level ? kMaxParts : tuned_parts > 0 ? (uint64_t)tuned_parts : kMaxParts2;
      const uint32_t per = (uint32_t)std::max<uint64_t>(1, ceil_div(n_src, max_parts));
      const uint64_t n_parts = ceil_div(n_src, per);
      const uint64_t nb1 = two_level ? nbk >> kFineBits : nbk;
      DBuf<unsigned long long> part(pool, n_parts * per * kInsTile);
      DBuf<uint32_t> starts(pool, n_parts * (nb1 + 1));
      stage = DBuf<unsigned long long>(pool, nbk * kRepStride);
      const Parts p1{part.get(), (uint64_t)per * kInsTile, nullptr, starts.get(),
                     (uint32_t)nb1, (uint32_t)n_parts, (uint32_t)(nbk - 1),
                     two_level ? kFineBits : 0u};
      // Second-level regions per coarse bucket ("count_l2parts", at most kWave:
      // a counting block takes that many segments in its few-segment path).
      const int tuned_l2 = ctx.tuned("count_l2parts");
      const uint32_t l2parts =
          tuned_l2 > 0 ? std::min<uint32_t>((uint32_t)tuned_l2, kWave) : kL2Parts;
      const uint64_t n_l2 = two_level ? nb1 * l2parts : 0;
      DBuf<uint32_t> l2_sizes, l2_starts;
      DBuf<uint64_t> l2_base;
      DBuf<unsigned long long> part2;
      Parts pt = p1;
      if (two_level) {
        l2_sizes = DBuf<uint32_t>(pool, n_l2);
        l2_base = DBuf<uint64_t>(pool, n_l2 + 1);
        l2_starts = DBuf<uint32_t>(pool, n_l2 * (kFineBuckets + 1));
      }
      ctx.segment(seg_key("count"), [&] {
      fill_runs(ctx, {{sums.get(), 9, 0, 8}});
      ctx.timed("pair_insert", [&] {
        if (mostly_distinct) {
          hipLaunchKernelGGL(k_bucket_part, dim3((uint32_t)n_parts), dim3(kBlock), nb1 * 4,
                             ctx.stream(), cur, heads.get(), L, span, fresh, n_ins, per, p1);
        } else {
          // Hot keys: aggregate each tile first, so a key enters a bucket at most
          // once per tile.
          aggregate();
          hipLaunchKernelGGL(k_bucket_part_tiles, dim3((uint32_t)n_parts), dim3(kBlock), nb1 * 4,
                             ctx.stream(), tiles.get(), tiles_n.get(), n_agg, per, p1);
        }
        GFZ_CHECK_LAUNCH();
        if (two_level) {
          hipLaunchKernelGGL(k_part2_sizes, dim3((uint32_t)ceil_div(n_l2, kBlock)), dim3(kBlock),
                             0, ctx.stream(), p1, l2parts, (uint32_t)nb1, l2_sizes.get());
          GFZ_CHECK_LAUNCH();
          exclusive_scan(ctx, l2_sizes.get(), l2_base.get(), n_l2);
          // The second level's regions hold the first level's entries: the
          // aggregated tiles (read by then) have room for them.
          unsigned long long *e2 = tiles.get();
          if (!e2) {
            part2 = DBuf<unsigned long long>(pool, n_parts * per * kInsTile);
            e2 = part2.get();
          }
          pt = Parts{e2, 0, l2_base.get(), l2_starts.get(), kFineBuckets, l2parts,
                     (uint32_t)(nbk - 1), 0u};
          hipLaunchKernelGGL(k_bucket_part2, dim3((uint32_t)n_l2), dim3(kBlock),
                             kFineBuckets * 4, ctx.stream(), p1, pt);
          GFZ_CHECK_LAUNCH();
        }
        hipLaunchKernelGGL(k_bucket_count, dim3((uint32_t)nbk), dim3(kBktThreads), 0,
                           ctx.stream(), pt, stage.get(), bcounts.get(), sums.get(),
                           reinterpret_cast<uint32_t *>(sums.get() + 2), rep_cap);
        GFZ_CHECK_LAUNCH();
      });
      });
      ctx.add_bytes("pair_insert", 4.0 * L);
      const uint64_t *hc = ctx.fetch({{sums.get(), 9, 0}});
      if (!hc[2]) {
        distinct = hc[0] >> 32;
        repeated = hc[0] & 0xffffffffull;
        regions = nbk;
        region_max = hc[8];
        counted = true;
        stage_stride = kRepStride;
        stage_blocks = nbk;
        cap = nbk * kBktSlots;
      } else {
        info.table_retries += 1;  // a bucket overflowed: count with the table instead
        if (two_level && !mostly_distinct)
          tiles = DBuf<unsigned long long>();  // its room held the second level: aggregate again
      }
    }
    if (!counted)
      stage = DBuf<unsigned long long>(pool, stage_blocks * kStageCap);
    for (; !counted;) {
      ctab = DBuf<unsigned long long>(pool, cap);
      fill_runs(ctx, {{ctab.get(), cap, kFree, 8}, {sums.get(), 4, 0, 8}});
      CountTable t{ctab.get(), cap - 1, reinterpret_cast<uint32_t *>(sums.get() + 2),
                   bcounts.get(), stage.get(), reinterpret_cast<uint32_t *>(sums.get() + 3)};
      ctx.timed("pair_insert", [&] {
        if (direct) {
          hipLaunchKernelGGL(k_pair_insert_direct, dim3((uint32_t)n_ins), dim3(kBlock), 0,
                             ctx.stream(), cur, heads.get(), L, span, fresh, t);
        } else {
          if (!tiles.get()) {  // the aggregation does not depend on the table size
            aggregate();
          }
          hipLaunchKernelGGL(k_pair_insert_staged, dim3((uint32_t)n_agg), dim3(kBlock), 0,
                             ctx.stream(), tiles.get(), tiles_n.get(), t);
        }
      });
      GFZ_CHECK_LAUNCH();
      ctx.add_bytes("pair_insert", 4.0 * L);
      hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(1024), 0, ctx.stream(), bcounts.get(),
                         stage_blocks, sums.get());
      GFZ_CHECK_LAUNCH();
      const uint64_t *hc = ctx.fetch({{sums.get(), 4, 0}});
      stage_overflow = hc[3] != 0;
      if (!hc[2] && hc[0] <= cap / 2) {
        distinct = hc[0];
        repeated = hc[1];
        counted = true;
        break;
      }
      cap *= 4;
      info.table_retries += 1;
    }
    info.distinct_pairs = distinct;
    info.table_slots = cap;
    prev_distinct = distinct;
    prev_len = L;
    prev_elig = elig;
    if (repeated == 0)
      break;  // no pair occurs twice: the reference stops here (compression_workflow.cpp:576)

    // K1b: rule table of the repeated keys (load <= 1/4, whole 64-byte lines);
    // a bucket-counted round takes one region per bucket of whole lines, 9/4 of
    // its longest list ("rule_region_pct" gives another percentage above 100;
    // load <= 4/9 in the fullest region, lower on average). Lookups of absent
    // keys (candidate marking) want a sparse table, the arrays indexed by slot
    // (values, used flags; emit and remap) a small one: on plain_10m 150 %,
    // 200 %, 250 % and 400 % took 139.0, 134.2, 133.7 and 135.6 ms of kernels
    // per compress, twice the longest list rounded up to a power of two 137.
    const int region_pct = ctx.tuned("rule_region_pct");
    const uint64_t region_lines =
        regions ? std::min<uint64_t>(
                      kBktSlots / kLine,
                      std::max<uint64_t>(1, ceil_div(region_max * (region_pct > 100 ? region_pct
                                                                                    : 225),
                                                     100 * (uint64_t)kLine)))
                : 0;
    const uint64_t rcap =
        regions ? regions * region_lines * kLine
                : next_pow2(std::max<uint64_t>(1024, 4 * (uint64_t)repeated));
    if ((uint64_t)next + rcap >= (uint64_t)INT32_MAX)
      throw Error(kErrOverflow, "rule-ID space exhausted (next id " + std::to_string(next) +
                                    ", rule table " + std::to_string(rcap) + ")");
    DBuf<unsigned long long> rkeys(pool, rcap);  // pool blocks are whole hipMalloc blocks
    DBuf<uint32_t> rvals(pool, rcap);
    // The absence filter pays once a quarter of the fresh pairs are no rule (each
    // of the distinct - repeated keys is one such pair): about 16 bits per repeated
    // key, at most 2^27 bits (16 MB, cache-resident).
    const bool use_filter = 4 * (distinct - std::min(distinct, repeated)) > elig;
    const uint32_t fbits =
        use_filter ? (uint32_t)std::min<uint64_t>(
                         27, std::max<uint64_t>(12, 64 - __builtin_clzll(
                                                             std::max<uint64_t>(repeated, 1)) + 4))
                   : 0u;
    DBuf<uint32_t> filter;
    if (use_filter)
      filter = DBuf<uint32_t>(pool, (uint64_t(1) << fbits) / 32);
    const RuleTable rt{rkeys.get(), rvals.get(),
                       (uint32_t)(regions ? region_lines : rcap / kLine),
                       regions ? (uint32_t)(regions - 1) : 0u, filter.get(), fbits};
    // Round scalars (read back once, after K4): [0] next length, [1] used keys,
    // [2] the longest traversal of the next stream.
    // One launch clears everything the round needs cleared: the rule table, the
    // scalars, the used flags and the next round's head bits (up to L: the next
    // length is at most L).
    DBuf<unsigned long long> scal(pool, 3);
    DBuf<uint8_t> used(pool, rcap);
    const uint64_t fwords = use_filter ? (uint64_t(1) << fbits) / 32 : 0;
    DBuf<HotEntry> hot(pool, kHotSlots);
    DBuf<HotPair> top(pool, kHotTop);
    // K2 + K2b: candidates, tile carries, next length.
    DBuf<uint64_t> cand(pool, n_words), coldw(pool, n_words);
    DBuf<uint32_t> tile_fn(pool, n_tiles);
    DBuf<uint64_t> tile_info(pool, n_tiles);
    DBuf<uint32_t> slots(pool, n_tiles * kRtTile);  // whole tiles: each wave's words at wave * kWavePos
    // Up to kTileScanSingle tiles one block scans them all (one launch); beyond,
    // blocks of kTileScanThreads tiles while that makes at most kTileScanThreads
    // blocks (then each thread takes a chunk), in three launches.
    const bool single = n_tiles <= kTileScanSingle;
    const uint64_t n_blocks =
        single ? 1 : std::min<uint64_t>(ceil_div(n_tiles, kTileScanThreads), kTileScanThreads);
    const uint64_t per_block = ceil_div(n_tiles, n_blocks);
    DBuf<FnWide> block_fn(pool, n_blocks);
    uint64_t *d_total = reinterpret_cast<uint64_t *>(scal.get());
    DBuf<uint64_t> ukeys(pool, repeated), utmp(pool, repeated);  // K4's used keys
    // Candidate marking in the cross-traversal tile order while its grid pads the
    // tiles by at most a quarter (traversals of similar length; "cand_order" = 0
    // keeps the launch order).
    TileOrder order{nullptr, L, 0, 0};
    uint64_t cand_grid = n_tiles;
    if (n_start >= 2 && ctx.tuned("cand_order") != 0) {
      uint32_t cols = 0;
      const uint64_t g = order_grid(n_start, longest, cols);
      if (4 * g <= 5 * n_tiles && g <= UINT32_MAX) {
        order = TileOrder{tstart.get(), L, n_start, cols};
        cand_grid = g;
      }
    }
    // Everything from the rule table to the used keys runs without a host step
    // in between: one graph segment.
    ctx.segment(seg_key("tail"), [&] {
    if (regions)  // the regions are written whole: the table needs no clearing
      fill_runs(ctx, {{scal.get(), 3, 0, 8},
                      {used.get(), rcap, 0, 1},
                      {heads_next.get(), ceil_div(L, 32), 0, 4},
                      {filter.get(), fwords, 0, 4}});
    else
      fill_runs(ctx, {{rkeys.get(), rcap, kFree, 8},
                      {scal.get(), 3, 0, 8},
                      {used.get(), rcap, 0, 1},
                      {heads_next.get(), ceil_div(L, 32), 0, 4},
                      {filter.get(), fwords, 0, 4}});
    ctx.timed("extract_repeated", [&] {
      if (regions)
        hipLaunchKernelGGL(k_region_build, dim3((uint32_t)regions), dim3(kBktThreads), 0,
                           ctx.stream(), stage.get(), stage_stride, bcounts.get(), rt,
                           d_err.get());
      else if (stage_overflow)  // some block staged too many: scan the whole table
        hipLaunchKernelGGL(k_extract_repeated,
                           dim3((uint32_t)ceil_div(cap, (uint64_t)kBlock * kExtractItems)),
                           dim3(kBlock), 0, ctx.stream(), ctab.get(), cap, rt, d_err.get());
      else
        hipLaunchKernelGGL(k_extract_staged, dim3((uint32_t)stage_blocks), dim3(kBlock), 0,
                           ctx.stream(), stage.get(), stage_stride, bcounts.get(), rt,
                           d_err.get());
    });
    GFZ_CHECK_LAUNCH();
    ctx.timed("cand_tiles", [&] {
      hipLaunchKernelGGL(k_hot_sample, dim3(1), dim3(kHotThreads), 0, ctx.stream(), cur,
                         heads.get(), L, span, fresh, rt, hot.get(), top.get());
    });
    GFZ_CHECK_LAUNCH();
    ctx.timed("cand_tiles", [&] {
      hipLaunchKernelGGL(k_cand, dim3((uint32_t)cand_grid), dim3(kBlock), 0, ctx.stream(), cur,
                         heads.get(), L, span, fresh, rt, hot.get(), top.get(), cand.get(),
                         coldw.get(), n_words, tile_fn.get(), slots.get(), order);
    });
    GFZ_CHECK_LAUNCH();
    ctx.timed("tile_scan", [&] {
      if (single) {
        hipLaunchKernelGGL(k_tile_apply, dim3(1), dim3(kTileScanThreads), 0, ctx.stream(),
                           tile_fn.get(), n_tiles, per_block, L, (const FnWide *)nullptr,
                           tile_info.get(), d_total);
      } else {
        hipLaunchKernelGGL(k_tile_reduce, dim3((uint32_t)n_blocks), dim3(kTileScanThreads), 0,
                           ctx.stream(), tile_fn.get(), n_tiles, per_block, L, block_fn.get());
        hipLaunchKernelGGL(k_tile_blocks, dim3(1), dim3(kTileScanThreads), 0, ctx.stream(),
                           block_fn.get(), n_blocks, d_total);
        hipLaunchKernelGGL(k_tile_apply, dim3((uint32_t)n_blocks), dim3(kTileScanThreads), 0,
                           ctx.stream(), tile_fn.get(), n_tiles, per_block, L, block_fn.get(),
                           tile_info.get(), (uint64_t *)nullptr);
      }
    });
    GFZ_CHECK_LAUNCH();

    // K3: emit.
    ctx.timed("emit", [&] {
      hipLaunchKernelGGL(k_emit, dim3((uint32_t)n_tiles), dim3(kBlock), 0, ctx.stream(), cur,
                         L, cand.get(), coldw.get(), n_words, tile_info.get(), slots.get(), rcap,
                         top.get(), used.get(), next, dst, d_err.get());
    });
    GFZ_CHECK_LAUNCH();
    if (n_start) {
      hipLaunchKernelGGL(k_starts_next, dim3((uint32_t)ceil_div(n_start, kBlock / kWave)),
                         dim3(kBlock), 0, ctx.stream(), tstart.get(), n_start, cand.get(),
                         n_words, tile_info.get(), L, tstart_next.get());
      // The next stream's head bits, at its traversal starts (all below L).
      set_head_bits(ctx, tstart_next.get(), n_start, heads_next.get(), L);
      hipLaunchKernelGGL(k_trav_longest, dim3((n_start + kBlock - 1) / kBlock), dim3(kBlock), 0,
                         ctx.stream(), tstart_next.get(), n_start, d_total, scal.get() + 2);
      GFZ_CHECK_LAUNCH();
    }

    // K4: used keys.
    ctx.timed("collect_used", [&] {
      hipLaunchKernelGGL(k_collect_used, dim3((uint32_t)ceil_div(rcap, kCollectTile)),
                         dim3(kBlock), 0, ctx.stream(), rkeys.get(), used.get(), rcap, ukeys.get(),
                         scal.get());
    });
    GFZ_CHECK_LAUNCH();
    });
    ctx.add_bytes("cand_tiles", 4.0 * L);
    stage.reset();
    ctab.reset();
    cand.reset();
    coldw.reset();
    tile_fn.reset();
    tile_info.reset();
    const uint64_t *hs = ctx.fetch({{scal.get(), 3, 0}, {d_err.get(), 1, 4}});
    const uint64_t L_next = hs[0], R = hs[1];
    const uint64_t longest_next = hs[2];
    if ((uint32_t)hs[4])
      throw Error(kErrCorrupt, "pair table lost a key (internal error)");
    if (L_next >= L)
      throw Error(kErrCorrupt, "repeated pairs found but none picked (internal error)");
    ctx.add_bytes("emit", 4.0 * L + 4.0 * L_next);
    if (R == 0)
      break;  // cannot happen when some pair repeats (compression_workflow.cpp:611)

    if (out.n_rules + R > rules_cap) {
      // Room for four times the first round's repeated keys up front: later
      // rounds add fewer rules, so the columns rarely have to grow.
      const uint64_t ncap = std::max<uint64_t>(std::max<uint64_t>(2 * rules_cap, out.n_rules + R),
                                               4 * (uint64_t)repeated);
      DBuf<int32_t> nf(pool, ncap), ns(pool, ncap);
      if (out.n_rules) {
        GFZ_HIP(hipMemcpyAsync(nf.get(), rf.get(), out.n_rules * 4, hipMemcpyDeviceToDevice,
                               ctx.stream()));
        GFZ_HIP(hipMemcpyAsync(ns.get(), rs.get(), out.n_rules * 4, hipMemcpyDeviceToDevice,
                               ctx.stream()));
      }
      rf = std::move(nf);
      rs = std::move(ns);
      rules_cap = ncap;
    }
    // Remap in the cross-traversal tile order of the next stream (its traversal
    // starts are tstart_next), on the same terms as candidate marking.
    const uint64_t next_tiles = ceil_div(L_next, kRtTile);
    TileOrder remap_order{nullptr, L_next, 0, 0};
    uint64_t remap_grid = std::max<uint64_t>(next_tiles, 1);
    if (n_start >= 2 && ctx.tuned("cand_order") != 0) {
      uint32_t cols = 0;
      const uint64_t g = order_grid(n_start, longest_next, cols);
      if (4 * g <= 5 * next_tiles && g <= UINT32_MAX) {
        remap_order = TileOrder{tstart_next.get(), L_next, n_start, cols};
        remap_grid = g;
      }
    }
    if (remap_grid > UINT32_MAX)
      throw Error(kErrInvalid, "stream too long for one remap launch");
    ctx.segment(seg_key("ids"), [&] {
    // K5: sort by signed key (compacted keys sort as unsigned); K6: IDs, columns.
    sort_unique_u64(ctx, ukeys.get(), utmp.get(), R, d_err.get(), 0,
                    64 - (uint32_t)__builtin_clzll(span * span));  // keys < span^2
    ctx.timed("assign_ids", [&] {
      hipLaunchKernelGGL(k_assign, dim3(stride_grid(ctx, R)), dim3(kBlock), 0, ctx.stream(),
                         ukeys.get(), R, span, rt, next, rf.get() + out.n_rules,
                         rs.get() + out.n_rules, d_err.get());
    });
    GFZ_CHECK_LAUNCH();
    // K7: remap provisional IDs (errors of K6 surface at the next readback).
    ctx.timed("remap", [&] {
      hipLaunchKernelGGL(k_remap, dim3((uint32_t)remap_grid), dim3(kBlock), 0, ctx.stream(),
                         dst, L_next, rvals.get(), next, top.get(), remap_order);
    });
    GFZ_CHECK_LAUNCH();
    });
    ctx.add_bytes("remap", 8.0 * L_next);

    fresh = next;
    next += (uint32_t)R;
    out.n_rules += R;
    info.length_out = L_next;
    info.rules = R;
    out.rounds.push_back(info);
    L = L_next;
    cur = dst;
    cur_buf = dst;
    heads.swap(heads_next);
    tstart.swap(tstart_next);
    longest = longest_next;
  }
  out.end_id = next;
  if (next_wide) {
    const uint64_t end = (uint64_t)next_orig + (next - next_wide);
    if (end > (uint64_t)INT32_MAX + 1)
      throw Error(kErrOverflow, "rule IDs would pass INT32_MAX (first " +
                                    std::to_string(next_orig) + ", " +
                                    std::to_string(next - next_wide) + " rules)");
    if (L)
      hipLaunchKernelGGL(k_wide_unmap, dim3(stride_grid(ctx, L)), dim3(kBlock), 0, ctx.stream(),
                         cur_buf, L, d_wide.get(), next_wide, next_orig);
    if (out.n_rules) {
      hipLaunchKernelGGL(k_wide_unmap, dim3(stride_grid(ctx, out.n_rules)), dim3(kBlock), 0,
                         ctx.stream(), rf.get(), out.n_rules, d_wide.get(), next_wide, next_orig);
      hipLaunchKernelGGL(k_wide_unmap, dim3(stride_grid(ctx, out.n_rules)), dim3(kBlock), 0,
                         ctx.stream(), rs.get(), out.n_rules, d_wide.get(), next_wide, next_orig);
    }
    GFZ_CHECK_LAUNCH();
    out.end_id = (uint32_t)end;
  }
  ctx.stage_scalars(d_err.get(), 1);
  ctx.sync();
  if (const uint32_t code = (uint32_t)ctx.scalars()[0])
    throw Error(kErrCorrupt, "grammar induction failed an internal check (code " +
                                 std::to_string(code) + ": 1 picked pair or 2 sorted key "
                                 "missing from the rule table, 3 rule table full, 4 sort "
                                 "look-back timed out)");

  // Encoded lengths from the traversal starts of the final stream.
  if (n_start) {
    std::vector<uint64_t> h(n_start);
    GFZ_HIP(hipMemcpyAsync(h.data(), tstart.get(), h.size() * 8, hipMemcpyDeviceToHost,
                           ctx.stream()));
    ctx.sync();
    for (uint64_t k = 0; k < n_start; ++k) {
      const uint64_t end = k + 1 < n_start ? h[k + 1] : L;
      if (h[k] >= end)
        throw Error(kErrCorrupt, "traversal starts lost during encoding (internal error)");
      out.encoded_lengths[ne_index[k]] = (uint32_t)(end - h[k]);
    }
  }
  out.n_encoded = L;
  if (cur_buf == xa.get()) {
    out.encoded = std::move(xa);
  } else if (cur_buf == xb.get()) {
    out.encoded = std::move(xb);
  } else {
    out.encoded = DBuf<int32_t>(pool, std::max<uint64_t>(L, 1));
    if (L)
      GFZ_HIP(hipMemcpyAsync(out.encoded.get(), d_nodes, L * 4, hipMemcpyDeviceToDevice,
                             ctx.stream()));
  }
  out.rules_first = std::move(rf);
  out.rules_second = std::move(rs);
  ctx.sync();
}

double autotune_compress(Ctx &ctx, const int32_t *d_nodes, uint64_t N, const uint32_t *lengths,
                         uint64_t T, const CompressParams &p, std::vector<TuneTrial> &trials) {
  const std::string cls = std::to_string(Ctx::size_class(N));
  const std::string k_scale = "count_bucket_scale/" + cls, k_slots = "count_agg_slots/" + cls;
  const int old_scale = ctx.tuned(k_scale), old_slots = ctx.tuned(k_slots);
  trials.clear();
  CompressResult first;
  bool have_first = false;
  double best = -1;
  int best_scale = 2, best_slots = (int)kInsTile;
  try {
    for (int scale : {2, 1, 4}) {
      for (int slots : {(int)kInsTile, 2 * (int)kInsTile}) {
        ctx.set_tuned(k_scale, scale);
        ctx.set_tuned(k_slots, slots);
        double ms = -1;
        for (int rep = 0; rep < 2; ++rep) {
          CompressResult r;
          ctx.sync();
          const auto t0 = std::chrono::steady_clock::now();
          compress_device(ctx, d_nodes, N, lengths, T, p, r);
          ctx.sync();
          const double dt =
              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                  .count();
          ms = ms < 0 ? dt : std::min(ms, dt);
          if (!have_first) {
            first = std::move(r);
            have_first = true;
          } else if (rep == 0) {
            const bool same =
                r.n_encoded == first.n_encoded && r.n_rules == first.n_rules &&
                r.start_id == first.start_id && r.end_id == first.end_id &&
                r.encoded_lengths == first.encoded_lengths &&
                count_mismatch_i32(ctx, r.encoded.get(), first.encoded.get(), r.n_encoded) == 0 &&
                count_mismatch_i32(ctx, r.rules_first.get(), first.rules_first.get(),
                                   r.n_rules) == 0 &&
                count_mismatch_i32(ctx, r.rules_second.get(), first.rules_second.get(),
                                   r.n_rules) == 0;
            if (!same)
              throw Error(kErrCorrupt, "autotune: counting setting (scale " +
                                           std::to_string(scale) + ", slots " +
                                           std::to_string(slots) +
                                           ") gave a different grammar (internal error)");
          }
        }
        trials.push_back({scale, slots, ms});
        if (best < 0 || ms < best) {
          best = ms;
          best_scale = scale;
          best_slots = slots;
        }
      }
    }
  } catch (...) {
    ctx.set_tuned(k_scale, old_scale);
    ctx.set_tuned(k_slots, old_slots);
    throw;
  }
  ctx.set_tuned(k_scale, best_scale);
  ctx.set_tuned(k_slots, best_slots);
  return best;
}

}  // namespace gfz
