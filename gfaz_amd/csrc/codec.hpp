// Host-side entry points of the grammar codec on one device.
#pragma once

#include "ctx.hpp"

#include <vector>

namespace gfz {

struct CompressParams {
  int rounds = 8;
  int threshold = 2;  // count >= 2 whatever the value (the reference's two-set scheme)
  int delta_round = 1;
};

struct RoundInfo {
  uint64_t length_in = 0;
  uint64_t length_out = 0;
  uint64_t distinct_pairs = 0;
  uint64_t rules = 0;
  uint64_t table_slots = 0;
  int table_retries = 0;
};

// Result of compressing one batch of traversals. Device arrays are owned by
// the Ctx pool (DBuf); host vectors are filled on request.
struct CompressResult {
  DBuf<int32_t> encoded;       // flat encoded traversals (device)
  uint64_t n_encoded = 0;
  std::vector<uint32_t> encoded_lengths;  // one per input traversal (host)
  DBuf<int32_t> rules_first;   // raw, rule-ID order (device)
  DBuf<int32_t> rules_second;
  uint64_t n_rules = 0;
  uint32_t start_id = 0;
  uint32_t end_id = 0;
  std::vector<RoundInfo> rounds;
};

// nodes: device pointer to the flat traversals (paths first, then walks);
// lengths: host array, one per traversal.
void compress_device(Ctx &ctx, const int32_t *d_nodes, uint64_t n_nodes,
                     const uint32_t *lengths, uint64_t n_trav, const CompressParams &p,
                     CompressResult &out);

// Autotuning of the pair-counting knobs on one input: every combination of
// "count_bucket_scale" (1, 2, 4) and "count_agg_slots" (2048, 4096) for the
// input's size class is run twice; each must give the first one's grammar
// (else kErrCorrupt), and the fastest is kept in the ctx's tuned table. Returns
// the fastest compress time (ms) and fills `trials` with (scale, slots, ms).
struct TuneTrial {
  int scale, slots;
  double ms;
};
double autotune_compress(Ctx &ctx, const int32_t *d_nodes, uint64_t n_nodes,
                         const uint32_t *lengths, uint64_t n_trav, const CompressParams &p,
                         std::vector<TuneTrial> &trials);

// Device-resident rulebook for expansion: per rule {first, second, |first|, |second|}
// (expanded sizes of the two children) plus the rule-ID window.
struct RuleRec {
  int32_t a, b;
  uint32_t sa, sb;
};

struct Rulebook {
  DBuf<RuleRec> rec;
  // Per rule: x = expanded size, y = sum of the expansion (mod 2^32); one
  // 8-byte gather gives both.
  DBuf<uint2> ss;
  uint64_t n_rules = 0;
  uint32_t min_id = 0;
  uint32_t depth = 0;       // deepest rule
  // Prefix-summed bodies with a leading zero: rule r at [zoff[r], zoff[r+1]),
  // size[r] + 1 entries. Built when the rule depths are non-decreasing in rule
  // ID and the total fits the budget; otherwise expansion descends the tree.
  bool has_bodies = false;
  DBuf<uint64_t> zoff;
  // With the bodies: per rule {expanded size, sum, zoff (below 2^31), 0}, so a
  // rule symbol's size, sum and body position come in one 16-byte gather.
  DBuf<uint4> szo;
  DBuf<int32_t> z;          // + two trailing zeros that node-ID symbols point at
  uint64_t z_elems = 0;
  uint32_t z_raw = 0;       // index of the second trailing zero
};

// first/second: device pointers (raw rule columns, rule-ID order).
void rulebook_build(Ctx &ctx, const int32_t *d_first, const int32_t *d_second, uint64_t n_rules,
                    uint32_t min_id, Rulebook &rb);

namespace zx {
// What one encoded symbol writes at its local position q of an expansion
// through the rule bodies (device side: zexpand.hpp).
struct alignas(8) SymInfo {
  uint32_t base;
  uint32_t zz;
};
}  // namespace zx

struct DecodePlan {
  uint64_t n_out = 0;                 // total decoded symbols
  std::vector<uint64_t> out_offsets;  // per traversal (n_trav + 1), host
  DBuf<zx::SymInfo> syms;             // per encoded symbol (rulebooks with bodies, decode)
  bool length_mismatch = false;       // decoded lengths differ from the stored originals
};

// Expand the encoded traversals and undo delta_round prefix sums.
// d_enc: device; enc_lengths / orig_lengths: host (orig may be null).
// d_out must hold the decoded total; call decode_size first to learn it.
// delta_round < 0: sizes only (no symbol records); otherwise the records for a
// decode_run with that delta_round.
void decode_plan(Ctx &ctx, const Rulebook &rb, const int32_t *d_enc, uint64_t n_enc,
                 const uint32_t *enc_lengths, const uint32_t *orig_lengths, uint64_t n_trav,
                 int delta_round, DecodePlan &plan, DBuf<uint64_t> &sym_off);
void decode_run(Ctx &ctx, const Rulebook &rb, const int32_t *d_enc, uint64_t n_enc,
                const uint32_t *enc_lengths, uint64_t n_trav, int delta_round,
                const DecodePlan &plan, const uint64_t *d_sym_off, int32_t *d_out);

// Segmented inclusive prefix sum (int32, wrapping) of x[0..n) in place, with
// segments starting at the set bits of `heads`.
void segmented_prefix_sum(Ctx &ctx, int32_t *x, uint64_t n, const uint32_t *heads);

}  // namespace gfz
