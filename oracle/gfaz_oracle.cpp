// TEST INFRASTRUCTURE ONLY - CPU oracle for the grammar codec; see gfaz_oracle.h
// for who may call it and what pins each function. Every step cites the
// reference line it restates; the reference's robin_hood sets are replaced by a
// concurrent open-addressing table because only set membership, never the
// iteration order, reaches the output (the used rules are re-sorted by key,
// src/workflows/compression_workflow.cpp:647-663).
#include "gfaz_oracle.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

using Seq = std::vector<int32_t>;
using Seqs = std::vector<Seq>;

// include/grammar/packed_2mer.hpp:15-17
inline int64_t pack_2mer(int32_t a, int32_t b) {
  return (int64_t)(((uint64_t)(uint32_t)a << 32) | (uint64_t)(uint32_t)b);
}
inline int32_t unpack_first(int64_t k) { return (int32_t)(k >> 32); }
inline int32_t unpack_second(int64_t k) { return (int32_t)k; }
// packed_2mer.hpp:37-48 (reverse = (-second, -first); canonical = signed min)
inline int64_t reverse_2mer(int64_t k) { return pack_2mer(-unpack_second(k), -unpack_first(k)); }
inline int64_t canonical_2mer(int64_t k) { return std::min(k, reverse_2mer(k)); }

inline uint32_t abs_u32(int32_t v) { return (uint32_t)std::abs(v); }

Seqs split(const int32_t *nodes, const uint32_t *lengths, uint64_t n_trav) {
  Seqs s(n_trav);
  uint64_t o = 0;
  for (uint64_t t = 0; t < n_trav; ++t) {
    s[t].assign(nodes + o, nodes + o + lengths[t]);
    o += lengths[t];
  }
  return s;
}

// src/codec/codec.cpp:45-76 (delta_transform_and_max_abs)
uint32_t delta_transform_and_max_abs(Seqs &seqs) {
  uint32_t max_abs = 0;
#pragma omp parallel for schedule(dynamic) reduction(max : max_abs)
  for (int64_t p = 0; p < (int64_t)seqs.size(); ++p) {
    Seq &s = seqs[p];
    if (s.empty())
      continue;
    uint32_t local = abs_u32(s[0]);
    for (size_t i = s.size() - 1; i > 0; --i) {
      s[i] = s[i] - s[i - 1];
      local = std::max(local, abs_u32(s[i]));
    }
    max_abs = std::max(max_abs, local);
  }
  return max_abs;
}

// src/workflows/compression_workflow.cpp:217-233 (max_abs_symbol)
uint32_t max_abs_symbol(const Seqs &seqs) {
  uint32_t max_abs = 0;
#pragma omp parallel for schedule(dynamic) reduction(max : max_abs)
  for (int64_t p = 0; p < (int64_t)seqs.size(); ++p)
    for (int32_t v : seqs[p])
      max_abs = std::max(max_abs, abs_u32(v));
  return max_abs;
}

// src/codec/codec.cpp:29-41 (inverse_delta_transform, one traversal)
void inverse_delta(Seq &s) {
  for (size_t i = 1; i < s.size(); ++i)
    s[i] = s[i] + s[i - 1];
}

// Concurrent set of canonical keys with a seen/repeated state per key. This is
// the two-set scheme of src/grammar/rule_generator.cpp:159-216: a key becomes
// "repeated" on its second occurrence, overlapping occurrences included.
struct PairTable {
  static constexpr int64_t kEmpty = INT64_MIN;  // pack(INT32_MIN, 0): never a key
  uint64_t cap = 0, mask = 0;
  std::unique_ptr<std::atomic<int64_t>[]> keys;
  std::unique_ptr<std::atomic<uint8_t>[]> repeated;
  std::unique_ptr<std::atomic<uint8_t>[]> used;

  explicit PairTable(uint64_t pairs) {
    cap = 1024;
    while (cap < 2 * pairs + 64)
      cap <<= 1;
    mask = cap - 1;
    keys.reset(new std::atomic<int64_t>[cap]);
    repeated.reset(new std::atomic<uint8_t>[cap]);
    used.reset(new std::atomic<uint8_t>[cap]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)cap; ++i) {
      keys[i].store(kEmpty, std::memory_order_relaxed);
      repeated[i].store(0, std::memory_order_relaxed);
      used[i].store(0, std::memory_order_relaxed);
    }
  }
  uint64_t home(int64_t k) const { return ((uint64_t)k * 0x9E3779B97F4A7C15ull >> 17) & mask; }
  void insert(int64_t k) {
    for (uint64_t s = home(k);; s = (s + 1) & mask) {
      int64_t cur = keys[s].load(std::memory_order_relaxed);
      if (cur == kEmpty) {
        if (keys[s].compare_exchange_strong(cur, k))
          return;  // first occurrence: seen
      }
      if (cur == k) {
        repeated[s].store(1, std::memory_order_relaxed);
        return;
      }
    }
  }
  // Slot of k if k is a repeated key, else -1.
  int64_t find_repeated(int64_t k) const {
    for (uint64_t s = home(k);; s = (s + 1) & mask) {
      const int64_t cur = keys[s].load(std::memory_order_relaxed);
      if (cur == kEmpty)
        return -1;
      if (cur == k)
        return repeated[s].load(std::memory_order_relaxed) ? (int64_t)s : -1;
    }
  }
};

// One grammar round over all traversals; returns the number of rules used
// (0 = stop). src/workflows/compression_workflow.cpp:567-690.
uint64_t grammar_round(Seqs &seqs, uint32_t &next_id, std::vector<int64_t> &rulebook,
                       int *err) {
  uint64_t pairs = 0;
  for (const Seq &s : seqs)
    pairs += s.size() > 1 ? s.size() - 1 : 0;
  if (pairs == 0)
    return 0;
  PairTable table(pairs);
  // generate_rules_2mer_combined (rule_generator.cpp:148-260)
#pragma omp parallel for schedule(dynamic)
  for (int64_t p = 0; p < (int64_t)seqs.size(); ++p) {
    const Seq &s = seqs[p];
    for (size_t i = 0; i + 1 < s.size(); ++i)
      table.insert(canonical_2mer(pack_2mer(s[i], s[i + 1])));
  }
  // Provisional IDs next_id + k, k dense over the repeated keys in table order,
  // like the hash-order IDs next_id + k of rule_generator.cpp:249-255 (any
  // order gives the same output: the used rules are renumbered by key below).
  // prov[slot] = k for a repeated slot; chunk counts first, then offsets.
  std::unique_ptr<uint32_t[]> prov(new uint32_t[table.cap]);
  const uint64_t chunk = 1 << 20, n_chunks = (table.cap + chunk - 1) / chunk;
  std::vector<uint64_t> chunk_rep(n_chunks + 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < (int64_t)n_chunks; ++c) {
    uint64_t n = 0;
    for (uint64_t i = c * chunk; i < std::min<uint64_t>(table.cap, (c + 1) * chunk); ++i)
      n += table.repeated[i].load(std::memory_order_relaxed);
    chunk_rep[c + 1] = n;
  }
  for (uint64_t c = 0; c < n_chunks; ++c)
    chunk_rep[c + 1] += chunk_rep[c];
  const uint64_t n_rep = chunk_rep[n_chunks];
  if (n_rep == 0)
    return 0;  // compression_workflow.cpp:576-577
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < (int64_t)n_chunks; ++c) {
    uint64_t k = chunk_rep[c];
    for (uint64_t i = c * chunk; i < std::min<uint64_t>(table.cap, (c + 1) * chunk); ++i)
      if (table.repeated[i].load(std::memory_order_relaxed))
        prov[i] = (uint32_t)k++;
  }
  // The reference's symbols are int32: IDs must stay below 2^31.
  if ((uint64_t)next_id + n_rep > (uint64_t)INT32_MAX) {
    *err = -2;
    return 0;
  }
  const uint32_t start = next_id;
  // encode_paths_2mer (src/grammar/path_encoder.cpp:29-83): greedy stack
  // encoder; forward key first (+id), then reverse key (-id).
#pragma omp parallel for schedule(dynamic)
  for (int64_t p = 0; p < (int64_t)seqs.size(); ++p) {
    Seq &path = seqs[p];
    Seq stack;
    stack.reserve(path.size());
    for (int32_t node : path) {
      stack.push_back(node);
      if (stack.size() < 2)
        continue;
      const int64_t top = pack_2mer(stack[stack.size() - 2], stack.back());
      int64_t slot = table.find_repeated(top);
      int32_t sym = 0;
      if (slot >= 0) {
        sym = (int32_t)(start + prov[slot]);
      } else {
        slot = table.find_repeated(reverse_2mer(top));
        if (slot >= 0)
          sym = -(int32_t)(start + prov[slot]);
      }
      if (slot >= 0) {
        stack.pop_back();
        stack.pop_back();
        stack.push_back(sym);
        table.used[slot].store(1, std::memory_order_relaxed);
      }
    }
    path.swap(stack);
  }
  // Used rules, sorted by signed key (compression_workflow.cpp:601-663).
  std::vector<std::pair<int64_t, uint64_t>> used;
  for (uint64_t i = 0; i < table.cap; ++i)
    if (table.used[i].load(std::memory_order_relaxed))
      used.emplace_back(table.keys[i].load(std::memory_order_relaxed), i);
  if (used.empty())
    return 0;  // compression_workflow.cpp:611-627
  std::sort(used.begin(), used.end());
  std::vector<uint32_t> id_map(n_rep, 0);  // provisional offset -> final ID
  for (size_t r = 0; r < used.size(); ++r)
    id_map[prov[used[r].second]] = start + (uint32_t)r;
  // remap_rule_ids (src/workflows/compression_utils.cpp:201-221)
#pragma omp parallel for schedule(dynamic)
  for (int64_t p = 0; p < (int64_t)seqs.size(); ++p) {
    for (int32_t &node : seqs[p]) {
      const uint32_t a = abs_u32(node);
      if (a < start)
        continue;
      const uint64_t off = a - start;
      if (off >= id_map.size() || id_map[off] == 0)
        continue;
      node = node > 0 ? (int32_t)id_map[off] : -(int32_t)id_map[off];
    }
  }
  for (const auto &u : used)
    rulebook.push_back(u.first);  // rulebook[final_id - layer_start] (:672-678)
  next_id = start + (uint32_t)used.size();
  return used.size();
}

// Recursive rule expansion (src/workflows/decompression_workflow.cpp:70-106).
void expand_rule(uint32_t rule_id, bool reverse, const int32_t *first, const int32_t *second,
                 uint32_t min_id, uint32_t max_id, Seq &out) {
  const uint32_t idx = rule_id - min_id;
  const int32_t a = first[idx], b = second[idx];
  if (!reverse) {
    const uint32_t aa = abs_u32(a), ab = abs_u32(b);
    if (aa >= min_id && aa < max_id)
      expand_rule(aa, a < 0, first, second, min_id, max_id, out);
    else
      out.push_back(a);
    if (ab >= min_id && ab < max_id)
      expand_rule(ab, b < 0, first, second, min_id, max_id, out);
    else
      out.push_back(b);
  } else {
    const uint32_t aa = abs_u32(a), ab = abs_u32(b);
    if (ab >= min_id && ab < max_id)
      expand_rule(ab, b >= 0, first, second, min_id, max_id, out);
    else
      out.push_back(-b);
    if (aa >= min_id && aa < max_id)
      expand_rule(aa, a >= 0, first, second, min_id, max_id, out);
    else
      out.push_back(-a);
  }
}

// src/grammar/path_encoder.cpp:29-83 with an exact-key table (the
// add-haplotypes layers of add_haplotypes_workflow.cpp:336-369).
void encode_with_table(Seqs &seqs, const std::unordered_map<int64_t, uint32_t> &table) {
  for (Seq &path : seqs) {
    Seq stack;
    stack.reserve(path.size());
    for (int32_t node : path) {
      stack.push_back(node);
      if (stack.size() < 2)
        continue;
      const int64_t top = pack_2mer(stack[stack.size() - 2], stack.back());
      auto it = table.find(top);
      int32_t sym = 0;
      bool found = false;
      if (it != table.end()) {
        sym = (int32_t)it->second;
        found = true;
      } else {
        auto jt = table.find(reverse_2mer(top));
        if (jt != table.end()) {
          sym = -(int32_t)jt->second;
          found = true;
        }
      }
      if (found) {
        stack.pop_back();
        stack.pop_back();
        stack.push_back(sym);
      }
    }
    path.swap(stack);
  }
}

}  // namespace

extern "C" {

int oracle_compress(const int32_t *nodes, uint64_t n_nodes, const uint32_t *lengths,
                    uint64_t n_trav, int rounds, int threshold, int delta_round,
                    int threads, oracle_compress_result *out) {
  (void)n_nodes;
  (void)threshold;  // count >= 2 regardless (rule_generator.cpp:242-246)
  if (threads > 0)
    omp_set_num_threads(threads);
  std::memset(out, 0, sizeof(*out));
  Seqs seqs = split(nodes, lengths, n_trav);
  if (delta_round < 0)
    delta_round = 0;  // compression_workflow.cpp:189-193
  // apply_delta_transform (compression_workflow.cpp:235-265)
  uint32_t max_abs = 0;
  if (delta_round == 0)
    max_abs = max_abs_symbol(seqs);
  for (int i = 0; i < delta_round; ++i)
    max_abs = std::max(max_abs, delta_transform_and_max_abs(seqs));
  uint32_t next_id = 0;
  if (max_abs >= next_id) {
    if (max_abs == UINT32_MAX)
      return -1;
    next_id = max_abs + 1;
  }
  const uint32_t layer_start = next_id;
  std::vector<int64_t> rulebook;
  int err = 0;
  int ran = 0;
  for (int r = 0; r < rounds; ++r) {
    const uint64_t used = grammar_round(seqs, next_id, rulebook, &err);
    if (err)
      return err;
    if (used == 0)
      break;
    if (ran < 64)
      out->rules_per_round[ran] = used;
    ++ran;
  }
  out->rounds_run = ran;
  out->start_id = layer_start;
  out->end_id = next_id;
  out->n_trav = n_trav;
  uint64_t total = 0;
  for (const Seq &s : seqs)
    total += s.size();
  out->n_encoded = total;
  out->encoded = (int32_t *)std::malloc(std::max<uint64_t>(1, total) * 4);
  out->encoded_lengths = (uint32_t *)std::malloc(std::max<uint64_t>(1, n_trav) * 4);
  out->n_rules = rulebook.size();
  out->rules_first = (int32_t *)std::malloc(std::max<size_t>(1, rulebook.size()) * 4);
  out->rules_second = (int32_t *)std::malloc(std::max<size_t>(1, rulebook.size()) * 4);
  uint64_t o = 0;
  for (uint64_t t = 0; t < n_trav; ++t) {
    std::memcpy(out->encoded + o, seqs[t].data(), seqs[t].size() * 4);
    out->encoded_lengths[t] = (uint32_t)seqs[t].size();
    o += seqs[t].size();
  }
  // process_rules (src/workflows/compression_utils.cpp:165-193) before its
  // column delta: first/second in rule-ID order.
  for (size_t i = 0; i < rulebook.size(); ++i) {
    out->rules_first[i] = unpack_first(rulebook[i]);
    out->rules_second[i] = unpack_second(rulebook[i]);
  }
  return 0;
}

void oracle_free_result(oracle_compress_result *r) {
  std::free(r->encoded);
  std::free(r->encoded_lengths);
  std::free(r->rules_first);
  std::free(r->rules_second);
  std::memset(r, 0, sizeof(*r));
}

int oracle_decompress(const int32_t *encoded, uint64_t n_encoded,
                      const uint32_t *encoded_lengths, uint64_t n_trav,
                      const int32_t *rules_first, const int32_t *rules_second,
                      uint64_t n_rules, uint32_t min_rule_id, int delta_round,
                      int threads, int32_t *out, uint64_t out_cap, uint64_t *out_n) {
  (void)n_encoded;
  if (threads > 0)
    omp_set_num_threads(threads);
  std::vector<uint64_t> eoff(n_trav + 1, 0);
  for (uint64_t t = 0; t < n_trav; ++t)
    eoff[t + 1] = eoff[t] + encoded_lengths[t];
  const uint32_t max_id = min_rule_id + (uint32_t)n_rules;
  Seqs dec(n_trav);
  // expand_sequences (decompression_workflow.cpp:110-154)
#pragma omp parallel for schedule(dynamic)
  for (int64_t t = 0; t < (int64_t)n_trav; ++t) {
    Seq &o = dec[t];
    for (uint64_t i = eoff[t]; i < eoff[t + 1]; ++i) {
      const int32_t node = encoded[i];
      const uint32_t a = abs_u32(node);
      if (n_rules && a >= min_rule_id && a < max_id)
        expand_rule(a, node < 0, rules_first, rules_second, min_rule_id, max_id, o);
      else
        o.push_back(node);
    }
    for (int r = 0; r < delta_round; ++r)
      inverse_delta(o);
  }
  uint64_t total = 0;
  for (const Seq &s : dec)
    total += s.size();
  *out_n = total;
  if (total > out_cap)
    return -1;
  uint64_t o = 0;
  for (const Seq &s : dec) {
    std::memcpy(out + o, s.data(), s.size() * 4);
    o += s.size();
  }
  return 0;
}

int oracle_encode_existing(const int32_t *nodes, uint64_t n_nodes, const uint32_t *lengths,
                           uint64_t n_trav, const int32_t *rules_first,
                           const int32_t *rules_second, uint64_t n_rules,
                           const uint32_t *layer_start, const uint32_t *layer_end,
                           const uint64_t *layer_row_offset, uint64_t n_layers,
                           int delta_round, int32_t *out, uint32_t *out_lengths,
                           uint64_t *out_n) {
  (void)n_nodes;
  Seqs seqs = split(nodes, lengths, n_trav);
  // apply_delta_and_validate (add_haplotypes_workflow.cpp:447-465)
  uint32_t max_abs = 0;
  for (int i = 0; i < delta_round; ++i)
    max_abs = std::max(max_abs, delta_transform_and_max_abs(seqs));
  const uint32_t min_rule_id = n_layers ? layer_start[0] : 0;
  if (n_layers && max_abs >= min_rule_id)
    return -1;
  // build_layered_rules + encode_with_existing_rules (:336-379)
  for (uint64_t l = 0; l < n_layers; ++l) {
    std::unordered_map<int64_t, uint32_t> table;
    const uint32_t count = layer_end[l] - layer_start[l];
    for (uint32_t i = 0; i < count; ++i) {
      const uint64_t idx = layer_row_offset[l] + i;
      if (idx >= n_rules)
        return -2;  // "rule column is truncated"
      table.emplace(pack_2mer(rules_first[idx], rules_second[idx]), layer_start[l] + i);
    }
    encode_with_table(seqs, table);
  }
  uint64_t o = 0;
  for (uint64_t t = 0; t < n_trav; ++t) {
    std::memcpy(out + o, seqs[t].data(), seqs[t].size() * 4);
    out_lengths[t] = (uint32_t)seqs[t].size();
    o += seqs[t].size();
  }
  *out_n = o;
  return 0;
}

}  // extern "C"
