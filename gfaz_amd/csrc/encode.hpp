// Encoding new traversals with a stored rulebook (add-haplotypes).
#pragma once

#include "ctx.hpp"

#include <vector>

namespace gfz {

// One stored layer: rule IDs [start_id, end_id) are rows row_offset .. of the
// rule columns (LayerRuleRange, row = flattened_offset / k).
struct EncodeLayer {
  uint32_t start_id;
  uint32_t end_id;
  uint64_t row_offset;
};

struct EncodeResult {
  DBuf<int32_t> encoded;        // device
  uint64_t n_encoded = 0;
  std::vector<uint32_t> lengths;  // encoded length per traversal (host)
};

// The per-layer lookup tables pack(first, second) -> rule ID of a stored
// rulebook, built once and reused by every encode against it.
struct EncodeTables {
  std::vector<EncodeLayer> layers;
  std::vector<DBuf<uint8_t>> slots;  // per layer: cap 16-byte slots (empty when no rows)
  std::vector<uint64_t> caps;
};

// d_first, d_second: device; layers: host. Checks the layers against n_rules.
void encode_tables_build(Ctx &ctx, const int32_t *d_first, const int32_t *d_second,
                         uint64_t n_rules, const EncodeLayer *layers, uint64_t n_layers,
                         EncodeTables &out);

// Encodes with tables built by encode_tables_build (d_nodes device, lengths host).
void encode_with_tables(Ctx &ctx, const int32_t *d_nodes, uint64_t n, const uint32_t *lengths,
                        uint64_t n_trav, const EncodeTables &tables, int delta_round,
                        EncodeResult &out);

// Both steps in one call. d_nodes, d_first, d_second: device; lengths, layers: host.
void encode_with_rules(Ctx &ctx, const int32_t *d_nodes, uint64_t n, const uint32_t *lengths,
                       uint64_t n_trav, const int32_t *d_first, const int32_t *d_second,
                       uint64_t n_rules, const EncodeLayer *layers, uint64_t n_layers,
                       int delta_round, EncodeResult &out);

}  // namespace gfz
