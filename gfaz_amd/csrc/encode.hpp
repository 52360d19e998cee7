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

// d_nodes, d_first, d_second: device; lengths, layers: host.
void encode_with_rules(Ctx &ctx, const int32_t *d_nodes, uint64_t n, const uint32_t *lengths,
                       uint64_t n_trav, const int32_t *d_first, const int32_t *d_second,
                       uint64_t n_rules, const EncodeLayer *layers, uint64_t n_layers,
                       int delta_round, EncodeResult &out);

}  // namespace gfz
