// Node coverage of decoded traversals by groups (pangenome growth, PAV).
#pragma once

#include "ctx.hpp"

#include <vector>

namespace gfz {

struct Coverage {
  uint32_t n_groups = 0, num_nodes = 0;
  uint64_t words = 0;     // 32-bit words per group row (bits 0 .. num_nodes)
  DBuf<uint32_t> bits;    // n_groups rows
};

// Marks, for every step of d_nodes (device; lengths and group_of host, n_trav),
// node |step| in its traversal's group row (0 and IDs above num_nodes skipped).
void coverage_build(Ctx &ctx, const int32_t *d_nodes, uint64_t n, const uint32_t *lengths,
                    uint64_t n_trav, const uint32_t *group_of, uint32_t n_groups,
                    uint32_t num_nodes, Coverage &cov);
// hist (host, n_groups + 1): nodes 1 .. num_nodes by the number of groups visiting them.
void coverage_hist(Ctx &ctx, const Coverage &cov, uint64_t *hist);
// Per window w (nodes win_nodes[win_off[w] .. win_off[w + 1]), each 1 .. num_nodes):
// denominators[w] = sum of their segment lengths (host, may be null) and
// numerators[w * n_groups + g] = the sum over the nodes group g visits (host).
void coverage_windows(Ctx &ctx, const Coverage &cov, const uint64_t *win_off,
                      const uint32_t *win_nodes, uint64_t n_windows,
                      const uint32_t *segment_lengths, uint64_t *denominators,
                      uint64_t *numerators);

}  // namespace gfz
