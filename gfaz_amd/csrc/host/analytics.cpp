// Host steps of the growth and PAV analytics around the GPU coverage counts
// (coverage.hip); plain C++ compiled by the host compiler.
#include "gfaz_amd.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

#include "abi_error.hpp"

namespace {

// growth_workflow.cpp frac_covered_union: the chance that a random size-k subset
// of N groups contains at least one of the c groups covering a node.
double frac_covered_union(uint32_t N, uint32_t c, uint32_t k) {
  if (c == 0 || k == 0)
    return 0.0;
  if (c >= N)
    return 1.0;
  if (N - c < k)
    return 1.0;
  double ratio = 1.0;
  for (uint32_t i = 0; i < c; ++i)
    ratio *= static_cast<double>(N - k - i) / static_cast<double>(N - i);
  return 1.0 - ratio;
}

}  // namespace

extern "C" {

int gfz_growth_curve(const uint64_t *hist, uint32_t n_groups, double *growth) {
  if (!hist || !growth)
    return gfz::fail(GFZ_E_INVALID, "Growth workflow error: null argument");
  // compute_growth's closed form (src/workflows/growth_workflow.cpp:486-500), the
  // same operations in the same order, so the doubles are the reference's.
  growth[0] = 0.0;
  for (uint32_t k = 1; k <= n_groups; ++k) {
    double sum = 0.0;
    for (uint32_t c = 1; c <= n_groups; ++c) {
      const uint64_t hc = hist[c];
      if (hc == 0)
        continue;
      sum += static_cast<double>(hc) * frac_covered_union(n_groups, c, k);
    }
    growth[k] = sum;
  }
  return GFZ_OK;
}

int gfz_pav_window_nodes(const int32_t *ref_nodes, uint64_t n_ref,
                         const uint32_t *segment_lengths, uint32_t num_nodes,
                         const uint64_t *starts, const uint64_t *ends, uint64_t n_ranges,
                         uint64_t *window_off, uint32_t *window_nodes, uint64_t nodes_cap) {
  if ((n_ref && !ref_nodes) || (num_nodes && !segment_lengths) ||
      (n_ranges && (!starts || !ends)) || !window_off)
    return gfz::fail(GFZ_E_INVALID, "PAV workflow error: null argument");
  return gfz::guarded("PAV workflow error: ", [&] {
    // compute_pav's window walk (src/workflows/pav_workflow.cpp:655-710): ranges in
    // (start, end, index) order, the reference traversal's nodes 1 .. num_nodes in
    // path order with a running coordinate; a node overlapping a range joins it
    // once (pairs sorted and made unique).
    std::vector<uint32_t> order(n_ranges);
    for (uint64_t i = 0; i < n_ranges; ++i)
      order[i] = (uint32_t)i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      return std::tie(starts[a], ends[a], a) < std::tie(starts[b], ends[b], b);
    });
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    size_t next_range = 0;
    uint64_t offset = 0;
    for (uint64_t i = 0; i < n_ref; ++i) {
      const int64_t v = ref_nodes[i];
      const uint32_t node = (uint32_t)(v < 0 ? -v : v);
      if (node == 0 || node > num_nodes)
        continue;
      const uint64_t len = segment_lengths[node - 1];
      const uint64_t node_start = offset, node_end = offset + len;
      while (next_range < order.size() && ends[order[next_range]] <= node_start)
        ++next_range;
      for (size_t j = next_range; j < order.size(); ++j) {
        const uint32_t r = order[j];
        if (starts[r] >= node_end)
          break;
        if (ends[r] > node_start && starts[r] < node_end)
          pairs.emplace_back(r, node);
      }
      offset = node_end;
    }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    std::memset(window_off, 0, (n_ranges + 1) * sizeof(uint64_t));
    for (const auto &p : pairs)
      ++window_off[p.first + 1];
    for (uint64_t w = 0; w < n_ranges; ++w)
      window_off[w + 1] += window_off[w];
    if (!window_nodes)
      return;
    if (pairs.size() > nodes_cap)
      throw gfz::Error(GFZ_E_INVALID, "node capacity " + std::to_string(nodes_cap) + " < " +
                                          std::to_string(pairs.size()) + " window nodes");
    for (size_t k = 0; k < pairs.size(); ++k)
      window_nodes[k] = pairs[k].second;
  });
}

}  // extern "C"
