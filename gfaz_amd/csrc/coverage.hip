// Node coverage of decoded traversals on gfx950: which groups of traversals
// (haplotypes) visit which nodes, for pangenome growth and presence/absence
// variation.
//
// The reference computes both on the host from traversals it decodes itself:
// compute_growth (src/workflows/growth_workflow.cpp:261-503) counts, per node,
// the groups that visit it (a per-thread stamp array) and histograms the counts;
// compute_pav (src/workflows/pav_workflow.cpp:475-760) sums, per BED window and
// group, the lengths of the window's nodes that the group visits. Here both start
// from one membership bitmap per group (bit a of row g: group g visits node a):
//   * marking: each thread takes 16 consecutive steps and ORs its bits into the
//     row of the step's group, one atomic per run of steps in the same 32-node
//     word (traversals mostly walk nodes in ID order);
//   * coverage counts: one thread per 32-node word adds the groups' words as
//     bit-sliced counters and bins each node's count in an LDS histogram;
//   * windows: one block per window, a thread per group summing the lengths of
//     the window's nodes in its row.
// Node IDs are |step| as uint32; 0 and IDs above num_nodes are skipped, as both
// reference workflows do.
#include "codec.hpp"
#include "coverage.hpp"
#include "prims.hpp"

#include <algorithm>

namespace gfz {

namespace {

constexpr uint32_t kCovItems = 16;
constexpr uint32_t kCovTile = kBlock * kCovItems;
constexpr uint32_t kPlanes = 16;          // bit-sliced counters: up to 2^16 - 1 groups
constexpr uint32_t kLdsHistBins = 8192;   // LDS histogram when groups + 1 fit
constexpr uint32_t kWinChunk = 2048;      // window nodes staged in LDS at a time

__global__ __launch_bounds__(kBlock) void k_cov_mark(const int32_t *__restrict__ x, uint64_t n,
                                                     const uint64_t *__restrict__ tstart,
                                                     uint64_t n_trav,
                                                     const uint32_t *__restrict__ group_of,
                                                     uint32_t num_nodes, uint64_t words,
                                                     uint32_t *__restrict__ bits) {
  const uint64_t i0 = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * kCovItems;
  if (i0 >= n)
    return;
  // Traversal of step i0: last t with tstart[t] <= i0 (tstart[n_trav] = n > i0).
  uint64_t lo = 0, hi = n_trav;
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) / 2;
    if (tstart[mid] <= i0)
      lo = mid;
    else
      hi = mid;
  }
  uint64_t t = lo, next = tstart[t + 1];
  uint32_t *row = bits + (uint64_t)group_of[t] * words;
  uint64_t cur_w = ~0ull;
  uint32_t mask = 0;
  const uint64_t i1 = i0 + kCovItems < n ? i0 + kCovItems : n;
  for (uint64_t i = i0; i < i1; ++i) {
    if (i >= next) {  // a new traversal (skipping empty ones): flush, change row
      if (mask)
        atomicOr(&row[cur_w], mask);
      mask = 0;
      cur_w = ~0ull;
      while (tstart[t + 1] <= i)
        ++t;
      next = tstart[t + 1];
      row = bits + (uint64_t)group_of[t] * words;
    }
    const uint32_t a = uabs(x[i]);
    if (a == 0 || a > num_nodes)
      continue;
    const uint64_t w = a >> 5;
    if (w != cur_w) {
      if (mask)
        atomicOr(&row[cur_w], mask);
      cur_w = w;
      mask = 0;
    }
    mask |= 1u << (a & 31);
  }
  if (mask)
    atomicOr(&row[cur_w], mask);
}

// hist[c] += nodes of [1, num_nodes] visited by exactly c groups.
__global__ __launch_bounds__(kBlock) void k_cov_hist(const uint32_t *__restrict__ bits,
                                                     uint32_t n_groups, uint64_t words,
                                                     uint32_t num_nodes,
                                                     unsigned long long *__restrict__ hist) {
  __shared__ uint32_t sh[kLdsHistBins];
  const bool in_lds = n_groups + 1 <= kLdsHistBins;
  if (in_lds) {
    for (uint32_t i = threadIdx.x; i <= n_groups; i += kBlock)
      sh[i] = 0;
    __syncthreads();
  }
  const uint64_t w = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (w < words) {
    uint32_t plane[kPlanes];
#pragma unroll
    for (uint32_t k = 0; k < kPlanes; ++k)
      plane[k] = 0;
    for (uint32_t g = 0; g < n_groups; ++g) {
      uint32_t carry = bits[(uint64_t)g * words + w];
#pragma unroll
      for (uint32_t k = 0; k < kPlanes; ++k) {
        const uint32_t c2 = plane[k] & carry;
        plane[k] ^= carry;
        carry = c2;
      }
    }
    for (uint32_t b = 0; b < 32; ++b) {
      const uint64_t node = w * 32 + b;
      if (node == 0 || node > num_nodes)
        continue;
      uint32_t c = 0;
#pragma unroll
      for (uint32_t k = 0; k < kPlanes; ++k)
        c |= ((plane[k] >> b) & 1u) << k;
      if (in_lds)
        atomicAdd(&sh[c], 1u);
      else
        atomicAdd(&hist[c], 1ull);
    }
  }
  if (in_lds) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i <= n_groups; i += kBlock)
      if (sh[i])
        atomicAdd(&hist[i], (unsigned long long)sh[i]);
  }
}

// numerators[w * n_groups + g] = sum of lens of window w's nodes that group g
// visits; window w's nodes are nodes[off[w] .. off[w + 1]), lens alongside.
__global__ __launch_bounds__(kBlock) void k_cov_windows(const uint32_t *__restrict__ bits,
                                                        uint32_t n_groups, uint64_t words,
                                                        const uint64_t *__restrict__ off,
                                                        const uint32_t *__restrict__ nodes,
                                                        const uint32_t *__restrict__ lens,
                                                        unsigned long long *__restrict__ num) {
  __shared__ uint32_t sn[kWinChunk], sl[kWinChunk];
  const uint64_t win = blockIdx.x;
  const uint64_t b = off[win], e = off[win + 1];
  for (uint32_t g0 = 0; g0 < n_groups; g0 += kBlock) {
    const uint32_t g = g0 + threadIdx.x;
    const uint32_t *row = bits + (uint64_t)(g < n_groups ? g : 0) * words;
    unsigned long long sum = 0;
    for (uint64_t c0 = b; c0 < e; c0 += kWinChunk) {
      const uint32_t m = (uint32_t)(e - c0 < kWinChunk ? e - c0 : kWinChunk);
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < m; i += kBlock) {
        sn[i] = nodes[c0 + i];
        sl[i] = lens[c0 + i];
      }
      __syncthreads();
      if (g < n_groups)
        for (uint32_t i = 0; i < m; ++i) {
          const uint32_t a = sn[i];
          if ((row[a >> 5] >> (a & 31)) & 1u)
            sum += sl[i];
        }
    }
    if (g < n_groups)
      num[win * n_groups + g] = sum;
  }
}

}  // namespace

void coverage_build(Ctx &ctx, const int32_t *d_nodes, uint64_t n, const uint32_t *lengths,
                    uint64_t n_trav, const uint32_t *group_of, uint32_t n_groups,
                    uint32_t num_nodes, Coverage &cov) {
  if (n_groups == 0 || n_groups >= (1u << kPlanes))
    throw Error(kErrInvalid, "groups must number 1 .. " + std::to_string((1u << kPlanes) - 1));
  std::vector<uint64_t> tstart(n_trav + 1, 0);
  for (uint64_t t = 0; t < n_trav; ++t) {
    if (group_of[t] >= n_groups)
      throw Error(kErrInvalid, "traversal " + std::to_string(t) + " names group " +
                                   std::to_string(group_of[t]) + " of " +
                                   std::to_string(n_groups));
    tstart[t + 1] = tstart[t] + lengths[t];
  }
  if (tstart[n_trav] != n)
    throw Error(kErrInvalid, "sum of traversal lengths (" + std::to_string(tstart[n_trav]) +
                                 ") != number of steps (" + std::to_string(n) + ")");
  Pool &pool = ctx.pool();
  cov = Coverage();
  cov.n_groups = n_groups;
  cov.num_nodes = num_nodes;
  cov.words = (uint64_t)num_nodes / 32 + 1;  // bits 0 .. num_nodes
  cov.bits = DBuf<uint32_t>(pool, (uint64_t)n_groups * cov.words);
  fill_u32(ctx, cov.bits.get(), 0, (uint64_t)n_groups * cov.words);
  if (n && n_trav) {
    DBuf<uint64_t> d_tstart(pool, n_trav + 1);
    DBuf<uint32_t> d_group(pool, n_trav);
    GFZ_HIP(hipMemcpyAsync(d_tstart.get(), tstart.data(), (n_trav + 1) * 8,
                           hipMemcpyHostToDevice, ctx.stream()));
    GFZ_HIP(hipMemcpyAsync(d_group.get(), group_of, n_trav * 4, hipMemcpyHostToDevice,
                           ctx.stream()));
    ctx.timed("coverage", [&] {
      hipLaunchKernelGGL(k_cov_mark, dim3((uint32_t)ceil_div(n, (uint64_t)kCovTile)),
                         dim3(kBlock), 0, ctx.stream(), d_nodes, n, d_tstart.get(), n_trav,
                         d_group.get(), num_nodes, cov.words, cov.bits.get());
    });
    GFZ_CHECK_LAUNCH();
    ctx.add_bytes("coverage", 4.0 * n);
    ctx.sync();  // the pageable tables go out of scope
  }
}

void coverage_hist(Ctx &ctx, const Coverage &cov, uint64_t *hist) {
  DBuf<unsigned long long> d_hist(ctx.pool(), (uint64_t)cov.n_groups + 1);
  fill_u64(ctx, reinterpret_cast<uint64_t *>(d_hist.get()), 0, (uint64_t)cov.n_groups + 1);
  hipLaunchKernelGGL(k_cov_hist, dim3((uint32_t)ceil_div(cov.words, (uint64_t)kBlock)),
                     dim3(kBlock), 0, ctx.stream(), cov.bits.get(), cov.n_groups, cov.words,
                     cov.num_nodes, d_hist.get());
  GFZ_CHECK_LAUNCH();
  GFZ_HIP(hipMemcpyAsync(hist, d_hist.get(), ((uint64_t)cov.n_groups + 1) * 8,
                         hipMemcpyDeviceToHost, ctx.stream()));
  ctx.sync();
}

void coverage_windows(Ctx &ctx, const Coverage &cov, const uint64_t *win_off,
                      const uint32_t *win_nodes, uint64_t n_windows,
                      const uint32_t *segment_lengths, uint64_t *denominators,
                      uint64_t *numerators) {
  const uint64_t total = win_off[n_windows];
  std::vector<uint32_t> lens(std::max<uint64_t>(total, 1));
  for (uint64_t w = 0; w < n_windows; ++w) {
    if (win_off[w + 1] < win_off[w])
      throw Error(kErrInvalid, "window offsets must be non-decreasing");
    uint64_t den = 0;
    for (uint64_t k = win_off[w]; k < win_off[w + 1]; ++k) {
      const uint32_t a = win_nodes[k];
      if (a == 0 || a > cov.num_nodes)
        throw Error(kErrInvalid, "window node " + std::to_string(a) + " outside 1 .. " +
                                     std::to_string(cov.num_nodes));
      lens[k] = segment_lengths[a - 1];
      den += lens[k];
    }
    if (denominators)
      denominators[w] = den;
  }
  if (!n_windows)
    return;
  Pool &pool = ctx.pool();
  DBuf<uint64_t> d_off(pool, n_windows + 1);
  DBuf<uint32_t> d_nodes(pool, std::max<uint64_t>(total, 1)), d_lens(pool, std::max<uint64_t>(total, 1));
  DBuf<unsigned long long> d_num(pool, n_windows * cov.n_groups);
  GFZ_HIP(hipMemcpyAsync(d_off.get(), win_off, (n_windows + 1) * 8, hipMemcpyHostToDevice,
                         ctx.stream()));
  if (total) {
    GFZ_HIP(hipMemcpyAsync(d_nodes.get(), win_nodes, total * 4, hipMemcpyHostToDevice,
                           ctx.stream()));
    GFZ_HIP(hipMemcpyAsync(d_lens.get(), lens.data(), total * 4, hipMemcpyHostToDevice,
                           ctx.stream()));
  }
  hipLaunchKernelGGL(k_cov_windows, dim3((uint32_t)n_windows), dim3(kBlock), 0, ctx.stream(),
                     cov.bits.get(), cov.n_groups, cov.words, d_off.get(), d_nodes.get(),
                     d_lens.get(), d_num.get());
  GFZ_CHECK_LAUNCH();
  GFZ_HIP(hipMemcpyAsync(numerators, d_num.get(), n_windows * cov.n_groups * 8,
                         hipMemcpyDeviceToHost, ctx.stream()));
  ctx.sync();
}

}  // namespace gfz
