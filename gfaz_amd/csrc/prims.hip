#include "prims.hpp"

namespace gfz {

namespace {

constexpr uint32_t kScanItems = 16;
constexpr uint32_t kScanTile = kBlock * kScanItems;
constexpr uint32_t kPartialThreads = 1024;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const T *__restrict__ in, uint64_t n,
                                                        uint64_t *__restrict__ partial) {
  __shared__ uint64_t lds[kBlock / kWave + 1];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint64_t s = 0;
#pragma unroll
  for (uint32_t j = 0; j < kScanItems; ++j) {
    const uint64_t i = base + j;
    if (i < n)
      s += (uint64_t)in[i];
  }
  uint64_t total;
  block_exclusive_scan<kBlock>(s, lds, total);
  if (threadIdx.x == 0)
    partial[blockIdx.x] = total;
}

// In-place exclusive scan of partial[0..nb); partial[nb] = total. One block.
__global__ __launch_bounds__(kPartialThreads) void k_scan_partials(uint64_t *partial,
                                                                    uint64_t nb) {
  __shared__ uint64_t lds[kPartialThreads / kWave + 1];
  const uint64_t chunk = (nb + kPartialThreads - 1) / kPartialThreads;
  const uint64_t begin = (uint64_t)threadIdx.x * chunk;
  const uint64_t end = begin + chunk < nb ? begin + chunk : nb;
  uint64_t s = 0;
  for (uint64_t i = begin; i < end; ++i)
    s += partial[i];
  uint64_t total;
  uint64_t run = block_exclusive_scan<kPartialThreads>(s, lds, total);
  for (uint64_t i = begin; i < end; ++i) {
    const uint64_t t = partial[i];
    partial[i] = run;
    run += t;
  }
  if (threadIdx.x == 0)
    partial[nb] = total;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_apply(const T *__restrict__ in, uint64_t n,
                                                       const uint64_t *__restrict__ partial,
                                                       uint64_t *__restrict__ out) {
  __shared__ uint64_t lds[kBlock / kWave + 1];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint64_t v[kScanItems];
  uint64_t s = 0;
#pragma unroll
  for (uint32_t j = 0; j < kScanItems; ++j) {
    const uint64_t i = base + j;
    v[j] = i < n ? (uint64_t)in[i] : 0;
    s += v[j];
  }
  uint64_t total;
  uint64_t run = block_exclusive_scan<kBlock>(s, lds, total) + partial[blockIdx.x];
#pragma unroll
  for (uint32_t j = 0; j < kScanItems; ++j) {
    const uint64_t i = base + j;
    if (i < n)
      out[i] = run;
    run += v[j];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    out[n] = partial[gridDim.x];
}

template <typename T>
void exclusive_scan_impl(Ctx &ctx, const T *in, uint64_t *out, uint64_t n) {
  if (n == 0) {
    fill_u64(ctx, out, 0, 1);
    return;
  }
  const uint64_t nb = ceil_div(n, kScanTile);
  DBuf<uint64_t> partial(ctx.pool(), nb + 1);
  hipLaunchKernelGGL(k_scan_reduce<T>, dim3((uint32_t)nb), dim3(kBlock), 0, ctx.stream(), in, n,
                     partial.get());
  GFZ_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kPartialThreads), 0, ctx.stream(),
                     partial.get(), nb);
  GFZ_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_apply<T>, dim3((uint32_t)nb), dim3(kBlock), 0, ctx.stream(), in, n,
                     partial.get(), out);
  GFZ_CHECK_LAUNCH();
}

// ---- radix sort -------------------------------------------------------------
// Keys per thread. Sorts of at most kInlineOffsetBlocks tiles let each block
// sum the per-block histograms itself (2 launches per digit instead of 5).
// Larger tiles (64 keys per thread) to reach that threshold were measured
// slower: too few blocks and four barriers per key step.
constexpr uint32_t kSortItemsLarge = 16;
constexpr uint32_t kSortItemsSmall = 16;
constexpr uint64_t kSign = 0x8000000000000000ull;
constexpr uint32_t kInlineOffsetBlocks = 16;

__global__ __launch_bounds__(kBlock) void k_or_diff(const uint64_t *__restrict__ keys, uint64_t n,
                                                    unsigned long long *acc) {
  const uint64_t first = keys[0];
  uint64_t x = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    x |= keys[i] ^ first;
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1)
    x |= __shfl_xor(x, d);
  if ((threadIdx.x & (kWave - 1)) == 0 && x)
    atomicOr(acc, (unsigned long long)x);
}

template <uint32_t kSortItems>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint64_t *__restrict__ keys,
                                                       uint64_t n, uint32_t shift,
                                                       uint32_t *__restrict__ hist, uint32_t nb) {
  constexpr uint32_t kSortTile = kBlock * kSortItems;
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (uint32_t j = 0; j < kSortItems; ++j) {
    const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
    if (i < n)
      atomicAdd(&h[((keys[i] ^ kSign) >> shift) & 255], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: keys keep their input order within a digit. Ranks inside a
// wave come from eight ballots (lanes whose digit agrees on every bit).
// With offs == nullptr the block derives its digit offsets from the per-block
// histograms itself (cheap while nb is small), saving the three scan launches.
template <uint32_t kSortItems>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint64_t *__restrict__ in,
                                                          uint64_t *__restrict__ out, uint64_t n,
                                                          uint32_t shift,
                                                          const uint64_t *__restrict__ offs,
                                                          const uint32_t *__restrict__ hist,
                                                          uint32_t nb) {
  constexpr uint32_t kSortTile = kBlock * kSortItems;
  __shared__ uint64_t base_off[256];
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[kBlock / kWave][256];
  __shared__ uint64_t lds[kBlock / kWave + 1];
  const uint32_t tid = threadIdx.x, wid = tid / kWave;
  if (offs) {
    base_off[tid] = offs[(uint64_t)tid * nb + blockIdx.x];
  } else {
    uint64_t tot = 0, pre = 0;
    for (uint32_t b = 0; b < nb; ++b) {
      const uint32_t h = hist[(uint64_t)tid * nb + b];
      tot += h;
      if (b < blockIdx.x)
        pre += h;
    }
    uint64_t total;
    base_off[tid] = block_exclusive_scan<kBlock>(tot, lds, total) + pre;
  }
  run[tid] = 0;
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
  for (uint32_t j = 0; j < kSortItems; ++j) {
    const uint64_t i = base + (uint64_t)j * kBlock + tid;
    const bool valid = i < n;
    const uint64_t k = valid ? in[i] : 0;
    const uint32_t d = (uint32_t)(((k ^ kSign) >> shift) & 255);
#pragma unroll
    for (uint32_t w = 0; w < kBlock / kWave; ++w)
      wcnt[w][tid] = 0;
    __syncthreads();
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint64_t below = peers & lanemask_lt();
    const uint32_t rank = (uint32_t)__popcll(below);
    if (valid && below == 0)
      wcnt[wid][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
      uint32_t r = run[tid];
#pragma unroll
      for (uint32_t w = 0; w < kBlock / kWave; ++w) {
        const uint32_t t = wcnt[w][tid];
        wcnt[w][tid] = r;
        r += t;
      }
      run[tid] = r;
    }
    __syncthreads();
    if (valid)
      out[base_off[d] + wcnt[wid][d] + rank] = k;
    __syncthreads();
  }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u8(uint8_t *p, uint8_t v, uint64_t n) {
  // 16-byte stores for the aligned body, bytes for the tail.
  const uint64_t nv = n / 16;
  uint4 *q = reinterpret_cast<uint4 *>(p);
  const uint32_t w = v * 0x01010101u;
  const uint4 vv = make_uint4(w, w, w, w);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += (uint64_t)gridDim.x * blockDim.x)
    q[i] = vv;
  for (uint64_t i = nv * 16 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_set_heads(const uint64_t *__restrict__ off, uint64_t n_off, uint32_t *heads,
                            uint64_t n_pos) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_off;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t o = off[t];
    if (o < n_pos)
      atomicOr(&heads[o >> 5], 1u << (o & 31));
  }
}

__global__ void k_bit_counts(const uint32_t *__restrict__ bits, uint64_t n_bits,
                             uint32_t *__restrict__ cnt, uint64_t nw) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t v = bits[w];
    const uint64_t lo = w * 32;
    if (lo + 32 > n_bits)
      v &= (uint32_t)((1ull << (n_bits - lo)) - 1);
    cnt[w] = (uint32_t)__popc(v);
  }
}

__global__ void k_bit_positions(const uint32_t *__restrict__ bits, uint64_t n_bits,
                                const uint64_t *__restrict__ rank, uint64_t nw,
                                uint64_t *__restrict__ pos, uint64_t pos_cap) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t v = bits[w];
    const uint64_t lo = w * 32;
    if (lo + 32 > n_bits)
      v &= (uint32_t)((1ull << (n_bits - lo)) - 1);
    uint64_t r = rank[w];
    while (v && r < pos_cap) {
      const uint32_t b = (uint32_t)__ffs(v) - 1;
      pos[r++] = lo + b;
      v &= v - 1;
    }
  }
}

uint32_t grid_for(Ctx &ctx, uint64_t n, uint32_t per_block) {
  const uint64_t want = ceil_div(n, per_block);
  const uint64_t cap = (uint64_t)ctx.num_cus() * 16;
  return (uint32_t)std::max<uint64_t>(1, std::min(want, cap));
}

}  // namespace

void exclusive_scan(Ctx &ctx, const uint32_t *in, uint64_t *out, uint64_t n) {
  exclusive_scan_impl<uint32_t>(ctx, in, out, n);
}
void exclusive_scan(Ctx &ctx, const uint64_t *in, uint64_t *out, uint64_t n) {
  exclusive_scan_impl<uint64_t>(ctx, in, out, n);
}

void fill_u32(Ctx &ctx, uint32_t *p, uint32_t v, uint64_t n) {
  if (!n)
    return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, n, kBlock * 4)), dim3(kBlock), 0,
                     ctx.stream(), p, v, n);
  GFZ_CHECK_LAUNCH();
}
void fill_u64(Ctx &ctx, uint64_t *p, uint64_t v, uint64_t n) {
  if (!n)
    return;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ctx, n, kBlock * 4)), dim3(kBlock), 0,
                     ctx.stream(), p, v, n);
  GFZ_CHECK_LAUNCH();
}
void fill_u8(Ctx &ctx, uint8_t *p, uint8_t v, uint64_t n) {
  if (!n)
    return;
  hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(ctx, n / 16 + 1, kBlock * 4)), dim3(kBlock), 0,
                     ctx.stream(), p, v, n);
  GFZ_CHECK_LAUNCH();
}

void set_head_bits(Ctx &ctx, const uint64_t *off, uint64_t n_off, uint32_t *heads,
                   uint64_t n_pos) {
  if (!n_off)
    return;
  hipLaunchKernelGGL(k_set_heads, dim3(grid_for(ctx, n_off, kBlock)), dim3(kBlock), 0,
                     ctx.stream(), off, n_off, heads, n_pos);
  GFZ_CHECK_LAUNCH();
}

uint64_t bitmap_positions(Ctx &ctx, const uint32_t *bits, uint64_t n_bits, uint64_t *pos,
                          uint64_t pos_cap) {
  if (!n_bits)
    return 0;
  const uint64_t nw = ceil_div(n_bits, 32);
  DBuf<uint32_t> cnt(ctx.pool(), nw);
  DBuf<uint64_t> rank(ctx.pool(), nw + 1);
  hipLaunchKernelGGL(k_bit_counts, dim3(grid_for(ctx, nw, kBlock)), dim3(kBlock), 0, ctx.stream(),
                     bits, n_bits, cnt.get(), nw);
  GFZ_CHECK_LAUNCH();
  exclusive_scan(ctx, cnt.get(), rank.get(), nw);
  hipLaunchKernelGGL(k_bit_positions, dim3(grid_for(ctx, nw, kBlock)), dim3(kBlock), 0,
                     ctx.stream(), bits, n_bits, rank.get(), nw, pos, pos_cap);
  GFZ_CHECK_LAUNCH();
  uint64_t total = 0;
  GFZ_HIP(hipMemcpyAsync(&total, rank.get() + nw, sizeof(uint64_t), hipMemcpyDeviceToHost,
                         ctx.stream()));
  ctx.sync();
  return total;
}

void sort_i64(Ctx &ctx, uint64_t *keys, uint64_t *tmp, uint64_t n) {
  if (n < 2)
    return;
  DBuf<unsigned long long> acc(ctx.pool(), 1);
  fill_u64(ctx, reinterpret_cast<uint64_t *>(acc.get()), 0, 1);
  hipLaunchKernelGGL(k_or_diff, dim3(grid_for(ctx, n, kBlock * 8)), dim3(kBlock), 0,
                     ctx.stream(), keys, n, acc.get());
  GFZ_CHECK_LAUNCH();
  ctx.stage_scalars(acc.get(), 1);
  ctx.sync();
  sort_i64_varying(ctx, keys, tmp, n, ctx.scalars()[0]);
}

template <uint32_t kSortItems>
void radix_passes(Ctx &ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t varying) {
  constexpr uint32_t kSortTile = kBlock * kSortItems;
  const uint32_t nb = (uint32_t)ceil_div(n, kSortTile);
  DBuf<uint32_t> hist(ctx.pool(), (uint64_t)256 * nb);
  DBuf<uint64_t> offs(ctx.pool(), (uint64_t)256 * nb + 1);
  uint64_t *src = keys, *dst = tmp;
  for (uint32_t shift = 0; shift < 64; shift += 8) {
    if (((varying >> shift) & 255) == 0)
      continue;
    ctx.timed("radix_sort", [&] {
      hipLaunchKernelGGL(k_radix_hist<kSortItems>, dim3(nb), dim3(kBlock), 0, ctx.stream(), src,
                         n, shift, hist.get(), nb);
      GFZ_CHECK_LAUNCH();
      const bool inline_offsets = nb <= kInlineOffsetBlocks;
      if (!inline_offsets)
        exclusive_scan(ctx, hist.get(), offs.get(), (uint64_t)256 * nb);
      hipLaunchKernelGGL(k_radix_scatter<kSortItems>, dim3(nb), dim3(kBlock), 0, ctx.stream(),
                         src, dst, n, shift,
                         inline_offsets ? (const uint64_t *)nullptr : offs.get(), hist.get(), nb);
      GFZ_CHECK_LAUNCH();
    });
    std::swap(src, dst);
  }
  if (src != keys)
    GFZ_HIP(hipMemcpyAsync(keys, src, n * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                           ctx.stream()));
}

void sort_i64_varying(Ctx &ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t varying) {
  if (n < 2 || !varying)
    return;
  if (n <= (uint64_t)kInlineOffsetBlocks * kBlock * kSortItemsSmall)
    radix_passes<kSortItemsSmall>(ctx, keys, tmp, n, varying);
  else
    radix_passes<kSortItemsLarge>(ctx, keys, tmp, n, varying);
}

}  // namespace gfz
