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
constexpr uint64_t kSign = 0x8000000000000000ull;  // signed order as unsigned

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

__global__ __launch_bounds__(kBlock) void k_fill_runs(FillRuns r) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  for (uint32_t q = 0; q < r.count; ++q) {
    const FillRun f = r.run[q];
    if (f.width == 8) {
      uint64_t *p = static_cast<uint64_t *>(f.ptr);
      for (uint64_t i = g; i < f.n; i += stride)
        p[i] = f.value;
    } else if (f.width == 4) {
      uint32_t *p = static_cast<uint32_t *>(f.ptr);
      for (uint64_t i = g; i < f.n; i += stride)
        p[i] = (uint32_t)f.value;
    } else {
      uint8_t *p = static_cast<uint8_t *>(f.ptr);
      for (uint64_t i = g; i < f.n; i += stride)
        p[i] = (uint8_t)f.value;
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

void fill_runs(Ctx &ctx, std::initializer_list<FillRun> runs) {
  FillRuns r{};
  uint64_t most = 0;
  for (const FillRun &f : runs) {
    if (!f.n)
      continue;
    if (r.count == kMaxFillRuns)
      throw Error(kErrInvalid, "fill_runs: more than " + std::to_string(kMaxFillRuns) + " runs");
    if (f.width != 1 && f.width != 4 && f.width != 8)
      throw Error(kErrInvalid, "fill_runs: element width must be 1, 4 or 8 bytes");
    r.run[r.count++] = f;
    most = std::max(most, f.n);
  }
  if (!r.count)
    return;
  hipLaunchKernelGGL(k_fill_runs, dim3(grid_for(ctx, most, kBlock * 4)), dim3(kBlock), 0,
                     ctx.stream(), r);
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
  const uint64_t varying = ctx.scalars()[0];
  DBuf<uint32_t> err(ctx.pool(), 2);
  fill_u32(ctx, err.get(), 0, 2);
  sort_i64_varying(ctx, keys, tmp, n, varying, err.get());
  ctx.stage_scalars(err.get(), 1);
  ctx.sync();
  if ((uint32_t)ctx.scalars()[0])
    throw Error(kErrInvalid, "radix sort look-back did not complete (internal error)");
}

// ---- one-sweep LSD radix sort -------------------------------------------------
// One histogram kernel for every digit, then one kernel per digit: each block
// takes the next tile by ticket, ranks its keys stably, publishes its per-digit
// counts, and finds the counts of all earlier tiles by decoupled look-back on
// per-(tile, digit) status words:
//   bits 63..62 = 1: this tile's own count, 2: inclusive prefix up to it.
// A block only waits on tiles with smaller tickets, which belong to blocks
// that already run and publish before they wait, so the waits end; each is
// still bounded (kLookbackSpin polls) and raises an error flag if exceeded.
// A status word carries its own count, so relaxed agent-scope atomics suffice;
// acquire/release at agent scope would add L2 write-back and invalidation
// (one L2 per XCD) to every publish and poll.
// Within a tile each wave owns kOsItems * 64 consecutive keys, loaded at once;
// a key's rank among its wave's keys of the same digit comes from ballots over
// the digit bits and one LDS add per digit group, so ranking needs no block
// barrier, and one barrier turns the per-wave counts into wave offsets.
constexpr uint32_t kOsItems = 16;
constexpr uint32_t kOsTile = kBlock * kOsItems;
constexpr uint64_t kStFlagAgg = 1ull << 62, kStFlagPre = 2ull << 62,
                   kStCount = (1ull << 62) - 1;
constexpr uint32_t kLookbackSpin = 1u << 22;
constexpr uint32_t kWindow = 16;

__global__ __launch_bounds__(kBlock) void k_sort_hist_all(const uint64_t *__restrict__ keys,
                                                          uint64_t n, uint64_t varying,
                                                          unsigned long long *__restrict__ ghist) {
  __shared__ uint32_t h[8][256];
  for (uint32_t i = threadIdx.x; i < 8 * 256; i += kBlock)
    (&h[0][0])[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kOsTile;
  uint64_t key[kOsItems];
#pragma unroll
  for (uint32_t j = 0; j < kOsItems; ++j) {
    const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
    key[j] = i < n ? keys[i] ^ kSign : 0;
  }
  for (uint32_t j = 0; j < kOsItems; ++j) {
    if (base + (uint64_t)j * kBlock + threadIdx.x >= n)
      break;
#pragma unroll
    for (uint32_t p = 0; p < 8; ++p)
      if ((varying >> (8 * p)) & 255)
        atomicAdd(&h[p][(key[j] >> (8 * p)) & 255], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 8 * 256; i += kBlock) {
    const uint32_t c = (&h[0][0])[i];
    if (c)
      atomicAdd(&ghist[i], (unsigned long long)c);
  }
}

__global__ __launch_bounds__(kBlock) void k_onesweep(const uint64_t *__restrict__ in,
                                                     uint64_t *__restrict__ out, uint64_t n,
                                                     uint32_t shift,
                                                     const unsigned long long *__restrict__ ghist,
                                                     unsigned long long *status,
                                                     unsigned long long *ticket, uint32_t *err) {
  __shared__ uint32_t wbin[kBlock / kWave][256];  // per-wave digit counts, then wave offsets
  __shared__ uint64_t base_off[256];
  __shared__ uint64_t lds[kBlock / kWave + 1];
  __shared__ unsigned long long s_tile;
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
  if (tid == 0)
    s_tile = atomicAdd(ticket, 1ull);
#pragma unroll
  for (uint32_t w = 0; w < kBlock / kWave; ++w)
    wbin[w][tid] = 0;
  __syncthreads();
  const uint64_t t = s_tile;
  const uint64_t wbase = t * kOsTile + (uint64_t)wid * (kOsItems * kWave);
  uint64_t key[kOsItems];
#pragma unroll
  for (uint32_t j = 0; j < kOsItems; ++j) {
    const uint64_t i = wbase + (uint64_t)j * kWave + lane;
    key[j] = i < n ? in[i] : 0;
  }
  // Phase A: rank among the wave's keys of the same digit (item, then lane).
  const uint64_t below = lanemask_lt();
  uint32_t rank[kOsItems];
#pragma unroll
  for (uint32_t j = 0; j < kOsItems; ++j) {
    const bool valid = wbase + (uint64_t)j * kWave + lane < n;
    const uint32_t d = (uint32_t)(((key[j] ^ kSign) >> shift) & 255);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t leader = peers ? (uint32_t)__ffsll((long long)peers) - 1 : lane;
    uint32_t old = 0;
    if (valid && lane == leader)
      old = atomicAdd(&wbin[wid][d], (uint32_t)__popcll(peers));
    old = __shfl(old, (int)leader);
    rank[j] = old + (uint32_t)__popcll(peers & below);
  }
  __syncthreads();
  uint32_t r = 0;
#pragma unroll
  for (uint32_t w = 0; w < kBlock / kWave; ++w) {
    const uint32_t c = wbin[w][tid];
    wbin[w][tid] = r;
    r += c;
  }
  // Phase B: digit tid - global start, then this tile's place among earlier tiles.
  uint64_t total;
  const uint64_t gstart = block_exclusive_scan<kBlock>((uint64_t)ghist[tid], lds, total);
  const uint64_t mine = r;
  unsigned long long *st = status + t * 256 + tid;
  uint64_t excl = 0;
  if (t == 0) {
    __hip_atomic_store(st, kStFlagPre | mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(st, kStFlagAgg | mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Windowed look-back: kWindow predecessors per round trip, nearest first.
    int64_t k = (int64_t)t - 1;  // tile 0 always publishes a prefix: k never passes it
    uint32_t spins = 0;
    for (;;) {
      uint64_t v[kWindow];
#pragma unroll
      for (uint32_t w = 0; w < kWindow; ++w)
        v[w] = k - (int64_t)w >= 0
                   ? __hip_atomic_load(status + (uint64_t)(k - (int64_t)w) * 256 + tid,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                   : kStFlagPre;  // below tile 0: nothing more to add
      bool done = false;
      uint32_t used = 0;
#pragma unroll
      for (uint32_t w = 0; w < kWindow; ++w) {
        if (done || used < w)
          continue;
        if (v[w] & kStFlagPre) {
          excl += v[w] & kStCount;
          done = true;
        } else if (v[w] & kStFlagAgg) {
          excl += v[w] & kStCount;
          used = w + 1;
        }
      }
      if (done)
        break;
      k -= used;  // resume at the first tile that has not published yet
      if (used == 0) {
        if (++spins > kLookbackSpin) {
          atomicExch(err, 4u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __hip_atomic_store(st, kStFlagPre | (excl + mine), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  base_off[tid] = gstart + excl;
  __syncthreads();
  // Phase C: scatter.
#pragma unroll
  for (uint32_t j = 0; j < kOsItems; ++j) {
    if (wbase + (uint64_t)j * kWave + lane < n) {
      const uint32_t d = (uint32_t)(((key[j] ^ kSign) >> shift) & 255);
      out[base_off[d] + wbin[wid][d] + rank[j]] = key[j];
    }
  }
}

void sort_i64_varying(Ctx &ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t varying,
                      uint32_t *err) {
  if (n < 2 || !varying)
    return;
  uint32_t passes = 0;
  for (uint32_t p = 0; p < 8; ++p)
    passes += ((varying >> (8 * p)) & 255) ? 1 : 0;
  const uint64_t tiles = ceil_div(n, kOsTile);
  // ghist (8 x 256) | tickets (8) | status (passes x tiles x 256), cleared by one fill.
  const uint64_t words = 8 * 256 + 8 + (uint64_t)passes * tiles * 256;
  DBuf<unsigned long long> scratch(ctx.pool(), words);
  unsigned long long *ghist = scratch.get(), *tickets = ghist + 8 * 256,
                     *status = tickets + 8;
  fill_u64(ctx, reinterpret_cast<uint64_t *>(scratch.get()), 0, words);
  ctx.timed("radix_sort", [&] {
    hipLaunchKernelGGL(k_sort_hist_all, dim3((uint32_t)tiles), dim3(kBlock), 0, ctx.stream(),
                       keys, n, varying, ghist);
    GFZ_CHECK_LAUNCH();
    uint64_t *src = keys, *dst = tmp;
    uint32_t q = 0;
    for (uint32_t p = 0; p < 8; ++p) {
      if (((varying >> (8 * p)) & 255) == 0)
        continue;
      hipLaunchKernelGGL(k_onesweep, dim3((uint32_t)tiles), dim3(kBlock), 0, ctx.stream(), src,
                         dst, n, 8 * p, ghist + 256 * p, status + (uint64_t)q * tiles * 256,
                         tickets + p, err);
      GFZ_CHECK_LAUNCH();
      std::swap(src, dst);
      ++q;
    }
    if (src != keys)
      GFZ_HIP(hipMemcpyAsync(keys, src, n * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                             ctx.stream()));
  });
}

}  // namespace gfz
