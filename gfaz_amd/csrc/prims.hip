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

namespace {
__global__ void k_count_mismatch(const int32_t *__restrict__ a, const int32_t *__restrict__ b,
                                 uint64_t n, unsigned long long *__restrict__ out) {
  unsigned long long c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    c += a[i] != b[i] ? 1ull : 0ull;
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1)
    c += __shfl_xor(c, d);
  if ((threadIdx.x & (kWave - 1)) == 0 && c)
    atomicAdd(out, c);
}
}  // namespace

uint64_t count_mismatch_i32(Ctx &ctx, const int32_t *a, const int32_t *b, uint64_t n) {
  if (!n)
    return 0;
  DBuf<unsigned long long> cnt(ctx.pool(), 1);
  fill_u64(ctx, reinterpret_cast<uint64_t *>(cnt.get()), 0, 1);
  hipLaunchKernelGGL(k_count_mismatch, dim3(grid_for(ctx, n, kBlock * 8)), dim3(kBlock), 0,
                     ctx.stream(), a, b, n, cnt.get());
  GFZ_CHECK_LAUNCH();
  ctx.stage_scalars(cnt.get(), 1);
  ctx.sync();
  return ctx.scalars()[0];
}

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

// ---- one-sweep LSD radix sort (inputs too large for the sample sort) ----------
namespace {
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
    key[j] = i < n ? keys[i] : 0;
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
    const uint32_t d = (uint32_t)((key[j] >> shift) & 255);
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
      const uint32_t d = (uint32_t)((key[j] >> shift) & 255);
      out[base_off[d] + wbin[wid][d] + rank[j]] = key[j];
    }
  }
}


void sort_radix_large(Ctx &ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t varying,
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
  {
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
  }
}

}  // namespace

// ---- sample sort of distinct 64-bit keys ---------------------------------------
// For keys that are all distinct (stability is moot): one block sorts a sample
// of 8 B keys (at most 4096) and takes B - 1 evenly spaced splitters;
// keys are counted per bucket, then placed bucket by bucket (each block
// reserves its share of every bucket with one atomic); finally one block per
// bucket sorts that bucket with a bitonic network held in registers and writes
// it at the bucket's offset: four launches, where an LSD radix sort of these
// keys takes a histogram and six digit passes. Buckets average about kSsAvg
// keys (eight samples each), so one beyond kSsLocalCap is a once-in-a-billion
// event; such a bucket is sorted by its block in global memory, radix by radix.
namespace {
constexpr uint32_t kSsMaxBuckets = 512;
constexpr uint32_t kSsAvg = 1024;        // target keys per bucket
constexpr uint32_t kSsLocalCap = 16 * kBlock;  // keys one block sorts in registers
constexpr uint32_t kSsSingle = 4 * kBlock;     // inputs sorted by one block, unbucketed
constexpr uint32_t kSsItems = 16;
constexpr uint32_t kSsTile = kBlock * kSsItems;

// Ascending bitonic sort of K * NT keys held K per thread (thread t holds
// indices t*K .. t*K + K - 1). Stages of stride below K compare within a
// thread's registers, strides below kWave * K pair lanes of one wave through
// shuffles, and only strides across waves go through LDS (s, K * NT keys) with
// block barriers. The stage loops unroll fully, so register indices are static.
template <int K, int NT>
__device__ inline void bitonic_regs(uint64_t (&v)[K], uint64_t *s) {
  constexpr int kN = K * NT;
  constexpr int kLog = __builtin_ctz(kN);
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (int lk = 1; lk <= kLog; ++lk) {
    const uint32_t k = 1u << lk;
#pragma unroll
    for (int lj = lk - 1; lj >= 0; --lj) {
      const uint32_t j = 1u << lj;
      if (j < (uint32_t)K) {
#pragma unroll
        for (int r = 0; r < K; ++r) {
          if (r & j)
            continue;
          const uint32_t idx = t * K + r;
          const uint64_t a = v[r], b = v[r ^ j];
          if ((a > b) == ((idx & k) == 0)) {
            v[r] = b;
            v[r ^ j] = a;
          }
        }
      } else if (j / K < (uint32_t)kWave) {
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const uint32_t idx = t * K + r;
          const uint64_t b = __shfl_xor(v[r], (int)(j / K));
          const bool keep_min = ((idx & j) == 0) == ((idx & k) == 0);
          v[r] = keep_min ? (v[r] < b ? v[r] : b) : (v[r] > b ? v[r] : b);
        }
      } else {
#pragma unroll
        for (int r = 0; r < K; ++r)
          s[t * K + r] = v[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const uint32_t idx = t * K + r;
          const uint64_t b = s[idx ^ j];
          const bool keep_min = ((idx & j) == 0) == ((idx & k) == 0);
          v[r] = keep_min ? (v[r] < b ? v[r] : b) : (v[r] > b ? v[r] : b);
        }
        __syncthreads();
      }
    }
  }
}

// Bucket of k: the number of splitters <= k (spl holds B - 1 splitters and a
// final ~0, B a power of two).
__device__ inline uint32_t ss_bucket(const uint64_t *spl, uint32_t B, uint64_t k) {
  uint32_t pos = 0;
  for (uint32_t step = B >> 1; step >= 1; step >>= 1)
    if (spl[pos + step - 1] <= k)
      pos += step;
  return pos;
}

// Splitters from a sample of 4 * NT keys (eight per bucket), sorted in registers;
// also clears the bucket counters (fill, then cursor: 2 B words).
template <int NT>
__global__ __launch_bounds__(NT) void k_ss_splitters(const uint64_t *__restrict__ keys, uint64_t n,
                                                    uint32_t B, uint64_t *__restrict__ spl,
                                                    unsigned int *__restrict__ counters) {
  constexpr int kPer = 4, kS = kPer * NT;
  __shared__ uint64_t s[kS];
  for (uint32_t j = threadIdx.x; j < 2 * B; j += NT)
    counters[j] = 0;
  uint64_t v[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r)  // n > kS: distinct positions
    v[r] = keys[(uint64_t)(threadIdx.x * kPer + r) * n / kS];
  bitonic_regs<kPer, NT>(v, s);
#pragma unroll
  for (int r = 0; r < kPer; ++r)
    s[threadIdx.x * kPer + r] = v[r];
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < B; j += NT)
    spl[j] = j + 1 < B ? s[(uint64_t)(j + 1) * kS / B] : ~0ull;
}

__global__ __launch_bounds__(kBlock) void k_ss_count(const uint64_t *__restrict__ keys, uint64_t n,
                                                     const uint64_t *__restrict__ spl, uint32_t B,
                                                     unsigned int *__restrict__ fill) {
  __shared__ uint64_t s_spl[kSsMaxBuckets];
  __shared__ uint32_t hist[kSsMaxBuckets];
  for (uint32_t b = threadIdx.x; b < B; b += kBlock) {
    s_spl[b] = spl[b];
    hist[b] = 0;
  }
  const uint64_t base = (uint64_t)blockIdx.x * kSsTile;
  uint64_t k[kSsItems];
#pragma unroll
  for (uint32_t j = 0; j < kSsItems; ++j) {
    const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
    k[j] = i < n ? keys[i] : 0;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kSsItems; ++j)
    if (base + (uint64_t)j * kBlock + threadIdx.x < n)
      atomicAdd(&hist[ss_bucket(s_spl, B, k[j])], 1u);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < B; b += kBlock)
    if (hist[b])
      atomicAdd(&fill[b], hist[b]);
}

// Exclusive prefix of fill[0, B) into s_base (B <= 2 * kBlock).
__device__ inline void ss_bases(const unsigned int *__restrict__ fill, uint32_t B,
                                uint64_t *s_base, uint64_t *lds) {
  const uint32_t b0 = 2 * threadIdx.x;
  const uint64_t c0 = b0 < B ? fill[b0] : 0, c1 = b0 + 1 < B ? fill[b0 + 1] : 0;
  uint64_t total;
  const uint64_t pre = block_exclusive_scan<kBlock>(c0 + c1, lds, total);
  if (b0 < B)
    s_base[b0] = pre;
  if (b0 + 1 < B)
    s_base[b0 + 1] = pre + c0;
}

__global__ __launch_bounds__(kBlock) void k_ss_place(const uint64_t *__restrict__ keys, uint64_t n,
                                                     const uint64_t *__restrict__ spl, uint32_t B,
                                                     const unsigned int *__restrict__ fill,
                                                     unsigned int *__restrict__ cursor,
                                                     uint64_t *__restrict__ out) {
  __shared__ uint64_t s_spl[kSsMaxBuckets];
  __shared__ uint64_t s_base[kSsMaxBuckets];
  __shared__ uint32_t hist[kSsMaxBuckets];
  __shared__ uint64_t lds[kBlock / kWave + 1];
  for (uint32_t b = threadIdx.x; b < B; b += kBlock) {
    s_spl[b] = spl[b];
    hist[b] = 0;
  }
  ss_bases(fill, B, s_base, lds);  // has block barriers
  const uint64_t base = (uint64_t)blockIdx.x * kSsTile;
  uint64_t k[kSsItems];
  uint32_t bk[kSsItems], loc[kSsItems];
#pragma unroll
  for (uint32_t j = 0; j < kSsItems; ++j) {
    const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
    k[j] = i < n ? keys[i] : 0;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kSsItems; ++j) {
    const bool v = base + (uint64_t)j * kBlock + threadIdx.x < n;
    bk[j] = v ? ss_bucket(s_spl, B, k[j]) : 0;
    loc[j] = v ? atomicAdd(&hist[bk[j]], 1u) : 0;
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < B; b += kBlock) {
    const uint32_t c = hist[b];
    hist[b] = c ? atomicAdd(&cursor[b], c) : 0u;  // this block's start in bucket b
  }
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kSsItems; ++j)
    if (base + (uint64_t)j * kBlock + threadIdx.x < n)
      out[s_base[bk[j]] + hist[bk[j]] + loc[j]] = k[j];
}

// Block-wide LSD radix sort of a[0, n) in global memory, 8 bits per pass with
// b as the other buffer; the result is copied into b. Ranks within each chunk
// of kBlock keys are stable (digit-bit ballots within the wave, then the wave
// counts before it), chunks are taken in order.
__device__ void ss_radix_global(uint64_t *a, uint64_t *b, uint64_t n) {
  __shared__ uint32_t s_cnt[256];
  __shared__ uint32_t s_w[kBlock / kWave][256];
  __shared__ uint64_t s_off[256];
  const uint32_t tid = threadIdx.x, wid = tid / kWave;
  const uint64_t below = lanemask_lt();
  uint64_t *src = a, *dst = b;
  for (uint32_t shift = 0; shift < 64; shift += 8) {
    s_cnt[tid] = 0;
    __syncthreads();
    for (uint64_t i = tid; i < n; i += kBlock)
      atomicAdd(&s_cnt[(src[i] >> shift) & 255], 1u);
    __syncthreads();
    if (tid == 0) {
      uint64_t run = 0;
      for (uint32_t d = 0; d < 256; ++d) {
        s_off[d] = run;
        run += s_cnt[d];
      }
    }
    __syncthreads();
    for (uint64_t c0 = 0; c0 < n; c0 += kBlock) {
      const uint64_t i = c0 + tid;
      const bool valid = i < n;
      const uint64_t k = valid ? src[i] : 0;
      const uint32_t d = (uint32_t)((k >> shift) & 255);
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (uint32_t bit = 0; bit < 8; ++bit) {
        const bool on = (d >> bit) & 1;
        const uint64_t m = __ballot(on);
        peers &= on ? m : ~m;
      }
#pragma unroll
      for (uint32_t w = 0; w < kBlock / kWave; ++w)
        s_w[w][tid] = 0;
      __syncthreads();
      if (valid && (peers & below) == 0)
        s_w[wid][d] = (uint32_t)__popcll(peers);
      __syncthreads();
      if (valid) {
        uint64_t pre = s_off[d];
        for (uint32_t w = 0; w < wid; ++w)
          pre += s_w[w][d];
        dst[pre + (uint64_t)__popcll(peers & below)] = k;
      }
      __syncthreads();
      uint32_t tot = 0;
#pragma unroll
      for (uint32_t w = 0; w < kBlock / kWave; ++w)
        tot += s_w[w][tid];
      s_off[tid] += tot;
      __syncthreads();
    }
    uint64_t *t = src;
    src = dst;
    dst = t;
  }
  // Eight passes end where they began (in a).
  for (uint64_t i = tid; i < n; i += kBlock)
    b[i] = a[i];
}

// nb <= K * kBlock keys from in, sorted into out (in == out allowed: every key
// is read before the first barrier of the sort).
template <int K>
__device__ inline void ss_local_sort(const uint64_t *in, uint64_t *out, uint32_t nb,
                                     uint64_t *s) {
  uint64_t v[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const uint32_t idx = threadIdx.x * K + r;
    v[r] = idx < nb ? in[idx] : ~0ull;
  }
  bitonic_regs<K, kBlock>(v, s);
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const uint32_t idx = threadIdx.x * K + r;
    if (idx < nb)
      out[idx] = v[r];
  }
}

// One block per bucket (B == 1: the whole input, n keys, in place).
__global__ __launch_bounds__(kBlock) void k_ss_local(uint64_t *in,
                                                     const unsigned int *__restrict__ fill,
                                                     uint32_t B, uint64_t *out, uint64_t n,
                                                     uint32_t local_cap) {
  __shared__ uint64_t s[kSsLocalCap];
  __shared__ uint64_t lds[kBlock / kWave + 1];
  uint64_t nb = n, base = 0;
  if (B > 1) {
    uint64_t part = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kBlock)
      part += fill[b];
    uint64_t total;
    block_exclusive_scan<kBlock>(part, lds, total);
    base = total;
    nb = fill[blockIdx.x];
  }
  if (nb <= local_cap) {
    if (nb <= kBlock)
      ss_local_sort<1>(in + base, out + base, (uint32_t)nb, s);
    else if (nb <= 2 * kBlock)
      ss_local_sort<2>(in + base, out + base, (uint32_t)nb, s);
    else if (nb <= 4 * kBlock)
      ss_local_sort<4>(in + base, out + base, (uint32_t)nb, s);
    else if (nb <= 8 * kBlock)
      ss_local_sort<8>(in + base, out + base, (uint32_t)nb, s);
    else
      ss_local_sort<16>(in + base, out + base, (uint32_t)nb, s);
  } else {
    ss_radix_global(in + base, out + base, nb);
  }
}
}  // namespace

void sort_unique_u64(Ctx &ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, uint32_t *err,
                     uint32_t local_cap, uint32_t key_bits) {
  if (n < 2)
    return;
  if (local_cap == 0 || local_cap > kSsLocalCap)
    local_cap = kSsLocalCap;
  if (n > (uint64_t)kSsMaxBuckets * kSsAvg) {
    // More keys than 512 buckets of about kSsAvg hold: radix passes over the
    // bytes a key of key_bits bits can have.
    const uint64_t varying = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1);
    ctx.timed("rule_sort", [&] { sort_radix_large(ctx, keys, tmp, n, varying, err); });
    return;
  }
  ctx.timed("rule_sort", [&] {
    if (n <= local_cap && n <= kSsSingle) {
      hipLaunchKernelGGL(k_ss_local, dim3(1), dim3(kBlock), 0, ctx.stream(), keys,
                         (const unsigned int *)nullptr, 1u, keys, n, local_cap);
      GFZ_CHECK_LAUNCH();
      return;
    }
    if (n <= 64) {  // too few keys to sample: one bucket, sorted by the fallback
      hipLaunchKernelGGL(k_ss_local, dim3(1), dim3(kBlock), 0, ctx.stream(), keys,
                         (const unsigned int *)nullptr, 1u, tmp, n, local_cap);
      GFZ_CHECK_LAUNCH();
      GFZ_HIP(hipMemcpyAsync(keys, tmp, n * 8, hipMemcpyDeviceToDevice, ctx.stream()));
      return;
    }
    const uint32_t B = (uint32_t)std::min<uint64_t>(
        kSsMaxBuckets, std::max<uint64_t>(2, next_pow2(ceil_div(n, kSsAvg))));
    DBuf<uint64_t> scratch(ctx.pool(), 2 * (uint64_t)B);  // splitters | fill, cursor (u32)
    uint64_t *spl = scratch.get();
    unsigned int *fill = reinterpret_cast<unsigned int *>(scratch.get() + B), *cursor = fill + B;
    // Eight samples per bucket, four per thread.
    switch (B) {
      case 2: case 4: case 8: case 16: case 32:
        hipLaunchKernelGGL(k_ss_splitters<64>, dim3(1), dim3(64), 0, ctx.stream(), keys, n, B, spl,
                           fill);
        break;
      case 64:
        hipLaunchKernelGGL(k_ss_splitters<128>, dim3(1), dim3(128), 0, ctx.stream(), keys, n, B, spl,
                           fill);
        break;
      case 128:
        hipLaunchKernelGGL(k_ss_splitters<256>, dim3(1), dim3(256), 0, ctx.stream(), keys, n, B, spl,
                           fill);
        break;
      case 256:
        hipLaunchKernelGGL(k_ss_splitters<512>, dim3(1), dim3(512), 0, ctx.stream(), keys, n, B, spl,
                           fill);
        break;
      default:
        hipLaunchKernelGGL(k_ss_splitters<1024>, dim3(1), dim3(1024), 0, ctx.stream(), keys, n, B,
                           spl, fill);
        break;
    }
    GFZ_CHECK_LAUNCH();
    const uint32_t tiles = (uint32_t)ceil_div(n, kSsTile);
    hipLaunchKernelGGL(k_ss_count, dim3(tiles), dim3(kBlock), 0, ctx.stream(), keys, n, spl, B,
                       fill);
    GFZ_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ss_place, dim3(tiles), dim3(kBlock), 0, ctx.stream(), keys, n, spl, B,
                       fill, cursor, tmp);
    GFZ_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ss_local, dim3(B), dim3(kBlock), 0, ctx.stream(), tmp, fill, B, keys,
                       (uint64_t)0, local_cap);
    GFZ_CHECK_LAUNCH();
  });
}

}  // namespace gfz
