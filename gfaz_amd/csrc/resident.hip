// Resident decoder: a rulebook with its Z bodies plus encoded traversals kept in
// HBM, decoded per request - any subset of traversals in one persistent,
// work-stealing launch. The reference decodes extract requests on the host
// (src/workflows/extraction_workflow.cpp:106-133,206-424, which zstd-decodes the
// whole column per call); here the per-symbol state is computed once at load:
//   sym_off[k]  decoded start of encoded symbol k (whole stream, exclusive scan);
//   syms[k]     base and Z start of symbol k, with the traversal's first inverse
//               delta folded in (see expand.hip);
//   grid[c]     the symbol covering decoded position c * kGrid (coarse index).
// A request is a list of segments (one per traversal); segment tiles are numbered
// consecutively and blocks take the next tile from an atomic counter. A block
// finds its segment and the symbols covering its tile with wave-wide 64-way
// searches (tile prefix, then the grid cell, then sym_off), then expands the tile
// exactly as the batch decoder does.
#include "resident.hpp"

#include "prims.hpp"
#include "zexpand.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace gfz {

namespace {

using namespace zx;

constexpr uint64_t kGrid = 4096;  // decoded positions per grid cell

// Largest k in [lo, hi) with a[k] <= target, given a non-decreasing and
// a[lo] <= target. Called by all 64 lanes of one wave; every lane returns it.
// An empty range (hi <= lo) or a broken precondition returns lo: the search
// always ends within 64-ary steps, whatever the inputs.
__device__ inline uint64_t wave_last_le(const uint64_t *__restrict__ a, uint64_t lo, uint64_t hi,
                                        uint64_t target) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  while (hi > lo + 1) {
    const uint64_t step = (hi - lo + kWave - 1) / kWave;
    const uint64_t k = lo + lane * step;
    const bool ok = k < hi && a[k] <= target;
    const uint64_t m = __ballot(ok);  // lane 0 holds when the precondition does
    if (m == 0)
      break;
    const uint32_t last = 63u - (uint32_t)__clzll((long long)m);
    lo = lo + (uint64_t)last * step;
    hi = hi < lo + step ? hi : lo + step;
  }
  return lo;
}

// grid[c] = index of the symbol covering decoded position c * kGrid, for the
// symbols base .. base + n_sym - 1 whose decoded starts are sym_off[0 .. n_sym]
// (sym_off points at the first of them; indices stored are whole-stream ones).
__global__ void k_grid(const uint64_t *__restrict__ sym_off, uint64_t n_sym, uint64_t base,
                       uint64_t *__restrict__ grid) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_sym;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = sym_off[k], hi = sym_off[k + 1];
    for (uint64_t c = (lo + kGrid - 1) / kGrid; c * kGrid < hi; ++c)
      grid[c] = base + k;
  }
}

template <uint32_t kItems, bool kDiff>
__global__ __launch_bounds__(kBlock) void k_expand_batch(
    const SegDesc *__restrict__ segs, uint64_t n_segs, const uint64_t *__restrict__ tile_prefix,
    uint64_t total_tiles, const SymInfo *__restrict__ syms, const uint64_t *__restrict__ sym_off,
    const uint64_t *__restrict__ grid, uint64_t n_cells, const int32_t *Z, uint32_t zraw,
    int32_t *out, unsigned long long *counter) {
  constexpr uint32_t kTile = kBlock * kItems;
  __shared__ ZTileSmem<kItems> sm;
  __shared__ unsigned long long s_tile;
  __shared__ uint64_t s_k0, s_kl, s_o0, s_dst;
  __shared__ uint32_t s_n;
  for (;;) {
    if (threadIdx.x == 0)
      s_tile = atomicAdd(counter, 1ull);
    __syncthreads();
    const uint64_t t = s_tile;
    if (t >= total_tiles)
      break;  // uniform across the block
    if (threadIdx.x < kWave) {
      const uint64_t seg = wave_last_le(tile_prefix, 0, n_segs, t);
      const SegDesc sd = segs[seg];
      const uint64_t lt = t - tile_prefix[seg];
      const uint64_t o0 = sd.dec_begin + lt * kTile;
      const uint64_t seg_end = sd.dec_begin + sd.len;
      const uint64_t end = o0 + kTile < seg_end ? o0 + kTile : seg_end;
      const uint64_t c0 = o0 / kGrid, c1 = (end - 1) / kGrid + 1;
      const uint64_t g0 = grid[c0], g1 = (c1 < n_cells ? grid[c1] : sd.k_hi - 1) + 1;
      const uint64_t lo = sd.k_lo > g0 ? sd.k_lo : g0;
      const uint64_t hi = sd.k_hi < g1 ? sd.k_hi : g1;
      const uint64_t k0 = wave_last_le(sym_off, lo, hi, o0);
      const uint64_t kl = wave_last_le(sym_off, k0, hi, end - 1);
      if (threadIdx.x == 0) {
        s_k0 = k0;
        s_kl = kl;
        s_o0 = o0;
        s_n = (uint32_t)(end - o0);
        s_dst = sd.dst + lt * kTile;
      }
    }
    __syncthreads();
    expand_tile<kItems, kDiff>(sm, syms, sym_off, s_k0, (uint32_t)(s_kl - s_k0 + 1), s_o0, s_n,
                               Z, zraw, out + s_dst);
  }
}

__global__ void k_sym_plan_r(const int32_t *__restrict__ enc, uint64_t n,
                             const uint2 *__restrict__ rss,
                             uint64_t R, uint32_t min_id, uint32_t *__restrict__ sz,
                             int32_t *__restrict__ ssum) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const int32_t v = enc[i];
    if (is_rule(v, min_id, R)) {
      const uint2 w = rss[uabs(v) - min_id];
      sz[i] = w.x;
      ssum[i] = (int32_t)(v > 0 ? w.y : (uint32_t)0 - w.y);
    } else {
      sz[i] = 1u;
      ssum[i] = v;
    }
  }
}

__global__ void k_sym_info_r(const int32_t *__restrict__ enc, uint64_t n,
                             const int32_t *__restrict__ incl, const int32_t *__restrict__ ssum,
                             bool delta, const uint4 *__restrict__ szo,
                             uint64_t R, uint32_t min_id, uint32_t zraw,
                             SymInfo *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const int32_t v = enc[i];
    const uint32_t pre = delta ? (uint32_t)incl[i] - (uint32_t)ssum[i] : 0u;
    out[i] = is_rule(v, min_id, R) ? rule_sym(v, pre, delta, szo, min_id)
                                   : SymInfo{pre + (uint32_t)v, zraw};
  }
}

uint32_t grid_for(Ctx &ctx, uint64_t n) {
  return (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(ceil_div(n, (uint64_t)kBlock * 4), (uint64_t)ctx.num_cus() * 8));
}

template <uint32_t kItems>
void launch_batch(Ctx &ctx, bool diff, const Resident &res, const SegDesc *segs, uint64_t n_segs,
                  const uint64_t *tile_prefix, uint64_t total_tiles, int32_t *out,
                  unsigned long long *counter, uint32_t blocks) {
  const Rulebook &rb = *res.rb;
  if (diff)
    hipLaunchKernelGGL((k_expand_batch<kItems, true>), dim3(blocks), dim3(kBlock), 0,
                       ctx.stream(), segs, n_segs, tile_prefix, total_tiles, res.syms.get(),
                       res.sym_off.get(), res.grid.get(), res.n_cells, rb.z.get(), rb.z_raw, out,
                       counter);
  else
    hipLaunchKernelGGL((k_expand_batch<kItems, false>), dim3(blocks), dim3(kBlock), 0,
                       ctx.stream(), segs, n_segs, tile_prefix, total_tiles, res.syms.get(),
                       res.sym_off.get(), res.grid.get(), res.n_cells, rb.z.get(), rb.z_raw, out,
                       counter);
  GFZ_CHECK_LAUNCH();
}

}  // namespace

void resident_load(Ctx &ctx, const Rulebook &rb, const int32_t *d_enc, uint64_t n_enc,
                   const uint32_t *enc_lengths, uint64_t n_trav, int delta_round, Resident &res) {
  res = Resident();
  res.rb = &rb;
  res.descent = !rb.has_bodies && rb.n_rules;
  res.delta_round = delta_round < 0 ? 0 : delta_round;
  res.n_trav = n_trav;
  res.n_enc = n_enc;
  res.enc_start.resize(n_trav);
  res.enc_len.assign(enc_lengths, enc_lengths + n_trav);
  uint64_t total = 0;
  std::vector<uint64_t> starts;
  for (uint64_t t = 0; t < n_trav; ++t) {
    res.enc_start[t] = total;
    if (enc_lengths[t])
      starts.push_back(total);
    total += enc_lengths[t];
  }
  if (total != n_enc)
    throw Error(kErrCorrupt, "flattened sequence block is truncated (lengths sum " +
                                 std::to_string(total) + ", block holds " +
                                 std::to_string(n_enc) + ")");
  res.sym_off = DBuf<uint64_t>(ctx.pool(), n_enc + 1);
  if (res.descent) {
    res.enc = DBuf<int32_t>(ctx.pool(), std::max<uint64_t>(n_enc, 1));
    if (n_enc)
      GFZ_HIP(hipMemcpyAsync(res.enc.get(), d_enc, n_enc * 4, hipMemcpyDeviceToDevice,
                             ctx.stream()));
  } else {
    res.syms = DBuf<SymInfo>(ctx.pool(), std::max<uint64_t>(n_enc, 1));
  }
  const bool delta = res.delta_round > 0;
  {
    DBuf<uint32_t> sz(ctx.pool(), std::max<uint64_t>(n_enc, 1));
    DBuf<int32_t> ssum(ctx.pool(), std::max<uint64_t>(n_enc, 1));
    DBuf<int32_t> incl(ctx.pool(), std::max<uint64_t>(n_enc, 1));
    if (n_enc) {
      hipLaunchKernelGGL(k_sym_plan_r, dim3(grid_for(ctx, n_enc)), dim3(kBlock), 0, ctx.stream(),
                         d_enc, n_enc, rb.ss.get(), rb.n_rules, rb.min_id,
                         sz.get(), ssum.get());
      GFZ_CHECK_LAUNCH();
    }
    exclusive_scan(ctx, sz.get(), res.sym_off.get(), n_enc);
    if (delta && n_enc && !res.descent) {
      const uint64_t w = ceil_div(n_enc, 32);
      DBuf<uint32_t> heads(ctx.pool(), w);
      fill_u32(ctx, heads.get(), 0, w);
      DBuf<uint64_t> d_st(ctx.pool(), std::max<size_t>(starts.size(), 1));
      if (!starts.empty()) {
        GFZ_HIP(hipMemcpyAsync(d_st.get(), starts.data(), starts.size() * 8,
                               hipMemcpyHostToDevice, ctx.stream()));
        set_head_bits(ctx, d_st.get(), starts.size(), heads.get(), n_enc);
      }
      GFZ_HIP(hipMemcpyAsync(incl.get(), ssum.get(), n_enc * 4, hipMemcpyDeviceToDevice,
                             ctx.stream()));
      segmented_prefix_sum(ctx, incl.get(), n_enc, heads.get());
      ctx.sync();  // `starts` (pageable source) and the scratch buffers go out of scope
    }
    if (n_enc && !res.descent) {
      hipLaunchKernelGGL(k_sym_info_r, dim3(grid_for(ctx, n_enc)), dim3(kBlock), 0, ctx.stream(),
                         d_enc, n_enc, incl.get(), ssum.get(), delta, rb.szo.get(),
                         rb.n_rules, rb.min_id, rb.z_raw, res.syms.get());
      GFZ_CHECK_LAUNCH();
    }
    ctx.sync();
  }
  // Decoded starts and lengths per traversal (host), and the total.
  std::vector<uint64_t> h_off(n_enc + 1);
  GFZ_HIP(hipMemcpyAsync(h_off.data(), res.sym_off.get(), (n_enc + 1) * 8, hipMemcpyDeviceToHost,
                         ctx.stream()));
  ctx.sync();
  res.n_out = h_off[n_enc];
  res.dec_start.resize(n_trav);
  res.dec_len.resize(n_trav);
  for (uint64_t t = 0; t < n_trav; ++t) {
    res.dec_start[t] = h_off[res.enc_start[t]];
    res.dec_len[t] = h_off[res.enc_start[t] + res.enc_len[t]] - res.dec_start[t];
  }
  if (res.descent)
    return;  // no coarse index: extraction decodes whole traversals by descent
  // Coarse index; the last entry is a sentinel for ranges that end the stream.
  res.n_cells = ceil_div(std::max<uint64_t>(res.n_out, 1), kGrid) + 1;
  res.grid = DBuf<uint64_t>(ctx.pool(), res.n_cells);
  fill_u64(ctx, res.grid.get(), n_enc ? n_enc - 1 : 0, res.n_cells);
  if (n_enc) {
    hipLaunchKernelGGL(k_grid, dim3(grid_for(ctx, n_enc)), dim3(kBlock), 0, ctx.stream(),
                       res.sym_off.get(), n_enc, (uint64_t)0, res.grid.get());
    GFZ_CHECK_LAUNCH();
  }
  ctx.sync();
}

uint64_t extract_plan(const Resident &res, const uint64_t *ids, uint64_t n_ids, uint32_t items,
                      ExtractPlan &plan, uint64_t *offsets) {
  plan = ExtractPlan();
  plan.items = items;
  plan.tile_prefix.push_back(0);
  const uint64_t tile = (uint64_t)kBlock * items;
  uint64_t total = 0;
  for (uint64_t i = 0; i < n_ids; ++i) {
    const uint64_t t = ids[i];
    if (t >= res.n_trav)
      throw Error(kErrInvalid, "traversal index " + std::to_string(t) + " out of range (" +
                                   std::to_string(res.n_trav) + " stored)");
    if (offsets)
      offsets[i] = total;
    const uint64_t len = res.dec_len[t];
    if (len) {
      plan.segs.push_back(SegDesc{res.dec_start[t], len, total, res.enc_start[t],
                                  res.enc_start[t] + res.enc_len[t]});
      plan.tile_prefix.push_back(plan.tile_prefix.back() + ceil_div(len, tile));
      plan.seg_dst.push_back(total);
    }
    total += len;
  }
  if (offsets)
    offsets[n_ids] = total;
  plan.total = total;
  return total;
}

size_t extract_meta_bytes(const ExtractPlan &plan) {
  return plan.segs.size() * sizeof(SegDesc) + plan.tile_prefix.size() * 8 + 8 +
         plan.seg_dst.size() * 8;
}

namespace {

// Extraction from a store without rule bodies: the planned traversals' encoded
// symbols are gathered in output order (one copy per run of consecutive stored
// traversals) and decoded by the descent expansion, inverse deltas included.
void descent_enqueue(Ctx &ctx, const Resident &res, const ExtractPlan &plan, int32_t *d_out) {
  std::vector<uint32_t> lens(plan.segs.size());
  uint64_t n_sub = 0;
  for (size_t i = 0; i < plan.segs.size(); ++i) {
    lens[i] = (uint32_t)(plan.segs[i].k_hi - plan.segs[i].k_lo);
    n_sub += lens[i];
  }
  DBuf<int32_t> sub(ctx.pool(), std::max<uint64_t>(n_sub, 1));
  uint64_t at = 0;
  for (size_t i = 0; i < plan.segs.size();) {
    const uint64_t lo = plan.segs[i].k_lo;
    uint64_t hi = plan.segs[i].k_hi;
    size_t j = i + 1;
    while (j < plan.segs.size() && plan.segs[j].k_lo == hi)
      hi = plan.segs[j++].k_hi;
    if (hi > lo)
      GFZ_HIP(hipMemcpyAsync(sub.get() + at, res.enc.get() + lo, (hi - lo) * 4,
                             hipMemcpyDeviceToDevice, ctx.stream()));
    at += hi - lo;
    i = j;
  }
  DecodePlan dp;
  DBuf<uint64_t> sym_off;
  decode_plan(ctx, *res.rb, sub.get(), n_sub, lens.data(), nullptr, lens.size(), res.delta_round,
              dp, sym_off);
  if (dp.n_out != plan.total)
    throw Error(kErrCorrupt, "descent extraction decoded " + std::to_string(dp.n_out) +
                                 " symbols where the store records " +
                                 std::to_string(plan.total));
  decode_run(ctx, *res.rb, sub.get(), n_sub, lens.data(), lens.size(), res.delta_round, dp,
             sym_off.get(), d_out);
  ctx.add_bytes("extract", 4.0 * plan.total);
}

}  // namespace

void extract_enqueue(Ctx &ctx, const Resident &res, const ExtractPlan &plan, uint8_t *h_stage,
                     uint8_t *d_meta, int32_t *d_out) {
  if (!plan.total)
    return;
  if (res.descent) {
    descent_enqueue(ctx, res, plan, d_out);
    return;
  }
  const uint64_t n_segs = plan.segs.size(), total_tiles = plan.tile_prefix.back();
  const size_t seg_bytes = n_segs * sizeof(SegDesc), pre_bytes = (n_segs + 1) * 8;
  const size_t dst_off = seg_bytes + pre_bytes + 8;
  std::memcpy(h_stage, plan.segs.data(), seg_bytes);
  std::memcpy(h_stage + seg_bytes, plan.tile_prefix.data(), pre_bytes);
  std::memset(h_stage + seg_bytes + pre_bytes, 0, 8);  // work-stealing counter
  std::memcpy(h_stage + dst_off, plan.seg_dst.data(), plan.seg_dst.size() * 8);
  GFZ_HIP(hipMemcpyAsync(d_meta, h_stage, extract_meta_bytes(plan), hipMemcpyHostToDevice,
                         ctx.stream()));
  const SegDesc *d_segs = reinterpret_cast<const SegDesc *>(d_meta);
  const uint64_t *d_pre = reinterpret_cast<const uint64_t *>(d_meta + seg_bytes);
  auto *counter = reinterpret_cast<unsigned long long *>(d_meta + seg_bytes + pre_bytes);
  const uint32_t items = plan.items;
  const uint32_t per_cu = items >= 32 ? 2 : 4;
  const uint32_t blocks =
      (uint32_t)std::min<uint64_t>(total_tiles, (uint64_t)ctx.num_cus() * per_cu);
  const bool diff = res.delta_round == 0;
  ctx.timed("extract", [&] {
    switch (items) {
      case 8: launch_batch<8>(ctx, diff, res, d_segs, n_segs, d_pre, total_tiles, d_out, counter, blocks); break;
      case 32: launch_batch<32>(ctx, diff, res, d_segs, n_segs, d_pre, total_tiles, d_out, counter, blocks); break;
      default: launch_batch<16>(ctx, diff, res, d_segs, n_segs, d_pre, total_tiles, d_out, counter, blocks); break;
    }
  });
  ctx.add_bytes("extract", 4.0 * plan.total);
  if (res.delta_round >= 2) {
    // Further inverse deltas: segmented scans restarted at every traversal.
    const uint64_t words = ceil_div(plan.total, 32);
    DBuf<uint32_t> heads(ctx.pool(), words);  // stream-ordered reuse: no sync needed
    fill_u32(ctx, heads.get(), 0, words);
    set_head_bits(ctx, reinterpret_cast<const uint64_t *>(d_meta + dst_off), plan.seg_dst.size(),
                  heads.get(), plan.total);
    for (int r = 1; r < res.delta_round; ++r)
      segmented_prefix_sum(ctx, d_out, plan.total, heads.get());
  }
}

uint64_t resident_extract(Ctx &ctx, const Resident &res, const uint64_t *ids, uint64_t n_ids,
                          int32_t *d_out, uint64_t out_cap, uint64_t *offsets, uint32_t items) {
  ExtractPlan plan;
  const uint64_t total = extract_plan(res, ids, n_ids, items, plan, offsets);
  if (total > out_cap)
    throw Error(kErrInvalid, "output capacity " + std::to_string(out_cap) + " < extracted size " +
                                 std::to_string(total));
  if (!total)
    return 0;
  const size_t bytes = extract_meta_bytes(plan);
  DBuf<uint8_t> meta(ctx.pool(), bytes);
  std::vector<uint8_t> stage(bytes);
  extract_enqueue(ctx, res, plan, stage.data(), meta.get(), d_out);
  ctx.sync();  // the pageable staging vector goes out of scope
  return total;
}

// ---- extraction without allocation ------------------------------------------

namespace {

// One per device; the host must never pass this name to a kernel or a copy
// (in host code it is only a stand-in): noalloc_scratch() gives the address.
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

void noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw Error(kErrInvalid, "device index " + std::to_string(device) + " beyond " +
                                 std::to_string(kMaxDevices));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device])
    return;
  DeviceGuard g(device);
  void *ptr = nullptr;
  GFZ_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr)
    throw Error(kErrHip, "no device address for the extraction scratch array");
  g_scratch_ptr[device] = static_cast<uint8_t *>(ptr);
}

uint8_t *noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  uint8_t *p = device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
  if (!p)
    throw Error(kErrInvalid, "extraction scratch of device " + std::to_string(device) +
                                 " is not set up (create a context on it first)");
  return p;
}

uint64_t resident_extract_noalloc(Ctx &ctx, const Resident &res, const uint64_t *ids,
                                  uint64_t n_ids, int32_t *d_out, uint64_t out_cap,
                                  uint64_t *offsets, uint32_t items) {
  if (res.descent)
    throw Error(kErrInvalid, "extraction without allocation needs the rule bodies in HBM (this "
                             "store's rulebook is expanded by descent)");
  if (res.delta_round >= 2)
    throw Error(kErrInvalid, "extraction without allocation supports delta_round 0 and 1 "
                             "(this store has " + std::to_string(res.delta_round) + ")");
  ExtractPlan plan;
  const uint64_t total = extract_plan(res, ids, n_ids, items, plan, offsets);
  if (total > out_cap)
    throw Error(kErrInvalid, "output capacity " + std::to_string(out_cap) + " < extracted size " +
                                 std::to_string(total));
  if (!total)
    return 0;
  const size_t bytes = extract_meta_bytes(plan);
  if (bytes > kNoAllocScratchBytes)
    throw Error(kErrInvalid, "request plan needs " + std::to_string(bytes) +
                                 " bytes of scratch, more than the fixed " +
                                 std::to_string(kNoAllocScratchBytes) + " (" +
                                 std::to_string(plan.segs.size()) + " non-empty traversals)");
  uint8_t *d_meta = noalloc_scratch(ctx.device());
  std::vector<uint8_t> stage(bytes);
  extract_enqueue(ctx, res, plan, stage.data(), d_meta, d_out);  // copies the zeroed counter too
  ctx.sync();  // the host staging vector goes out of scope
  return total;
}

namespace {

__global__ void k_add_offset(const uint64_t *__restrict__ in, uint64_t n, uint64_t add,
                             uint64_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[i] + add;
}

template <class T>
void grow(Ctx &ctx, DBuf<T> &buf, uint64_t used, uint64_t need) {
  if (need <= buf.size())
    return;
  DBuf<T> nb(ctx.pool(), std::max<uint64_t>(need, 2 * (uint64_t)buf.size()));
  if (used)
    GFZ_HIP(hipMemcpyAsync(nb.get(), buf.get(), used * sizeof(T), hipMemcpyDeviceToDevice,
                           ctx.stream()));
  buf = std::move(nb);
}

}  // namespace

void resident_append(Ctx &ctx, Resident &res, const int32_t *d_enc, uint64_t n_new,
                     const uint32_t *enc_lengths, uint64_t n_trav_new) {
  const Rulebook &rb = *res.rb;
  uint64_t total = 0;
  std::vector<uint64_t> starts;
  for (uint64_t t = 0; t < n_trav_new; ++t) {
    if (enc_lengths[t])
      starts.push_back(total);
    total += enc_lengths[t];
  }
  if (total != n_new)
    throw Error(kErrInvalid, "appended lengths sum " + std::to_string(total) + " != " +
                                 std::to_string(n_new) + " symbols");
  const uint64_t e0 = res.n_enc, out0 = res.n_out;
  if (res.descent) {
    grow(ctx, res.enc, e0, e0 + n_new);
    if (n_new)
      GFZ_HIP(hipMemcpyAsync(res.enc.get() + e0, d_enc, n_new * 4, hipMemcpyDeviceToDevice,
                             ctx.stream()));
  } else {
    grow(ctx, res.syms, e0, e0 + n_new);
  }
  grow(ctx, res.sym_off, e0 + 1, e0 + n_new + 1);
  std::vector<uint64_t> h_off(n_new + 1, 0);
  if (n_new) {
    DBuf<uint32_t> sz(ctx.pool(), n_new);
    DBuf<int32_t> ssum(ctx.pool(), n_new), incl(ctx.pool(), n_new);
    DBuf<uint64_t> off(ctx.pool(), n_new + 1);
    hipLaunchKernelGGL(k_sym_plan_r, dim3(grid_for(ctx, n_new)), dim3(kBlock), 0, ctx.stream(),
                       d_enc, n_new, rb.ss.get(), rb.n_rules, rb.min_id, sz.get(),
                       ssum.get());
    GFZ_CHECK_LAUNCH();
    exclusive_scan(ctx, sz.get(), off.get(), n_new);
    hipLaunchKernelGGL(k_add_offset, dim3(grid_for(ctx, n_new + 1)), dim3(kBlock), 0,
                       ctx.stream(), off.get(), n_new + 1, out0, res.sym_off.get() + e0);
    GFZ_CHECK_LAUNCH();
    const bool delta = res.delta_round > 0;
    if (delta && !res.descent) {
      const uint64_t w = ceil_div(n_new, 32);
      DBuf<uint32_t> heads(ctx.pool(), w);
      fill_u32(ctx, heads.get(), 0, w);
      DBuf<uint64_t> d_st(ctx.pool(), std::max<size_t>(starts.size(), 1));
      GFZ_HIP(hipMemcpyAsync(d_st.get(), starts.data(), starts.size() * 8, hipMemcpyHostToDevice,
                             ctx.stream()));
      set_head_bits(ctx, d_st.get(), starts.size(), heads.get(), n_new);
      GFZ_HIP(hipMemcpyAsync(incl.get(), ssum.get(), n_new * 4, hipMemcpyDeviceToDevice,
                             ctx.stream()));
      segmented_prefix_sum(ctx, incl.get(), n_new, heads.get());
      ctx.sync();  // `starts` is pageable
    }
    if (!res.descent) {
      hipLaunchKernelGGL(k_sym_info_r, dim3(grid_for(ctx, n_new)), dim3(kBlock), 0,
                         ctx.stream(), d_enc, n_new, incl.get(), ssum.get(), delta,
                         rb.szo.get(), rb.n_rules, rb.min_id, rb.z_raw, res.syms.get() + e0);
      GFZ_CHECK_LAUNCH();
    }
    GFZ_HIP(hipMemcpyAsync(h_off.data(), off.get(), (n_new + 1) * 8, hipMemcpyDeviceToHost,
                           ctx.stream()));
    ctx.sync();
  }
  // Host bookkeeping of the new traversals.
  uint64_t k = 0;
  for (uint64_t t = 0; t < n_trav_new; ++t) {
    res.enc_start.push_back(e0 + k);
    res.enc_len.push_back(enc_lengths[t]);
    res.dec_start.push_back(out0 + h_off[k]);
    res.dec_len.push_back(h_off[k + enc_lengths[t]] - h_off[k]);
    k += enc_lengths[t];
  }
  res.n_trav += n_trav_new;
  res.n_enc = e0 + n_new;
  res.n_out = out0 + h_off[n_new];
  if (res.descent)
    return;
  // Grid: cells from the old sentinel on are set again (sentinel, then the new symbols).
  const uint64_t old_cells = res.n_cells;
  res.n_cells = ceil_div(std::max<uint64_t>(res.n_out, 1), kGrid) + 1;
  grow(ctx, res.grid, old_cells, res.n_cells);
  const uint64_t first = old_cells ? old_cells - 1 : 0;
  fill_u64(ctx, res.grid.get() + first, res.n_enc ? res.n_enc - 1 : 0, res.n_cells - first);
  if (n_new) {
    hipLaunchKernelGGL(k_grid, dim3(grid_for(ctx, n_new)), dim3(kBlock), 0, ctx.stream(),
                       res.sym_off.get() + e0, n_new, e0, res.grid.get());
    GFZ_CHECK_LAUNCH();
  }
  ctx.sync();
}

}  // namespace gfz

namespace gfz {

void decompress_stream(Ctx &ctx, const Resident &res, uint64_t chunk_bytes,
                       const std::function<bool(const StreamChunk &)> &consume) {
  // Chunks of whole traversals, at most chunk_bytes of output each unless a
  // single traversal is larger (then it is a chunk of its own).
  struct Range {
    uint64_t b, e, total;
  };
  std::vector<Range> chunks;
  const uint64_t cap_elems = std::max<uint64_t>(chunk_bytes / 4, 1);
  uint64_t biggest = 0;
  for (uint64_t t = 0; t < res.n_trav;) {
    uint64_t e = t, total = 0;
    while (e < res.n_trav && (e == t || total + res.dec_len[e] <= cap_elems))
      total += res.dec_len[e++];
    chunks.push_back({t, e, total});
    biggest = std::max(biggest, total);
    t = e;
  }
  if (chunks.empty())
    return;
  const uint64_t buf_elems = std::max<uint64_t>(biggest, 1);
  // Two device outputs and two pinned host buffers, used alternately.
  DBuf<int32_t> d_out[2] = {DBuf<int32_t>(ctx.pool(), buf_elems),
                            DBuf<int32_t>(ctx.pool(), buf_elems)};
  int32_t *h_buf[2] = {nullptr, nullptr};
  hipEvent_t ready[2] = {nullptr, nullptr};
  struct Cleanup {
    int32_t **h;
    hipEvent_t *ev;
    hipStream_t s;
    ~Cleanup() {
      (void)hipStreamSynchronize(s);
      for (int i = 0; i < 2; ++i) {
        if (h[i])
          (void)hipHostFree(h[i]);
        if (ev[i])
          (void)hipEventDestroy(ev[i]);
      }
    }
  } cleanup{h_buf, ready, ctx.stream()};
  for (int i = 0; i < 2; ++i) {
    void *p = nullptr;
    GFZ_HIP(hipHostMalloc(&p, buf_elems * 4, hipHostMallocDefault));
    h_buf[i] = static_cast<int32_t *>(p);
    GFZ_HIP(hipEventCreateWithFlags(&ready[i], hipEventDisableTiming));
  }
  std::vector<uint64_t> ids;
  std::vector<uint8_t> stage[2];  // reused only after its chunk's event was waited on
  uint64_t expanded = 0;
  std::vector<uint64_t> expanded_at(chunks.size());
  auto enqueue = [&](size_t k) {
    const Range &r = chunks[k];
    ids.resize(r.e - r.b);
    for (uint64_t t = r.b; t < r.e; ++t)
      ids[t - r.b] = t;
    ExtractPlan plan;
    const int tuned = ctx.tuned("stream_items");
    extract_plan(res, ids.data(), ids.size(), tuned > 0 ? (uint32_t)tuned : 16u, plan, nullptr);
    if (plan.total) {
      const size_t mb = extract_meta_bytes(plan);
      stage[k & 1].resize(mb);
      DBuf<uint8_t> d_meta(ctx.pool(), mb);  // stream-ordered reuse after return
      extract_enqueue(ctx, res, plan, stage[k & 1].data(), d_meta.get(), d_out[k & 1].get());
      GFZ_HIP(hipMemcpyAsync(h_buf[k & 1], d_out[k & 1].get(), plan.total * 4,
                             hipMemcpyDeviceToHost, ctx.stream()));
    }
    GFZ_HIP(hipEventRecord(ready[k & 1], ctx.stream()));
    expanded_at[k] = expanded;
    expanded += r.total;
  };
  std::vector<uint32_t> lens;
  auto deliver = [&](size_t k) {
    GFZ_HIP(hipEventSynchronize(ready[k & 1]));
    const Range &r = chunks[k];
    lens.resize(r.e - r.b);
    for (uint64_t t = r.b; t < r.e; ++t)
      lens[t - r.b] = (uint32_t)res.dec_len[t];
    StreamChunk c{h_buf[k & 1], r.total, r.b, r.e, expanded_at[k], lens.data()};
    if (!consume(c))
      throw Error(kErrInvalid, "stream stopped by the consumer at chunk " + std::to_string(k));
  };
  enqueue(0);
  for (size_t k = 1; k < chunks.size(); ++k) {
    enqueue(k);      // the device decodes chunk k ...
    deliver(k - 1);  // ... while the consumer reads chunk k - 1
  }
  deliver(chunks.size() - 1);
}

}  // namespace gfz
