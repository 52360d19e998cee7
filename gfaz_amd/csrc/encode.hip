// Encoding new traversals with a stored rulebook (add-haplotypes) on gfx950.
//
// Restates src/workflows/add_haplotypes_workflow.cpp:336-379,447-465 with the
// encoder of src/grammar/path_encoder.cpp:13-84:
//   * delta_round transforms; the largest |value| must stay below the first
//     layer's start ID when layers exist;
//   * per stored layer, a table pack(first, second) -> start_id + i over the
//     layer's rows (the key as stored, not canonicalised; on duplicate keys the
//     first row wins, as emplace does);
//   * per traversal and layer a shift-reduce: push x; if the top two (s, x) form
//     a key -> replace them by +id, else if pack(-x, -s) is a key -> -id.
// Only the top pair is checked after a push, so the stack top s_i after
// position i depends only on s_{i-1} and x_i:
//   s_i = T(s_{i-1}, x_i) if a rule matches ("reduction at i"), else x_i,
// and the encoded traversal is every s_{i-1} with no reduction at i, then the
// last s. On the GPU each chunk of 256 positions is first run assuming no
// reduction at its first position; then, in passes until nothing changes,
// every chunk re-runs from the previous chunk's end state until its states
// agree with the stored ones (from there on they are equal). A pass whose
// re-run crosses a whole chunk flags another pass. Emission is a stream
// compaction.
#include "codec.hpp"
#include "encode.hpp"
#include "prims.hpp"

#include <algorithm>

namespace gfz {

namespace {

constexpr uint32_t kChunk = 256;       // positions per speculative run
constexpr uint32_t kMaxEncProbe = 512;  // probe bound of the rule table

struct EncSlot {
  unsigned long long key;
  uint32_t val;  // rule ID; the smallest row wins
  uint32_t pad;
};

__device__ inline uint64_t pack_u(int32_t a, int32_t b) { return (uint64_t)pack2(a, b); }

__global__ void k_enc_fill(EncSlot *slots, uint64_t cap) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap;
       i += (uint64_t)gridDim.x * blockDim.x)
    slots[i] = EncSlot{(unsigned long long)kEmptyKey, 0xFFFFFFFFu, 0u};
}

// Rows [row0, row0 + count) of the rule columns, IDs id0 + i.
__global__ void k_enc_insert(const int32_t *__restrict__ first, const int32_t *__restrict__ second,
                             uint64_t row0, uint32_t count, uint32_t id0, EncSlot *slots,
                             uint64_t mask, uint32_t *err) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = pack_u(first[row0 + i], second[row0 + i]);
    if (key == kEmptyKey)
      continue;  // holds INT32_MIN: no valid pair equals it
    uint64_t s = key_hash(key) & mask;
    bool placed = false;
    for (uint32_t p = 0; p < kMaxEncProbe && !placed; ++p) {
      const unsigned long long old =
          atomicCAS(&slots[s].key, (unsigned long long)kEmptyKey, (unsigned long long)key);
      if (old == kEmptyKey || old == key) {
        atomicMin(&slots[s].val, id0 + (uint32_t)i);
        placed = true;
      } else {
        s = (s + 1) & mask;
      }
    }
    if (!placed)
      atomicExch(err, 1u);
  }
}

__device__ inline uint32_t enc_find(const EncSlot *__restrict__ slots, uint64_t mask,
                                    uint64_t key) {
  uint64_t s = key_hash(key) & mask;
  for (uint32_t p = 0; p < kMaxEncProbe; ++p) {
    const uint64_t k = slots[s].key;
    if (k == key)
      return slots[s].val;
    if (k == kEmptyKey)
      return 0xFFFFFFFFu;
    s = (s + 1) & mask;
  }
  return 0xFFFFFFFFu;  // the insert pass bounds every key's distance, so a miss
}

// One shift-reduce step: state s, pushed x -> new top; *red = a rule matched.
__device__ inline int32_t enc_step(const EncSlot *slots, uint64_t mask, int32_t s, int32_t x,
                                   bool *red) {
  const uint64_t k = pack_u(s, x);
  uint32_t v = enc_find(slots, mask, k);
  if (v != 0xFFFFFFFFu) {
    *red = true;
    return (int32_t)v;
  }
  v = enc_find(slots, mask, (uint64_t)reverse2((int64_t)k));
  if (v != 0xFFFFFFFFu) {
    *red = true;
    return -(int32_t)v;
  }
  *red = false;
  return x;
}

__device__ inline bool head_bit(const uint32_t *heads, uint64_t i) {
  return (heads[i >> 5] >> (i & 31)) & 1u;
}

// Speculative run of each chunk: no reduction at its first position.
__global__ void k_enc_spec(const int32_t *__restrict__ x, uint64_t n,
                           const uint32_t *__restrict__ heads, const EncSlot *__restrict__ slots,
                           uint64_t mask, int32_t *__restrict__ st, uint8_t *__restrict__ red) {
  const uint64_t n_chunks = (n + kChunk - 1) / kChunk;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n_chunks;
       c += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = c * kChunk, e = b + kChunk < n ? b + kChunk : n;
    int32_t s = x[b];
    st[b] = s;
    red[b] = 0;
    for (uint64_t i = b + 1; i < e; ++i) {
      const int32_t xi = x[i];
      bool r = false;
      if (!head_bit(heads, i))
        s = enc_step(slots, mask, s, xi, &r);
      else
        s = xi;
      st[i] = s;
      red[i] = r ? 1 : 0;
    }
  }
}

__global__ void k_enc_ends(const int32_t *__restrict__ st, uint64_t n, int32_t *__restrict__ ends) {
  const uint64_t n_chunks = (n + kChunk - 1) / kChunk;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n_chunks;
       c += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = (c + 1) * kChunk < n ? (c + 1) * kChunk : n;
    ends[c] = st[e - 1];
  }
}

// Re-run each chunk from the previous chunk's end state (ends[], taken before
// the pass) until the states agree with the stored ones.
__global__ void k_enc_fix(const int32_t *__restrict__ x, uint64_t n,
                          const uint32_t *__restrict__ heads, const EncSlot *__restrict__ slots,
                          uint64_t mask, const int32_t *__restrict__ ends, int32_t *__restrict__ st,
                          uint8_t *__restrict__ red, uint32_t *changed) {
  const uint64_t n_chunks = (n + kChunk - 1) / kChunk;
  for (uint64_t c = 1 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n_chunks;
       c += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = c * kChunk, e = b + kChunk < n ? b + kChunk : n;
    if (head_bit(heads, b))
      continue;  // a traversal starts here: the speculation was exact
    int32_t s = ends[c - 1];
    bool agreed = false;
    for (uint64_t i = b; i < e; ++i) {
      if (i > b && head_bit(heads, i)) {
        agreed = true;  // later positions do not depend on earlier ones
        break;
      }
      bool r = false;
      s = enc_step(slots, mask, s, x[i], &r);
      if (s == st[i] && (uint8_t)r == red[i]) {
        agreed = true;
        break;
      }
      st[i] = s;
      red[i] = r ? 1 : 0;
    }
    if (!agreed)
      atomicExch(changed, 1u);
  }
}

// Position i is emitted when position i + 1 is not a reduction (or ends the
// traversal / stream).
__global__ void k_enc_flags(const uint8_t *__restrict__ red, const uint32_t *__restrict__ heads,
                            uint64_t n, uint32_t *__restrict__ flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    flag[i] = (i + 1 == n || head_bit(heads, i + 1) || !red[i + 1]) ? 1u : 0u;
}

__global__ void k_enc_compact(const int32_t *__restrict__ st, const uint32_t *__restrict__ flag,
                              const uint64_t *__restrict__ pos, uint64_t n,
                              int32_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (flag[i])
      out[pos[i]] = st[i];
}

__global__ void k_enc_delta(const int32_t *__restrict__ in, int32_t *__restrict__ out,
                            const uint32_t *__restrict__ heads, uint64_t n, uint32_t *max_abs) {
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    int32_t v = in[i];
    if (!head_bit(heads, i))
      v = (int32_t)((uint32_t)v - (uint32_t)in[i - 1]);
    out[i] = v;
    const uint32_t a = uabs(v);
    m = a > m ? a : m;
  }
  m = wave_reduce_max(m);
  if ((threadIdx.x & (kWave - 1)) == 0 && m)
    atomicMax(max_abs, m);
}

uint32_t grid_for(Ctx &ctx, uint64_t n, uint32_t per = kBlock) {
  return (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(ceil_div(n, (uint64_t)per), (uint64_t)ctx.num_cus() * 8));
}

void upload_heads(Ctx &ctx, const std::vector<uint64_t> &starts, uint64_t n, DBuf<uint32_t> &heads) {
  const uint64_t words = ceil_div(std::max<uint64_t>(n, 1), 32);
  heads = DBuf<uint32_t>(ctx.pool(), words);
  fill_u32(ctx, heads.get(), 0, words);
  if (!starts.empty()) {
    DBuf<uint64_t> d(ctx.pool(), starts.size());
    GFZ_HIP(hipMemcpyAsync(d.get(), starts.data(), starts.size() * 8, hipMemcpyHostToDevice,
                           ctx.stream()));
    set_head_bits(ctx, d.get(), starts.size(), heads.get(), n);
    ctx.sync();  // `starts` is pageable host memory
  }
}

}  // namespace

void encode_tables_build(Ctx &ctx, const int32_t *d_first, const int32_t *d_second,
                         uint64_t n_rules, const EncodeLayer *layers, uint64_t n_layers,
                         EncodeTables &out) {
  out = EncodeTables();
  for (uint64_t l = 0; l < n_layers; ++l) {
    const uint64_t count = (uint64_t)layers[l].end_id - layers[l].start_id;
    if (layers[l].end_id < layers[l].start_id || layers[l].row_offset + count > n_rules)
      throw Error(kErrCorrupt, "rule column is truncated");
  }
  Pool &pool = ctx.pool();
  DBuf<uint32_t> err(pool, 1);
  fill_u32(ctx, err.get(), 0, 1);
  for (uint64_t l = 0; l < n_layers; ++l) {
    out.layers.push_back(layers[l]);
    const uint32_t count = layers[l].end_id - layers[l].start_id;
    if (!count) {
      out.slots.emplace_back();
      out.caps.push_back(0);
      continue;
    }
    const uint64_t cap = next_pow2(std::max<uint64_t>(1024, 2 * (uint64_t)count));
    DBuf<uint8_t> buf(pool, cap * sizeof(EncSlot));
    EncSlot *slots = reinterpret_cast<EncSlot *>(buf.get());
    hipLaunchKernelGGL(k_enc_fill, dim3(grid_for(ctx, cap)), dim3(kBlock), 0, ctx.stream(),
                       slots, cap);
    GFZ_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_enc_insert, dim3(grid_for(ctx, count)), dim3(kBlock), 0, ctx.stream(),
                       d_first, d_second, layers[l].row_offset, count, layers[l].start_id,
                       slots, cap - 1, err.get());
    GFZ_CHECK_LAUNCH();
    out.slots.push_back(std::move(buf));
    out.caps.push_back(cap);
  }
  ctx.stage_scalars(err.get(), 1);
  ctx.sync();
  if (reinterpret_cast<const uint32_t *>(ctx.scalars())[0])
    throw Error(kErrCorrupt, "rule table probe bound exceeded (internal error)");
}

void encode_with_rules(Ctx &ctx, const int32_t *d_nodes, uint64_t n, const uint32_t *lengths,
                       uint64_t n_trav, const int32_t *d_first, const int32_t *d_second,
                       uint64_t n_rules, const EncodeLayer *layers, uint64_t n_layers,
                       int delta_round, EncodeResult &out) {
  uint64_t total = 0;
  for (uint64_t t = 0; t < n_trav; ++t)
    total += lengths[t];
  if (total != n)
    throw Error(kErrInvalid, "sum of traversal lengths (" + std::to_string(total) +
                                 ") != number of nodes (" + std::to_string(n) + ")");
  EncodeTables tables;
  encode_tables_build(ctx, d_first, d_second, n_rules, layers, n_layers, tables);
  encode_with_tables(ctx, d_nodes, n, lengths, n_trav, tables, delta_round, out);
}

void encode_with_tables(Ctx &ctx, const int32_t *d_nodes, uint64_t n, const uint32_t *lengths,
                        uint64_t n_trav, const EncodeTables &tables, int delta_round,
                        EncodeResult &out) {
  out = EncodeResult();
  out.lengths.assign(lengths, lengths + n_trav);
  std::vector<uint64_t> starts;
  uint64_t total = 0;
  for (uint64_t t = 0; t < n_trav; ++t) {
    if (lengths[t])
      starts.push_back(total);
    total += lengths[t];
  }
  if (total != n)
    throw Error(kErrInvalid, "sum of traversal lengths (" + std::to_string(total) +
                                 ") != number of nodes (" + std::to_string(n) + ")");
  const uint64_t n_layers = tables.layers.size();
  const EncodeLayer *layers = tables.layers.data();
  Pool &pool = ctx.pool();
  DBuf<int32_t> cur(pool, std::max<uint64_t>(n, 1));
  if (n)
    GFZ_HIP(hipMemcpyAsync(cur.get(), d_nodes, n * 4, hipMemcpyDeviceToDevice, ctx.stream()));
  DBuf<uint32_t> heads;
  upload_heads(ctx, starts, n, heads);

  // Delta transforms and the collision check (add_haplotypes_workflow.cpp:447-465).
  if (delta_round > 0 && n) {
    DBuf<int32_t> tmp(pool, n);
    DBuf<uint32_t> d_max(pool, 1);
    fill_u32(ctx, d_max.get(), 0, 1);
    for (int r = 0; r < delta_round; ++r) {
      hipLaunchKernelGGL(k_enc_delta, dim3(grid_for(ctx, n)), dim3(kBlock), 0, ctx.stream(),
                         cur.get(), tmp.get(), heads.get(), n, d_max.get());
      GFZ_CHECK_LAUNCH();
      std::swap(cur, tmp);
    }
    ctx.stage_scalars(d_max.get(), 1);
    ctx.sync();
    const uint32_t max_abs = (uint32_t)ctx.scalars()[0];
    if (n_layers && max_abs >= layers[0].start_id)
      throw Error(kErrOverflow, "delta-encoded appended sequence reaches ID " +
                                    std::to_string(max_abs) +
                                    ", which collides with the stored rule region starting at " +
                                    std::to_string(layers[0].start_id));
  }

  uint64_t L = n;
  std::vector<uint32_t> cur_len(lengths, lengths + n_trav);
  DBuf<uint32_t> flags(pool, 1);  // changed
  for (uint64_t l = 0; l < n_layers && L; ++l) {
    const uint64_t cap = tables.caps[l];
    if (!cap)
      continue;
    const EncSlot *slots = reinterpret_cast<const EncSlot *>(tables.slots[l].get());
    const uint64_t n_chunks = ceil_div(L, kChunk);
    DBuf<int32_t> st(pool, L), ends(pool, n_chunks);
    DBuf<uint8_t> red(pool, L);
    ctx.timed("encode_rules", [&] {
      hipLaunchKernelGGL(k_enc_spec, dim3(grid_for(ctx, n_chunks, 64)), dim3(64), 0, ctx.stream(),
                         cur.get(), L, heads.get(), slots, cap - 1, st.get(), red.get());
    });
    GFZ_CHECK_LAUNCH();
    // Fix-up passes; each pass makes at least one more chunk exact, so at most
    // n_chunks passes (typically one or two).
    for (uint64_t pass = 0;; ++pass) {
      if (pass > n_chunks)
        throw Error(kErrCorrupt, "add-haplotypes encoding did not settle (internal error)");
      fill_u32(ctx, flags.get(), 0, 1);
      hipLaunchKernelGGL(k_enc_ends, dim3(grid_for(ctx, n_chunks)), dim3(kBlock), 0,
                         ctx.stream(), st.get(), L, ends.get());
      GFZ_CHECK_LAUNCH();
      ctx.timed("encode_rules", [&] {
        hipLaunchKernelGGL(k_enc_fix, dim3(grid_for(ctx, n_chunks, 64)), dim3(64), 0,
                           ctx.stream(), cur.get(), L, heads.get(), slots, cap - 1,
                           ends.get(), st.get(), red.get(), flags.get());
      });
      GFZ_CHECK_LAUNCH();
      ctx.stage_scalars(flags.get(), 1);
      ctx.sync();
      if (!reinterpret_cast<const uint32_t *>(ctx.scalars())[0])
        break;
    }
    // Emission: compaction of the stack tops that survive.
    DBuf<uint32_t> flag(pool, L);
    DBuf<uint64_t> pos(pool, L + 1);
    hipLaunchKernelGGL(k_enc_flags, dim3(grid_for(ctx, L)), dim3(kBlock), 0, ctx.stream(),
                       red.get(), heads.get(), L, flag.get());
    GFZ_CHECK_LAUNCH();
    exclusive_scan(ctx, flag.get(), pos.get(), L);
    // New lengths: emitted count per traversal, from the scan at traversal starts.
    std::vector<uint64_t> h_pos(L + 1);
    GFZ_HIP(hipMemcpyAsync(h_pos.data(), pos.get(), (L + 1) * 8, hipMemcpyDeviceToHost,
                           ctx.stream()));
    ctx.sync();
    const uint64_t L_next = h_pos[L];
    DBuf<int32_t> nxt(pool, std::max<uint64_t>(L_next, 1));
    hipLaunchKernelGGL(k_enc_compact, dim3(grid_for(ctx, L)), dim3(kBlock), 0, ctx.stream(),
                       st.get(), flag.get(), pos.get(), L, nxt.get());
    GFZ_CHECK_LAUNCH();
    std::vector<uint64_t> new_starts;
    uint64_t o = 0;
    for (uint64_t t = 0; t < n_trav; ++t) {
      const uint64_t b = o, e = o + cur_len[t];
      const uint32_t nl = (uint32_t)(h_pos[e] - h_pos[b]);
      if (nl)
        new_starts.push_back(h_pos[b]);
      cur_len[t] = nl;
      o = e;
    }
    ctx.sync();
    cur = std::move(nxt);
    L = L_next;
    upload_heads(ctx, new_starts, L, heads);
  }
  out.n_encoded = L;
  out.encoded = std::move(cur);
  out.lengths = std::move(cur_len);
  ctx.sync();
}

}  // namespace gfz
