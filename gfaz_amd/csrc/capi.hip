// C ABI of include/gfaz_amd.h over the gfz:: device codec. Nothing throws
// across this boundary: every entry point catches, records the message for
// gfz_last_error() and returns a status code.
#include "gfaz_amd.h"

#include "codec.hpp"
#include "encode.hpp"
#include "prims.hpp"
#include "resident.hpp"
#include "text.hpp"
#include "coverage.hpp"
#include "host/catalog.hpp"

#include <optional>
#include <cctype>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <set>
#include <string>

#include "abi_error.hpp"
#include "abi_types.hpp"

namespace gfz {
thread_local std::string g_last_error;
}  // namespace gfz

namespace {

using gfz::fail;
using gfz::guarded;

constexpr const char *kCompressPrefix = "GPU compression error: ";
constexpr const char *kDecompressPrefix = "GPU decompression error: ";
constexpr const char *kRuntimePrefix = "GPU runtime error: ";
constexpr const char *kExtractPrefix = "Extraction workflow error: ";

struct CompressInternal {
  gfz::CompressResult res;
};

template <class T>
T *host_copy(gfz::Ctx &ctx, const T *d, uint64_t n) {
  T *h = static_cast<T *>(std::malloc(std::max<uint64_t>(n, 1) * sizeof(T)));
  if (!h)
    throw std::bad_alloc();
  if (n) {
    GFZ_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, ctx.stream()));
    ctx.sync();
  }
  return h;
}

// Decoded size of traversals ids[0..n) of a resident store (range-checked).
uint64_t extracted_size(const gfz_resident *r, const uint64_t *ids, uint64_t n_ids) {
  uint64_t total = 0;
  for (uint64_t i = 0; i < n_ids; ++i) {
    if (ids[i] >= r->res.n_trav)
      throw gfz::Error(GFZ_E_INVALID, "traversal index " + std::to_string(ids[i]) +
                                          " out of range (" + std::to_string(r->res.n_trav) +
                                          " stored)");
    total += r->res.dec_len[ids[i]];
  }
  return total;
}

// Resident decode of ids into d_out (device); the tile size is autotuned per
// size class, like the batch decoder's.
void extract_device(gfz::Ctx &c, const gfz_resident *r, const uint64_t *ids, uint64_t n_ids,
                    int32_t *d_out, uint64_t out_cap, uint64_t *offsets, uint64_t total) {
  uint32_t cls = 0;
  for (uint64_t n = total; n >= 4; n >>= 2)
    ++cls;
  const std::string key = "extract_items/" + std::to_string(cls);
  int items = c.tuned(key);
  if (items < 0 && c.autotune() && total >= (uint64_t(1) << 20)) {
    float best = 0;
    for (uint32_t cand : {8u, 16u, 32u}) {
      gfz::resident_extract(c, r->res, ids, n_ids, d_out, out_cap, offsets, cand);  // warm
      const float ms = c.time_ms([&] {
        gfz::resident_extract(c, r->res, ids, n_ids, d_out, out_cap, offsets, cand);
      });
      if (items < 0 || ms < best) {
        best = ms;
        items = (int)cand;
      }
    }
    c.set_tuned(key, items);
  }
  gfz::resident_extract(c, r->res, ids, n_ids, d_out, out_cap, offsets,
                        items < 0 ? 16u : (uint32_t)items);
}
}  // namespace

extern "C" {

const char *gfz_last_error(void) { return gfz::g_last_error.c_str(); }
const char *gfz_version(void) { return "gfaz_amd 0.1 (gfx950)"; }

int gfz_ctx_create(int device, gfz_ctx **out) {
  if (!out)
    return fail(GFZ_E_INVALID, "gfz_ctx_create: null output");
  *out = nullptr;
  return guarded(kRuntimePrefix, [&] {
    int n = 0;
    GFZ_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
      throw gfz::Error(GFZ_E_INVALID, "device " + std::to_string(device) + " not present (" +
                                          std::to_string(n) + " visible)");
    gfz::DeviceGuard g(device);
    auto c = std::make_unique<gfz_ctx>();
    c->ctx = std::make_unique<gfz::Ctx>(device);
    gfz::noalloc_scratch_init(device);  // the fixed scratch of gfz_extract_noalloc
    *out = c.release();
  });
}

void gfz_ctx_destroy(gfz_ctx *ctx) { delete ctx; }

int gfz_ctx_sync(gfz_ctx *ctx) {
  if (!ctx)
    return fail(GFZ_E_INVALID, "null ctx");
  return guarded(kRuntimePrefix, [&] {
    gfz::DeviceGuard g(ctx->ctx->device());
    ctx->ctx->sync();
  });
}

int gfz_ctx_set_memory_limit(gfz_ctx *ctx, uint64_t bytes) {
  if (!ctx)
    return fail(GFZ_E_INVALID, "null ctx");
  ctx->ctx->pool().set_limit(bytes);
  return GFZ_OK;
}

int gfz_ctx_set_profiling(gfz_ctx *ctx, int on) {
  if (!ctx)
    return fail(GFZ_E_INVALID, "null ctx");
  ctx->ctx->set_profiling(on != 0);
  return GFZ_OK;
}

int gfz_ctx_set_profile_groups(gfz_ctx *ctx, const char *groups) {
  if (!ctx)
    return fail(GFZ_E_INVALID, "null ctx");
  std::set<std::string> g;
  std::string cur;
  for (const char *c = groups ? groups : ""; ; ++c) {
    if (*c == ',' || *c == 0) {
      if (!cur.empty())
        g.insert(cur);
      cur.clear();
      if (*c == 0)
        break;
    } else {
      cur.push_back(*c);
    }
  }
  ctx->ctx->set_profile_groups(std::move(g));
  return GFZ_OK;
}

int gfz_ctx_reset_stats(gfz_ctx *ctx) {
  if (!ctx)
    return fail(GFZ_E_INVALID, "null ctx");
  ctx->ctx->reset_stats();
  return GFZ_OK;
}

int gfz_ctx_kernel_stat(gfz_ctx *ctx, const char *name, uint64_t *launches, double *total_ms,
                        double *bytes) {
  if (!ctx || !name)
    return fail(GFZ_E_INVALID, "null argument");
  const auto &st = ctx->ctx->stats();
  auto it = st.find(name);
  if (launches)
    *launches = it == st.end() ? 0 : it->second.launches;
  if (total_ms)
    *total_ms = it == st.end() ? 0 : it->second.total_ms;
  if (bytes)
    *bytes = it == st.end() ? 0 : it->second.bytes;
  return GFZ_OK;
}

int gfz_ctx_num_cus(gfz_ctx *ctx) { return ctx ? ctx->ctx->num_cus() : 0; }

int gfz_ctx_set_autotune(gfz_ctx *ctx, int on) {
  if (!ctx)
    return fail(GFZ_E_INVALID, "null ctx");
  ctx->ctx->set_autotune(on != 0);
  return GFZ_OK;
}

int gfz_ctx_tuned(gfz_ctx *ctx, const char *key, int *value) {
  if (!ctx || !key || !value)
    return fail(GFZ_E_INVALID, "null argument");
  *value = ctx->ctx->tuned(key);
  return GFZ_OK;
}

int gfz_ctx_set_tuned(gfz_ctx *ctx, const char *key, int value) {
  if (!ctx || !key)
    return fail(GFZ_E_INVALID, "null argument");
  ctx->ctx->set_tuned(key, value);
  return GFZ_OK;
}

int gfz_ctx_graph_stats(gfz_ctx *ctx, uint64_t out[3]) {
  if (!ctx || !out)
    return fail(GFZ_E_INVALID, "null argument");
  const gfz::Ctx::GraphStats &g = ctx->ctx->graph_stats();
  out[0] = g.launches;
  out[1] = g.instantiated;
  out[2] = g.updated;
  return GFZ_OK;
}


int gfz_autotune_compress(gfz_ctx *ctx, const int32_t *nodes, uint64_t n_nodes,
                          const uint32_t *lengths, uint64_t n_trav, int rounds, int threshold,
                          int delta_round, uint32_t flags, double *best_ms) {
  if (!ctx || (n_nodes && !nodes) || (n_trav && !lengths))
    return fail(GFZ_E_INVALID, std::string(kCompressPrefix) + "null argument");
  return guarded(kCompressPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> staged;
    const int32_t *d_nodes = nodes;
    if (!(flags & GFZ_INPUT_DEVICE)) {
      staged = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(n_nodes, 1));
      if (n_nodes)
        GFZ_HIP(hipMemcpyAsync(staged.get(), nodes, n_nodes * 4, hipMemcpyHostToDevice,
                               c.stream()));
      d_nodes = staged.get();
    }
    gfz::CompressParams p;
    p.rounds = rounds;
    p.threshold = threshold;
    p.delta_round = delta_round;
    std::vector<gfz::TuneTrial> trials;
    const double ms = gfz::autotune_compress(c, d_nodes, n_nodes, lengths, n_trav, p, trials);
    if (best_ms)
      *best_ms = ms;
  });
}

int gfz_ctx_save_tuning(gfz_ctx *ctx, const char *path) {
  if (!ctx || !path)
    return fail(GFZ_E_INVALID, std::string(kRuntimePrefix) + "null argument");
  return guarded(kRuntimePrefix, [&] {
    const gfz::Ctx &c = *ctx->ctx;
    std::string js = "{\"arch\": \"" + c.arch() + "\", \"cus\": " + std::to_string(c.num_cus()) +
                     ", \"tuned\": {";
    bool first = true;
    for (const auto &kv : c.tuned_all()) {
      js += (first ? "" : ", ") + std::string("\"") + kv.first + "\": " + std::to_string(kv.second);
      first = false;
    }
    js += "}}\n";
    FILE *f = std::fopen(path, "wb");
    if (!f)
      throw gfz::Error(GFZ_E_INVALID, std::string("cannot write ") + path);
    const bool ok = std::fwrite(js.data(), 1, js.size(), f) == js.size();
    if (std::fclose(f) != 0 || !ok)
      throw gfz::Error(GFZ_E_INVALID, std::string("cannot write ") + path);
  });
}

namespace {
// Reader of the flat JSON that gfz_ctx_save_tuning writes: "arch", "cus" and a
// "tuned" object of integer values whose keys hold no quotes or backslashes.
struct TuningFile {
  std::string arch;
  long long cus = -1;
  std::vector<std::pair<std::string, int>> tuned;
};

bool parse_tuning(const std::string &s, TuningFile &out) {
  size_t i = 0;
  auto ws = [&] {
    while (i < s.size() && std::isspace((unsigned char)s[i]))
      ++i;
  };
  auto lit = [&](char ch) {
    ws();
    if (i < s.size() && s[i] == ch) {
      ++i;
      return true;
    }
    return false;
  };
  auto str = [&](std::string &v) {
    if (!lit('"'))
      return false;
    const size_t e = s.find('"', i);
    if (e == std::string::npos || s.find('\\', i) < e)
      return false;
    v = s.substr(i, e - i);
    i = e + 1;
    return true;
  };
  auto num = [&](long long &v) {
    ws();
    size_t used = 0;
    try {
      v = std::stoll(s.substr(i, 24), &used);
    } catch (...) {
      return false;
    }
    i += used;
    return used > 0;
  };
  if (!lit('{'))
    return false;
  bool more = !lit('}');
  while (more) {
    std::string key;
    if (!str(key) || !lit(':'))
      return false;
    if (key == "arch") {
      if (!str(out.arch))
        return false;
    } else if (key == "cus") {
      if (!num(out.cus))
        return false;
    } else if (key == "tuned") {
      if (!lit('{'))
        return false;
      bool m = !lit('}');
      while (m) {
        std::string k;
        long long v;
        if (!str(k) || !lit(':') || !num(v) || v < INT32_MIN || v > INT32_MAX)
          return false;
        out.tuned.emplace_back(k, (int)v);
        if (lit(','))
          continue;
        if (!lit('}'))
          return false;
        m = false;
      }
    } else {
      return false;
    }
    if (lit(','))
      continue;
    if (!lit('}'))
      return false;
    more = false;
  }
  ws();
  return i == s.size();
}
}  // namespace

int gfz_ctx_load_tuning(gfz_ctx *ctx, const char *path) {
  if (!ctx || !path)
    return fail(GFZ_E_INVALID, std::string(kRuntimePrefix) + "null argument");
  return guarded(kRuntimePrefix, [&] {
    FILE *f = std::fopen(path, "rb");
    if (!f)
      throw gfz::Error(GFZ_E_INVALID, std::string("cannot read ") + path);
    std::string s;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0)
      s.append(buf, n);
    std::fclose(f);
    TuningFile tf;
    if (!parse_tuning(s, tf))
      throw gfz::Error(GFZ_E_INVALID, std::string("malformed tuning file ") + path);
    gfz::Ctx &c = *ctx->ctx;
    if (tf.arch != c.arch() || tf.cus != c.num_cus())
      throw gfz::Error(GFZ_E_INVALID, "tuning file is for " + tf.arch + " with " +
                                          std::to_string(tf.cus) + " CUs, this device is " +
                                          c.arch() + " with " + std::to_string(c.num_cus()));
    for (const auto &kv : tf.tuned)
      c.set_tuned(kv.first, kv.second);
  });
}

int gfz_dev_alloc(gfz_ctx *ctx, uint64_t bytes, void **out) {
  if (!ctx || !out)
    return fail(GFZ_E_INVALID, "null argument");
  return guarded(kRuntimePrefix, [&] {
    gfz::DeviceGuard g(ctx->ctx->device());
    *out = ctx->ctx->pool().get(bytes ? bytes : 1);
  });
}

int gfz_dev_free(gfz_ctx *ctx, void *ptr) {
  if (!ctx)
    return fail(GFZ_E_INVALID, "null ctx");
  ctx->ctx->pool().put(ptr);
  return GFZ_OK;
}

int gfz_memcpy_h2d(gfz_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx)
    return fail(GFZ_E_INVALID, "null ctx");
  return guarded(kRuntimePrefix, [&] {
    gfz::DeviceGuard g(ctx->ctx->device());
    if (bytes)
      GFZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->ctx->stream()));
    ctx->ctx->sync();
  });
}

int gfz_memcpy_d2h(gfz_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx)
    return fail(GFZ_E_INVALID, "null ctx");
  return guarded(kRuntimePrefix, [&] {
    gfz::DeviceGuard g(ctx->ctx->device());
    if (bytes)
      GFZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->ctx->stream()));
    ctx->ctx->sync();
  });
}

int gfz_debug_sort_u64(gfz_ctx *ctx, uint64_t *keys, uint64_t n, uint32_t local_cap) {
  if (!ctx || (n && !keys))
    return fail(GFZ_E_INVALID, "null argument");
  return guarded(kRuntimePrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<uint64_t> tmp(c.pool(), std::max<uint64_t>(n, 1));
    gfz::DBuf<uint32_t> err(c.pool(), 2);
    gfz::fill_u32(c, err.get(), 0, 2);
    gfz::sort_unique_u64(c, keys, tmp.get(), n, err.get(), local_cap);
    if ((uint32_t)c.fetch({{err.get(), 1, 0}})[0])
      throw gfz::Error(GFZ_E_INVALID, "sort look-back did not complete (internal error)");
    c.sync();
  });
}

int gfz_compress_traversals(gfz_ctx *ctx, const int32_t *nodes, uint64_t n_nodes,
                            const uint32_t *lengths, uint64_t n_trav, uint32_t n_paths,
                            int rounds, int threshold, int delta_round, uint32_t flags,
                            gfz_compress_out *out) {
  if (!ctx || !out || (n_trav && !lengths) || (n_nodes && !nodes))
    return fail(GFZ_E_INVALID, std::string(kCompressPrefix) + "null argument");
  if (n_paths > n_trav)
    return fail(GFZ_E_INVALID, std::string(kCompressPrefix) + "n_paths exceeds n_trav");
  std::memset(out, 0, sizeof(*out));
  return guarded(kCompressPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> staged;
    const int32_t *d_nodes = nodes;
    if (!(flags & GFZ_INPUT_DEVICE)) {
      staged = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(n_nodes, 1));
      if (n_nodes)
        GFZ_HIP(hipMemcpyAsync(staged.get(), nodes, n_nodes * 4, hipMemcpyHostToDevice,
                               c.stream()));
      d_nodes = staged.get();
    }
    auto in = std::make_unique<CompressInternal>();
    gfz::CompressParams p;
    p.rounds = rounds;
    p.threshold = threshold;
    p.delta_round = delta_round;
    gfz::compress_device(c, d_nodes, n_nodes, lengths, n_trav, p, in->res);
    staged.reset();
    gfz::CompressResult &r = in->res;
    out->n_encoded = r.n_encoded;
    out->n_trav = n_trav;
    out->n_rules = r.n_rules;
    out->start_id = r.start_id;
    out->end_id = r.end_id;
    out->rounds_run = (int32_t)r.rounds.size();
    out->flags = flags;
    for (size_t i = 0; i < r.rounds.size() && i < 32; ++i) {
      out->round_rules[i] = r.rounds[i].rules;
      out->round_lengths[i] = r.rounds[i].length_out;
      out->round_distinct[i] = r.rounds[i].distinct_pairs;
      out->round_slots[i] = r.rounds[i].table_slots;
    }
    out->encoded_lengths =
        static_cast<uint32_t *>(std::malloc(std::max<uint64_t>(n_trav, 1) * sizeof(uint32_t)));
    if (!out->encoded_lengths)
      throw std::bad_alloc();
    if (n_trav)
      std::memcpy(out->encoded_lengths, r.encoded_lengths.data(), n_trav * sizeof(uint32_t));
    if (flags & GFZ_OUTPUT_DEVICE) {
      out->encoded = r.encoded.get();
      out->rules_first = r.rules_first.get();
      out->rules_second = r.rules_second.get();
    } else {
      out->encoded = host_copy(c, r.encoded.get(), r.n_encoded);
      out->rules_first = host_copy(c, r.rules_first.get(), r.n_rules);
      out->rules_second = host_copy(c, r.rules_second.get(), r.n_rules);
      r.encoded.reset();
      r.rules_first.reset();
      r.rules_second.reset();
    }
    out->internal = in.release();
  });
}

void gfz_compress_free(gfz_compress_out *out) {
  if (!out)
    return;
  if (!(out->flags & GFZ_OUTPUT_DEVICE)) {
    std::free(out->encoded);
    std::free(out->rules_first);
    std::free(out->rules_second);
  }
  std::free(out->encoded_lengths);
  delete static_cast<CompressInternal *>(out->internal);
  std::memset(out, 0, sizeof(*out));
}

int gfz_rulebook_load(gfz_ctx *ctx, const int32_t *rules_first, const int32_t *rules_second,
                      uint64_t n_rules, uint32_t min_rule_id, uint32_t flags,
                      gfz_rulebook **out) {
  if (!ctx || !out || (n_rules && (!rules_first || !rules_second)))
    return fail(GFZ_E_INVALID, std::string(kDecompressPrefix) + "null argument");
  *out = nullptr;
  return guarded(kDecompressPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> f, s;
    const int32_t *df = rules_first, *ds = rules_second;
    if (!(flags & GFZ_INPUT_DEVICE)) {
      f = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(n_rules, 1));
      s = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(n_rules, 1));
      if (n_rules) {
        GFZ_HIP(hipMemcpyAsync(f.get(), rules_first, n_rules * 4, hipMemcpyHostToDevice,
                               c.stream()));
        GFZ_HIP(hipMemcpyAsync(s.get(), rules_second, n_rules * 4, hipMemcpyHostToDevice,
                               c.stream()));
      }
      df = f.get();
      ds = s.get();
    }
    auto rb = std::make_unique<gfz_rulebook>();
    rb->owner = ctx;
    // The build's first readback follows the copies of host columns, so on
    // return the caller's buffers are no longer read; the bodies finish on the
    // ctx stream, which every use of the rulebook runs on.
    gfz::rulebook_build(c, df, ds, n_rules, min_rule_id, rb->rb);
    *out = rb.release();
  });
}

void gfz_rulebook_free(gfz_rulebook *rb) { delete rb; }

int gfz_rulebook_rule_size(gfz_rulebook *rb, uint64_t rule_index, uint32_t *size) {
  if (!rb || !size)
    return fail(GFZ_E_INVALID, "null argument");
  if (rule_index >= rb->rb.n_rules)
    return fail(GFZ_E_INVALID, "rule index out of range");
  return guarded(kDecompressPrefix, [&] {
    gfz::Ctx &c = *rb->owner->ctx;
    gfz::DeviceGuard g(c.device());
    GFZ_HIP(hipMemcpyAsync(size, &rb->rb.ss.get()[rule_index].x, 4, hipMemcpyDeviceToHost,
                           c.stream()));
    c.sync();
  });
}

int gfz_rulebook_info(gfz_rulebook *rb, uint32_t *max_depth, int *has_bodies,
                      uint64_t *body_elems) {
  if (!rb)
    return fail(GFZ_E_INVALID, "null rulebook");
  if (max_depth)
    *max_depth = rb->rb.depth;
  if (has_bodies)
    *has_bodies = rb->rb.has_bodies ? 1 : 0;
  if (body_elems)
    *body_elems = rb->rb.z_elems;
  return GFZ_OK;
}

static void stage_encoded(gfz::Ctx &c, const int32_t *encoded, uint64_t n, uint32_t flags,
                          gfz::DBuf<int32_t> &staged, const int32_t *&d_enc) {
  d_enc = encoded;
  if (!(flags & GFZ_INPUT_DEVICE)) {
    staged = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(n, 1));
    if (n)
      GFZ_HIP(hipMemcpyAsync(staged.get(), encoded, n * 4, hipMemcpyHostToDevice, c.stream()));
    d_enc = staged.get();
  }
}

int gfz_decoded_size(gfz_ctx *ctx, const gfz_rulebook *rb, const int32_t *encoded,
                     uint64_t n_encoded, const uint32_t *encoded_lengths, uint64_t n_trav,
                     uint32_t flags, uint64_t *n_out) {
  if (!ctx || !rb || !n_out || (n_trav && !encoded_lengths) || (n_encoded && !encoded))
    return fail(GFZ_E_INVALID, std::string(kDecompressPrefix) + "null argument");
  if (rb->owner != ctx)
    return fail(GFZ_E_INVALID, std::string(kDecompressPrefix) + "rulebook belongs to another ctx");
  return guarded(kDecompressPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> staged;
    const int32_t *d_enc;
    stage_encoded(c, encoded, n_encoded, flags, staged, d_enc);
    gfz::DecodePlan plan;
    gfz::DBuf<uint64_t> sym_off;
    gfz::decode_plan(c, rb->rb, d_enc, n_encoded, encoded_lengths, nullptr, n_trav, -1, plan,
                     sym_off);
    *n_out = plan.n_out;
  });
}

int gfz_decompress_traversals(gfz_ctx *ctx, const gfz_rulebook *rb, const int32_t *encoded,
                              uint64_t n_encoded, const uint32_t *encoded_lengths,
                              const uint32_t *orig_lengths, uint64_t n_trav, int delta_round,
                              int32_t *out, uint64_t out_cap, uint32_t flags, uint64_t *n_out,
                              uint32_t *out_lengths, int *length_mismatch) {
  if (!ctx || !rb || (n_trav && !encoded_lengths) || (n_encoded && !encoded))
    return fail(GFZ_E_INVALID, std::string(kDecompressPrefix) + "null argument");
  if (rb->owner != ctx)
    return fail(GFZ_E_INVALID, std::string(kDecompressPrefix) + "rulebook belongs to another ctx");
  return guarded(kDecompressPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> staged;
    const int32_t *d_enc;
    stage_encoded(c, encoded, n_encoded, flags, staged, d_enc);
    gfz::DecodePlan plan;
    gfz::DBuf<uint64_t> sym_off;
    const int dr = delta_round < 0 ? 0 : delta_round;
    gfz::decode_plan(c, rb->rb, d_enc, n_encoded, encoded_lengths, orig_lengths, n_trav, dr,
                     plan, sym_off);
    if (n_out)
      *n_out = plan.n_out;
    if (length_mismatch)
      *length_mismatch = plan.length_mismatch ? 1 : 0;
    if (out_lengths)
      for (uint64_t t = 0; t < n_trav; ++t)
        out_lengths[t] = (uint32_t)(plan.out_offsets[t + 1] - plan.out_offsets[t]);
    if (plan.n_out > out_cap)
      throw gfz::Error(GFZ_E_INVALID, "output capacity " + std::to_string(out_cap) +
                                          " < decoded size " + std::to_string(plan.n_out));
    if (plan.n_out && !out)
      throw gfz::Error(GFZ_E_INVALID, "null output buffer");
    if (flags & GFZ_OUTPUT_DEVICE) {
      gfz::decode_run(c, rb->rb, d_enc, n_encoded, encoded_lengths, n_trav, dr, plan,
                      sym_off.get(), out);
      c.sync();
    } else {
      gfz::DBuf<int32_t> d_out(c.pool(), std::max<uint64_t>(plan.n_out, 1));
      gfz::decode_run(c, rb->rb, d_enc, n_encoded, encoded_lengths, n_trav, dr, plan,
                      sym_off.get(), d_out.get());
      if (plan.n_out)
        GFZ_HIP(hipMemcpyAsync(out, d_out.get(), plan.n_out * 4, hipMemcpyDeviceToHost,
                               c.stream()));
      c.sync();
    }
  });
}

// ---- streaming decompress ------------------------------------------------------

int gfz_decompress_stream(gfz_ctx *ctx, const gfz_rulebook *rb, const int32_t *encoded,
                          uint64_t n_encoded, const uint32_t *encoded_lengths, uint64_t n_trav,
                          int delta_round, uint64_t chunk_bytes, uint32_t flags,
                          gfz_chunk_consumer consumer, void *user) {
  if (!ctx || !rb || !consumer || (n_trav && !encoded_lengths) || (n_encoded && !encoded))
    return fail(GFZ_E_INVALID, std::string(kDecompressPrefix) + "null argument");
  if (rb->owner != ctx)
    return fail(GFZ_E_INVALID, std::string(kDecompressPrefix) + "rulebook belongs to another ctx");
  return guarded(kDecompressPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> staged;
    const int32_t *d_enc;
    stage_encoded(c, encoded, n_encoded, flags, staged, d_enc);
    gfz::Resident res;
    gfz::resident_load(c, rb->rb, d_enc, n_encoded, encoded_lengths, n_trav, delta_round, res);
    gfz::decompress_stream(c, res, chunk_bytes ? chunk_bytes : (uint64_t(1) << 30),
                           [&](const gfz::StreamChunk &k) {
                             gfz_chunk ch{k.values, k.n_values, k.trav_begin, k.trav_end,
                                          k.expanded_begin, k.lengths};
                             return consumer(&ch, user) == 0;
                           });
  });
}

// ---- add-haplotypes encoding ---------------------------------------------------

namespace {
constexpr const char *kAddPrefix = "Add-haplotypes workflow error: ";
struct EncodeInternal {
  gfz::EncodeResult res;
};
}  // namespace

int gfz_encode_with_rules(gfz_ctx *ctx, const int32_t *nodes, uint64_t n_nodes,
                          const uint32_t *lengths, uint64_t n_trav, const int32_t *rules_first,
                          const int32_t *rules_second, uint64_t n_rules, const gfz_layer *layers,
                          uint64_t n_layers, int delta_round, uint32_t flags,
                          gfz_encode_out *out) {
  if (!ctx || !out || (n_trav && !lengths) || (n_nodes && !nodes) ||
      (n_rules && (!rules_first || !rules_second)) || (n_layers && !layers))
    return fail(GFZ_E_INVALID, std::string(kAddPrefix) + "null argument");
  std::memset(out, 0, sizeof(*out));
  return guarded(kAddPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> sn, sf, ss;
    const int32_t *dn = nodes, *df = rules_first, *ds = rules_second;
    if (!(flags & GFZ_INPUT_DEVICE)) {
      sn = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(n_nodes, 1));
      sf = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(n_rules, 1));
      ss = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(n_rules, 1));
      if (n_nodes)
        GFZ_HIP(hipMemcpyAsync(sn.get(), nodes, n_nodes * 4, hipMemcpyHostToDevice, c.stream()));
      if (n_rules) {
        GFZ_HIP(hipMemcpyAsync(sf.get(), rules_first, n_rules * 4, hipMemcpyHostToDevice,
                               c.stream()));
        GFZ_HIP(hipMemcpyAsync(ss.get(), rules_second, n_rules * 4, hipMemcpyHostToDevice,
                               c.stream()));
      }
      dn = sn.get();
      df = sf.get();
      ds = ss.get();
    }
    std::vector<gfz::EncodeLayer> ls(n_layers);
    for (uint64_t l = 0; l < n_layers; ++l)
      ls[l] = gfz::EncodeLayer{layers[l].start_id, layers[l].end_id, layers[l].row_offset};
    auto in = std::make_unique<EncodeInternal>();
    gfz::encode_with_rules(c, dn, n_nodes, lengths, n_trav, df, ds, n_rules, ls.data(), n_layers,
                           delta_round < 0 ? 0 : delta_round, in->res);
    out->n_encoded = in->res.n_encoded;
    out->n_trav = n_trav;
    out->flags = flags;
    out->encoded_lengths = static_cast<uint32_t *>(std::malloc(std::max<uint64_t>(n_trav, 1) * 4));
    if (!out->encoded_lengths)
      throw std::bad_alloc();
    std::memcpy(out->encoded_lengths, in->res.lengths.data(), n_trav * 4);
    if (flags & GFZ_OUTPUT_DEVICE) {
      out->encoded = in->res.encoded.get();
      out->internal = in.release();
    } else {
      out->encoded = host_copy(c, in->res.encoded.get(), in->res.n_encoded);
    }
  });
}

void gfz_encode_free(gfz_encode_out *out) {
  if (!out)
    return;
  std::free(out->encoded_lengths);
  if (out->internal)
    delete static_cast<EncodeInternal *>(out->internal);
  else
    std::free(out->encoded);
  std::memset(out, 0, sizeof(*out));
}

// ---- resident decoding / extract -------------------------------------------

int gfz_resident_load(gfz_ctx *ctx, const gfz_rulebook *rb, const int32_t *encoded,
                      uint64_t n_encoded, const uint32_t *encoded_lengths, uint64_t n_trav,
                      int delta_round, uint32_t flags, gfz_resident **out) {
  if (!ctx || !rb || !out || (n_trav && !encoded_lengths) || (n_encoded && !encoded))
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "null argument");
  if (rb->owner != ctx)
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "rulebook belongs to another ctx");
  *out = nullptr;
  return guarded(kExtractPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> staged;
    const int32_t *d_enc;
    stage_encoded(c, encoded, n_encoded, flags, staged, d_enc);
    auto r = std::make_unique<gfz_resident>();
    r->owner = ctx;
    r->rulebook = rb;
    gfz::resident_load(c, rb->rb, d_enc, n_encoded, encoded_lengths, n_trav, delta_round, r->res);
    *out = r.release();
  });
}

void gfz_resident_free(gfz_resident *r) {
  if (!r)
    return;
  // Never throws: at process exit the runtime may be gone (the device can no
  // longer be made current); then only the host side is released.
  std::optional<gfz::DeviceGuard> g;
  try {
    g.emplace(r->owner->ctx->device());
    (void)hipStreamSynchronize(r->owner->ctx->stream());
  } catch (const gfz::Error &) {
  }
  delete r;
}

int gfz_resident_append(gfz_ctx *ctx, gfz_resident *r, const int32_t *encoded,
                        uint64_t n_encoded, const uint32_t *encoded_lengths, uint64_t n_trav,
                        uint32_t flags, uint64_t *first_new_id) {
  if (!ctx || !r || (n_trav && !encoded_lengths) || (n_encoded && !encoded))
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "null argument");
  if (r->owner != ctx)
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "resident state belongs to another ctx");
  return guarded(kExtractPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> staged;
    const int32_t *d_enc;
    stage_encoded(c, encoded, n_encoded, flags, staged, d_enc);
    if (first_new_id)
      *first_new_id = r->res.n_trav;
    gfz::resident_append(c, r->res, d_enc, n_encoded, encoded_lengths, n_trav);
  });
}

int gfz_resident_info(const gfz_resident *r, uint64_t *n_trav, uint64_t *n_decoded) {
  if (!r)
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "null resident");
  if (n_trav)
    *n_trav = r->res.n_trav;
  if (n_decoded)
    *n_decoded = r->res.n_out;
  return GFZ_OK;
}

int gfz_resident_lengths(const gfz_resident *r, uint32_t *decoded_lengths) {
  if (!r || !decoded_lengths)
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "null argument");
  for (uint64_t t = 0; t < r->res.n_trav; ++t)
    decoded_lengths[t] = (uint32_t)r->res.dec_len[t];
  return GFZ_OK;
}

int gfz_extract(gfz_ctx *ctx, const gfz_resident *r, const uint64_t *ids, uint64_t n_ids,
                int32_t *out, uint64_t out_cap, uint32_t flags, uint64_t *offsets,
                uint64_t *n_out) {
  if (!ctx || !r || (n_ids && !ids))
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "null argument");
  if (r->owner != ctx)
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "resident state belongs to another ctx");
  return guarded(kExtractPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    const uint64_t total = extracted_size(r, ids, n_ids);
    if (n_out)
      *n_out = total;
    if (total > out_cap)
      throw gfz::Error(GFZ_E_INVALID, "output capacity " + std::to_string(out_cap) +
                                          " < extracted size " + std::to_string(total));
    if (total && !out)
      throw gfz::Error(GFZ_E_INVALID, "null output buffer");
    gfz::DBuf<int32_t> tmp;
    int32_t *d_out = out;
    if (!(flags & GFZ_OUTPUT_DEVICE)) {
      tmp = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(total, 1));
      d_out = tmp.get();
    }
    extract_device(c, r, ids, n_ids, d_out, out_cap, offsets, total);
    if (!(flags & GFZ_OUTPUT_DEVICE) && total) {
      GFZ_HIP(hipMemcpyAsync(out, d_out, total * 4, hipMemcpyDeviceToHost, c.stream()));
      c.sync();
    }
  });
}

// ---- extraction front-end: catalog lookups + resident decode + line text ------

}  // extern "C"

namespace {

// Decodes the specs' traversals and formats their lines on the GPU.
void extract_lines(gfz_ctx *ctx, const gfz_resident *r, int kind,
                   const std::vector<gfz::LineSpec> &specs, gfz_lines *out) {
  gfz::Ctx &c = *ctx->ctx;
  gfz::DeviceGuard g(c.device());
  const uint64_t n = specs.size();
  if (!n) {
    out->text = static_cast<uint8_t *>(std::malloc(1));
    out->line_off = static_cast<uint64_t *>(std::calloc(1, 8));
    if (!out->text || !out->line_off)
      throw std::bad_alloc();
    return;
  }
  std::vector<uint64_t> ids(n), head_off(n + 1, 0), tail_off(n + 1, 0);
  std::vector<uint32_t> lens(n);
  std::string heads, tails;
  for (uint64_t i = 0; i < n; ++i) {
    ids[i] = specs[i].trav_id;
    heads += specs[i].head;
    tails += specs[i].tail;
    head_off[i + 1] = heads.size();
    tail_off[i + 1] = tails.size();
  }
  const uint64_t total = extracted_size(r, ids.data(), n);
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t len = r->res.dec_len[ids[i]];
    if (len > UINT32_MAX)
      throw gfz::Error(GFZ_E_INVALID, "traversal " + std::to_string(ids[i]) +
                                          " is too long for a line (" + std::to_string(len) +
                                          " steps)");
    lens[i] = (uint32_t)len;
  }
  gfz::DBuf<int32_t> d_nodes(c.pool(), std::max<uint64_t>(total, 1));
  extract_device(c, r, ids.data(), n, d_nodes.get(), total, nullptr, total);
  gfz::FormatPlan plan;
  gfz::format_plan(c, kind, d_nodes.get(), total, lens.data(), n, head_off.data(),
                   tail_off.data(), plan);
  gfz::DBuf<uint8_t> d_txt(c.pool(), std::max<uint64_t>(plan.total, 1));
  gfz::format_write(c, plan, reinterpret_cast<const uint8_t *>(heads.data()), head_off.data(),
                    reinterpret_cast<const uint8_t *>(tails.data()), tail_off.data(),
                    d_txt.get(), plan.total);
  out->text = static_cast<uint8_t *>(std::malloc(std::max<uint64_t>(plan.total, 1)));
  out->line_off = static_cast<uint64_t *>(std::malloc((n + 1) * 8));
  if (!out->text || !out->line_off)
    throw std::bad_alloc();
  if (plan.total)
    GFZ_HIP(hipMemcpyAsync(out->text, d_txt.get(), plan.total, hipMemcpyDeviceToHost,
                           c.stream()));
  std::memcpy(out->line_off, plan.line_off.data(), (n + 1) * 8);
  out->n_bytes = plan.total;
  out->n_lines = n;
  c.sync();
}

template <class F>
int lines_entry(gfz_ctx *ctx, const gfz_resident *r, const gfz_catalog *cat, gfz_lines *out,
                int kind, F specs) {
  if (!ctx || !r || !cat || !out)
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "null argument");
  if (r->owner != ctx)
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "resident state belongs to another ctx");
  std::memset(out, 0, sizeof(*out));
  return guarded(kExtractPrefix, [&] {
    try {
      extract_lines(ctx, r, kind, specs(), out);
    } catch (...) {
      gfz_lines_free(out);
      throw;
    }
  });
}

}  // namespace

extern "C" {

int gfz_extract_path_lines(gfz_ctx *ctx, const gfz_resident *r, const gfz_catalog *c,
                           const gfz_strings *names, gfz_lines *out) {
  return lines_entry(ctx, r, c, out, gfz::kTextP, [&] {
    return c->cat.path_lines(gfz::strings_from_abi(names, "path names"));
  });
}

int gfz_extract_walk_lines(gfz_ctx *ctx, const gfz_resident *r, const gfz_catalog *c,
                           const gfz_walk_keys *keys, gfz_lines *out) {
  return lines_entry(ctx, r, c, out, gfz::kTextW,
                     [&] { return c->cat.walk_lines(gfz::walk_keys_from_abi(keys)); });
}

int gfz_extract_walk_lines_by_name(gfz_ctx *ctx, const gfz_resident *r, const gfz_catalog *c,
                                   const gfz_strings *names, gfz_lines *out) {
  return lines_entry(ctx, r, c, out, gfz::kTextW, [&] {
    return c->cat.walk_lines_by_name(gfz::strings_from_abi(names, "walk names"));
  });
}

void gfz_lines_free(gfz_lines *l) {
  if (!l)
    return;
  std::free(l->text);
  std::free(l->line_off);
  std::memset(l, 0, sizeof(*l));
}

// ---- P/W line step text ------------------------------------------------------

static_assert(gfz::kTextP == GFZ_TEXT_P && gfz::kTextW == GFZ_TEXT_W, "line kinds");

int gfz_format_steps(gfz_ctx *ctx, int kind, const int32_t *nodes, uint64_t n_nodes,
                     const uint32_t *lengths, uint64_t n_trav, const uint8_t *heads,
                     const uint64_t *head_off, const uint8_t *tails, const uint64_t *tail_off,
                     uint32_t flags, uint8_t *out, uint64_t out_cap, uint64_t *n_out,
                     uint64_t *line_off) {
  constexpr const char *kPrefix = "GFA writer error: ";
  if (!ctx || (n_nodes && !nodes) || (n_trav && !lengths) || !head_off || !tail_off ||
      (head_off[n_trav] && !heads) || (tail_off[n_trav] && !tails))
    return fail(GFZ_E_INVALID, std::string(kPrefix) + "null argument");
  return guarded(kPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> staged;
    const int32_t *d_nodes;
    stage_encoded(c, nodes, n_nodes, flags, staged, d_nodes);
    gfz::FormatPlan plan;
    gfz::format_plan(c, kind, d_nodes, n_nodes, lengths, n_trav, head_off, tail_off, plan);
    const uint64_t total = plan.total;
    if (n_out)
      *n_out = total;
    if (line_off)
      std::memcpy(line_off, plan.line_off.data(), (n_trav + 1) * 8);
    if (!out)
      return;
    if (total > out_cap)
      throw gfz::Error(GFZ_E_INVALID, "output capacity " + std::to_string(out_cap) +
                                          " < text size " + std::to_string(total));
    if (flags & GFZ_OUTPUT_DEVICE) {
      gfz::format_write(c, plan, heads, head_off, tails, tail_off, out, out_cap);
    } else {
      gfz::DBuf<uint8_t> d_out(c.pool(), std::max<uint64_t>(total, 1));
      gfz::format_write(c, plan, heads, head_off, tails, tail_off, d_out.get(), total);
      if (total)
        GFZ_HIP(hipMemcpyAsync(out, d_out.get(), total, hipMemcpyDeviceToHost, c.stream()));
      c.sync();
    }
  });
}

int gfz_parse_steps(gfz_ctx *ctx, int kind, const uint8_t *text, uint64_t n_bytes,
                    const uint64_t *field_begin, const uint64_t *field_end, uint64_t n_lines,
                    uint32_t flags, int32_t *nodes, uint64_t nodes_cap, uint32_t *lengths,
                    uint64_t *n_nodes) {
  constexpr const char *kPrefix = "GFA parser error: ";
  if (!ctx || (n_bytes && !text) || (n_lines && (!field_begin || !field_end)))
    return fail(GFZ_E_INVALID, std::string(kPrefix) + "null argument");
  return guarded(kPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<uint8_t> staged;
    const uint8_t *d_txt = text;
    if (!(flags & GFZ_INPUT_DEVICE)) {
      staged = gfz::DBuf<uint8_t>(c.pool(), std::max<uint64_t>(n_bytes, 1));
      if (n_bytes)
        GFZ_HIP(hipMemcpyAsync(staged.get(), text, n_bytes, hipMemcpyHostToDevice, c.stream()));
      d_txt = staged.get();
    }
    gfz::ParsePlan plan;
    gfz::parse_plan(c, kind, d_txt, n_bytes, field_begin, field_end, n_lines, lengths, plan);
    const uint64_t total = plan.total;
    if (n_nodes)
      *n_nodes = total;
    if (!nodes)
      return;
    if (total > nodes_cap)
      throw gfz::Error(GFZ_E_INVALID, "node capacity " + std::to_string(nodes_cap) + " < steps " +
                                          std::to_string(total));
    if (flags & GFZ_OUTPUT_DEVICE) {
      gfz::parse_write(c, plan, nodes, nodes_cap);
    } else {
      gfz::DBuf<int32_t> d_nodes(c.pool(), std::max<uint64_t>(total, 1));
      gfz::parse_write(c, plan, d_nodes.get(), total);
      if (total)
        GFZ_HIP(hipMemcpyAsync(nodes, d_nodes.get(), total * 4, hipMemcpyDeviceToHost,
                               c.stream()));
      c.sync();
    }
  });
}

// ---- node coverage ---------------------------------------------------------------

}  // extern "C"

struct gfz_coverage {
  gfz_ctx *owner = nullptr;
  gfz::Coverage cov;
};

extern "C" {

int gfz_coverage_build(gfz_ctx *ctx, const int32_t *nodes, uint64_t n_nodes,
                       const uint32_t *lengths, uint64_t n_trav, const uint32_t *group_of,
                       uint32_t n_groups, uint32_t num_nodes, uint32_t flags,
                       gfz_coverage **out) {
  constexpr const char *kPrefix = "Coverage error: ";
  if (!ctx || !out || (n_nodes && !nodes) || (n_trav && (!lengths || !group_of)))
    return fail(GFZ_E_INVALID, std::string(kPrefix) + "null argument");
  *out = nullptr;
  return guarded(kPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::DBuf<int32_t> staged;
    const int32_t *d_nodes;
    stage_encoded(c, nodes, n_nodes, flags, staged, d_nodes);
    auto cv = std::make_unique<gfz_coverage>();
    cv->owner = ctx;
    gfz::coverage_build(c, d_nodes, n_nodes, lengths, n_trav, group_of, n_groups, num_nodes,
                        cv->cov);
    *out = cv.release();
  });
}

void gfz_coverage_free(gfz_coverage *cov) {
  if (!cov)
    return;
  // Never throws: at process exit the runtime may be gone (the device can no
  // longer be made current); then only the host side is released.
  std::optional<gfz::DeviceGuard> g;
  try {
    g.emplace(cov->owner->ctx->device());
    (void)hipStreamSynchronize(cov->owner->ctx->stream());
  } catch (const gfz::Error &) {
  }
  delete cov;
}

int gfz_coverage_hist(gfz_coverage *cov, uint64_t *hist) {
  if (!cov || !hist)
    return fail(GFZ_E_INVALID, "Growth workflow error: null argument");
  return guarded("Growth workflow error: ", [&] {
    gfz::Ctx &c = *cov->owner->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::coverage_hist(c, cov->cov, hist);
  });
}

int gfz_coverage_windows(gfz_coverage *cov, const uint64_t *window_off,
                         const uint32_t *window_nodes, uint64_t n_windows,
                         const uint32_t *segment_lengths, uint64_t *denominators,
                         uint64_t *numerators) {
  if (!cov || !window_off || !numerators || (n_windows && window_off[n_windows] &&
                                             (!window_nodes || !segment_lengths)))
    return fail(GFZ_E_INVALID, "PAV workflow error: null argument");
  return guarded("PAV workflow error: ", [&] {
    gfz::Ctx &c = *cov->owner->ctx;
    gfz::DeviceGuard g(c.device());
    gfz::coverage_windows(c, cov->cov, window_off, window_nodes, n_windows, segment_lengths,
                          denominators, numerators);
  });
}

static_assert(gfz::kNoAllocScratchBytes == GFZ_NOALLOC_SCRATCH_BYTES,
              "the header states the scratch array's size");

int gfz_extract_noalloc(gfz_ctx *ctx, const gfz_resident *r, const uint64_t *ids,
                        uint64_t n_ids, int32_t *out, uint64_t out_cap, uint64_t *offsets,
                        uint64_t *n_out) {
  if (!ctx || !r || (n_ids && !ids))
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "null argument");
  if (r->owner != ctx)
    return fail(GFZ_E_INVALID, std::string(kExtractPrefix) + "resident state belongs to another ctx");
  return guarded(kExtractPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    uint64_t total = 0;
    for (uint64_t i = 0; i < n_ids; ++i) {
      if (ids[i] >= r->res.n_trav)
        throw gfz::Error(GFZ_E_INVALID, "traversal index " + std::to_string(ids[i]) +
                                            " out of range (" + std::to_string(r->res.n_trav) +
                                            " stored)");
      total += r->res.dec_len[ids[i]];
    }
    if (n_out)
      *n_out = total;
    if (total && !out)
      throw gfz::Error(GFZ_E_INVALID, "null output buffer");
    // The tile size the autotuner chose for gfz_extract in this size class, if any;
    // no timing runs here.
    uint32_t cls = 0;
    for (uint64_t n = total; n >= 4; n >>= 2)
      ++cls;
    const int items = c.tuned("extract_items/" + std::to_string(cls));
    gfz::resident_extract_noalloc(c, r->res, ids, n_ids, out, out_cap, offsets,
                                  items < 0 ? 16u : (uint32_t)items);
  });
}

}  // extern "C"
