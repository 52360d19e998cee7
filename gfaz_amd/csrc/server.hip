// Serving loop: a rulebook and encoded traversals resident in HBM, extract and
// add-haplotypes requests batched as they arrive.
//
// The reference has no serving path: extract and add-haplotypes run as host
// workflows per call (src/workflows/extraction_workflow.cpp:206-424,
// src/workflows/add_haplotypes_workflow.cpp:336-592). Here one worker thread
// forms batches continuously: whenever one of the two output slots is free it
// admits queued requests in arrival order while their decoded bytes fit the
// slot (half the HBM budget), applies the batch's add requests (GPU encode with
// the stored rules, then appended to the resident state), and enqueues one
// persistent work-stealing extraction for all the batch's extract requests,
// followed by a device-to-pinned-host copy on a second stream. The worker does
// not wait for that copy: the next batch's decode runs on the compute stream
// while the previous batch's results stream to the host. A completion thread
// waits for each slot's copy in launch order, hands results to their requests
// and frees the slot.
#include "abi_error.hpp"
#include "codec.hpp"
#include "encode.hpp"
#include "resident.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;
constexpr const char *kServePrefix = "Serving error: ";
constexpr uint32_t kDefaultWaitMs = 120000;  // gfz_wait's deadline when the config gives 0

// Where a request is; gfz_wait names it when its deadline passes.
enum Stage : int { kQueued = 0, kAdding, kDecoding, kCopying, kDone };
const char *stage_name(int st) {
  switch (st) {
    case kQueued: return "queued";
    case kAdding: return "encoding and appending (add)";
    case kDecoding: return "admitted, decode being enqueued";
    case kCopying: return "decode launched, copy to host pending";
    default: return "done";
  }
}

struct Request {
  uint64_t ticket = 0;
  bool add = false;
  std::vector<uint64_t> ids;                                // extract
  std::vector<int32_t> nodes;                               // add
  std::vector<uint32_t> lengths;                            // add
  int status = GFZ_OK;
  std::string error;
  std::vector<int32_t> values;                              // extract result
  std::vector<uint64_t> offsets;                            // extract result, ids + 1
  uint64_t first_new_id = 0;                                // add result
  std::vector<uint32_t> enc_lengths;                        // add result
  uint64_t out_elems = 0, slot_off = 0;
  Clock::time_point submitted, finished;
  bool done = false;
  int stage = kQueued;  // under the server's mu
};
using ReqPtr = std::shared_ptr<Request>;

struct Slot {
  gfz::DBuf<int32_t> d_out;
  uint64_t cap = 0;  // int32 elements
  int32_t *h_out = nullptr;
  uint8_t *h_stage = nullptr;
  size_t stage_cap = 0;
  hipEvent_t computed = nullptr, copied = nullptr;
  bool busy = false;
  std::vector<ReqPtr> reqs;
  uint64_t total = 0;
};

void pinned_grow(uint8_t *&p, size_t &cap, size_t need) {
  if (need <= cap)
    return;
  if (p)
    GFZ_HIP(hipHostFree(p));
  p = nullptr;
  cap = std::max(need, 2 * cap);
  void *q = nullptr;
  GFZ_HIP(hipHostMalloc(&q, cap, hipHostMallocDefault));
  p = static_cast<uint8_t *>(q);
}

}  // namespace

struct gfz_server {
  gfz_server_config cfg{};
  std::unique_ptr<gfz::Ctx> ctx;
  hipStream_t copy_stream = nullptr;
  gfz::Rulebook rb;
  gfz::Resident res;
  gfz::DBuf<int32_t> d_first, d_second;
  std::vector<gfz::EncodeLayer> layers;
  gfz::EncodeTables enc_tables;  // add-haplotypes lookup tables, built once
  Slot slots[2];

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ReqPtr> queue;
  std::deque<int> inflight;  // slots in launch order
  std::unordered_map<uint64_t, ReqPtr> tickets;
  std::vector<uint64_t> known_len;  // decoded length per stored traversal, under mu
                                    // (its size is the stored traversal count)
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker, completer;

  // statistics
  std::vector<double> latencies_ms;
  uint64_t completed = 0, rejected = 0, batches = 0, bytes_out = 0;
  Clock::time_point first_submit{}, last_finish{};
  bool any_submit = false;

  ~gfz_server() { shutdown(); }

  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu);
      if (stop && !worker.joinable() && !completer.joinable())
        return;
      stop = true;
    }
    cv_work.notify_all();
    if (worker.joinable())
      worker.join();
    cv_work.notify_all();
    if (completer.joinable())
      completer.join();
    if (ctx) {
      gfz::DeviceGuard g(ctx->device());
      (void)hipStreamSynchronize(ctx->stream());
      if (copy_stream)
        (void)hipStreamSynchronize(copy_stream);
      for (Slot &s : slots) {
        if (s.h_out)
          (void)hipHostFree(s.h_out);
        if (s.h_stage)
          (void)hipHostFree(s.h_stage);
        if (s.computed)
          (void)hipEventDestroy(s.computed);
        if (s.copied)
          (void)hipEventDestroy(s.copied);
        s = Slot();
      }
      if (copy_stream)
        (void)hipStreamDestroy(copy_stream);
      copy_stream = nullptr;
    }
  }

  void finish(const ReqPtr &r, Clock::time_point t) {
    // caller holds mu
    r->finished = t;
    r->done = true;
    r->stage = kDone;
    ++completed;
    latencies_ms.push_back(std::chrono::duration<double, std::milli>(t - r->submitted).count());
    last_finish = t;
  }

  void run_worker() {
    gfz::DeviceGuard g(ctx->device());
    for (;;) {
      std::vector<ReqPtr> batch;
      int si = -1;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] {
          return (stop && queue.empty()) ||
                 (!queue.empty() && (!slots[0].busy || !slots[1].busy));
        });
        if (stop && queue.empty())
          return;
        si = !slots[0].busy ? 0 : 1;
        uint64_t used = 0;
        while (!queue.empty() && batch.size() < cfg.max_batch_requests) {
          ReqPtr r = queue.front();
          if (!r->add && r->out_elems > slots[si].cap) {
            if (batch.empty()) {  // alone it can never fit: refuse it
              queue.pop_front();
              r->status = GFZ_E_NOMEM;
              r->error = std::string(kServePrefix) + "request needs " +
                         std::to_string(r->out_elems * 4) +
                         " bytes of output, more than one slot of the HBM budget (" +
                         std::to_string(slots[si].cap * 4) + ")";
              ++rejected;
              finish(r, Clock::now());
              cv_done.notify_all();
              continue;
            }
            break;
          }
          if (!r->add && used + r->out_elems > slots[si].cap)
            break;  // waits for the next slot
          queue.pop_front();
          if (!r->add) {
            r->slot_off = used;
            used += r->out_elems;
          }
          r->stage = r->add ? kAdding : kDecoding;
          batch.push_back(r);
        }
        if (batch.empty())
          continue;
        slots[si].busy = true;
      }
      run_batch(si, batch);
    }
  }

  // One batch: its extract requests are planned and launched first (decode on the
  // compute stream, copy to pinned host memory on the copy stream), then its add
  // requests are encoded and appended in arrival order, stream-ordered after that
  // decode, so an extract never waits for the adds of its own batch. An extract
  // naming traversals that are not stored yet is put back in the queue when the
  // batch holds adds (a later batch sees them), and refused otherwise.
  void run_batch(int si, std::vector<ReqPtr> &batch) {
    Slot &s = slots[si];
    gfz::Ctx &c = *ctx;
    bool has_add = false;
    for (const ReqPtr &r : batch)
      has_add = has_add || r->add;
    // Extracts: validate per request, then one plan for all of them.
    std::vector<ReqPtr> ex, deferred;
    std::vector<uint64_t> ids;
    uint64_t off = 0;
    for (const ReqPtr &r : batch) {
      if (r->add)
        continue;
      bool ok = true;
      for (uint64_t id : r->ids)
        ok = ok && id < res.n_trav;
      if (!ok && has_add) {
        deferred.push_back(r);
        continue;
      }
      if (!ok) {
        std::lock_guard<std::mutex> lk(mu);
        r->status = GFZ_E_INVALID;
        r->error = std::string(kServePrefix) + "traversal index out of range (" +
                   std::to_string(res.n_trav) + " stored)";
        ++rejected;
        finish(r, Clock::now());
        continue;
      }
      uint64_t len = 0;
      for (uint64_t id : r->ids)
        len += res.dec_len[id];
      if (off + len > s.cap) {
        // Sized at submission before traversals it names were added: next batch.
        r->out_elems = len;
        deferred.push_back(r);
        continue;
      }
      r->slot_off = off;
      r->offsets.resize(r->ids.size() + 1);
      uint64_t o = 0;
      for (size_t i = 0; i < r->ids.size(); ++i) {
        r->offsets[i] = o;
        o += res.dec_len[r->ids[i]];
      }
      r->offsets[r->ids.size()] = o;
      r->out_elems = len;
      off += len;
      ids.insert(ids.end(), r->ids.begin(), r->ids.end());
      ex.push_back(r);
    }
    s.reqs = ex;
    s.total = off;
    bool launched = false;
    if (!ex.empty() && off) {
      try {
        gfz::ExtractPlan plan;
        int items = c.tuned("serve_items");
        gfz::extract_plan(res, ids.data(), ids.size(), items > 0 ? (uint32_t)items : 16u, plan,
                          nullptr);
        const size_t mb = gfz::extract_meta_bytes(plan);
        pinned_grow(s.h_stage, s.stage_cap, mb);
        gfz::DBuf<uint8_t> d_meta(c.pool(), mb);  // stream-ordered reuse after return
        gfz::extract_enqueue(c, res, plan, s.h_stage, d_meta.get(), s.d_out.get());
        GFZ_HIP(hipEventRecord(s.computed, c.stream()));
        GFZ_HIP(hipStreamWaitEvent(copy_stream, s.computed, 0));
        GFZ_HIP(hipMemcpyAsync(s.h_out, s.d_out.get(), off * 4, hipMemcpyDeviceToHost,
                               copy_stream));
        GFZ_HIP(hipEventRecord(s.copied, copy_stream));
        launched = true;
      } catch (const gfz::Error &e) {
        std::lock_guard<std::mutex> lk(mu);
        for (const ReqPtr &r : ex) {
          r->status = e.code;
          r->error = std::string(kServePrefix) + e.what();
          ++rejected;
          finish(r, Clock::now());
        }
        s.reqs.clear();
      }
    }
    {
      // The completion thread may take this slot's results while the adds run.
      std::lock_guard<std::mutex> lk(mu);
      if (launched) {
        for (const ReqPtr &r : s.reqs)
          r->stage = kCopying;
        inflight.push_back(si);
        cv_work.notify_all();
      }
    }
    // Adds, in arrival order: encode with the stored rules (tables built once) and
    // append; stream-ordered after the decode launched above.
    for (const ReqPtr &r : batch) {
      if (!r->add)
        continue;
      try {
        const uint64_t n = r->nodes.size();
        gfz::DBuf<int32_t> d_nodes(c.pool(), std::max<uint64_t>(n, 1));
        if (n)
          GFZ_HIP(hipMemcpyAsync(d_nodes.get(), r->nodes.data(), n * 4, hipMemcpyHostToDevice,
                                 c.stream()));
        gfz::EncodeResult er;
        gfz::encode_with_tables(c, d_nodes.get(), n, r->lengths.data(), r->lengths.size(),
                                enc_tables, res.delta_round, er);
        const uint64_t first = res.n_trav;
        gfz::resident_append(c, res, er.encoded.get(), er.n_encoded, er.lengths.data(),
                             er.lengths.size());
        std::lock_guard<std::mutex> lk(mu);
        known_len.insert(known_len.end(), res.dec_len.begin() + first, res.dec_len.end());
        r->first_new_id = first;
        r->enc_lengths = er.lengths;
        finish(r, Clock::now());
      } catch (const gfz::Error &e) {
        std::lock_guard<std::mutex> lk(mu);
        r->status = e.code;
        r->error = std::string(kServePrefix) + e.what();
        ++rejected;
        finish(r, Clock::now());
      }
    }
    std::lock_guard<std::mutex> lk(mu);
    ++batches;
    for (auto it = deferred.rbegin(); it != deferred.rend(); ++it) {
      (*it)->stage = kQueued;
      queue.push_front(*it);
    }
    if (!launched) {
      for (const ReqPtr &r : s.reqs)  // empty extracts (no output)
        finish(r, Clock::now());
      s.reqs.clear();
      s.busy = false;
    }
    cv_work.notify_all();
    cv_done.notify_all();
  }

  void run_completer() {
    gfz::DeviceGuard g(ctx->device());
    for (;;) {
      int si;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return !inflight.empty() || (stop && !worker_alive()); });
        if (inflight.empty())
          return;
        si = inflight.front();
      }
      Slot &s = slots[si];
      const hipError_t e = hipEventSynchronize(s.copied);
      const Clock::time_point t = Clock::now();
      std::lock_guard<std::mutex> lk(mu);
      for (const ReqPtr &r : s.reqs) {
        if (e != hipSuccess) {
          r->status = GFZ_E_HIP;
          r->error = std::string(kServePrefix) + hipGetErrorString(e);
          ++rejected;
        } else {
          r->values.assign(s.h_out + r->slot_off, s.h_out + r->slot_off + r->out_elems);
          bytes_out += r->out_elems * 4;
        }
        finish(r, t);
      }
      s.reqs.clear();
      s.busy = false;
      inflight.pop_front();
      cv_work.notify_all();
      cv_done.notify_all();
    }
  }

  bool worker_alive() const { return !worker_done; }
  bool worker_done = false;
};

extern "C" {

int gfz_server_create(const gfz_server_config *cfg, const int32_t *rules_first,
                      const int32_t *rules_second, uint64_t n_rules, uint32_t min_rule_id,
                      const gfz_layer *layers, uint64_t n_layers, const int32_t *encoded,
                      uint64_t n_encoded, const uint32_t *encoded_lengths, uint64_t n_trav,
                      int delta_round, gfz_server **out) {
  if (!cfg || !out || (n_rules && (!rules_first || !rules_second)) || (n_layers && !layers) ||
      (n_encoded && !encoded) || (n_trav && !encoded_lengths))
    return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "null argument");
  if (cfg->hbm_budget_bytes < 2 * 4096)
    return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "HBM budget too small");
  *out = nullptr;
  return gfz::guarded(kServePrefix, [&] {
    auto s = std::make_unique<gfz_server>();
    s->cfg = *cfg;
    if (!s->cfg.max_batch_requests)
      s->cfg.max_batch_requests = 256;
    s->ctx = std::make_unique<gfz::Ctx>(cfg->device);
    gfz::Ctx &c = *s->ctx;
    gfz::DeviceGuard g(c.device());
    GFZ_HIP(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    s->d_first = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(n_rules, 1));
    s->d_second = gfz::DBuf<int32_t>(c.pool(), std::max<uint64_t>(n_rules, 1));
    if (n_rules) {
      GFZ_HIP(hipMemcpyAsync(s->d_first.get(), rules_first, n_rules * 4, hipMemcpyHostToDevice,
                             c.stream()));
      GFZ_HIP(hipMemcpyAsync(s->d_second.get(), rules_second, n_rules * 4,
                             hipMemcpyHostToDevice, c.stream()));
    }
    for (uint64_t l = 0; l < n_layers; ++l)
      s->layers.push_back({layers[l].start_id, layers[l].end_id, layers[l].row_offset});
    gfz::rulebook_build(c, s->d_first.get(), s->d_second.get(), n_rules, min_rule_id, s->rb);
    gfz::encode_tables_build(c, s->d_first.get(), s->d_second.get(), n_rules,
                             s->layers.data(), s->layers.size(), s->enc_tables);
    gfz::DBuf<int32_t> d_enc(c.pool(), std::max<uint64_t>(n_encoded, 1));
    if (n_encoded)
      GFZ_HIP(hipMemcpyAsync(d_enc.get(), encoded, n_encoded * 4, hipMemcpyHostToDevice,
                             c.stream()));
    gfz::resident_load(c, s->rb, d_enc.get(), n_encoded, encoded_lengths, n_trav, delta_round,
                       s->res);
    s->known_len = s->res.dec_len;
    const uint64_t slot_elems = cfg->hbm_budget_bytes / 2 / 4;
    for (Slot &sl : s->slots) {
      sl.cap = slot_elems;
      sl.d_out = gfz::DBuf<int32_t>(c.pool(), slot_elems);
      void *p = nullptr;
      GFZ_HIP(hipHostMalloc(&p, slot_elems * 4, hipHostMallocDefault));
      sl.h_out = static_cast<int32_t *>(p);
      GFZ_HIP(hipEventCreateWithFlags(&sl.computed, hipEventDisableTiming));
      GFZ_HIP(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
    }
    c.sync();
    gfz_server *raw = s.get();
    raw->worker = std::thread([raw] {
      raw->run_worker();
      std::lock_guard<std::mutex> lk(raw->mu);
      raw->worker_done = true;
      raw->cv_work.notify_all();
    });
    raw->completer = std::thread([raw] { raw->run_completer(); });
    *out = s.release();
  });
}

void gfz_server_destroy(gfz_server *s) { delete s; }

static int submit(gfz_server *s, ReqPtr r, uint64_t *ticket) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop)
    return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "server is stopping");
  if (s->cfg.max_queue && s->queue.size() >= s->cfg.max_queue)
    return gfz::fail(GFZ_E_BUSY, std::string(kServePrefix) + "request queue is full");
  r->ticket = s->next_ticket++;
  r->submitted = Clock::now();
  if (!s->any_submit) {
    s->any_submit = true;
    s->first_submit = r->submitted;
  }
  s->tickets[r->ticket] = r;
  s->queue.push_back(std::move(r));
  *ticket = s->next_ticket - 1;
  s->cv_work.notify_all();
  return GFZ_OK;
}

int gfz_submit_extract(gfz_server *s, const uint64_t *ids, uint64_t n_ids, uint64_t *ticket) {
  if (!s || !ticket || (n_ids && !ids))
    return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "null argument");
  return gfz::guarded(kServePrefix, [&] {
    auto r = std::make_shared<Request>();
    r->ids.assign(ids, ids + n_ids);
    {
      // Output size for admission, from the lengths known now (adds only grow the store).
      std::lock_guard<std::mutex> lk(s->mu);
      for (uint64_t id : r->ids)
        r->out_elems += id < s->known_len.size() ? s->known_len[id] : 0;
    }
    const int rc = submit(s, std::move(r), ticket);
    if (rc != GFZ_OK)
      throw gfz::Error(rc, gfz::g_last_error.substr(std::strlen(kServePrefix)));
  });
}

int gfz_submit_add(gfz_server *s, const int32_t *nodes, uint64_t n_nodes,
                   const uint32_t *lengths, uint64_t n_trav, uint64_t *ticket) {
  if (!s || !ticket || (n_nodes && !nodes) || (n_trav && !lengths))
    return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "null argument");
  return gfz::guarded(kServePrefix, [&] {
    auto r = std::make_shared<Request>();
    r->add = true;
    r->nodes.assign(nodes, nodes + n_nodes);
    r->lengths.assign(lengths, lengths + n_trav);
    uint64_t sum = 0;
    for (uint32_t l : r->lengths)
      sum += l;
    if (sum != n_nodes)
      throw gfz::Error(GFZ_E_INVALID, "sum of traversal lengths != number of nodes");
    const int rc = submit(s, std::move(r), ticket);
    if (rc != GFZ_OK)
      throw gfz::Error(rc, gfz::g_last_error.substr(std::strlen(kServePrefix)));
  });
}

int gfz_poll(gfz_server *s, uint64_t ticket, int *done) {
  if (!s || !done)
    return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end())
    return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "unknown ticket");
  *done = it->second->done ? 1 : 0;
  return GFZ_OK;
}

int gfz_wait(gfz_server *s, uint64_t ticket, gfz_result *out) {
  if (!s || !out)
    return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "null argument");
  std::memset(out, 0, sizeof(*out));
  ReqPtr r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end())
      return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "unknown ticket");
    r = it->second;
    const uint32_t ms = s->cfg.wait_timeout_ms ? s->cfg.wait_timeout_ms : kDefaultWaitMs;
    if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(ms), [&] { return r->done; }))
      // The ticket stays valid: a later gfz_wait may still collect the result.
      return gfz::fail(GFZ_E_HIP, std::string(kServePrefix) + "request " +
                                      std::to_string(ticket) + " not finished after " +
                                      std::to_string(ms) + " ms (stage: " +
                                      stage_name(r->stage) + ")");
    s->tickets.erase(ticket);  // by key: submissions may have rehashed the map meanwhile
  }
  out->status = r->status;
  out->latency_ms = std::chrono::duration<double, std::milli>(r->finished - r->submitted).count();
  out->is_add = r->add ? 1 : 0;
  if (r->status != GFZ_OK)
    return gfz::fail(r->status, r->error);
  if (r->add) {
    out->first_new_id = r->first_new_id;
    out->n_items = r->enc_lengths.size();
    out->encoded_lengths = static_cast<uint32_t *>(std::malloc(std::max<size_t>(1, out->n_items) * 4));
    if (!out->encoded_lengths)
      return gfz::fail(GFZ_E_NOMEM, std::string(kServePrefix) + "host allocation failed");
    std::memcpy(out->encoded_lengths, r->enc_lengths.data(), out->n_items * 4);
    return GFZ_OK;
  }
  out->n_items = r->ids.size();
  out->n_values = r->values.size();
  out->values = static_cast<int32_t *>(std::malloc(std::max<size_t>(1, out->n_values) * 4));
  out->offsets = static_cast<uint64_t *>(std::malloc((out->n_items + 1) * 8));
  if (!out->values || !out->offsets) {
    gfz_result_free(out);
    return gfz::fail(GFZ_E_NOMEM, std::string(kServePrefix) + "host allocation failed");
  }
  std::memcpy(out->values, r->values.data(), out->n_values * 4);
  if (r->offsets.empty())
    out->offsets[0] = 0;
  else
    std::memcpy(out->offsets, r->offsets.data(), (out->n_items + 1) * 8);
  return GFZ_OK;
}

void gfz_result_free(gfz_result *r) {
  if (!r)
    return;
  std::free(r->values);
  std::free(r->offsets);
  std::free(r->encoded_lengths);
  r->values = nullptr;
  r->offsets = nullptr;
  r->encoded_lengths = nullptr;
}

int gfz_server_stats(gfz_server *s, gfz_serve_stats *out) {
  if (!s || !out)
    return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->completed = s->completed;
  out->rejected = s->rejected;
  out->batches = s->batches;
  out->bytes_out = s->bytes_out;
  out->n_traversals = s->known_len.size();  // the worker changes res without mu
  std::vector<double> l = s->latencies_ms;
  if (!l.empty()) {
    auto pct = [&](double q) {
      const size_t k = std::min(l.size() - 1, (size_t)(q * (double)(l.size() - 1) + 0.5));
      std::nth_element(l.begin(), l.begin() + k, l.end());
      return l[k];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    double sum = 0;
    for (double x : l)
      sum += x;
    out->mean_ms = sum / (double)l.size();
  }
  if (s->any_submit && s->completed) {
    const double secs = std::chrono::duration<double>(s->last_finish - s->first_submit).count();
    out->elapsed_s = secs;
    out->gbps = secs > 0 ? (double)s->bytes_out / secs / 1e9 : 0.0;
  }
  return GFZ_OK;
}

int gfz_server_reset_stats(gfz_server *s) {
  if (!s)
    return gfz::fail(GFZ_E_INVALID, std::string(kServePrefix) + "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  s->latencies_ms.clear();
  s->completed = s->rejected = s->batches = s->bytes_out = 0;
  s->any_submit = false;
  return GFZ_OK;
}

}  // extern "C"
