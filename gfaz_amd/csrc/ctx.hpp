// Device context: one per (device, caller). Owns a stream, a caching pool of
// device buffers (workspaces are reused across calls instead of hipMalloc'd per
// stage) and optional per-kernel HIP-event timing on that stream.
#pragma once

#include "common.hpp"

#include <initializer_list>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

namespace gfz {

class Pool {
 public:
  Pool() = default;
  ~Pool();
  void *get(size_t bytes);
  void put(void *p);
  void trim();
  size_t cached_bytes() const { return cached_; }
  size_t live_bytes() const { return live_; }
  // Optional cap on live + cached bytes (0 = none); get() throws kErrOutOfMemory
  // past it after trimming the cache.
  void set_limit(size_t bytes) { limit_ = bytes; }
  // While set, get() never frees cached blocks to make room (it throws
  // kErrOutOfMemory instead): during stream capture a cached block may still be
  // named by captured work that has not run yet.
  void set_keep_cache(bool on) { keep_cache_ = on; }

 private:
  void release_cache();  // caller holds mu_
  std::multimap<size_t, void *> free_;
  std::unordered_map<void *, size_t> size_of_;
  size_t cached_ = 0, live_ = 0, limit_ = 0;
  bool keep_cache_ = false;
  std::mutex mu_;
};

// RAII device buffer from a Pool.
template <class T>
class DBuf {
 public:
  DBuf() = default;
  DBuf(Pool &pool, size_t n) : pool_(&pool), n_(n) {
    ptr_ = static_cast<T *>(pool.get(std::max<size_t>(1, n) * sizeof(T)));
  }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept { swap(o); }
  DBuf &operator=(DBuf &&o) noexcept {
    reset();
    swap(o);
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() {
    if (ptr_ && pool_)
      pool_->put(ptr_);
    ptr_ = nullptr;
    n_ = 0;
  }
  void swap(DBuf &o) noexcept {
    std::swap(pool_, o.pool_);
    std::swap(ptr_, o.ptr_);
    std::swap(n_, o.n_);
  }
  T *get() const { return ptr_; }
  size_t size() const { return n_; }
  // Hand ownership to the caller (returned to the pool by Pool::put later).
  T *release() {
    T *p = ptr_;
    ptr_ = nullptr;
    n_ = 0;
    return p;
  }

 private:
  Pool *pool_ = nullptr;
  T *ptr_ = nullptr;
  size_t n_ = 0;
};

struct KernelStat {
  uint64_t launches = 0;
  double total_ms = 0;
  double bytes = 0;  // algorithmic bytes accumulated by the caller
};

class Ctx {
 public:
  explicit Ctx(int device);
  ~Ctx();
  int device() const { return device_; }
  const std::string &arch() const { return arch_; }  // e.g. "gfx950:sramecc+:xnack-"
  hipStream_t stream() const { return stream_; }
  Pool &pool() { return pool_; }
  int num_cus() const { return num_cus_; }

  // Per-kernel timing: when enabled, launches wrapped in timed() record a
  // start/stop event pair on the context stream; collect() accumulates the
  // elapsed times after the stream has been synchronised.
  // An event pair drains the stream around its launches (several us per
  // pair): set_profile_groups limits timing to the named groups (empty = all).
  void set_profiling(bool on) { profiling_ = on; }
  bool profiling() const { return profiling_; }
  void set_profile_groups(std::set<std::string> groups) { only_ = std::move(groups); }
  template <class F>
  void timed(const char *name, F &&launch) {
    if (!profiling_ || capturing_ || (!only_.empty() && !only_.count(name))) {
      launch();
      return;
    }
    hipEvent_t a = event(), b = event();
    GFZ_HIP(hipEventRecord(a, stream_));
    launch();
    GFZ_HIP(hipEventRecord(b, stream_));
    pending_.push_back({name, a, b});
  }
  void sync();  // hipStreamSynchronize + collect()
  void collect();
  void reset_stats() { stats_.clear(); }
  const std::map<std::string, KernelStat> &stats() const { return stats_; }
  void add_bytes(const char *name, double bytes) {
    if (profiling_)
      stats_[name].bytes += bytes;
  }

  // Device-to-host readback of a few 64-bit words without a stream
  // synchronisation: one single-wave kernel copies up to 4 runs into a
  // coherent pinned mailbox, then stores a sequence number after them, which
  // the host polls (querying the stream for errors while it waits). Returns
  // the mailbox: run r lands at words [at, at + n), at + n <= 512.
  struct FetchRun {
    const void *dev;
    uint32_t n;
    uint32_t at;
  };
  const uint64_t *fetch(std::initializer_list<FetchRun> runs);

  // Pinned host staging buffer of at least `bytes`, reused across calls.
  void *pinned(size_t bytes);
  // A second one, for columns handed between the host and the device in pieces
  // (framing); separate so the codec's own staging never lands in it.
  void *column_staging(size_t bytes);
  // Small pinned area (512 x 8 bytes) for device-to-host scalar readbacks.
  uint64_t *scalars();
  // Copy n 64-bit device words into scalars()[at..) (no sync).
  void stage_scalars(const void *dev, size_t n, size_t at = 0) {
    GFZ_HIP(hipMemcpyAsync(scalars() + at, dev, n * 8, hipMemcpyDeviceToHost, stream_));
  }

  // Autotuned launch choices (tile sizes and the like), keyed by kernel and
  // size class. When autotuning is on, the first call that needs an untuned
  // key times every candidate on its own input and keeps the fastest; all
  // candidates compute the same result.
  void set_autotune(bool on) { autotune_ = on; }
  bool autotune() const { return autotune_; }
  int tuned(const std::string &key) const {
    auto it = tuned_.find(key);
    return it == tuned_.end() ? -1 : it->second;
  }
  void set_tuned(const std::string &key, int v) { tuned_[key] = v; }
  // Size class of n elements (log4), the suffix of size-dependent tuned keys.
  static uint32_t size_class(uint64_t n) {
    uint32_t c = 0;
    for (; n >= 4; n >>= 2)
      ++c;
    return c;
  }
  const std::map<std::string, int> &tuned_all() const { return tuned_; }
  // HIP graphs of launch sequences. With graphs on, segment(key, fn) captures
  // the work fn puts on the context stream (no host synchronisation may happen
  // inside fn), updates the executable graph kept under `key` to it (or
  // instantiates one when the sequence's shape changed) and launches it; with
  // graphs off it just runs fn. Timed groups are not timed while capturing.
  void set_graphs(bool on) { graphs_ = on; }
  bool graphs() const { return graphs_; }
  template <class F>
  void segment(const std::string &key, F &&fn) {
    if (!graphs_) {
      fn();
      return;
    }
    capture_begin();
    try {
      fn();
    } catch (...) {
      capture_abort();
      throw;
    }
    capture_end_launch(key);
  }
  struct GraphStats {
    uint64_t launches = 0, instantiated = 0, updated = 0;
  };
  const GraphStats &graph_stats() const { return graph_stats_; }

  // Elapsed milliseconds of `launch` on the context stream (synchronises).
  template <class F>
  float time_ms(F &&launch) {
    hipEvent_t a = event(), b = event();
    GFZ_HIP(hipEventRecord(a, stream_));
    launch();
    GFZ_HIP(hipEventRecord(b, stream_));
    GFZ_HIP(hipEventSynchronize(b));
    float ms = 0;
    GFZ_HIP(hipEventElapsedTime(&ms, a, b));
    free_events_.push_back(a);
    free_events_.push_back(b);
    return ms;
  }

 private:
  hipEvent_t event();
  void capture_begin();
  void capture_abort();
  void capture_end_launch(const std::string &key);
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  int device_;
  std::string arch_;
  int num_cus_ = 256;
  hipStream_t stream_ = nullptr;
  Pool pool_;
  bool profiling_ = false;
  uint64_t *mail_ = nullptr;      // coherent pinned mailbox (host address)
  uint64_t *mail_dev_ = nullptr;  // the same memory as the device sees it
  uint64_t mail_seq_ = 0;
  std::set<std::string> only_;
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> free_events_;
  std::map<std::string, KernelStat> stats_;
  bool autotune_ = true;
  std::map<std::string, int> tuned_;
  void *pinned_ = nullptr;
  uint64_t *scal_ = nullptr;
  size_t pinned_bytes_ = 0;
  void *staging_ = nullptr;
  size_t staging_bytes_ = 0;
  bool graphs_ = false, capturing_ = false;
  std::map<std::string, hipGraphExec_t> execs_;
  GraphStats graph_stats_;
};

// RAII: make `ctx`'s device current for the duration of a call.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    GFZ_HIP(hipGetDevice(&prev));
    if (prev != dev)
      GFZ_HIP(hipSetDevice(dev));
  }
  ~DeviceGuard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && prev >= 0 && cur != prev)
      (void)hipSetDevice(prev);
  }
};

}  // namespace gfz
