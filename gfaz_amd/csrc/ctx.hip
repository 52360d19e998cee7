#include "ctx.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>

namespace gfz {

Pool::~Pool() { trim(); }

// Requests are rounded to size classes (256-byte multiples up to 4 KB, then
// quarter steps between powers of two: at most 25% slack), so blocks freed by
// one call serve the slightly different sizes of the next.
static size_t size_class(size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  if (bytes <= 4096)
    return bytes;
  size_t p = 1;
  while (p <= bytes / 2)
    p <<= 1;  // largest power of two <= bytes
  const size_t step = p / 4;
  return (bytes + step - 1) / step * step;
}

void *Pool::get(size_t bytes) {
  bytes = size_class(bytes);
  std::lock_guard<std::mutex> lk(mu_);
  // Smallest cached block that fits and wastes at most half of itself.
  auto it = free_.lower_bound(bytes);
  if (it != free_.end() && it->first <= 2 * bytes) {
    void *p = it->second;
    cached_ -= it->first;
    live_ += it->first;
    free_.erase(it);
    return p;
  }
  if (limit_ && live_ + cached_ + bytes > limit_) {
    if (keep_cache_)
      throw Error(kErrOutOfMemory, "device pool limit reached during graph capture");
    release_cache();
    if (live_ + bytes > limit_)
      throw Error(kErrOutOfMemory, "device pool limit reached (" + std::to_string(live_ + bytes) +
                                       " > " + std::to_string(limit_) + " bytes)");
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (keep_cache_)
      throw Error(kErrOutOfMemory,
                  "hipMalloc of " + std::to_string(bytes) + " bytes failed during graph capture");
    release_cache();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(kErrOutOfMemory, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
  }
  size_of_[p] = bytes;
  live_ += bytes;
  return p;
}

void Pool::put(void *p) {
  if (!p)
    return;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = size_of_.find(p);
  if (it == size_of_.end())
    return;
  live_ -= it->second;
  cached_ += it->second;
  free_.emplace(it->second, p);
}

void Pool::release_cache() {
  for (auto &kv : free_) {
    (void)hipFree(kv.second);
    size_of_.erase(kv.second);
  }
  free_.clear();
  cached_ = 0;
}

void Pool::trim() {
  std::lock_guard<std::mutex> lk(mu_);
  release_cache();
}

Ctx::Ctx(int device) : device_(device) {
  DeviceGuard g(device);
  GFZ_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    num_cus_ = std::max(1, prop.multiProcessorCount);
    arch_ = prop.gcnArchName;
  }
}

Ctx::~Ctx() {
  (void)hipSetDevice(device_);
  if (stream_)
    (void)hipStreamSynchronize(stream_);
  for (auto &kv : execs_)
    (void)hipGraphExecDestroy(kv.second);
  for (auto &p : pending_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : free_events_)
    (void)hipEventDestroy(e);
  if (pinned_)
    (void)hipHostFree(pinned_);
  if (staging_)
    (void)hipHostFree(staging_);
  if (scal_)
    (void)hipHostFree(scal_);
  if (mail_)
    (void)hipHostFree(mail_);
  pool_.trim();
  if (stream_)
    (void)hipStreamDestroy(stream_);
}

// Stream capture in relaxed mode: the pool may still allocate (hipMalloc is not
// stream work), but keeps its cached blocks, which captured work may name.
void Ctx::capture_begin() {
  GFZ_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
  capturing_ = true;
  pool_.set_keep_cache(true);
}

void Ctx::capture_abort() {
  hipGraph_t g = nullptr;
  (void)hipStreamEndCapture(stream_, &g);
  if (g)
    (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  capturing_ = false;
  pool_.set_keep_cache(false);
}

void Ctx::capture_end_launch(const std::string &key) {
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(stream_, &g);
  capturing_ = false;
  pool_.set_keep_cache(false);
  if (e != hipSuccess || !g) {
    (void)hipGetLastError();
    throw Error(kErrHip, std::string("graph capture failed: ") + hipGetErrorString(e));
  }
  hipGraphExec_t &x = execs_[key];
  bool ok = false;
  if (x) {
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    ok = hipGraphExecUpdate(x, g, &err_node, &res) == hipSuccess &&
         res == hipGraphExecUpdateSuccess;
    if (ok) {
      ++graph_stats_.updated;
    } else {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(x);
      x = nullptr;
    }
  }
  if (!ok) {
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      x = nullptr;
      (void)hipGraphDestroy(g);
      throw Error(kErrHip, std::string("graph instantiation failed: ") + hipGetErrorString(ei));
    }
    ++graph_stats_.instantiated;
  }
  (void)hipGraphDestroy(g);
  GFZ_HIP(hipGraphLaunch(x, stream_));
  ++graph_stats_.launches;
}

hipEvent_t Ctx::event() {
  if (!free_events_.empty()) {
    hipEvent_t e = free_events_.back();
    free_events_.pop_back();
    return e;
  }
  hipEvent_t e;
  GFZ_HIP(hipEventCreate(&e));
  return e;
}

void Ctx::sync() {
  GFZ_HIP(hipStreamSynchronize(stream_));
  collect();
}

void Ctx::collect() {
  for (auto &p : pending_) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto &s = stats_[p.name];
      s.launches += 1;
      s.total_ms += ms;
    }
    free_events_.push_back(p.a);
    free_events_.push_back(p.b);
  }
  pending_.clear();
}

namespace {
constexpr uint32_t kMailWords = 512;  // the sequence number follows them
constexpr uint32_t kMaxFetch = 4;
struct FetchArgs {
  const uint64_t *src[kMaxFetch];
  uint32_t n[kMaxFetch];
  uint32_t at[kMaxFetch];
  uint32_t count;
};

// One wave: its stores, then a system-scope fence, then the sequence number.
__global__ __launch_bounds__(kWave) void k_fetch(FetchArgs a, uint64_t *mail, uint64_t seq) {
  for (uint32_t r = 0; r < a.count; ++r)
    for (uint32_t i = threadIdx.x; i < a.n[r]; i += kWave)
      mail[a.at[r] + i] = a.src[r][i];
  __threadfence_system();
  if (threadIdx.x == 0)
    __hip_atomic_store(mail + kMailWords, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
}  // namespace

const uint64_t *Ctx::fetch(std::initializer_list<FetchRun> runs) {
  if (capturing_)  // the captured work has not run: its mailbox would never fill
    throw Error(kErrInvalid, "host readback inside a graph segment (internal error)");
  if (!mail_) {
    void *p = nullptr;
    GFZ_HIP(hipHostMalloc(&p, (kMailWords + 1) * 8, hipHostMallocMapped | hipHostMallocCoherent));
    mail_ = static_cast<uint64_t *>(p);
    std::memset(mail_, 0, (kMailWords + 1) * 8);
    void *d = nullptr;
    GFZ_HIP(hipHostGetDevicePointer(&d, p, 0));
    mail_dev_ = static_cast<uint64_t *>(d);
  }
  FetchArgs a{};
  for (const FetchRun &r : runs) {
    if (a.count == kMaxFetch || r.at + r.n > kMailWords)
      throw Error(kErrInvalid, "fetch: too many runs or words for the mailbox");
    a.src[a.count] = static_cast<const uint64_t *>(r.dev);
    a.n[a.count] = r.n;
    a.at[a.count] = r.at;
    ++a.count;
  }
  const uint64_t seq = ++mail_seq_;
  hipLaunchKernelGGL(k_fetch, dim3(1), dim3(kWave), 0, stream_, a, mail_dev_, seq);
  GFZ_HIP(hipGetLastError());
  const auto t0 = std::chrono::steady_clock::now();
  for (uint64_t polls = 1;; ++polls) {
    if (__atomic_load_n(mail_ + kMailWords, __ATOMIC_ACQUIRE) == seq)
      break;
    if (polls % 256 == 0) {
      const hipError_t e = hipStreamQuery(stream_);
      if (e == hipSuccess) {
        if (__atomic_load_n(mail_ + kMailWords, __ATOMIC_ACQUIRE) == seq)
          break;
        throw Error(kErrHip, "readback mailbox not written by a completed stream");
      }
      if (e != hipErrorNotReady) {
        (void)hipGetLastError();
        throw Error(kErrHip, std::string("stream failed during a readback: ") +
                                 hipGetErrorString(e));
      }
      if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(500))
        std::this_thread::yield();
    }
    __builtin_ia32_pause();
  }
  return mail_;
}

uint64_t *Ctx::scalars() {
  if (!scal_) {
    void *p = nullptr;
    GFZ_HIP(hipHostMalloc(&p, 512 * 8, hipHostMallocDefault));
    scal_ = static_cast<uint64_t *>(p);
  }
  return scal_;
}

void *Ctx::pinned(size_t bytes) {
  if (bytes <= pinned_bytes_)
    return pinned_;
  if (pinned_) {
    GFZ_HIP(hipStreamSynchronize(stream_));
    GFZ_HIP(hipHostFree(pinned_));
    pinned_ = nullptr;
    pinned_bytes_ = 0;
  }
  GFZ_HIP(hipHostMalloc(&pinned_, bytes, hipHostMallocDefault));
  pinned_bytes_ = bytes;
  return pinned_;
}

void *Ctx::column_staging(size_t bytes) {
  if (bytes <= staging_bytes_)
    return staging_;
  if (staging_) {
    // Calls that copy through it drain their copy streams before returning.
    GFZ_HIP(hipStreamSynchronize(stream_));
    GFZ_HIP(hipHostFree(staging_));
    staging_ = nullptr;
    staging_bytes_ = 0;
  }
  GFZ_HIP(hipHostMalloc(&staging_, bytes, hipHostMallocDefault));
  staging_bytes_ = bytes;
  return staging_;
}

}  // namespace gfz
