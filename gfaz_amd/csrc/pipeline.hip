// The traversal codec with its zstd framing overlapped: compress + frame and
// unframe + decompress as one call each.
//
// The reference frames the columns one after another on the host after the GPU
// codec (finalize_traversal_columns, src/gpu/compression/path_compression_gpu_rolling.cu:
// 29-56; CPU: src/workflows/compression_workflow.cpp:281-383) and unframes them all
// before the first upload (src/gpu/decompression/decompression_primitives_gpu.cu:18-56).
// Here a column crosses between host and device in 4 MiB pieces and every column
// has its own host thread:
//   compress: the four columns' pieces are copied device-to-host in one
//     interleaved sequence on the ctx stream, an event after each; each column's
//     thread frames its pieces as their events complete (ColumnFramer: whole zstd
//     blocks of the one pinned buffer, so the frame has the bytes of a single
//     ZSTD_compress2 call);
//   decompress: the four payloads decode at once; the rule columns go up first and
//     the rulebook builds on the GPU while the traversal payloads are still being
//     decoded, each thread uploading its column's pieces on its own copy stream as
//     they come out of zstd.
#include "abi_error.hpp"
#include "abi_types.hpp"
#include "codec.hpp"
#include "host/frame.hpp"

#include <chrono>
#include <cstring>
#include <exception>
#include <thread>
#include <vector>

namespace {

constexpr const char *kCompressPrefix = "GPU compression error: ";
constexpr const char *kDecompressPrefix = "GPU decompression error: ";

// Handoff piece: a whole number of zstd blocks.
constexpr uint64_t kPieceBytes = uint64_t(4) << 20;
static_assert(kPieceBytes % gfz::kFrameBlockBytes == 0, "pieces hold whole zstd blocks");

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// fn on a thread of its own; join() rethrows what it threw. The destructor joins,
// so a Task never outlives the locals its function refers to.
class Task {
 public:
  template <class F>
  explicit Task(F fn) {
    th_ = std::thread([this, fn] {
      try {
        fn();
      } catch (...) {
        err_ = std::current_exception();
      }
    });
  }
  Task(const Task &) = delete;
  Task &operator=(const Task &) = delete;
  ~Task() {
    if (th_.joinable())
      th_.join();
  }
  void join() {
    if (th_.joinable())
      th_.join();
    if (err_)
      std::rethrow_exception(err_);
  }

 private:
  std::thread th_;
  std::exception_ptr err_;
};

// Events and streams of one call, released when the call ends.
struct HipHandles {
  std::vector<hipEvent_t> events;
  std::vector<hipStream_t> streams;
  hipEvent_t event() {
    hipEvent_t e;
    GFZ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    events.push_back(e);
    return e;
  }
  hipStream_t stream() {
    hipStream_t s;
    GFZ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    streams.push_back(s);
    return s;
  }
  ~HipHandles() {
    for (hipStream_t s : streams) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    for (hipEvent_t e : events)
      (void)hipEventDestroy(e);
  }
};

// One column on its way to the host: device source, pinned destination, the
// event recorded after each of its pieces.
struct OutColumn {
  const int32_t *dev = nullptr;
  uint64_t n = 0;  // int32 values
  bool rule = false;
  int32_t *host = nullptr;
  gfz_block *block = nullptr;
  std::vector<hipEvent_t> piece_done;
  uint64_t pieces() const { return (n * 4 + kPieceBytes - 1) / kPieceBytes; }
};

void frame_column(OutColumn &col, int threads) {
  gfz::ColumnFramer framer(col.host, col.n * 4, threads);
  const uint64_t per = kPieceBytes / 4;
  int32_t carry = 0;  // raw value before the piece (rule columns)
  for (uint64_t p = 0; p < col.piece_done.size(); ++p) {
    GFZ_HIP(hipEventSynchronize(col.piece_done[p]));
    const uint64_t a = p * per, b = std::min(col.n, a + per);
    if (col.rule) {
      // process_rules' delta over the whole column (src/codec/codec.cpp:79-86),
      // one piece at a time.
      int32_t *h = col.host;
      const int32_t last = h[b - 1];
      for (uint64_t i = b - 1; i > a; --i)
        h[i] = (int32_t)((uint32_t)h[i] - (uint32_t)h[i - 1]);
      if (a > 0)
        h[a] = (int32_t)((uint32_t)h[a] - (uint32_t)carry);
      carry = last;
    }
    framer.ready(b * 4);
  }
  framer.finish(col.block);
}

void check_abi(int rc) {
  if (rc != GFZ_OK)
    throw gfz::Error(rc, gfz_last_error());
}

}  // namespace

extern "C" {

int gfz_compress_framed(gfz_ctx *ctx, const int32_t *nodes, uint64_t n_nodes,
                        const uint32_t *lengths, uint64_t n_trav, uint32_t n_paths, int rounds,
                        int threshold, int delta_round, uint32_t flags, int threads,
                        gfz_framed *out) {
  if (!ctx || !out || (n_trav && !lengths) || (n_nodes && !nodes))
    return gfz::fail(GFZ_E_INVALID, std::string(kCompressPrefix) + "null argument");
  if (n_paths > n_trav)
    return gfz::fail(GFZ_E_INVALID, std::string(kCompressPrefix) + "n_paths exceeds n_trav");
  std::memset(out, 0, sizeof(*out));
  return gfz::guarded(kCompressPrefix, [&] {
    try {
      gfz::Ctx &c = *ctx->ctx;
      gfz::DeviceGuard g(c.device());
      const auto t0 = Clock::now();
      gfz::CompressResult r;
      {
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
        gfz::compress_device(c, d_nodes, n_nodes, lengths, n_trav, p, r);
      }
      out->gpu_ms = ms_since(t0);
      out->start_id = r.start_id;
      out->end_id = r.end_id;
      out->n_rules = r.n_rules;
      out->rounds_run = (int32_t)r.rounds.size();
      out->delta_round = delta_round < 0 ? 0 : delta_round;
      out->n_trav = n_trav;
      out->n_paths = n_paths;
      out->encoded_lengths =
          static_cast<uint32_t *>(std::malloc(std::max<uint64_t>(n_trav, 1) * 4));
      if (!out->encoded_lengths)
        throw std::bad_alloc();
      if (n_trav)
        std::memcpy(out->encoded_lengths, r.encoded_lengths.data(), n_trav * 4);
      uint64_t split = 0;
      for (uint32_t t = 0; t < n_paths; ++t)
        split += r.encoded_lengths[t];

      OutColumn cols[4];
      cols[0] = {r.rules_first.get(), r.n_rules, true, nullptr, &out->rules_first, {}};
      cols[1] = {r.rules_second.get(), r.n_rules, true, nullptr, &out->rules_second, {}};
      cols[2] = {r.encoded.get(), split, false, nullptr, &out->paths, {}};
      cols[3] = {r.encoded.get() + split, r.n_encoded - split, false, nullptr, &out->walks, {}};
      uint64_t total = 0;
      for (auto &col : cols)
        total += col.n;
      int32_t *host = static_cast<int32_t *>(c.column_staging(std::max<uint64_t>(total, 1) * 4));
      uint64_t at = 0, max_pieces = 0;
      for (auto &col : cols) {
        col.host = host + at;
        at += col.n;
        max_pieces = std::max(max_pieces, col.pieces());
      }
      // Every copy and event is queued before a thread waits on one.
      HipHandles hip;
      for (uint64_t p = 0; p < max_pieces; ++p)
        for (auto &col : cols) {
          if (p >= col.pieces())
            continue;
          const uint64_t a = p * (kPieceBytes / 4),
                         b = std::min(col.n, a + kPieceBytes / 4);
          GFZ_HIP(hipMemcpyAsync(col.host + a, col.dev + a, (b - a) * 4,
                                 hipMemcpyDeviceToHost, c.stream()));
          hipEvent_t e = hip.event();
          GFZ_HIP(hipEventRecord(e, c.stream()));
          col.piece_done.push_back(e);
        }
      {
        Task t0([&] { frame_column(cols[0], threads); });
        Task t1([&] { frame_column(cols[1], threads); });
        Task t2([&] { frame_column(cols[2], threads); });
        Task t3([&] { frame_column(cols[3], threads); });
        t0.join();
        t1.join();
        t2.join();
        t3.join();
      }
      c.sync();
      out->total_ms = ms_since(t0);
    } catch (...) {
      gfz_framed_free(out);
      throw;
    }
  });
}

void gfz_framed_free(gfz_framed *f) {
  if (!f)
    return;
  gfz_block_free(&f->rules_first);
  gfz_block_free(&f->rules_second);
  gfz_block_free(&f->paths);
  gfz_block_free(&f->walks);
  std::free(f->encoded_lengths);
  f->encoded_lengths = nullptr;
}

int gfz_decompress_framed(gfz_ctx *ctx, const gfz_framed *in, const uint32_t *orig_lengths,
                          int32_t *out, uint64_t out_cap, uint32_t flags, uint64_t *n_out,
                          uint32_t *out_lengths, int *length_mismatch, double *timings_ms) {
  if (!ctx || !in || (in->n_trav && !in->encoded_lengths))
    return gfz::fail(GFZ_E_INVALID, std::string(kDecompressPrefix) + "null argument");
  if (in->n_paths > in->n_trav)
    return gfz::fail(GFZ_E_INVALID, std::string(kDecompressPrefix) + "n_paths exceeds n_trav");
  return gfz::guarded(kDecompressPrefix, [&] {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    const auto t0 = Clock::now();
    const uint64_t R = in->n_rules;
    if (in->rules_first.original_size != R * 4 || in->rules_second.original_size != R * 4)
      throw gfz::Error(GFZ_E_CORRUPT, "rule column sizes disagree with the rule count");
    if (in->paths.original_size % 4 || in->walks.original_size % 4)
      throw gfz::Error(GFZ_E_CORRUPT, "int32 column size is not a multiple of 4");
    const uint64_t np = in->paths.original_size / 4, nw = in->walks.original_size / 4;
    uint64_t total = 0, in_paths = 0;
    for (uint64_t t = 0; t < in->n_trav; ++t) {
      total += in->encoded_lengths[t];
      if (t < in->n_paths)
        in_paths += in->encoded_lengths[t];
    }
    if (in_paths != np || total != np + nw)
      throw gfz::Error(GFZ_E_CORRUPT, "encoded lengths disagree with the column sizes");

    int32_t *host =
        static_cast<int32_t *>(c.column_staging(std::max<uint64_t>(2 * R + total, 1) * 4));
    int32_t *h_rf = host, *h_rs = host + R, *h_enc = host + 2 * R;
    gfz::DBuf<int32_t> d_rf(c.pool(), std::max<uint64_t>(R, 1));
    gfz::DBuf<int32_t> d_rs(c.pool(), std::max<uint64_t>(R, 1));
    gfz::DBuf<int32_t> d_enc(c.pool(), std::max<uint64_t>(total, 1));
    HipHandles hip;
    hipStream_t s_paths = hip.stream(), s_walks = hip.stream();
    const int dev = c.device();

    auto rule_column = [](const gfz_block &b, int32_t *dst, uint64_t n) {
      check_abi(gfz_unframe_int32(b.payload, b.payload_size, b.original_size, dst, n));
      for (uint64_t i = 1; i < n; ++i)  // delta_decode_int32 (src/codec/codec.cpp:88-95)
        dst[i] = (int32_t)((uint32_t)dst[i] + (uint32_t)dst[i - 1]);
    };
    auto traversal_column = [&](const gfz_block &b, int32_t *h, int32_t *d, hipStream_t s) {
      gfz::DeviceGuard tg(dev);
      gfz::ColumnUnframer u(b.payload, b.payload_size, h, b.original_size);
      uint64_t sent = 0;
      const uint8_t *hb = reinterpret_cast<const uint8_t *>(h);
      uint8_t *db = reinterpret_cast<uint8_t *>(d);
      while (!u.done()) {
        const uint64_t got = u.next(kPieceBytes);
        if (got > sent) {
          GFZ_HIP(hipMemcpyAsync(db + sent, hb + sent, got - sent, hipMemcpyHostToDevice, s));
          sent = got;
        }
      }
    };

    double t_rules = 0, t_unframed = 0;
    {
      Task tp([&] { traversal_column(in->paths, h_enc, d_enc.get(), s_paths); });
      Task tw([&] { traversal_column(in->walks, h_enc + np, d_enc.get() + np, s_walks); });
      {
        Task ta([&] { rule_column(in->rules_first, h_rf, R); });
        Task tb([&] { rule_column(in->rules_second, h_rs, R); });
        ta.join();
        tb.join();
      }
      if (R) {
        GFZ_HIP(hipMemcpyAsync(d_rf.get(), h_rf, R * 4, hipMemcpyHostToDevice, c.stream()));
        GFZ_HIP(hipMemcpyAsync(d_rs.get(), h_rs, R * 4, hipMemcpyHostToDevice, c.stream()));
      }
      gfz::Rulebook rb_build;
      gfz::rulebook_build(c, d_rf.get(), d_rs.get(), R, in->start_id, rb_build);
      t_rules = ms_since(t0);
      tp.join();
      tw.join();
      t_unframed = ms_since(t0);
      GFZ_HIP(hipStreamSynchronize(s_paths));
      GFZ_HIP(hipStreamSynchronize(s_walks));

      const int dr = in->delta_round < 0 ? 0 : in->delta_round;
      gfz::DecodePlan plan;
      gfz::DBuf<uint64_t> sym_off;
      gfz::decode_plan(c, rb_build, d_enc.get(), total, in->encoded_lengths, orig_lengths,
                       in->n_trav, dr, plan, sym_off);
      if (n_out)
        *n_out = plan.n_out;
      if (length_mismatch)
        *length_mismatch = plan.length_mismatch ? 1 : 0;
      if (out_lengths)
        for (uint64_t t = 0; t < in->n_trav; ++t)
          out_lengths[t] = (uint32_t)(plan.out_offsets[t + 1] - plan.out_offsets[t]);
      if (plan.n_out > out_cap)
        throw gfz::Error(GFZ_E_INVALID, "output capacity " + std::to_string(out_cap) +
                                            " < decoded size " + std::to_string(plan.n_out));
      if (plan.n_out && !out)
        throw gfz::Error(GFZ_E_INVALID, "null output buffer");
      if (flags & GFZ_OUTPUT_DEVICE) {
        gfz::decode_run(c, rb_build, d_enc.get(), total, in->encoded_lengths, in->n_trav, dr,
                        plan, sym_off.get(), out);
        c.sync();
      } else {
        gfz::DBuf<int32_t> d_out(c.pool(), std::max<uint64_t>(plan.n_out, 1));
        gfz::decode_run(c, rb_build, d_enc.get(), total, in->encoded_lengths, in->n_trav, dr,
                        plan, sym_off.get(), d_out.get());
        if (plan.n_out)
          GFZ_HIP(hipMemcpyAsync(out, d_out.get(), plan.n_out * 4, hipMemcpyDeviceToHost,
                                 c.stream()));
        c.sync();
      }
    }
    if (timings_ms) {
      timings_ms[0] = t_rules;
      timings_ms[1] = t_unframed;
      timings_ms[2] = ms_since(t0);
    }
  });
}

}  // extern "C"
