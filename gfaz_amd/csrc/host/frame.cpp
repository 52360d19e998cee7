// Host framing of the traversal columns: what the reference CPU workflow does
// around the grammar before serialisation, with the same bytes.
//   rule columns:  process_rules (src/workflows/compression_utils.cpp:165-193)
//                  = first/second in ID order, each delta-encoded over the
//                  whole column (src/codec/codec.cpp:79-86), then zstd;
//   int32 columns: zstd_compress_int32_vector (src/codec/codec.cpp:180-217):
//                  ZSTD_compress2 at level 9 (or GFA_COMPRESSION_ZSTD_LEVEL in
//                  1..22) with nbWorkers = GFA_COMPRESSION_ZSTD_WORKERS if > 0,
//                  else the thread count (at least 1);
//   and the inverses (src/codec/codec.cpp:88-95,224-244).
// zstd is the system libzstd the reference links by default (CMakeLists.txt:14).
#include "abi_error.hpp"
#include "frame.hpp"

// The block-level (buffer-less) API used for piecewise framing.
#define ZSTD_STATIC_LINKING_ONLY
#include <zstd.h>

#include <algorithm>

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr const char *kFramePrefix = "Traversal framing error: ";

int zstd_level() {
  if (const char *e = std::getenv("GFA_COMPRESSION_ZSTD_LEVEL")) {
    const int v = std::atoi(e);
    if (v >= 1 && v <= 22)
      return v;
  }
  return 9;
}

int zstd_workers(int threads) {
  if (const char *e = std::getenv("GFA_COMPRESSION_ZSTD_WORKERS")) {
    const int v = std::atoi(e);
    if (v > 0)
      return v;
  }
  if (threads > 0)
    return threads;
#ifdef _OPENMP
  return std::max(1, omp_get_max_threads());
#else
  return 1;
#endif
}

void frame_bytes(const void *data, size_t bytes, int threads, gfz_block *out) {
  gfz::ColumnFramer f(data, bytes, threads);
  f.finish(out);
}

void unframe_bytes(const uint8_t *payload, uint64_t payload_size, uint64_t original_size,
                   void *dst, uint64_t cap_bytes) {
  if (!original_size)
    return;
  if (original_size > cap_bytes)
    throw gfz::Error(GFZ_E_INVALID, "output holds " + std::to_string(cap_bytes) + " < " +
                                        std::to_string(original_size) + " bytes");
  gfz::ColumnUnframer u(payload, payload_size, dst, original_size);
  u.next(original_size);
}

void delta_encode(std::vector<int32_t> &v) {
  for (size_t i = v.size(); i-- > 1;)
    v[i] = (int32_t)((uint32_t)v[i] - (uint32_t)v[i - 1]);
}

void delta_decode(int32_t *v, uint64_t n) {
  for (uint64_t i = 1; i < n; ++i)
    v[i] = (int32_t)((uint32_t)v[i] + (uint32_t)v[i - 1]);
}

}  // namespace

namespace gfz {

// ---- ColumnFramer -------------------------------------------------------------

struct ColumnFramer::Impl {
  const uint8_t *base = nullptr;
  uint64_t total = 0;
  uint64_t fed = 0;          // bytes handed to zstd so far
  bool blockwise = false;    // block-level API (else one ZSTD_compress2 at the end)
  bool ended = false;
  ZSTD_CCtx *cctx = nullptr;
  uint8_t *buf = nullptr;    // frame being written
  uint64_t cap = 0, pos = 0;

  ~Impl() {
    if (cctx)
      ZSTD_freeCCtx(cctx);
    std::free(buf);
  }

  void fail_zstd(size_t r, const char *what) {
    throw Error(GFZ_E_INVALID, std::string(what) + ZSTD_getErrorName(r));
  }
};

ColumnFramer::ColumnFramer(const void *base, uint64_t bytes, int threads)
    : impl_(std::make_unique<Impl>()) {
  Impl &m = *impl_;
  m.base = static_cast<const uint8_t *>(base);
  m.total = bytes;
  if (!bytes)
    return;
  m.cctx = ZSTD_createCCtx();
  if (!m.cctx)
    throw Error(GFZ_E_NOMEM, "unable to create a zstd compression context");
  const int level = zstd_level();
  const size_t r = ZSTD_CCtx_setParameter(m.cctx, ZSTD_c_compressionLevel, level);
  if (ZSTD_isError(r))
    m.fail_zstd(r, "zstd compression failed: ");
  // The worker count is refused by a libzstd built without threads; ZSTD_compress2
  // then compresses on the calling thread, which the block-level API reproduces.
  m.blockwise = ZSTD_isError(ZSTD_CCtx_setParameter(m.cctx, ZSTD_c_nbWorkers,
                                                    zstd_workers(threads)));
  m.cap = ZSTD_compressBound(bytes);
  m.buf = static_cast<uint8_t *>(std::malloc(m.cap));
  if (!m.buf)
    throw std::bad_alloc();
  if (m.blockwise) {
    // ZSTD_compress2 derives its parameters from the level and the known size
    // and writes the content size into the frame header; so do these.
    const ZSTD_parameters p = ZSTD_getParams(level, bytes, 0);
    const size_t b = ZSTD_compressBegin_advanced(m.cctx, nullptr, 0, p, bytes);
    if (ZSTD_isError(b))
      m.fail_zstd(b, "zstd compression failed: ");
  }
}

ColumnFramer::~ColumnFramer() = default;

void ColumnFramer::ready(uint64_t bytes_ready) {
  Impl &m = *impl_;
  if (!m.total || m.ended)
    return;
  bytes_ready = std::min(bytes_ready, m.total);
  if (!m.blockwise) {
    if (bytes_ready < m.total)
      return;
    const size_t n = ZSTD_compress2(m.cctx, m.buf, m.cap, m.base, m.total);
    if (ZSTD_isError(n))
      m.fail_zstd(n, "zstd compression failed: ");
    m.pos = n;
    m.fed = m.total;
    m.ended = true;
    return;
  }
  const uint64_t upto =
      bytes_ready == m.total ? m.total : bytes_ready / kFrameBlockBytes * kFrameBlockBytes;
  if (upto <= m.fed)
    return;
  const bool last = upto == m.total;
  const size_t n =
      last ? ZSTD_compressEnd(m.cctx, m.buf + m.pos, m.cap - m.pos, m.base + m.fed, upto - m.fed)
           : ZSTD_compressContinue(m.cctx, m.buf + m.pos, m.cap - m.pos, m.base + m.fed,
                                   upto - m.fed);
  if (ZSTD_isError(n))
    m.fail_zstd(n, "zstd compression failed: ");
  m.pos += n;
  m.fed = upto;
  m.ended = last;
}

void ColumnFramer::finish(gfz_block *out) {
  Impl &m = *impl_;
  ready(m.total);
  out->original_size = m.total;
  out->payload = nullptr;
  out->payload_size = 0;
  if (!m.total)
    return;
  out->payload = m.buf;
  out->payload_size = m.pos;
  m.buf = nullptr;
}

// ---- ColumnUnframer -----------------------------------------------------------

struct ColumnUnframer::Impl {
  const uint8_t *src = nullptr;
  uint64_t src_size = 0, ip = 0;
  uint8_t *dst = nullptr;
  uint64_t total = 0, op = 0;
  bool single_frame = false;
  bool finished = false;
  ZSTD_DCtx *dctx = nullptr;

  ~Impl() {
    if (dctx)
      ZSTD_freeDCtx(dctx);
  }
};

ColumnUnframer::ColumnUnframer(const uint8_t *payload, uint64_t payload_size, void *dst,
                               uint64_t original_size)
    : impl_(std::make_unique<Impl>()) {
  Impl &m = *impl_;
  m.src = payload;
  m.src_size = payload_size;
  m.dst = static_cast<uint8_t *>(dst);
  m.total = original_size;
  if (!original_size) {
    m.finished = true;
    return;
  }
  if (!payload)
    throw Error(GFZ_E_CORRUPT, "empty zstd payload for a non-empty column");
  // A payload of exactly one frame is decoded block by block straight into dst;
  // anything else (several frames) in one ZSTD_decompress call, as the reference.
  const size_t fsize = ZSTD_findFrameCompressedSize(payload, payload_size);
  m.single_frame = !ZSTD_isError(fsize) && fsize == payload_size;
  m.dctx = ZSTD_createDCtx();
  if (!m.dctx)
    throw Error(GFZ_E_NOMEM, "unable to create a zstd decompression context");
  if (m.single_frame) {
    const size_t r = ZSTD_decompressBegin(m.dctx);
    if (ZSTD_isError(r))
      throw Error(GFZ_E_CORRUPT, std::string("zstd decompression failed: ") +
                                     ZSTD_getErrorName(r));
  }
}

ColumnUnframer::~ColumnUnframer() = default;

bool ColumnUnframer::done() const { return impl_->finished; }

uint64_t ColumnUnframer::next(uint64_t min_bytes) {
  Impl &m = *impl_;
  if (m.finished)
    return m.op;
  auto bad = [](size_t r) {
    throw Error(GFZ_E_CORRUPT, std::string("zstd decompression failed: ") + ZSTD_getErrorName(r));
  };
  if (!m.single_frame) {
    const size_t n = ZSTD_decompressDCtx(m.dctx, m.dst, m.total, m.src, m.src_size);
    if (ZSTD_isError(n))
      bad(n);
    if (n != m.total)
      throw Error(GFZ_E_CORRUPT, "zstd payload decodes to " + std::to_string(n) +
                                     " bytes, block says " + std::to_string(m.total));
    m.op = m.total;
    m.finished = true;
    return m.op;
  }
  const uint64_t goal = m.op + std::max<uint64_t>(min_bytes, 1);
  size_t want = ZSTD_nextSrcSizeToDecompress(m.dctx);
  while (want && m.op < goal) {
    if (want > m.src_size - m.ip)
      throw Error(GFZ_E_CORRUPT, "zstd payload is truncated");
    const size_t n =
        ZSTD_decompressContinue(m.dctx, m.dst + m.op, m.total - m.op, m.src + m.ip, want);
    if (ZSTD_isError(n))
      bad(n);
    m.ip += want;
    m.op += n;
    want = ZSTD_nextSrcSizeToDecompress(m.dctx);
  }
  if (!want) {
    if (m.op != m.total)
      throw Error(GFZ_E_CORRUPT, "zstd payload decodes to " + std::to_string(m.op) +
                                     " bytes, block says " + std::to_string(m.total));
    m.finished = true;
  }
  return m.op;
}

}  // namespace gfz

extern "C" {

int gfz_frame_int32(const int32_t *data, uint64_t n, int threads, gfz_block *out) {
  if (!out || (n && !data))
    return gfz::fail(GFZ_E_INVALID, std::string(kFramePrefix) + "null argument");
  return gfz::guarded(kFramePrefix, [&] { frame_bytes(data, n * 4, threads, out); });
}

int gfz_frame_rule_columns(const int32_t *first, const int32_t *second, uint64_t n_rules,
                           int threads, gfz_block *out_first, gfz_block *out_second) {
  if (!out_first || !out_second || (n_rules && (!first || !second)))
    return gfz::fail(GFZ_E_INVALID, std::string(kFramePrefix) + "null argument");
  return gfz::guarded(kFramePrefix, [&] {
    out_first->payload = out_second->payload = nullptr;
    std::vector<int32_t> f(first, first + n_rules), s(second, second + n_rules);
    delta_encode(f);
    delta_encode(s);
    frame_bytes(f.data(), n_rules * 4, threads, out_first);
    try {
      frame_bytes(s.data(), n_rules * 4, threads, out_second);
    } catch (...) {
      std::free(out_first->payload);
      out_first->payload = nullptr;
      throw;
    }
  });
}

int gfz_unframe_int32(const uint8_t *payload, uint64_t payload_size, uint64_t original_size,
                      int32_t *out, uint64_t out_cap) {
  if (original_size % 4)
    return gfz::fail(GFZ_E_CORRUPT, std::string(kFramePrefix) +
                                        "int32 column size is not a multiple of 4");
  if (original_size && !out)
    return gfz::fail(GFZ_E_INVALID, std::string(kFramePrefix) + "null output");
  return gfz::guarded(kFramePrefix, [&] {
    unframe_bytes(payload, payload_size, original_size, out, out_cap * 4);
  });
}

int gfz_unframe_rule_columns(const uint8_t *first_payload, uint64_t first_payload_size,
                             uint64_t first_original_size, const uint8_t *second_payload,
                             uint64_t second_payload_size, uint64_t second_original_size,
                             int32_t *first, int32_t *second, uint64_t n_rules) {
  if (first_original_size != n_rules * 4 || second_original_size != n_rules * 4)
    return gfz::fail(GFZ_E_CORRUPT, std::string(kFramePrefix) +
                                        "rule column sizes disagree with the rule count");
  if (n_rules && (!first || !second))
    return gfz::fail(GFZ_E_INVALID, std::string(kFramePrefix) + "null output");
  return gfz::guarded(kFramePrefix, [&] {
    unframe_bytes(first_payload, first_payload_size, first_original_size, first, n_rules * 4);
    unframe_bytes(second_payload, second_payload_size, second_original_size, second,
                  n_rules * 4);
    delta_decode(first, n_rules);
    delta_decode(second, n_rules);
  });
}

struct gfz_framer {
  gfz::ColumnFramer f;
  uint64_t bytes;
  gfz_framer(const void *base, uint64_t n, int threads) : f(base, n, threads), bytes(n) {}
};

int gfz_framer_begin(const void *column, uint64_t bytes, int threads, gfz_framer **out) {
  if (!out || (bytes && !column))
    return gfz::fail(GFZ_E_INVALID, std::string(kFramePrefix) + "null argument");
  *out = nullptr;
  return gfz::guarded(kFramePrefix, [&] { *out = new gfz_framer(column, bytes, threads); });
}

int gfz_framer_ready(gfz_framer *f, uint64_t bytes_ready) {
  if (!f)
    return gfz::fail(GFZ_E_INVALID, std::string(kFramePrefix) + "null framer");
  if (bytes_ready > f->bytes)
    return gfz::fail(GFZ_E_INVALID, std::string(kFramePrefix) + "bytes_ready exceeds the column");
  return gfz::guarded(kFramePrefix, [&] { f->f.ready(bytes_ready); });
}

int gfz_framer_end(gfz_framer *f, gfz_block *out) {
  if (!f || !out)
    return gfz::fail(GFZ_E_INVALID, std::string(kFramePrefix) + "null argument");
  return gfz::guarded(kFramePrefix, [&] { f->f.finish(out); });
}

void gfz_framer_free(gfz_framer *f) { delete f; }

void gfz_block_free(gfz_block *b) {
  if (!b)
    return;
  std::free(b->payload);
  b->payload = nullptr;
  b->payload_size = 0;
}

}  // extern "C"
