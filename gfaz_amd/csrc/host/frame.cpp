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

#include <zstd.h>

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
  out->original_size = bytes;
  out->payload = nullptr;
  out->payload_size = 0;
  if (!bytes)
    return;
  ZSTD_CCtx *cctx = ZSTD_createCCtx();
  if (!cctx)
    throw gfz::Error(GFZ_E_NOMEM, "unable to create a zstd compression context");
  size_t r = ZSTD_CCtx_setParameter(cctx, ZSTD_c_compressionLevel, zstd_level());
  if (!ZSTD_isError(r))
    (void)ZSTD_CCtx_setParameter(cctx, ZSTD_c_nbWorkers, zstd_workers(threads));
  const size_t bound = ZSTD_compressBound(bytes);
  uint8_t *buf = static_cast<uint8_t *>(std::malloc(bound));
  if (!buf) {
    ZSTD_freeCCtx(cctx);
    throw std::bad_alloc();
  }
  const size_t n = ZSTD_isError(r) ? r : ZSTD_compress2(cctx, buf, bound, data, bytes);
  ZSTD_freeCCtx(cctx);
  if (ZSTD_isError(n)) {
    std::free(buf);
    throw gfz::Error(GFZ_E_INVALID, std::string("zstd compression failed: ") +
                                        ZSTD_getErrorName(n));
  }
  out->payload = buf;
  out->payload_size = n;
}

void unframe_bytes(const uint8_t *payload, uint64_t payload_size, uint64_t original_size,
                   void *dst, uint64_t cap_bytes) {
  if (!original_size)
    return;
  if (original_size > cap_bytes)
    throw gfz::Error(GFZ_E_INVALID, "output holds " + std::to_string(cap_bytes) + " < " +
                                        std::to_string(original_size) + " bytes");
  if (!payload)
    throw gfz::Error(GFZ_E_CORRUPT, "empty zstd payload for a non-empty column");
  const size_t n = ZSTD_decompress(dst, original_size, payload, payload_size);
  if (ZSTD_isError(n))
    throw gfz::Error(GFZ_E_CORRUPT, std::string("zstd decompression failed: ") +
                                        ZSTD_getErrorName(n));
  if (n != original_size)
    throw gfz::Error(GFZ_E_CORRUPT, "zstd payload decodes to " + std::to_string(n) +
                                        " bytes, block says " + std::to_string(original_size));
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

void gfz_block_free(gfz_block *b) {
  if (!b)
    return;
  std::free(b->payload);
  b->payload = nullptr;
  b->payload_size = 0;
}

}  // extern "C"
