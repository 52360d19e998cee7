// Piecewise zstd framing of one column held in a contiguous host buffer, for
// handing a column over in pieces (device-to-host copies, decompression into a
// staging buffer) while the other side of the handoff keeps working.
#pragma once

#include "gfaz_amd.h"

#include <cstdint>
#include <memory>

namespace gfz {

// Granularity of the framer: zstd's largest block (128 KiB). Feeding whole
// blocks of one contiguous buffer to zstd's block-level API keeps every block's
// window and entropy state as in a single ZSTD_compress2 call over the column,
// so the frame has the same bytes.
constexpr uint64_t kFrameBlockBytes = uint64_t(1) << 17;

// zstd frame of `bytes` bytes at `base`, built as the column becomes final from
// the front. Level and worker count as gfz_frame_int32. When the linked libzstd
// runs ZSTD_compress2 with worker threads (then the frame depends on the worker
// split), the framer waits for the whole column and makes that one call.
class ColumnFramer {
 public:
  ColumnFramer(const void *base, uint64_t bytes, int threads);
  ~ColumnFramer();
  ColumnFramer(const ColumnFramer &) = delete;
  ColumnFramer &operator=(const ColumnFramer &) = delete;
  // Bytes [0, bytes_ready) are final and stay unchanged until finish().
  void ready(uint64_t bytes_ready);
  // The whole column is final: completes the frame and hands it to `out`.
  void finish(gfz_block *out);

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

// Decodes one zstd column payload into `dst` (original_size bytes) front to back.
class ColumnUnframer {
 public:
  ColumnUnframer(const uint8_t *payload, uint64_t payload_size, void *dst,
                 uint64_t original_size);
  ~ColumnUnframer();
  ColumnUnframer(const ColumnUnframer &) = delete;
  ColumnUnframer &operator=(const ColumnUnframer &) = delete;
  // Decodes until at least `min_bytes` more bytes are out or the column is
  // complete; returns the bytes decoded so far. Throws Error(GFZ_E_CORRUPT) on a
  // bad payload or a size other than original_size.
  uint64_t next(uint64_t min_bytes);
  bool done() const;

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace gfz
