// Resident decoder state (rulebook bodies + encoded traversals in HBM) and the
// batched, work-stealing extraction over any subset of traversals.
#pragma once

#include "codec.hpp"
#include "zexpand.hpp"

#include <functional>
#include <vector>

namespace gfz {

// One traversal of a request: decoded range [dec_begin, dec_begin + len) of the
// resident stream, written at out[dst ..), its symbols in [k_lo, k_hi).
struct SegDesc {
  uint64_t dec_begin;
  uint64_t len;
  uint64_t dst;
  uint64_t k_lo;
  uint64_t k_hi;
};

struct Resident {
  const Rulebook *rb = nullptr;  // not owned; must outlive the resident state
  int delta_round = 1;
  uint64_t n_trav = 0, n_enc = 0, n_out = 0;
  // Rulebooks without bodies (depth not monotone in rule ID, or bodies over the
  // budget): the encoded symbols stay in HBM and extraction gathers the
  // requested traversals' symbols and expands them by descent; syms and grid
  // stay empty.
  bool descent = false;
  DBuf<int32_t> enc;         // encoded symbols (descent only)
  DBuf<zx::SymInfo> syms;    // per encoded symbol
  DBuf<uint64_t> sym_off;    // n_enc + 1 decoded starts
  DBuf<uint64_t> grid;       // symbol covering each kGrid-th decoded position
  uint64_t n_cells = 0;
  std::vector<uint64_t> enc_start, dec_start, dec_len;
  std::vector<uint32_t> enc_len;
};

// d_enc: device; enc_lengths: host. The encoded stream may be freed afterwards.
void resident_load(Ctx &ctx, const Rulebook &rb, const int32_t *d_enc, uint64_t n_enc,
                   const uint32_t *enc_lengths, uint64_t n_trav, int delta_round,
                   Resident &res);

// Host plan of one extraction: segments, tile numbering, where each starts.
struct ExtractPlan {
  uint32_t items = 16;
  std::vector<SegDesc> segs;
  std::vector<uint64_t> tile_prefix;  // segs + 1
  std::vector<uint64_t> seg_dst;
  uint64_t total = 0;
};

// Plans ids[0..n_ids) back to back (offsets: host, n_ids + 1, may be null);
// returns the total. Throws on an index out of range.
uint64_t extract_plan(const Resident &res, const uint64_t *ids, uint64_t n_ids, uint32_t items,
                      ExtractPlan &plan, uint64_t *offsets);
size_t extract_meta_bytes(const ExtractPlan &plan);
// Enqueues the extraction on the ctx stream without waiting: the plan is
// staged through h_stage (host, extract_meta_bytes; pinned for a truly
// asynchronous copy, and untouched until the stream has passed the copy) into
// d_meta (device, same size); results go to d_out[0..plan.total).
void extract_enqueue(Ctx &ctx, const Resident &res, const ExtractPlan &plan, uint8_t *h_stage,
                     uint8_t *d_meta, int32_t *d_out);

// Fixed device scratch for extraction without allocation: the plan of one
// request (segment table, tile prefix, work-stealing counter, segment starts)
// goes into a file-scope device array of kNoAllocScratchBytes per device,
// held until the process ends.
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
// Resolves (once per device, with that device current) the array's device
// address; throws on a HIP error. Called at context creation.
void noalloc_scratch_init(int device);
// The device address of the array on `device` (noalloc_scratch_init first).
uint8_t *noalloc_scratch(int device);
// Extraction into caller-provided device memory with the plan in the fixed
// scratch array: no device allocation. Rejects plans above the array's size
// and stores with delta_round >= 2 (their scans need more scratch); calls on
// one device share the array and must not overlap. Synchronises the stream.
uint64_t resident_extract_noalloc(Ctx &ctx, const Resident &res, const uint64_t *ids,
                                  uint64_t n_ids, int32_t *d_out, uint64_t out_cap,
                                  uint64_t *offsets, uint32_t items);

// Appends encoded traversals (d_enc device, enc_lengths host) to the resident
// state; they get the next traversal indices. Synchronises the ctx stream.
void resident_append(Ctx &ctx, Resident &res, const int32_t *d_enc, uint64_t n_new,
                     const uint32_t *enc_lengths, uint64_t n_trav_new);

// One chunk of a streamed decode: whole traversals [trav_begin, trav_end),
// decoded back to back in pinned host memory valid during the callback.
struct StreamChunk {
  const int32_t *values;
  uint64_t n_values;
  uint64_t trav_begin, trav_end;
  uint64_t expanded_begin;  // decoded offset of the chunk in the whole output
  const uint32_t *lengths;  // decoded length per traversal of the chunk
};

// Decodes every traversal in chunks of at most chunk_bytes of output (a larger
// traversal is a chunk of its own) through two alternating device and pinned
// host buffers, calling consume in order; consume for chunk k runs while the
// device decodes chunk k + 1. A false return stops the stream (Error).
void decompress_stream(Ctx &ctx, const Resident &res, uint64_t chunk_bytes,
                       const std::function<bool(const StreamChunk &)> &consume);

// Decode traversals ids[0..n_ids) back to back into d_out (device, out_cap
// int32); offsets (host, n_ids + 1, may be null) receives where each starts.
// Returns the total; synchronises the ctx stream. items = output elements per
// thread of the expansion tile (8, 16 or 32).
uint64_t resident_extract(Ctx &ctx, const Resident &res, const uint64_t *ids, uint64_t n_ids,
                          int32_t *d_out, uint64_t out_cap, uint64_t *offsets,
                          uint32_t items = 16);

}  // namespace gfz
