/* gfaz_amd - MI355X (gfx950) grammar codec for GFA path/walk traversals.
 *
 * C ABI: plain pointers and sizes, integer status codes, no exceptions and no
 * C++ or torch types across the boundary. Every entry point returns GFZ_OK (0)
 * or a negative GFZ_E* code; gfz_last_error() gives the calling thread's last
 * message (prefixed like the reference's exceptions, e.g. "GPU compression
 * error: ...").
 *
 * Traversal layout everywhere: one flat int32 array of signed node IDs (sign =
 * orientation), paths first then walks, plus one uint32 length per traversal -
 * the reference's FlattenedPaths (include/gpu/core/gfa_graph_gpu.hpp:20-26).
 *
 * Ownership: the caller owns every input pointer, which stays read-only for the
 * duration of the call. Buffers the library returns are released through the
 * matching gfz_*_free call. Device memory belongs to a gfz_ctx (one per device
 * and calling thread; calls on one ctx are serialised by the caller, calls on
 * different ctxs are independent). */
#ifndef GFAZ_AMD_H
#define GFAZ_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GFZ_OK = 0,
  GFZ_E_INVALID = -1,   /* bad argument */
  GFZ_E_HIP = -2,       /* HIP runtime error */
  GFZ_E_NOMEM = -3,     /* device or host allocation failed / budget exceeded */
  GFZ_E_OVERFLOW = -4,  /* ID space exhausted (codec.cpp overflow_error analogue) */
  GFZ_E_CORRUPT = -5,   /* malformed rulebook or encoded stream */
  GFZ_E_BUSY = -6       /* serving queue full or request not finished */
};

/* Flags for pointer arguments. */
enum {
  GFZ_INPUT_DEVICE = 1u,  /* node / encoded arrays are device pointers */
  GFZ_OUTPUT_DEVICE = 2u  /* results stay in device memory */
};

typedef struct gfz_ctx gfz_ctx;
typedef struct gfz_rulebook gfz_rulebook;

const char *gfz_last_error(void);
const char *gfz_version(void);

/* One context per device: a HIP stream, a caching device pool and optional
 * per-kernel event timing. */
int gfz_ctx_create(int device, gfz_ctx **out);
void gfz_ctx_destroy(gfz_ctx *ctx);
int gfz_ctx_sync(gfz_ctx *ctx);
/* Cap on device bytes the ctx may hold (0 = none): allocations beyond it fail
 * with GFZ_E_NOMEM instead of reaching hipMalloc. */
int gfz_ctx_set_memory_limit(gfz_ctx *ctx, uint64_t bytes);
int gfz_ctx_set_profiling(gfz_ctx *ctx, int on);
/* Limits profiling to the comma-separated kernel groups in `groups` (NULL or
 * "" = every group). Each timed group records an event pair that drains the
 * stream around its launches, so a throughput run times only the groups it
 * reports. No counterpart in the reference (measurement support). */
int gfz_ctx_set_profile_groups(gfz_ctx *ctx, const char *groups);
int gfz_ctx_reset_stats(gfz_ctx *ctx);
/* Accumulated launches / milliseconds / algorithmic bytes of a timed kernel
 * group ("expand", "pair_insert", "emit", ...); zeros if never timed. */
int gfz_ctx_kernel_stat(gfz_ctx *ctx, const char *name, uint64_t *launches, double *total_ms,
                        double *bytes);
/* Number of CUs of the ctx's device. */
int gfz_ctx_num_cus(gfz_ctx *ctx);
/* Launch-shape autotuning (no counterpart in the reference, which hard-codes
 * its chunk sizes, e.g. src/gpu/decompression/path_decompression_gpu_rolling.cu).
 * With autotuning on (the default), the first call that meets an untuned
 * kernel/size class times each candidate tile size on its own input and keeps
 * the fastest for the ctx; every candidate gives the same result. Keys look
 * like "expand_items/<size class>"; gfz_ctx_tuned gives -1 for an untuned key,
 * gfz_ctx_set_tuned pins a choice (e.g. one saved from an earlier process). */
int gfz_ctx_set_autotune(gfz_ctx *ctx, int on);
int gfz_ctx_tuned(gfz_ctx *ctx, const char *key, int *value);
int gfz_ctx_set_tuned(gfz_ctx *ctx, const char *key, int value);
/* Tunes the pair-counting knobs of compression on a representative input: each
 * setting of "count_bucket_scale/<size class>" (1, 2, 4: half, the estimate,
 * twice the estimated bucket count) and "count_agg_slots/<size class>" (2048,
 * 4096 LDS slots per aggregated tile) compresses the input twice; every setting
 * must give the same grammar (else GFZ_E_CORRUPT) and the fastest is kept in the
 * ctx. *best_ms (may be NULL) receives its time. Inputs as gfz_compress_traversals. */
int gfz_autotune_compress(gfz_ctx *ctx, const int32_t *nodes, uint64_t n_nodes,
                          const uint32_t *lengths, uint64_t n_trav, int rounds, int threshold,
                          int delta_round, uint32_t flags, double *best_ms);
/* HIP graphs of the induction rounds (no counterpart in the reference, whose
 * rolling scheduler launches every kernel from the host:
 * src/gpu/compression/path_compression_gpu_rolling.cu:325-625). With the tuned
 * key "round_graphs" = 1 each round's three stretches of launches between host
 * readbacks (counting; rule table to used keys; sort to remap) are captured and
 * launched as graphs, kept per ctx and round and updated in place by later
 * calls. out[0] = graph launches, out[1] = graphs instantiated, out[2] = graphs
 * updated in place, over the ctx's life. */
int gfz_ctx_graph_stats(gfz_ctx *ctx, uint64_t out[3]);
/* The ctx's tuned table as JSON {"arch": ..., "cus": N, "tuned": {key: value}},
 * and back: a file written for another GPU architecture or CU count is refused
 * (GFZ_E_INVALID) and leaves the table as it was. */
int gfz_ctx_save_tuning(gfz_ctx *ctx, const char *path);
int gfz_ctx_load_tuning(gfz_ctx *ctx, const char *path);

/* Device memory helpers (pool-backed) for callers without their own allocator. */
int gfz_dev_alloc(gfz_ctx *ctx, uint64_t bytes, void **out);
int gfz_dev_free(gfz_ctx *ctx, void *ptr);
int gfz_memcpy_h2d(gfz_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int gfz_memcpy_d2h(gfz_ctx *ctx, void *dst, const void *src, uint64_t bytes);
/* Sorts n distinct 64-bit keys in device memory ascending with the sort that
 * orders each round's rules; buckets above local_cap keys (0 = default) take
 * its in-memory path. No counterpart in the reference: a test and diagnostic
 * entry for the rule-ordering sort. */
int gfz_debug_sort_u64(gfz_ctx *ctx, uint64_t *keys, uint64_t n, uint32_t local_cap);

/* ---- compress ------------------------------------------------------------
 * Replaces run_path_compression_gpu (include/gpu/compression/
 * compression_workflow_gpu.hpp:75-78), whose result the reference turns into
 * CompressedData; here the traversal grammar itself: the same encoded stream,
 * encoded lengths, rule columns and rule-ID window that the reference CPU
 * compressor produces (src/workflows/compression_workflow.cpp:235-279,548-700),
 * honouring delta_round (the reference GPU path forces 1,
 * src/gpu/compression/compression_workflow_gpu.cu:268). threshold is accepted
 * for interface parity; as in the reference CPU path every pair seen twice
 * becomes a rule (src/grammar/rule_generator.cpp:242-246). */
typedef struct gfz_compress_out {
  int32_t *encoded;           /* n_encoded symbols; host or device per flags */
  uint64_t n_encoded;
  uint32_t *encoded_lengths;  /* host, one per input traversal */
  uint64_t n_trav;
  int32_t *rules_first;       /* n_rules, raw (not delta-coded), rule-ID order */
  int32_t *rules_second;      /* host or device per flags */
  uint64_t n_rules;
  uint32_t start_id;          /* LayerRuleRange{2, start_id, end_id, 0, 2*n_rules} */
  uint32_t end_id;
  int32_t rounds_run;
  uint32_t flags;             /* copy of the call's flags */
  uint64_t round_rules[32];   /* rules created per round */
  uint64_t round_lengths[32]; /* stream length after each round */
  uint64_t round_distinct[32];/* distinct canonical pairs seen per round */
  uint64_t round_slots[32];   /* pair-table slots used per round (after retries) */
  void *internal;             /* owned device buffers */
} gfz_compress_out;

int gfz_compress_traversals(gfz_ctx *ctx, const int32_t *nodes, uint64_t n_nodes,
                            const uint32_t *lengths, uint64_t n_trav, uint32_t n_paths,
                            int rounds, int threshold, int delta_round, uint32_t flags,
                            gfz_compress_out *out);
void gfz_compress_free(gfz_compress_out *out);

/* ---- decompress ----------------------------------------------------------
 * gfz_rulebook_load replaces prepare_gpu_traversal_rulebook
 * (src/gpu/decompression/decompression_primitives_gpu.cu:18-39): rule columns
 * already prefix-summed (raw), min_rule_id = layer_rule_ranges[0].start_id
 * (include/model/compressed_data.hpp:49-51). Expanded rule sizes are computed
 * on the device; a cyclic or out-of-range rulebook fails with GFZ_E_CORRUPT. */
int gfz_rulebook_load(gfz_ctx *ctx, const int32_t *rules_first, const int32_t *rules_second,
                      uint64_t n_rules, uint32_t min_rule_id, uint32_t flags,
                      gfz_rulebook **out);
void gfz_rulebook_free(gfz_rulebook *rb);
/* Expanded length of one rule (for tests and planning). */
int gfz_rulebook_rule_size(gfz_rulebook *rb, uint64_t rule_index, uint32_t *size);
/* Deepest rule, and whether the rule bodies were pre-expanded in HBM (and how
 * many int32 they take); without bodies, expansion descends the rule tree. */
int gfz_rulebook_info(gfz_rulebook *rb, uint32_t *max_depth, int *has_bodies,
                      uint64_t *body_elems);

/* Total decoded length of a batch of encoded traversals. */
int gfz_decoded_size(gfz_ctx *ctx, const gfz_rulebook *rb, const int32_t *encoded,
                     uint64_t n_encoded, const uint32_t *encoded_lengths, uint64_t n_trav,
                     uint32_t flags, uint64_t *n_out);

/* Replaces decompress_paths_gpu / decompress_gpu_traversal_materialized
 * (include/gpu/decompression/decompression_workflow_gpu.hpp:32-33,
 * include/gpu/decompression/traversal_decode_gpu.hpp:115-119): expand every
 * traversal, then undo delta_round prefix sums per traversal. out holds
 * out_cap int32 (host or device per flags); *n_out receives the decoded total
 * and out_lengths (host, n_trav, may be NULL) the decoded per-traversal
 * lengths. If orig_lengths is given and a decoded length differs, the call
 * still succeeds and sets *length_mismatch (the reference only warns,
 * src/workflows/decompression_workflow.cpp:144-150). */
int gfz_decompress_traversals(gfz_ctx *ctx, const gfz_rulebook *rb, const int32_t *encoded,
                              uint64_t n_encoded, const uint32_t *encoded_lengths,
                              const uint32_t *orig_lengths, uint64_t n_trav, int delta_round,
                              int32_t *out, uint64_t out_cap, uint32_t flags, uint64_t *n_out,
                              uint32_t *out_lengths, int *length_mismatch);

/* ---- resident decoding: extract -------------------------------------------
 * gfz_resident_load keeps a rulebook's encoded traversals decodable in HBM: the
 * per-symbol offsets and expansion state are computed once, so a request decodes
 * any subset of traversals in one launch. gfz_extract replaces the decode part of
 * extract_path_lines_by_name / extract_walk_lines(_by_name)
 * (src/workflows/extraction_workflow.cpp:106-133,206-424: decode_sequence on the
 * selected traversals after a zstd decode of the whole column on every call); the
 * name or tuple -> index lookup and the P/W line text stay with the caller.
 * Traversal indices count paths first, then walks (the flattened order). The
 * rulebook must outlive the resident state. Without its bodies in HBM
 * (gfz_rulebook_info has_bodies = 0) the store keeps the encoded symbols and
 * extracts by descending the rule tree (slower, same values). Errors carry the
 * reference's "Extraction workflow error: " prefix (extraction_workflow.cpp:15). */
typedef struct gfz_resident gfz_resident;

int gfz_resident_load(gfz_ctx *ctx, const gfz_rulebook *rb, const int32_t *encoded,
                      uint64_t n_encoded, const uint32_t *encoded_lengths, uint64_t n_trav,
                      int delta_round, uint32_t flags, gfz_resident **out);
void gfz_resident_free(gfz_resident *r);
int gfz_resident_info(const gfz_resident *r, uint64_t *n_trav, uint64_t *n_decoded);
/* Decoded length of every stored traversal (n_trav uint32). */
int gfz_resident_lengths(const gfz_resident *r, uint32_t *decoded_lengths);
/* Appends already-encoded traversals (encoded with the same rulebook, e.g. by
 * gfz_encode_with_rules) to the resident state; they get the next traversal
 * indices (*first_new_id receives the first). This is the resident half of
 * add_haplotypes (src/workflows/add_haplotypes_workflow.cpp:467-575), which
 * appends to the traversal columns and zstd-frames the whole column again; here
 * only the new symbols' decode state is computed, and nothing is re-framed.
 * encoded: host, or device with GFZ_INPUT_DEVICE. */
int gfz_resident_append(gfz_ctx *ctx, gfz_resident *r, const int32_t *encoded,
                        uint64_t n_encoded, const uint32_t *encoded_lengths, uint64_t n_trav,
                        uint32_t flags, uint64_t *first_new_id);
/* Decode traversals ids[0..n_ids) back to back into out (host, or device with
 * GFZ_OUTPUT_DEVICE; out_cap int32). offsets (host, n_ids + 1, may be NULL)
 * receives each one's start; *n_out the total. Repeated ids are decoded again. */
int gfz_extract(gfz_ctx *ctx, const gfz_resident *r, const uint64_t *ids, uint64_t n_ids,
                int32_t *out, uint64_t out_cap, uint32_t flags, uint64_t *offsets,
                uint64_t *n_out);

/* No counterpart in the reference: gfz_extract for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). out is device memory (out_cap int32); the request's plan
 * (segment table, tile prefix, work-stealing counter) goes into a fixed device
 * array of GFZ_NOALLOC_SCRATCH_BYTES that the library holds per device until the
 * process ends; a request whose plan needs more (about 18 K non-empty traversals)
 * fails with GFZ_E_INVALID, and so does a store with delta_round >= 2 or one whose
 * rulebook has no bodies in HBM (its descent plan allocates). Calls on
 * one device share that array and keep its last contents: calls that overlap on
 * one GPU race, so serialise them. Returns after the decode finished. */
#define GFZ_NOALLOC_SCRATCH_BYTES (1u << 20)
int gfz_extract_noalloc(gfz_ctx *ctx, const gfz_resident *r, const uint64_t *ids,
                        uint64_t n_ids, int32_t *out, uint64_t out_cap, uint64_t *offsets,
                        uint64_t *n_out);

/* ---- extraction front-end: a resident catalog and P/W lines ------------------
 * The reference's extract_path_lines_by_name, extract_walk_lines and
 * extract_walk_lines_by_name (src/workflows/extraction_workflow.cpp:206-424)
 * zstd-decode the name columns, build a name map (or scan the walk tuples
 * linearly) and zstd-decode the whole paths or walks column on every call. Here
 * a gfz_catalog indexes the names once (hash maps; appends extend them) and the
 * traversals stay decodable in a gfz_resident, so a request costs hash probes,
 * one resident decode of just the requested traversals and GPU formatting of
 * their lines (gfz_format_steps). Lookup semantics and messages are the
 * reference's: any repeated stored path name (walk sample ID) fails every lookup
 * by name with "path name is ambiguous: <first repeat>" ("walk name is
 * ambiguous: ..."); a missing one fails with "path not found: <name>" ("walk not
 * found: ..."); a walk tuple takes the first stored walk equal in all five
 * fields ("walk not found for the provided identifier: s\th\tseq\tstart\tend").
 * Lines: "P\t<name>\t<steps>\t<overlap or *>\n" and "W\t<sample>\t<hap>\t<seq>\t
 * <start or *>\t<end or *>\t<steps>\n" (format_path_line / format_walk_line,
 * :150-193), the requested fields, or for a walk by name the stored ones.
 * Catalog lookups and appends are thread-safe (lookups shared, appends
 * exclusive). Strings cross the ABI as concatenated bytes with n + 1 offsets. */
typedef struct gfz_strings {
  const char *bytes;
  const uint64_t *off; /* n + 1 */
  uint64_t n;
} gfz_strings;

typedef struct gfz_walk_keys {  /* WalkLookupKey (extraction_workflow.hpp) per walk */
  gfz_strings sample_ids;       /* n walks */
  const uint32_t *hap_indices;
  gfz_strings seq_ids;
  const int64_t *seq_starts;    /* -1 = "*" */
  const int64_t *seq_ends;
} gfz_walk_keys;

typedef struct gfz_catalog gfz_catalog;
int gfz_catalog_create(gfz_catalog **out);
void gfz_catalog_free(gfz_catalog *c);
/* Paths (walks) get traversal indices first_trav_id, first_trav_id + 1, ... of
 * the resident store: 0 for the paths of a loaded .gfaz, n_paths for its walks,
 * gfz_resident_append's first_new_id for appended ones. overlaps: NULL or one
 * per path (empty = "*"). */
int gfz_catalog_add_paths(gfz_catalog *c, const gfz_strings *names, const gfz_strings *overlaps,
                          uint64_t first_trav_id);
int gfz_catalog_add_walks(gfz_catalog *c, const gfz_walk_keys *walks, uint64_t first_trav_id);
int gfz_catalog_info(const gfz_catalog *c, uint64_t *n_paths, uint64_t *n_walks);
/* Resolve requests to traversal indices (trav_ids: one per request). */
int gfz_catalog_find_paths(const gfz_catalog *c, const gfz_strings *names, uint64_t *trav_ids);
int gfz_catalog_find_walks(const gfz_catalog *c, const gfz_walk_keys *keys, uint64_t *trav_ids);
int gfz_catalog_find_walks_by_name(const gfz_catalog *c, const gfz_strings *names,
                                   uint64_t *trav_ids);

/* Text of the requested lines, back to back (library-owned; gfz_lines_free). */
typedef struct gfz_lines {
  uint8_t *text;
  uint64_t n_bytes;
  uint64_t *line_off;  /* n_lines + 1 */
  uint64_t n_lines;
} gfz_lines;
int gfz_extract_path_lines(gfz_ctx *ctx, const gfz_resident *r, const gfz_catalog *c,
                           const gfz_strings *names, gfz_lines *out);
int gfz_extract_walk_lines(gfz_ctx *ctx, const gfz_resident *r, const gfz_catalog *c,
                           const gfz_walk_keys *keys, gfz_lines *out);
int gfz_extract_walk_lines_by_name(gfz_ctx *ctx, const gfz_resident *r, const gfz_catalog *c,
                                   const gfz_strings *names, gfz_lines *out);
void gfz_lines_free(gfz_lines *l);

/* ---- streaming decompress ---------------------------------------------------
 * Replaces decompress_gpu_traversal_rolling_direct_writer / the rolling stream
 * (src/gpu/decompression/traversal_decode_gpu.cu:122-156,
 * include/gpu/decompression/path_decompression_gpu_rolling.hpp:26-44,74-81):
 * every traversal is decoded in chunks of whole traversals of at most
 * chunk_bytes of output (a larger traversal is a chunk of its own; 0 = 1 GiB,
 * the reference's kDefaultRollingOutputChunkBytes) into two alternating pinned
 * host buffers; the consumer is called in order on the calling thread, and runs
 * for chunk k while the GPU decodes chunk k + 1. A nonzero return from the
 * consumer stops the stream (GFZ_E_INVALID). Rulebooks without bodies in HBM
 * decode each chunk by descent. */
typedef struct gfz_chunk {
  const int32_t *values;      /* pinned host memory, valid during the call */
  uint64_t n_values;
  uint64_t trav_begin;        /* traversals [trav_begin, trav_end) */
  uint64_t trav_end;
  uint64_t expanded_begin;    /* decoded offset of the chunk in the whole output */
  const uint32_t *lengths;    /* decoded length of each traversal of the chunk */
} gfz_chunk;
typedef int (*gfz_chunk_consumer)(const gfz_chunk *chunk, void *user);

int gfz_decompress_stream(gfz_ctx *ctx, const gfz_rulebook *rb, const int32_t *encoded,
                          uint64_t n_encoded, const uint32_t *encoded_lengths, uint64_t n_trav,
                          int delta_round, uint64_t chunk_bytes, uint32_t flags,
                          gfz_chunk_consumer consumer, void *user);

/* ---- add-haplotypes: encoding with a stored rulebook ----------------------
 * Replaces the encode step of add_haplotypes (src/workflows/
 * add_haplotypes_workflow.cpp:336-379,447-465 with src/grammar/path_encoder.cpp:
 * 13-84): delta_round transforms, the collision check (largest |value| must stay
 * below layers[0].start_id when layers exist, else GFZ_E_OVERFLOW with the
 * reference's message), then per layer the shift-reduce encoder over a table
 * pack(first, second) -> start_id + i of the layer's rows as stored (first row
 * wins on duplicate keys). Layer rows: row_offset = flattened_offset / k of the
 * LayerRuleRange (k = 0: row 0). Rows past the columns: GFZ_E_CORRUPT ("rule
 * column is truncated"). nodes and the rule columns are host or device arrays
 * per GFZ_INPUT_DEVICE; the encoded stream is returned per GFZ_OUTPUT_DEVICE. */
typedef struct gfz_layer {
  uint32_t start_id;
  uint32_t end_id;
  uint64_t row_offset;
} gfz_layer;

typedef struct gfz_encode_out {
  int32_t *encoded;           /* n_encoded symbols; host or device per flags */
  uint64_t n_encoded;
  uint32_t *encoded_lengths;  /* host, one per input traversal */
  uint64_t n_trav;
  uint32_t flags;
  void *internal;
} gfz_encode_out;

int gfz_encode_with_rules(gfz_ctx *ctx, const int32_t *nodes, uint64_t n_nodes,
                          const uint32_t *lengths, uint64_t n_trav, const int32_t *rules_first,
                          const int32_t *rules_second, uint64_t n_rules, const gfz_layer *layers,
                          uint64_t n_layers, int delta_round, uint32_t flags,
                          gfz_encode_out *out);
void gfz_encode_free(gfz_encode_out *out);

/* ---- serving ----------------------------------------------------------------
 * No counterpart in the reference, where extract and add-haplotypes are host
 * workflows per call (extraction_workflow.cpp:206-424,
 * add_haplotypes_workflow.cpp:336-592). A server keeps one rulebook and its
 * encoded traversals resident in HBM and serves requests as they arrive: a
 * worker thread admits queued requests in arrival order while their output fits
 * one of two output slots (half of hbm_budget_bytes each), applies the batch's
 * add requests (GPU encode with the stored rules, appended to the resident
 * store; new traversals get the next indices), decodes all of the batch's
 * extract requests in one persistent launch, and copies results to pinned host
 * memory on a second stream while the next batch decodes. A request whose
 * output exceeds one slot fails with GFZ_E_NOMEM; a full queue (max_queue > 0)
 * refuses submissions with GFZ_E_BUSY. Submit, poll and wait are thread-safe.
 * Latency = submission to result available on the host. */
typedef struct gfz_server gfz_server;

typedef struct gfz_server_config {
  int device;
  uint64_t hbm_budget_bytes;    /* device bytes for request outputs in flight */
  uint32_t max_batch_requests;  /* 0: 256 */
  uint32_t max_queue;           /* 0: unbounded */
  uint32_t wait_timeout_ms;     /* gfz_wait deadline; 0: 120000. On expiry gfz_wait returns
                                   GFZ_E_HIP naming the request's stage (queued, adding,
                                   decoding, copying) and the ticket stays waitable */
} gfz_server_config;

typedef struct gfz_result {
  int status;                 /* GFZ_OK or the request's error */
  int is_add;
  double latency_ms;
  int32_t *values;            /* extract: decoded traversals back to back */
  uint64_t n_values;
  uint64_t *offsets;          /* extract: n_items + 1 starts */
  uint64_t n_items;           /* extract: ids; add: traversals added */
  uint64_t first_new_id;      /* add: index of the first added traversal */
  uint32_t *encoded_lengths;  /* add: encoded length of each added traversal */
} gfz_result;

typedef struct gfz_server_stats {
  uint64_t completed, rejected, batches, bytes_out, n_traversals;
  double p50_ms, p99_ms, mean_ms;
  double elapsed_s;           /* first submission to last completion */
  double gbps;                /* decoded bytes delivered / elapsed_s */
} gfz_serve_stats;

int gfz_server_create(const gfz_server_config *cfg, const int32_t *rules_first,
                      const int32_t *rules_second, uint64_t n_rules, uint32_t min_rule_id,
                      const gfz_layer *layers, uint64_t n_layers, const int32_t *encoded,
                      uint64_t n_encoded, const uint32_t *encoded_lengths, uint64_t n_trav,
                      int delta_round, gfz_server **out);
/* Finishes every queued request, then stops. */
void gfz_server_destroy(gfz_server *s);
int gfz_submit_extract(gfz_server *s, const uint64_t *ids, uint64_t n_ids, uint64_t *ticket);
int gfz_submit_add(gfz_server *s, const int32_t *nodes, uint64_t n_nodes,
                   const uint32_t *lengths, uint64_t n_trav, uint64_t *ticket);
int gfz_poll(gfz_server *s, uint64_t ticket, int *done);
/* Blocks until the request is done; returns its status (result fields valid
 * on GFZ_OK; release with gfz_result_free). A ticket can be waited once. */
int gfz_wait(gfz_server *s, uint64_t ticket, gfz_result *out);
void gfz_result_free(gfz_result *r);
int gfz_server_stats(gfz_server *s, gfz_serve_stats *out);
int gfz_server_reset_stats(gfz_server *s);

/* ---- P/W line step text ------------------------------------------------------
 * gfz_format_steps writes traversals as GFA P or W lines, line t = head_t +
 * steps_t + tail_t, lines back to back: the step lists of
 * format_path_line_numeric (kind GFZ_TEXT_P: "12+,13-", src/io/
 * gfa_write_utils.cpp:238-260) and format_walk_line_numeric (GFZ_TEXT_W:
 * ">12<13", :262-290), formatted on the GPU; head_t / tail_t are the caller's
 * bytes around them ("P\tname\t" / "\t*\n", "W\tsample\thap\tseq\tstart\tend\t" /
 * "\n"), given as concatenated bytes with n_trav + 1 offsets each. nodes: host
 * or device (GFZ_INPUT_DEVICE); out: host or device (GFZ_OUTPUT_DEVICE) of
 * out_cap bytes, or NULL to only size the text (*n_out); line_off (host,
 * n_trav + 1, may be NULL) receives where each line starts.
 * gfz_parse_steps is the inverse for the reference parser's numeric-name path
 * (parse_p_line / parse_w_line with parse_uint32, src/io/gfa_parser.cpp:69-91,
 * 676-790, 1210-1219): the step-list field of each line,
 * text[field_begin[l] .. field_end[l]) (ordered, disjoint; other bytes, such as
 * the rest of each line, are ignored), into signed node IDs (steps with an empty
 * name or the value 0 are dropped, as there; a name without leading digits or
 * past 2^32 - 1 fails with the reference's "parse_uint32: ..." message).
 * text: host or device (GFZ_INPUT_DEVICE); nodes: host or device
 * (GFZ_OUTPUT_DEVICE), or NULL to only count; lengths (host, n_lines) may be NULL. */
enum { GFZ_TEXT_P = 0, GFZ_TEXT_W = 1 };
int gfz_format_steps(gfz_ctx *ctx, int kind, const int32_t *nodes, uint64_t n_nodes,
                     const uint32_t *lengths, uint64_t n_trav, const uint8_t *heads,
                     const uint64_t *head_off, const uint8_t *tails, const uint64_t *tail_off,
                     uint32_t flags, uint8_t *out, uint64_t out_cap, uint64_t *n_out,
                     uint64_t *line_off);
int gfz_parse_steps(gfz_ctx *ctx, int kind, const uint8_t *text, uint64_t n_bytes,
                    const uint64_t *field_begin, const uint64_t *field_end, uint64_t n_lines,
                    uint32_t flags, int32_t *nodes, uint64_t nodes_cap, uint32_t *lengths,
                    uint64_t *n_nodes);

/* ---- node coverage: pangenome growth and presence/absence (PAV) -------------
 * The GPU part of compute_growth (src/workflows/growth_workflow.cpp:261-503) and
 * compute_pav (src/workflows/pav_workflow.cpp:475-760), on traversals decoded by
 * gfz_decompress_traversals: gfz_coverage_build marks which of n_groups groups
 * (the caller's grouping of the traversals, group_of[t] per traversal, e.g. the
 * reference's PanSN sample / haplotype keys; up to 65535 groups) visit which
 * nodes 1 .. num_nodes (|step|; 0 and larger IDs are skipped as there). nodes:
 * host or device (GFZ_INPUT_DEVICE).
 * gfz_coverage_hist: hist[c] (host, n_groups + 1) = nodes visited by exactly c
 * groups; gfz_growth_curve (host) turns it into growth[k] = expected nodes covered
 * by k random groups, with the reference's closed form (k = 1 .. n_groups;
 * growth[0] = 0).
 * gfz_pav_window_nodes (host) finds, for BED ranges [starts[r], ends[r]) on one
 * reference traversal (ref_nodes, with segment lengths 1 .. num_nodes), each
 * range's distinct overlapping nodes as the reference does: window_off
 * (n_ranges + 1) always, window_nodes when not NULL (nodes_cap entries).
 * gfz_coverage_windows then gives denominators[w] (the windows' node lengths) and
 * numerators[w * n_groups + g] (the lengths of those nodes group g visits). */
typedef struct gfz_coverage gfz_coverage;
int gfz_coverage_build(gfz_ctx *ctx, const int32_t *nodes, uint64_t n_nodes,
                       const uint32_t *lengths, uint64_t n_trav, const uint32_t *group_of,
                       uint32_t n_groups, uint32_t num_nodes, uint32_t flags,
                       gfz_coverage **out);
void gfz_coverage_free(gfz_coverage *cov);
int gfz_coverage_hist(gfz_coverage *cov, uint64_t *hist);
int gfz_coverage_windows(gfz_coverage *cov, const uint64_t *window_off,
                         const uint32_t *window_nodes, uint64_t n_windows,
                         const uint32_t *segment_lengths, uint64_t *denominators,
                         uint64_t *numerators);
int gfz_growth_curve(const uint64_t *hist, uint32_t n_groups, double *growth);
int gfz_pav_window_nodes(const int32_t *ref_nodes, uint64_t n_ref,
                         const uint32_t *segment_lengths, uint32_t num_nodes,
                         const uint64_t *starts, const uint64_t *ends, uint64_t n_ranges,
                         uint64_t *window_off, uint32_t *window_nodes, uint64_t nodes_cap);

/* ---- traversal-column framing (host) -----------------------------------
 * The steps the reference CPU workflow applies around the grammar before
 * serialisation, producing the same bytes, so a caller can fill CompressedData
 * (include/model/compressed_data.hpp:43-66) from gfz_compress_traversals:
 *   rules_first_zstd / rules_second_zstd = gfz_frame_rule_columns: process_rules
 *     (src/workflows/compression_utils.cpp:165-193, one LayerRuleRange), i.e. each
 *     raw column delta-encoded over the whole column (src/codec/codec.cpp:79-86),
 *     then zstd_compress_int32_vector;
 *   paths_zstd / walks_zstd = gfz_frame_int32 over the flattened encoded paths,
 *     resp. walks (compression_utils.cpp:38-54; codec.cpp:180-217).
 * zstd runs at level 9 (GFA_COMPRESSION_ZSTD_LEVEL overrides, 1..22) with
 * nbWorkers = GFA_COMPRESSION_ZSTD_WORKERS if > 0, else `threads` (<= 0: the
 * OpenMP thread count), exactly as src/codec/codec.cpp:100-177.
 * The inverses replace zstd_decompress_int32_vector (codec.cpp:224-244) and
 * delta_decode_int32 (codec.cpp:88-95); where the reference prints a zstd error
 * and returns an empty column, these return GFZ_E_CORRUPT. */
typedef struct gfz_block {
  uint64_t original_size;  /* bytes before zstd (ZstdCompressedBlock::original_size) */
  uint8_t *payload;        /* zstd frame; library-owned, release with gfz_block_free */
  uint64_t payload_size;
} gfz_block;

int gfz_frame_int32(const int32_t *data, uint64_t n, int threads, gfz_block *out);
int gfz_frame_rule_columns(const int32_t *first, const int32_t *second, uint64_t n_rules,
                           int threads, gfz_block *out_first, gfz_block *out_second);
int gfz_unframe_int32(const uint8_t *payload, uint64_t payload_size, uint64_t original_size,
                      int32_t *out, uint64_t out_cap);
int gfz_unframe_rule_columns(const uint8_t *first_payload, uint64_t first_payload_size,
                             uint64_t first_original_size, const uint8_t *second_payload,
                             uint64_t second_payload_size, uint64_t second_original_size,
                             int32_t *first, int32_t *second, uint64_t n_rules);
void gfz_block_free(gfz_block *b);

/* Piecewise framing of one column in a contiguous host buffer: the chunked
 * handoff the pipelined calls below use, open to callers that produce a column in
 * pieces. begin() takes the column's address and full size; ready(n) says bytes
 * [0, n) are final (they must then stay unchanged until end()); end() completes
 * the frame. The payload has the bytes of gfz_frame_int32 over the same data, i.e.
 * of zstd_compress_bytes (src/codec/codec.cpp:180-216): zstd is fed whole 128 KiB
 * blocks of the one buffer, which keeps every block's window and entropy state as
 * in a single ZSTD_compress2 call. With a libzstd that accepts the worker count
 * (and so compresses with threads), the frame depends on the worker split and
 * end() makes that single call instead. free() releases a framer (ended or not). */
typedef struct gfz_framer gfz_framer;
int gfz_framer_begin(const void *column, uint64_t bytes, int threads, gfz_framer **out);
int gfz_framer_ready(gfz_framer *f, uint64_t bytes_ready);
int gfz_framer_end(gfz_framer *f, gfz_block *out);
void gfz_framer_free(gfz_framer *f);

/* ---- traversal codec with its framing overlapped -------------------------------
 * The GPU codec and the traversal-column framing in one call each way. They
 * replace, on the compress side, run_path_compression_gpu followed by
 * finalize_traversal_columns (src/gpu/compression/path_compression_gpu_rolling.cu:29-56,
 * 596; CPU: compress_rule_columns, compress_path_fields, compress_walk_fields,
 * src/workflows/compression_workflow.cpp:281-383), and on the decompress side
 * prepare_gpu_traversal_rulebook + prepare_gpu_traversal_payload + the decode
 * (src/gpu/decompression/decompression_primitives_gpu.cu:18-56). The reference
 * runs these steps one after the other; here:
 *   compress: the four columns (rule first/second, paths, walks) are copied to
 *     pinned host memory in 4 MiB pieces, interleaved, and each is framed on its
 *     own host thread as its pieces land (rule columns delta-encoded on the way);
 *   decompress: four host threads decode the four payloads at once; the rule
 *     columns go up first and the rulebook builds on the GPU while the traversal
 *     payloads are still being decoded, each handed to the device in 4 MiB pieces
 *     on its own copy stream as they come out.
 * Payload bytes equal gfz_frame_rule_columns / gfz_frame_int32 (so the reference's
 * .gfaz fields), and the decoded traversals equal gfz_decompress_traversals. Input
 * flags as gfz_compress_traversals (GFZ_INPUT_DEVICE) and gfz_decompress_traversals
 * (GFZ_OUTPUT_DEVICE). threads: zstd worker count as gfz_frame_int32. */
typedef struct gfz_framed {
  uint32_t start_id, end_id; /* the one LayerRuleRange of the rule columns */
  uint64_t n_rules;
  int32_t rounds_run;
  int32_t delta_round;       /* as given (negative stored as 0) */
  uint64_t n_trav;
  uint32_t n_paths;          /* traversals [0, n_paths) are paths, the rest walks */
  uint32_t *encoded_lengths; /* n_trav; library-owned */
  gfz_block rules_first, rules_second, paths, walks;
  double gpu_ms;             /* host clock: the codec, input upload included */
  double total_ms;           /* host clock: the whole call */
} gfz_framed;

int gfz_compress_framed(gfz_ctx *ctx, const int32_t *nodes, uint64_t n_nodes,
                        const uint32_t *lengths, uint64_t n_trav, uint32_t n_paths, int rounds,
                        int threshold, int delta_round, uint32_t flags, int threads,
                        gfz_framed *out);
void gfz_framed_free(gfz_framed *f);
/* orig_lengths (optional) and out_lengths (optional) as gfz_decompress_traversals.
 * timings_ms (optional, 3 values, host clock from the call's start): rulebook
 * built, payloads all decoded on the host, traversals decoded. */
int gfz_decompress_framed(gfz_ctx *ctx, const gfz_framed *in, const uint32_t *orig_lengths,
                          int32_t *out, uint64_t out_cap, uint32_t flags, uint64_t *n_out,
                          uint32_t *out_lengths, int *length_mismatch, double *timings_ms);

/* ---- measurement support -------------------------------------------------
 * No counterpart in the reference: a plain copy of `bytes` (a multiple of 16)
 * between device buffers, timed with HIP events on the ctx stream. Profiles use
 * it to scale the HBM traffic counters read for the codec kernels. */
int gfz_calibration_copy(gfz_ctx *ctx, void *dst, const void *src, uint64_t bytes, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* GFAZ_AMD_H */
