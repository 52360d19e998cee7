/* TEST INFRASTRUCTURE ONLY - the CPU oracle for the grammar codec.
 *
 * A plain C++/OpenMP restatement of the reference CPU algorithm for GFA
 * path/walk traversals. Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it, and only as the checker or the timed CPU
 * baseline; the product library under gfaz_amd/ never links or calls it.
 *
 * Traversals are one flat int32 array (paths first, then walks, as in the
 * reference's FlattenedPaths, include/gpu/core/gfa_graph_gpu.hpp:20-26) plus a
 * uint32 length per traversal. The grammar treats paths and walks as one
 * collection (src/grammar/rule_generator.cpp:236-240) and encodes each
 * traversal independently, so the flat form loses nothing.
 *
 * Pinning (what checks each function, beyond retyping it):
 *   oracle_compress   - golden .gfaz payloads written by the compiled reference
 *                       (tests/golden/*.npz), brute-force restatement on tiny
 *                       inputs, invariants (round trip, picks non-adjacent,
 *                       rules sorted by signed key, every used key repeated).
 *   oracle_decompress - round trip against the generator's input, golden
 *                       payloads, closed forms on hand-built rulebooks.
 *   oracle_encode_existing - golden add-haplotypes output of the compiled
 *                       reference and the CLI known answers
 *                       (tests/regression/test_cli_commands.py:157-160,194).
 */
#ifndef GFAZ_ORACLE_H
#define GFAZ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_compress_result {
  int32_t *encoded;          /* flat encoded traversals */
  uint64_t n_encoded;
  uint32_t *encoded_lengths; /* one per traversal */
  uint64_t n_trav;
  int32_t *rules_first;      /* raw (not delta-coded), in rule-ID order */
  int32_t *rules_second;
  uint64_t n_rules;
  uint32_t start_id;         /* first rule ID (= max |delta symbol| + 1) */
  uint32_t end_id;           /* one past the last rule ID */
  int32_t rounds_run;
  uint64_t rules_per_round[64];
} oracle_compress_result;

/* Compress traversals in place of compress_gfa's traversal stages
 * (src/workflows/compression_workflow.cpp:184-279, 548-700). Returns 0 on
 * success, <0 on error (overflow of the rule-ID space). */
int oracle_compress(const int32_t *nodes, uint64_t n_nodes, const uint32_t *lengths,
                    uint64_t n_trav, int rounds, int threshold, int delta_round,
                    int threads, oracle_compress_result *out);
void oracle_free_result(oracle_compress_result *r);

/* Expand + inverse delta (src/workflows/decompression_workflow.cpp:70-174).
 * out must hold sum(orig lengths); *out_n receives the decoded count.
 * Returns 0, or <0 if out_cap is too small. */
int oracle_decompress(const int32_t *encoded, uint64_t n_encoded,
                      const uint32_t *encoded_lengths, uint64_t n_trav,
                      const int32_t *rules_first, const int32_t *rules_second,
                      uint64_t n_rules, uint32_t min_rule_id, int delta_round,
                      int threads, int32_t *out, uint64_t out_cap, uint64_t *out_n);

/* add-haplotypes encoding of new traversals with a stored rulebook
 * (src/workflows/add_haplotypes_workflow.cpp:336-379, 447-465): delta,
 * collision check, then one greedy shift-reduce pass per layer with the
 * stored (first, second) pairs as keys (first inserted wins).
 * layer_* give each layer's [start_id, end_id) and its row offset into the
 * rule columns. Returns 0, -1 on an ID collision with the rule region. */
int oracle_encode_existing(const int32_t *nodes, uint64_t n_nodes, const uint32_t *lengths,
                           uint64_t n_trav, const int32_t *rules_first,
                           const int32_t *rules_second, uint64_t n_rules,
                           const uint32_t *layer_start, const uint32_t *layer_end,
                           const uint64_t *layer_row_offset, uint64_t n_layers,
                           int delta_round, int32_t *out, uint32_t *out_lengths,
                           uint64_t *out_n);

#ifdef __cplusplus
}
#endif
#endif
