"""Traversal fields of the reference's CompressedData, produced and consumed on the GPU.

Mirrors the reference's GPU boundary for this path:
  compress_traversals_gpu   ~ run_path_compression_gpu
                              (include/gpu/compression/compression_workflow_gpu.hpp:75-78),
                              which returns CompressedData with the traversal fields filled;
  decompress_traversals_gpu ~ decompress_paths_gpu
                              (include/gpu/decompression/decompression_workflow_gpu.hpp:32-33),
                              which returns FlattenedPaths (paths first, then walks).
The grammar comes from the HIP codec (gfz_compress_traversals / gfz_decompress_traversals);
the column framing (rule-column delta + zstd, paths and walks flattened separately) is the
library's host framing (gfz_frame_* / gfz_unframe_*), byte-identical to the reference CPU
workflow, so a .gfaz written from these fields matches the reference's and the reference
CPU path decodes it (and the reverse).

Field names follow include/model/compressed_data.hpp:43-66,108-119.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import codec


@dataclass
class ZstdBlock:
    """ZstdCompressedBlock (include/model/compressed_data.hpp:18-21)."""
    original_size: int = 0
    payload: bytes = b""

    @staticmethod
    def of(t: tuple[int, bytes]) -> "ZstdBlock":
        return ZstdBlock(int(t[0]), bytes(t[1]))

    def pair(self) -> tuple[int, bytes]:
        return self.original_size, self.payload


@dataclass
class LayerRuleRange:
    """include/model/compressed_data.hpp:10-16."""
    k: int
    start_id: int
    end_id: int
    flattened_offset: int
    element_count: int


@dataclass
class TraversalColumns:
    """The traversal part of CompressedData."""
    layer_rule_ranges: list[LayerRuleRange] = field(default_factory=list)
    rules_first_zstd: ZstdBlock = field(default_factory=ZstdBlock)
    rules_second_zstd: ZstdBlock = field(default_factory=ZstdBlock)
    delta_round: int = 1
    sequence_lengths: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    original_path_lengths: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    paths_zstd: ZstdBlock = field(default_factory=ZstdBlock)
    walk_lengths: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    original_walk_lengths: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    walks_zstd: ZstdBlock = field(default_factory=ZstdBlock)

    @property
    def min_rule_id(self) -> int:
        """compressed_data.hpp:49-51: the first range's start (0 without ranges)."""
        return self.layer_rule_ranges[0].start_id if self.layer_rule_ranges else 0

    @property
    def n_rules(self) -> int:
        return self.rules_first_zstd.original_size // 4


def compress_traversals_gpu(ctx: codec.Context, nodes, lengths, n_paths: int, rounds: int = 8,
                            threshold: int = 2, delta_round: int = 1, threads: int = 0,
                            overlapped: bool = True) -> TraversalColumns:
    """Grammar-compress paths (the first n_paths traversals) and walks on the GPU and
    frame the columns as the reference CPU workflow does
    (src/workflows/compression_workflow.cpp:548-700, compression_utils.cpp:38-54,165-193).
    overlapped: one gfz_compress_framed call (columns framed on host threads as their
    pieces leave HBM); else the grammar comes back whole and is framed column after
    column, as the reference does. Both give the same bytes."""
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    if not 0 <= n_paths <= lengths.size:
        raise ValueError(f"n_paths {n_paths} outside 0..{lengths.size}")
    out = TraversalColumns()
    out.delta_round = max(0, int(delta_round))
    if overlapped:
        f = codec.compress_framed(ctx, nodes, lengths, n_paths, rounds, threshold, delta_round,
                                  threads)
        enc_len = f.encoded_lengths.astype(np.uint32)
        out.layer_rule_ranges = [LayerRuleRange(2, f.start_id, f.end_id, 0, 2 * f.n_rules)]
        out.rules_first_zstd = ZstdBlock.of(f.rules_first)
        out.rules_second_zstd = ZstdBlock.of(f.rules_second)
        out.paths_zstd = ZstdBlock.of(f.paths)
        out.walks_zstd = ZstdBlock.of(f.walks)
    else:
        got = codec.compress_traversals(ctx, nodes, lengths, n_paths, rounds, threshold,
                                        delta_round)
        enc_len = got.encoded_lengths.astype(np.uint32)
        split = int(enc_len[:n_paths].sum(dtype=np.uint64))
        n_rules = int(got.rules_first.size)
        out.layer_rule_ranges = [LayerRuleRange(2, got.start_id, got.end_id, 0, 2 * n_rules)]
        f, s = codec.frame_rule_columns(got.rules_first, got.rules_second, threads)
        out.rules_first_zstd, out.rules_second_zstd = ZstdBlock.of(f), ZstdBlock.of(s)
        out.paths_zstd = ZstdBlock.of(codec.frame_int32(got.encoded[:split], threads))
        out.walks_zstd = ZstdBlock.of(codec.frame_int32(got.encoded[split:], threads))
    out.sequence_lengths = enc_len[:n_paths].copy()
    out.original_path_lengths = lengths[:n_paths].copy()
    out.walk_lengths = enc_len[n_paths:].copy()
    out.original_walk_lengths = lengths[n_paths:].copy()
    return out


def decompress_traversals_gpu(ctx: codec.Context, cols: TraversalColumns,
                              overlapped: bool = True):
    """Decode every path and walk on the GPU: (nodes, lengths, n_paths), paths first
    (decompression_workflow.cpp:110-174 semantics; lengths that disagree with the stored
    originals only warn there, and are returned as decoded here).
    overlapped: one gfz_decompress_framed call (payloads decoded on host threads,
    rulebook built meanwhile, columns uploaded in pieces); else unframe everything,
    then upload and decode, as the reference does."""
    enc_len = np.concatenate([cols.sequence_lengths, cols.walk_lengths]).astype(np.uint32)
    orig = np.concatenate([cols.original_path_lengths,
                           cols.original_walk_lengths]).astype(np.uint32)
    if overlapped:
        f = codec.Framed(cols.min_rule_id, cols.min_rule_id + cols.n_rules, cols.n_rules, 0,
                         cols.delta_round, int(cols.sequence_lengths.size), enc_len,
                         cols.rules_first_zstd.pair(), cols.rules_second_zstd.pair(),
                         cols.paths_zstd.pair(), cols.walks_zstd.pair())
        dec, lens, _, _ = codec.decompress_framed(ctx, f, orig)
        return dec, lens, int(cols.sequence_lengths.size)
    rf, rs = codec.unframe_rule_columns(cols.rules_first_zstd.pair(),
                                        cols.rules_second_zstd.pair(), cols.n_rules)
    paths = codec.unframe_int32(cols.paths_zstd.payload, cols.paths_zstd.original_size)
    walks = codec.unframe_int32(cols.walks_zstd.payload, cols.walks_zstd.original_size)
    enc = np.concatenate([paths, walks]).astype(np.int32)
    rb = codec.Rulebook(ctx, rf, rs, cols.min_rule_id)
    try:
        dec, lens, _ = codec.decompress_traversals(ctx, rb, enc, enc_len, orig, cols.delta_round)
    finally:
        rb.close()
    return dec, lens, int(cols.sequence_lengths.size)


def columns_from_gfaz(g) -> TraversalColumns:
    """TraversalColumns of a parsed .gfaz (any reader exposing the v5 fields)."""
    c = TraversalColumns()
    c.layer_rule_ranges = [LayerRuleRange(*r) for r in g.ranges]
    c.rules_first_zstd = ZstdBlock(g.rules_first.original_size, g.rules_first.payload)
    c.rules_second_zstd = ZstdBlock(g.rules_second.original_size, g.rules_second.payload)
    c.delta_round = g.delta_round
    c.sequence_lengths = np.asarray(g.sequence_lengths, np.uint32)
    c.original_path_lengths = np.asarray(g.original_path_lengths, np.uint32)
    c.paths_zstd = ZstdBlock(g.paths.original_size, g.paths.payload)
    c.walk_lengths = np.asarray(g.walk_lengths, np.uint32)
    c.original_walk_lengths = np.asarray(g.original_walk_lengths, np.uint32)
    c.walks_zstd = ZstdBlock(g.walks.original_size, g.walks.payload)
    return c
