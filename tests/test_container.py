"""Traversal columns of CompressedData against .gfaz files written by the reference CPU
path (tests/golden/*.gfaz): the host framing reproduces their bytes, and the GPU codec
plus framing writes and reads them (cross-backend parity)."""
import os
import sys

import numpy as np
import pytest

import goldens
from gfaz_amd import codec, container

sys.path.insert(0, goldens.GOLDEN_DIR)
import gfaz_py  # noqa: E402

GFAZ_CASES = sorted(n[:-5] for n in os.listdir(goldens.GOLDEN_DIR)
                    if n.endswith(".gfaz") and os.path.exists(
                        os.path.join(goldens.GOLDEN_DIR, n[:-5] + ".npz")))


def _gfaz(name):
    return gfaz_py.read_gfaz(os.path.join(goldens.GOLDEN_DIR, name + ".gfaz"))


def test_cases_present():
    assert len(GFAZ_CASES) >= 5, GFAZ_CASES


@pytest.mark.parametrize("name", GFAZ_CASES)
@pytest.mark.parametrize("threads", [1, 4])
def test_framing_matches_reference_bytes(name, threads):
    """Rule-column delta + zstd, and the paths/walks columns, framed from the golden
    grammar give the reference file's payload bytes (for any zstd worker count >= 1)."""
    g = goldens.load(name)
    z = _gfaz(name)
    f, s = codec.frame_rule_columns(g["rules_first"], g["rules_second"], threads)
    assert f == (z.rules_first.original_size, z.rules_first.payload)
    assert s == (z.rules_second.original_size, z.rules_second.payload)
    assert codec.frame_int32(g["paths_encoded"], threads) == (z.paths.original_size, z.paths.payload)
    assert codec.frame_int32(g["walks_encoded"], threads) == (z.walks.original_size, z.walks.payload)


@pytest.mark.parametrize("name", GFAZ_CASES)
def test_unframing_reference_columns(name):
    g = goldens.load(name)
    z = _gfaz(name)
    n = int(g["rules_first"].size)
    f, s = codec.unframe_rule_columns((z.rules_first.original_size, z.rules_first.payload),
                                      (z.rules_second.original_size, z.rules_second.payload), n)
    np.testing.assert_array_equal(f, g["rules_first"])
    np.testing.assert_array_equal(s, g["rules_second"])
    np.testing.assert_array_equal(codec.unframe_int32(z.paths.payload, z.paths.original_size),
                                  g["paths_encoded"])


def test_unframing_rejects_bad_payloads():
    with pytest.raises(RuntimeError, match="zstd decompression failed"):
        codec.unframe_int32(b"not a zstd frame", 16)
    with pytest.raises(RuntimeError, match="multiple of 4"):
        codec.unframe_int32(b"", 6)
    with pytest.raises(RuntimeError, match="disagree"):
        codec.unframe_rule_columns((8, b""), (8, b""), 3)
    orig, payload = codec.frame_int32(np.arange(10, dtype=np.int32))
    with pytest.raises(RuntimeError, match="zstd decompression failed"):
        codec.unframe_int32(payload, orig - 4)   # shorter destination than the frame holds


def test_empty_columns_frame_to_empty_blocks():
    assert codec.frame_int32(np.zeros(0, np.int32)) == (0, b"")
    assert codec.frame_rule_columns([], []) == ((0, b""), (0, b""))
    assert codec.unframe_int32(b"", 0).size == 0


@pytest.mark.gpu
@pytest.mark.parametrize("name", GFAZ_CASES)
def test_gpu_written_columns_equal_reference_file(gpu_ctx, name):
    """MI355X compress + host framing == the reference CPU .gfaz traversal fields."""
    g = goldens.load(name)
    nodes, lengths, n_paths = goldens.inputs(g)
    m = g["meta"]
    z = _gfaz(name)
    cols = container.compress_traversals_gpu(gpu_ctx, nodes, lengths, n_paths, m["rounds"],
                                             m["threshold"], m["delta_round"], threads=4)
    assert [tuple(vars(r).values()) for r in cols.layer_rule_ranges] == [tuple(r) for r in z.ranges]
    assert cols.rules_first_zstd.pair() == (z.rules_first.original_size, z.rules_first.payload)
    assert cols.rules_second_zstd.pair() == (z.rules_second.original_size, z.rules_second.payload)
    assert cols.paths_zstd.pair() == (z.paths.original_size, z.paths.payload)
    assert cols.walks_zstd.pair() == (z.walks.original_size, z.walks.payload)
    assert cols.delta_round == z.delta_round
    np.testing.assert_array_equal(cols.sequence_lengths, z.sequence_lengths)
    np.testing.assert_array_equal(cols.original_path_lengths, z.original_path_lengths)
    np.testing.assert_array_equal(cols.walk_lengths, z.walk_lengths)
    np.testing.assert_array_equal(cols.original_walk_lengths, z.original_walk_lengths)


@pytest.mark.gpu
@pytest.mark.parametrize("name", GFAZ_CASES)
def test_gpu_decodes_reference_file(gpu_ctx, name):
    """A .gfaz written by the reference CPU path decodes bit-exactly on the MI355X."""
    g = goldens.load(name)
    nodes, lengths, n_paths = goldens.inputs(g)
    cols = container.columns_from_gfaz(_gfaz(name))
    dec, lens, np_ = container.decompress_traversals_gpu(gpu_ctx, cols)
    assert np_ == n_paths
    np.testing.assert_array_equal(lens, lengths)
    np.testing.assert_array_equal(dec, nodes)
