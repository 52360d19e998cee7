"""gfz_compress_framed / gfz_decompress_framed: the codec with its framing
overlapped. Held to the step-by-step path (gfz_compress_traversals, then
gfz_frame_* column after column; gfz_unframe_*, then the decode), whose bytes the
.gfaz goldens pin (test_container.py runs those goldens through the overlapped
path too), at sizes where every column spans several 4 MiB handoff pieces."""
import numpy as np
import pytest

from gfaz_amd import codec, container, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plain_500k():
    tr = synth.generate(synth.CONFIGS["plain_500k"])
    return tr.nodes, tr.lengths


def _cols_equal(a: container.TraversalColumns, b: container.TraversalColumns):
    assert a.layer_rule_ranges == b.layer_rule_ranges
    for f in ("rules_first_zstd", "rules_second_zstd", "paths_zstd", "walks_zstd"):
        assert getattr(a, f).pair() == getattr(b, f).pair(), f
    for f in ("sequence_lengths", "original_path_lengths", "walk_lengths",
              "original_walk_lengths"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))


@pytest.mark.parametrize("path_share", [1.0, 0.5, 0.0])
def test_framed_equals_step_by_step(gpu_ctx, plain_500k, path_share):
    nodes, lengths = plain_500k
    n_paths = int(round(lengths.size * path_share))
    print(f"framed vs step-by-step: {nodes.size} steps, {n_paths} paths", flush=True)
    over = container.compress_traversals_gpu(gpu_ctx, nodes, lengths, n_paths, threads=4)
    step = container.compress_traversals_gpu(gpu_ctx, nodes, lengths, n_paths, threads=4,
                                             overlapped=False)
    _cols_equal(over, step)
    assert over.paths_zstd.original_size + over.walks_zstd.original_size > (8 << 20)
    dec, lens, np_ = container.decompress_traversals_gpu(gpu_ctx, over)
    assert np_ == n_paths
    np.testing.assert_array_equal(lens, lengths)
    np.testing.assert_array_equal(dec, nodes)


def test_framed_timings_and_device_input(gpu_ctx, plain_500k):
    nodes, lengths = plain_500k
    d = codec.DeviceArray.from_host(gpu_ctx, nodes)
    try:
        f = codec.compress_framed(gpu_ctx, d, lengths, lengths.size, threads=4)
    finally:
        d.free()
    g = codec.compress_framed(gpu_ctx, nodes, lengths, lengths.size, threads=4)
    assert (f.paths, f.rules_first, f.rules_second) == (g.paths, g.rules_first, g.rules_second)
    assert f.walks == (0, b"")
    assert 0 < f.gpu_ms <= f.total_ms
    dec, lens, mism, t = codec.decompress_framed(gpu_ctx, f, lengths)
    assert not mism
    assert 0 < t[0] and t[0] <= t[2] and t[1] <= t[2]
    np.testing.assert_array_equal(dec, nodes)


@pytest.mark.parametrize("case", ["no_repeats", "one_step", "empty_walk_and_path", "none"])
def test_framed_edge_cases(gpu_ctx, case):
    if case == "no_repeats":
        nodes, lengths, n_paths = np.arange(1, 5001, dtype=np.int32), np.array([2000, 3000]), 1
    elif case == "one_step":
        nodes, lengths, n_paths = np.array([-7], np.int32), np.array([1]), 0
    elif case == "empty_walk_and_path":
        nodes = np.array([1, 2, 1, 2, 1, 2, 3], np.int32)
        lengths, n_paths = np.array([0, 7, 0]), 1
    else:
        nodes, lengths, n_paths = np.zeros(0, np.int32), np.zeros(0), 0
    lengths = np.asarray(lengths, np.uint32)
    print(f"framed edge case {case}", flush=True)
    over = container.compress_traversals_gpu(gpu_ctx, nodes, lengths, n_paths)
    step = container.compress_traversals_gpu(gpu_ctx, nodes, lengths, n_paths, overlapped=False)
    _cols_equal(over, step)
    if nodes.size:
        dec, lens, _ = container.decompress_traversals_gpu(gpu_ctx, over)
        np.testing.assert_array_equal(dec, nodes)
        np.testing.assert_array_equal(lens, lengths)


def test_framed_decompress_rejects_inconsistent_columns(gpu_ctx):
    nodes = np.tile(np.array([1, 2, 3, 4], np.int32), 500)
    lengths = np.array([1000, 1000], np.uint32)
    f = codec.compress_framed(gpu_ctx, nodes, lengths, 1)
    bad = codec.Framed(**{**vars(f), "encoded_lengths": f.encoded_lengths + 1})
    with pytest.raises(RuntimeError, match="disagree"):
        codec.decompress_framed(gpu_ctx, bad, lengths)
    bad = codec.Framed(**{**vars(f), "paths": (f.paths[0], f.paths[1][:-3])})
    with pytest.raises(RuntimeError, match="zstd"):
        codec.decompress_framed(gpu_ctx, bad, lengths)
    dec, _, _, _ = codec.decompress_framed(gpu_ctx, f, lengths)  # ctx still usable
    np.testing.assert_array_equal(dec, nodes)
