"""Cross-traversal tile order of the decode kernels: the decode plan and the
expansion take their tiles column by column over one row per traversal (the same
stretch of every haplotype at once). The order is a permutation of tiles, so the
output must equal the plain order's and the input, also when traversal lengths
are far from equal (rows then cut across traversals) and when the row count is
not a multiple of the XCD count."""
import numpy as np
import pytest

from gfaz_amd import codec, synth
from test_gpu_codec import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph():
    cfg = synth.CONFIGS["plain_500k"].scaled(backbone=1_500_000, n_haps=20)
    return synth.generate(cfg)


def _decode(ctx, comp, lengths, delta):
    rb = codec.Rulebook(ctx, comp.rules_first, comp.rules_second, comp.start_id)
    try:
        return codec.decompress_traversals(ctx, rb, comp.encoded, comp.encoded_lengths,
                                           lengths, delta)
    finally:
        rb.close()


def _uneven(lengths):
    """The same steps cut into traversals of very different lengths, two empty."""
    n = int(lengths.sum())
    cuts = [n // 2, 1, 0, 5, n // 8, n // 16, 100, 0, n // 5]
    cuts.append(n - sum(cuts))
    assert min(cuts) >= 0
    return np.array(cuts, dtype=np.uint32)


@pytest.mark.parametrize("shape", ["haplotypes", "uneven"])
@pytest.mark.parametrize("delta", [0, 1, 2])
def test_tile_order_matches_plain_order(graph, shape, delta):
    nodes = graph.nodes
    lengths = graph.lengths if shape == "haplotypes" else _uneven(graph.lengths)
    say(f"tile order {shape} delta={delta} n={nodes.size} trav={lengths.size}")
    with codec.Context(0) as ordered, codec.Context(0) as plain:
        plain.set_tuned("expand_rows", 0)
        comp = codec.compress_traversals(ordered, nodes, lengths, 0, 8, 2, delta)
        # enough tiles for one row per traversal in both the plan and the expansion
        assert comp.encoded.size >= 2048 * 32 * lengths.size
        got, lens, mism = _decode(ordered, comp, lengths, delta)
        ref, rlens, rmism = _decode(plain, comp, lengths, delta)
    assert not mism and not rmism
    np.testing.assert_array_equal(lens, lengths)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got, nodes)
