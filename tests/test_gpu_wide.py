"""Node IDs near 2^31: the compressor renumbers symbol magnitudes of 2^28 and up when
the rule-ID floor reaches 2^29 (compact pair keys hold magnitudes below 2^30) and
maps the grammar back. Held bit-exactly to the oracle (itself pinned at these
magnitudes by brute force, tests/test_oracle.py), then decoded back to the input."""
import numpy as np
import pytest

from gfaz_amd import codec, synth
from test_gpu_codec import check_compress, check_decompress, say

pytestmark = pytest.mark.gpu


def _shift(nodes, offset):
    """Every node ID moved up by offset, orientation kept."""
    a = np.abs(nodes.astype(np.int64)) + offset
    assert a.max() < 2 ** 31
    return (np.sign(nodes) * a).astype(np.int32)


@pytest.fixture(scope="module")
def base():
    cfg = synth.CONFIGS["plain_20k"]
    return synth.generate(cfg)


# offset 1.5e9: deltas at orientation flips wrap past 2^31 as in the reference;
# 6e8: the floor lands in [2^29, 2^30), where plain compact keys would still fit.
@pytest.mark.parametrize("offset", [1_500_000_000, 600_000_000])
@pytest.mark.parametrize("delta", [0, 1, 2])
def test_large_node_ids_match_oracle(gpu_ctx, base, offset, delta):
    say(f"wide symbols offset={offset} delta={delta}")
    nodes = _shift(base.nodes, offset)
    got = check_compress(gpu_ctx, nodes, base.lengths, 8, delta)
    assert got.start_id >= 1 << 29
    comp = {"encoded": got.encoded, "encoded_lengths": got.encoded_lengths,
            "rules_first": got.rules_first, "rules_second": got.rules_second,
            "start_id": got.start_id}
    check_decompress(gpu_ctx, nodes, base.lengths, comp, delta)


def test_one_huge_id_among_small_ones(gpu_ctx, base):
    """A single step of magnitude 2^31 - 1 lifts the rule-ID floor to 2^31: the
    first rule would be 2^31, which no int32 holds, so compress fails cleanly
    (the reference would wrap), and the context stays usable."""
    nodes = base.nodes.copy()
    nodes[5] = 2 ** 31 - 1
    with pytest.raises(RuntimeError, match="INT32_MAX"):
        codec.compress_traversals(gpu_ctx, nodes, base.lengths, 0, 8, 2, 0)
    nodes[5] = 2 ** 31 - 2_000_000  # room for the rules: same grammar as the oracle's
    check_compress(gpu_ctx, nodes, base.lengths, 8, 0)
