"""Add-haplotypes encoding with a stored rulebook on the GPU (gfz_encode_with_rules)
against the reference's own add-haplotypes output (golden) and the CPU oracle."""
import sys

import numpy as np
import pytest

import goldens
from gfaz_amd import codec, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_encode_matches_reference_add_haplotypes(gpu_ctx):
    say("encode golden add6")
    g = goldens.load("plain_20k_add6")
    exp = goldens.expected(goldens.load("plain_20k_r8_t2_d1"))
    enc, lens = codec.encode_with_rules(gpu_ctx, g["appended_nodes"], g["appended_lengths"],
                                        exp["rules_first"], exp["rules_second"],
                                        [(exp["start_id"], exp["end_id"], 0)], 1)
    np.testing.assert_array_equal(lens, g["appended_encoded_lengths"])
    np.testing.assert_array_equal(enc, g["appended_encoded"])


def _layers_by_depth(first, second, start):
    """Split one layer into per-depth layers (depth = creating round for CPU files)."""
    R = first.size
    depth = np.zeros(R, np.int64)
    for r in range(R):
        d = 0
        for c in (first[r], second[r]):
            k = abs(int(c)) - start
            if 0 <= k < R:
                d = max(d, depth[k])
        depth[r] = d + 1
    layers, r0 = [], 0
    for r in range(1, R + 1):
        if r == R or depth[r] != depth[r0]:
            layers.append((start + r0, start + r, r0))
            r0 = r
    return layers


@pytest.mark.parametrize("seed,delta,multi", [(11, 1, False), (12, 1, True), (13, 2, False),
                                              (14, 0, False), (15, 1, True)])
def test_encode_matches_oracle(gpu_ctx, seed, delta, multi):
    say(f"encode seed={seed} delta={delta} multi={multi}")
    base = synth.CONFIGS["plain_20k"]
    tr = synth.generate(base)
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, delta, threads=8)
    new = synth.generate(base.scaled(n_haps=5, seed=base.seed + seed))
    if multi:
        layers = _layers_by_depth(ref["rules_first"], ref["rules_second"], ref["start_id"])
        assert len(layers) > 1
    else:
        layers = [(ref["start_id"], ref["end_id"], 0)]
    want_enc, want_len = oracle.encode_existing(new.nodes, new.lengths, ref["rules_first"],
                                                ref["rules_second"], layers, delta)
    enc, lens = codec.encode_with_rules(gpu_ctx, new.nodes, new.lengths, ref["rules_first"],
                                        ref["rules_second"], layers, delta)
    np.testing.assert_array_equal(lens, want_len)
    np.testing.assert_array_equal(enc, want_enc)
    # and the result decodes back to the new traversals
    rb = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
    dec, dl, mism = codec.decompress_traversals(gpu_ctx, rb, enc, lens, new.lengths, delta)
    rb.close()
    assert not mism
    np.testing.assert_array_equal(dec, new.nodes)


def test_encode_long_chains_across_chunks(gpu_ctx):
    """Long runs of one pair make reductions chain across the 256-position chunks,
    so the fix-up passes must carry states across several chunks."""
    say("encode chains")
    nodes = np.tile(np.array([4, 4, 4, 5], np.int32), 3000)
    lengths = np.array([nodes.size // 2, nodes.size // 2], np.uint32)
    ref = oracle.compress(nodes, lengths, 8, 2, 0, threads=4)
    for extra in ([4] * 1001, [4, 5] * 700 + [4], [5] + [4] * 2047):
        new = np.array(extra, np.int32)
        nl = np.array([new.size], np.uint32)
        want = oracle.encode_existing(new, nl, ref["rules_first"], ref["rules_second"],
                                      [(ref["start_id"], ref["end_id"], 0)], 0)
        got = codec.encode_with_rules(gpu_ctx, new, nl, ref["rules_first"], ref["rules_second"],
                                      [(ref["start_id"], ref["end_id"], 0)], 0)
        np.testing.assert_array_equal(got[1], want[1])
        np.testing.assert_array_equal(got[0], want[0])


def test_encode_errors_and_edges(gpu_ctx):
    say("encode errors")
    rf = np.array([1, 2], np.int32)
    rs = np.array([2, 3], np.int32)
    with pytest.raises(RuntimeError, match="collides with the stored rule region"):
        codec.encode_with_rules(gpu_ctx, np.array([1, 50], np.int32), np.array([2], np.uint32),
                                rf, rs, [(10, 12, 0)], 1)
    with pytest.raises(RuntimeError, match="rule column is truncated"):
        codec.encode_with_rules(gpu_ctx, np.array([1, 2], np.int32), np.array([2], np.uint32),
                                rf, rs, [(10, 13, 0)], 1)
    # no layers: only the delta transform
    enc, lens = codec.encode_with_rules(gpu_ctx, np.array([3, 5, 4], np.int32),
                                        np.array([3], np.uint32), rf, rs, [], 1)
    assert enc.tolist() == [3, 2, -1] and lens.tolist() == [3]
    # empty traversals and duplicate keys (first row wins)
    enc, lens = codec.encode_with_rules(gpu_ctx, np.array([1, 2, 1, 2], np.int32),
                                        np.array([0, 2, 0, 2], np.uint32),
                                        np.array([1, 1], np.int32), np.array([2, 2], np.int32),
                                        [(10, 12, 0)], 0)
    assert enc.tolist() == [10, 10] and lens.tolist() == [0, 1, 0, 1]
