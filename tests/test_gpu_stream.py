"""Streaming decompress on the GPU (gfz_decompress_stream): chunks of whole traversals
through two alternating pinned buffers reassemble the input exactly."""
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


def _collect(ctx, rb, enc, enc_len, delta, chunk_bytes):
    parts, meta = [], []

    def consume(values, b, e, expanded_begin, lengths):
        parts.append(values.copy())
        meta.append((b, e, expanded_begin, lengths.copy()))
        return False

    codec.decompress_stream(ctx, rb, enc, enc_len, delta, chunk_bytes, consume)
    return parts, meta


@pytest.mark.parametrize("delta,chunk_bytes", [(1, 64 << 10), (0, 200 << 10), (2, 1 << 30)])
def test_stream_reassembles_input(gpu_ctx, delta, chunk_bytes):
    say(f"stream delta={delta} chunk={chunk_bytes}")
    tr = synth.generate(synth.CONFIGS["plain_20k"])
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, delta, threads=8)
    rb = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
    parts, meta = _collect(gpu_ctx, rb, ref["encoded"], ref["encoded_lengths"], delta, chunk_bytes)
    rb.close()
    np.testing.assert_array_equal(np.concatenate(parts), tr.nodes)
    # chunk bookkeeping: contiguous traversal ranges, offsets and lengths
    t, off = 0, 0
    for (b, e, eb, lens), vals in zip(meta, parts):
        assert b == t and e > b and eb == off
        np.testing.assert_array_equal(lens, tr.lengths[b:e])
        assert vals.size == int(lens.sum())
        # a chunk exceeds the budget only when it is a single traversal
        assert vals.size * 4 <= chunk_bytes or e - b == 1
        t, off = e, off + vals.size
    assert t == tr.lengths.size
    if chunk_bytes == 64 << 10:
        assert len(meta) > 4


def test_stream_goldens_and_stop(gpu_ctx):
    for name in goldens.names():
        g = goldens.load(name)
        nodes, lengths, _ = goldens.inputs(g)
        exp = goldens.expected(g)
        say(f"stream golden {name}")
        rb = codec.Rulebook(gpu_ctx, exp["rules_first"], exp["rules_second"], exp["start_id"])
        parts, _ = _collect(gpu_ctx, rb, exp["encoded"], exp["encoded_lengths"],
                            exp["delta_round"], 4096)
        got = np.concatenate(parts) if parts else np.zeros(0, np.int32)
        np.testing.assert_array_equal(got, nodes)
        rb.close()
    tr = synth.generate(synth.CONFIGS["plain_20k"])
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=8)
    rb = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
    seen = []
    with pytest.raises(RuntimeError, match="stopped by the consumer"):
        codec.decompress_stream(gpu_ctx, rb, ref["encoded"], ref["encoded_lengths"], 1, 64 << 10,
                                lambda v, b, e, x, l: seen.append(b) or len(seen) == 2)
    assert len(seen) == 2
    rb.close()
