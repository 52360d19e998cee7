"""Compression with each induction round's stretches of launches captured as HIP
graphs (the ctx tuned value "round_graphs" = 1): the grammar equals the oracle's on
every call, while the graphs kept per round are updated in place by repeated calls
and re-made when another input changes their shape, with the program's own eager
work (decompressing each result) and synchronisations between the replays."""
import sys

import numpy as np
import pytest

from gfaz_amd import codec, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(backbone, haps):
    tr = synth.generate(synth.CONFIGS["plain_500k"].scaled(backbone=backbone, n_haps=haps))
    return tr, oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=16)


def _check(ctx, tr, ref):
    got = codec.compress_traversals(ctx, tr.nodes, tr.lengths, tr.n_paths, 8, 2, 1)
    assert (got.start_id, got.end_id) == (ref["start_id"], ref["end_id"])
    np.testing.assert_array_equal(got.encoded_lengths, ref["encoded_lengths"])
    np.testing.assert_array_equal(got.encoded, ref["encoded"])
    np.testing.assert_array_equal(got.rules_first, ref["rules_first"])
    np.testing.assert_array_equal(got.rules_second, ref["rules_second"])
    rb = codec.Rulebook(ctx, got.rules_first, got.rules_second, got.start_id)
    try:
        dec, lens, mism = codec.decompress_traversals(ctx, rb, got.encoded, got.encoded_lengths,
                                                      tr.lengths, 1)
    finally:
        rb.close()
    assert not mism
    np.testing.assert_array_equal(dec, tr.nodes)
    ctx.sync()


def test_round_graphs_give_the_oracle_grammar_on_every_replay():
    a, ref_a = _case(120_000, 24)
    b, ref_b = _case(40_000, 12)
    with codec.Context(0) as ctx:
        ctx.set_tuned("round_graphs", 1)
        for rep in range(3):
            say(f"round graphs: replay {rep}, input A ({a.n_steps} steps)")
            _check(ctx, a, ref_a)
            _check(ctx, a, ref_a)
            say(f"round graphs: replay {rep}, input B ({b.n_steps} steps)")
            _check(ctx, b, ref_b)
        st = ctx.graph_stats()
        say(f"round graphs: {st}")
        # three segments per round, eight rounds, nine compress calls
        assert st["launches"] >= 9 * 3 * 4
        assert st["updated"] > 0 and st["instantiated"] > 0
        ctx.set_tuned("round_graphs", 0)
        _check(ctx, a, ref_a)
        assert ctx.graph_stats()["launches"] == st["launches"]  # off: no graph launched
