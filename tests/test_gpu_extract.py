"""Resident decoding (gfz_resident_load / gfz_extract) on the GPU: any subset of
traversals, in any order and with repeats, decodes bit-exactly to the input
traversals (the CPU oracle and the reference goldens fix those)."""
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


def _split(nodes, lengths):
    offs = np.concatenate([[0], np.cumsum(lengths, dtype=np.int64)])
    return [nodes[offs[i]:offs[i + 1]] for i in range(lengths.size)]


def _check_subset(res, trav, ids, out=None):
    got, offs = res.extract(ids, out=out)
    if out is not None:
        got = out.to_host()[: int(offs[-1])]
    want = [trav[i] for i in ids]
    exp_offs = np.concatenate([[0], np.cumsum([w.size for w in want], dtype=np.int64)])
    np.testing.assert_array_equal(offs, exp_offs)
    np.testing.assert_array_equal(got, np.concatenate(want) if want else np.zeros(0, np.int32))


@pytest.mark.parametrize("delta", [0, 1, 2])
def test_extract_subsets_match_input(gpu_ctx, delta):
    say(f"extract subsets delta={delta}")
    tr = synth.generate(synth.CONFIGS["plain_20k"])
    lengths = tr.lengths.copy()
    ref = oracle.compress(tr.nodes, lengths, 8, 2, delta, threads=8)
    rb = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
    res = codec.Resident(gpu_ctx, rb, ref["encoded"], ref["encoded_lengths"], delta)
    trav = _split(tr.nodes, lengths)
    np.testing.assert_array_equal(res.lengths(), lengths)
    assert res.n_decoded == tr.n_steps
    rng = np.random.default_rng(delta)
    _check_subset(res, trav, np.arange(lengths.size))
    _check_subset(res, trav, np.arange(lengths.size)[::-1])
    _check_subset(res, trav, rng.integers(0, lengths.size, 40))     # repeats
    _check_subset(res, trav, [3])
    _check_subset(res, trav, [])
    out = codec.DeviceArray(gpu_ctx, tr.n_steps * 2, np.int32)
    _check_subset(res, trav, rng.integers(0, lengths.size, 20), out=out)
    with pytest.raises(RuntimeError, match="out of range"):
        res.extract([lengths.size])
    res.close()
    rb.close()


def test_extract_goldens_with_empty_traversals_and_walks(gpu_ctx):
    for name in goldens.names():
        g = goldens.load(name)
        nodes, lengths, _ = goldens.inputs(g)
        exp = goldens.expected(g)
        say(f"extract golden {name}")
        rb = codec.Rulebook(gpu_ctx, exp["rules_first"], exp["rules_second"], exp["start_id"])
        res = codec.Resident(gpu_ctx, rb, exp["encoded"], exp["encoded_lengths"], exp["delta_round"])
        trav = _split(nodes, lengths)
        _check_subset(res, trav, np.arange(lengths.size))
        _check_subset(res, trav, np.arange(lengths.size)[::-2])
        res.close()
        rb.close()


def test_extract_long_runs_and_tile_sizes(gpu_ctx):
    """A traversal of one repeated node compresses to a few symbols, each covering
    many tiles; every autotuner candidate decodes it exactly."""
    say("extract long runs")
    parts = [np.full(300_001, 7, np.int32), np.arange(1, 5000, dtype=np.int32),
             np.full(4097, -3, np.int32), np.zeros(0, np.int32), np.array([9], np.int32)]
    lengths = np.array([p.size for p in parts], np.uint32)
    nodes = np.concatenate(parts)
    ref = oracle.compress(nodes, lengths, 8, 2, 1, threads=8)
    rb = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
    res = codec.Resident(gpu_ctx, rb, ref["encoded"], ref["encoded_lengths"], 1)
    gpu_ctx.set_autotune(False)
    try:
        for items in (8, 16, 32):
            for c in range(12):
                gpu_ctx.set_tuned(f"extract_items/{c}", items)
            _check_subset(res, parts, [0, 1, 2, 3, 4, 0])
            _check_subset(res, parts, [2, 4, 1])
    finally:
        gpu_ctx.set_autotune(True)
        res.close()
        rb.close()


def test_extract_without_rules(gpu_ctx):
    """No pair repeats: the rulebook is empty and every symbol is a node ID."""
    nodes = np.array([5, -6, 7, 8, -9], np.int32)
    lengths = np.array([2, 0, 3], np.uint32)
    ref = oracle.compress(nodes, lengths, 8, 2, 1, threads=1)
    assert ref["rules_first"].size == 0
    rb = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
    res = codec.Resident(gpu_ctx, rb, ref["encoded"], ref["encoded_lengths"], 1)
    _check_subset(res, _split(nodes, lengths), [2, 1, 0])
    res.close()
    rb.close()


@pytest.mark.parametrize("delta", [0, 1, 2])
def test_resident_append_then_extract(gpu_ctx, delta):
    """Traversals appended to a resident store (gfz_resident_append) decode exactly,
    and so do the stored ones whose coarse-index cells the append rewrote: the
    traversals of one oracle compression are loaded in part, the rest appended in
    batches (one with an empty traversal), then every subset is extracted."""
    say(f"resident append delta={delta}")
    tr = synth.generate(synth.CONFIGS["plain_20k"])
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, delta, threads=8)
    trav = _split(tr.nodes, tr.lengths)
    enc = _split(ref["encoded"], ref["encoded_lengths"])
    rb = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
    base = 7
    res = codec.Resident(gpu_ctx, rb, np.concatenate(enc[:base]),
                         ref["encoded_lengths"][:base], delta)
    order = list(range(base))
    for batch in ([base], [base + 1, base + 2], [], [base + 3, base + 4, base + 5]):
        # an empty traversal rides along with the second batch
        parts = [enc[i] for i in batch]
        lens = [p.size for p in parts]
        if len(batch) == 2:
            parts.append(np.zeros(0, np.int32))
            lens.append(0)
        first = res.append(np.concatenate(parts) if parts else np.zeros(0, np.int32),
                           np.array(lens, np.uint32))
        assert first == len(order)
        order += batch + ([None] if len(batch) == 2 else [])
        stored = [trav[i] if i is not None else np.zeros(0, np.int32) for i in order]
        assert res.n_trav == len(stored)
        np.testing.assert_array_equal(res.lengths(), [s.size for s in stored])
        _check_subset(res, stored, np.arange(len(stored)))
        _check_subset(res, stored, np.arange(len(stored))[::-1])
        _check_subset(res, stored, np.random.default_rng(len(order)).integers(0, len(stored), 25))
    # appended to a store that started empty
    empty = codec.Resident(gpu_ctx, rb, np.zeros(0, np.int32), np.zeros(0, np.uint32), delta)
    assert empty.append(np.concatenate(enc[:3]), ref["encoded_lengths"][:3]) == 0
    _check_subset(empty, trav[:3], [2, 0, 1, 1])
    empty.close()
    res.close()
    rb.close()


@pytest.mark.parametrize("delta", [0, 1, 2])
def test_extract_append_stream_without_bodies(gpu_ctx, monkeypatch, delta):
    """A rulebook without bodies in HBM (budget 0 here; in use, bodies past the body
    budget or depths not monotone in rule ID) still backs a resident store: extraction
    gathers the requested traversals' symbols and expands them by descent, appends
    and streaming go through the same path, and every result equals the input
    traversals. The no-alloc entry point refuses such a store cleanly."""
    say(f"extract without bodies delta={delta}")
    tr = synth.generate(synth.CONFIGS["plain_20k"])
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, delta, threads=8)
    trav = _split(tr.nodes, tr.lengths)
    enc = _split(ref["encoded"], ref["encoded_lengths"])
    monkeypatch.setenv("GFAZ_BODY_BUDGET_MB", "0")
    rb = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
    assert not rb.info()["has_bodies"]
    base = 9
    res = codec.Resident(gpu_ctx, rb, np.concatenate(enc[:base]),
                         ref["encoded_lengths"][:base], delta)
    assert res.n_decoded == sum(t.size for t in trav[:base])
    rng = np.random.default_rng(40 + delta)
    _check_subset(res, trav[:base], np.arange(base))
    _check_subset(res, trav[:base], rng.integers(0, base, 30))
    _check_subset(res, trav[:base], [])
    res.append(np.concatenate(enc[base:]), ref["encoded_lengths"][base:])
    np.testing.assert_array_equal(res.lengths(), tr.lengths)
    _check_subset(res, trav, np.arange(len(trav))[::-1])
    _check_subset(res, trav, rng.integers(0, len(trav), 30))
    out = codec.DeviceArray(gpu_ctx, 16, np.int32)
    ids = np.array([0], np.uint64)
    offs = np.zeros(2, np.uint64)
    n = codec._u64()
    rc = codec.lib().gfz_extract_noalloc(gpu_ctx.handle, res.handle, ids.ctypes.data, 1, out.ptr,
                                         out.n, offs.ctypes.data, codec.ctypes.byref(n))
    assert rc != 0 and b"descent" in codec.lib().gfz_last_error()
    res.close()
    parts, meta = [], []

    def consume(values, b, e, expanded_begin, lengths):
        parts.append(values.copy())
        meta.append((b, e))
        return False

    codec.decompress_stream(gpu_ctx, rb, ref["encoded"], ref["encoded_lengths"], delta,
                            256 << 10, consume)
    assert len(parts) > 1 and meta[0][0] == 0 and meta[-1][1] == len(trav)
    np.testing.assert_array_equal(np.concatenate(parts), tr.nodes)
    rb.close()
