"""GPU parity tests: the HIP codec through the C ABI against the reference's golden
outputs and the CPU oracle, bit-exact (integer work, no tolerance)."""
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


def check_compress(ctx, nodes, lengths, rounds, delta, n_paths=0, threshold=2, ref=None):
    say(f"  compress n={nodes.size} trav={lengths.size} rounds={rounds} delta={delta}")
    got = codec.compress_traversals(ctx, nodes, lengths, n_paths, rounds, threshold, delta)
    if ref is None:
        ref = oracle.compress(nodes, lengths, rounds, threshold, delta, threads=8)
    assert (got.start_id, got.end_id) == (ref["start_id"], ref["end_id"])
    np.testing.assert_array_equal(got.encoded_lengths, ref["encoded_lengths"])
    np.testing.assert_array_equal(got.encoded, ref["encoded"])
    np.testing.assert_array_equal(got.rules_first, ref["rules_first"])
    np.testing.assert_array_equal(got.rules_second, ref["rules_second"])
    return got


def check_decompress(ctx, nodes, lengths, comp, delta):
    say(f"  decompress enc={comp['encoded'].size} rules={comp['rules_first'].size}")
    rb = codec.Rulebook(ctx, comp["rules_first"], comp["rules_second"], comp["start_id"])
    dec, lens, mism = codec.decompress_traversals(ctx, rb, comp["encoded"],
                                                  comp["encoded_lengths"], lengths, delta)
    rb.close()
    assert not mism
    np.testing.assert_array_equal(lens, lengths)
    np.testing.assert_array_equal(dec, nodes)


@pytest.mark.parametrize("name", goldens.names())
def test_compress_matches_reference_golden(gpu_ctx, name):
    say(f"golden compress {name}")
    g = goldens.load(name)
    nodes, lengths, n_paths = goldens.inputs(g)
    m = g["meta"]
    exp = goldens.expected(g)
    check_compress(gpu_ctx, nodes, lengths, m["rounds"], m["delta_round"], n_paths,
                   m["threshold"], ref=exp)


@pytest.mark.parametrize("name", goldens.names())
def test_decompress_reference_golden(gpu_ctx, name):
    say(f"golden decompress {name}")
    g = goldens.load(name)
    nodes, lengths, _ = goldens.inputs(g)
    exp = goldens.expected(g)
    check_decompress(gpu_ctx, nodes, lengths, exp, exp["delta_round"])


def _random_case(seed, n_trav, max_len, alphabet, zero_frac=0.1):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len + 1, size=n_trav).astype(np.uint32)
    lens[rng.random(n_trav) < zero_frac] = 0
    n = int(lens.sum(dtype=np.uint64))
    nodes = (rng.integers(1, alphabet + 1, size=n) * rng.choice([-1, 1], size=n)).astype(np.int32)
    return nodes, lens


@pytest.mark.parametrize("seed,n_trav,max_len,alphabet,rounds,delta", [
    (1, 1, 5000, 2, 8, 0),       # dense candidate runs crossing tile boundaries
    (2, 200, 3000, 3, 8, 1),
    (3, 50, 20000, 5, 8, 2),
    (4, 1000, 40, 2, 3, 0),      # many short traversals, empty ones, single steps
    (5, 7, 100000, 4, 8, 1),
    (6, 3, 1, 2, 8, 1),          # nothing to pair
    (7, 5000, 30, 3, 8, 1),      # thousands of traversal starts: several per decode tile
    (8, 3000, 60, 2, 8, 2),
])
def test_random_matches_oracle(gpu_ctx, seed, n_trav, max_len, alphabet, rounds, delta):
    say(f"random seed={seed}")
    nodes, lengths = _random_case(seed, n_trav, max_len, alphabet)
    ref = oracle.compress(nodes, lengths, rounds, 2, delta, threads=8)
    check_compress(gpu_ctx, nodes, lengths, rounds, delta, ref=ref)
    check_decompress(gpu_ctx, nodes, lengths, ref, delta)


def test_long_single_run_across_tiles(gpu_ctx):
    """One traversal of a repeated node: every pair is the same key, so the greedy
    parity must carry across thousands of tiles."""
    say("long run")
    for n in (4095, 4096, 4097, 8193, 300_001):
        nodes = np.full(n, 7, dtype=np.int32)
        lengths = np.array([n], dtype=np.uint32)
        ref = check_compress(gpu_ctx, nodes, lengths, 8, 1)
        comp = {"encoded": ref.encoded, "encoded_lengths": ref.encoded_lengths,
                "rules_first": ref.rules_first, "rules_second": ref.rules_second,
                "start_id": ref.start_id}
        check_decompress(gpu_ctx, nodes, lengths, comp, 1)


def test_empty_inputs(gpu_ctx):
    say("empty")
    for lengths in ([], [0], [0, 0, 0]):
        lengths = np.array(lengths, dtype=np.uint32)
        got = codec.compress_traversals(gpu_ctx, np.zeros(0, np.int32), lengths, 0, 8, 2, 1)
        ref = oracle.compress(np.zeros(0, np.int32), lengths, 8, 2, 1)
        assert got.encoded.size == 0 and got.n_rules == 0
        assert (got.start_id, got.end_id) == (ref["start_id"], ref["end_id"]) == (1, 1)
        np.testing.assert_array_equal(got.encoded_lengths, lengths)


# 200 K x 48 (about 11.9 M steps): the first rounds have more than 2048 x 2048
# pairs, so each bucket-partition block takes several tiles.
@pytest.mark.parametrize("cfg_name,backbone,haps", [("plain_500k", 100_000, 32),
                                                    ("plain_500k", 200_000, 48),
                                                    ("walks_2m", 200_000, 100)])
def test_synthetic_scaled_matches_oracle(gpu_ctx, cfg_name, backbone, haps):
    cfg = synth.CONFIGS[cfg_name].scaled(backbone=backbone, n_haps=haps, len_max=200_000)
    say(f"synthetic {cfg_name} scaled backbone={backbone}")
    tr = synth.generate(cfg)
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=8)
    check_compress(gpu_ctx, tr.nodes, tr.lengths, 8, 1, tr.n_paths, ref=ref)
    check_decompress(gpu_ctx, tr.nodes, tr.lengths, ref, 1)


def test_full_size_plain_500k_matches_oracle(gpu_ctx):
    """The benchmark workload itself (39.57 M steps, 64 P-lines, rounds=8,
    threshold=2, delta_round=1): every round's counting uses buckets and the
    first rounds give each bucket-partition block about ten tiles."""
    cfg = synth.CONFIGS["plain_500k"]
    say("full-size plain_500k")
    tr = synth.generate(cfg)
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=16)
    check_compress(gpu_ctx, tr.nodes, tr.lengths, 8, 1, tr.n_paths, ref=ref)
    check_decompress(gpu_ctx, tr.nodes, tr.lengths, ref, 1)


def test_large_walks_match_oracle(gpu_ctx):
    """W-lines of log-uniform length at 107 M steps (walks_2m with 400 walks):
    the largest walks case the oracle's 31-bit provisional IDs allow."""
    cfg = synth.CONFIGS["walks_2m"].scaled(n_haps=400)
    say("walks 107M")
    tr = synth.generate(cfg)
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=16)
    check_compress(gpu_ctx, tr.nodes, tr.lengths, 8, 1, tr.n_paths, ref=ref)


def test_device_resident_paths_and_determinism(gpu_ctx):
    say("device resident")
    tr = synth.generate(synth.CONFIGS["plain_20k"])
    d_nodes = codec.DeviceArray.from_host(gpu_ctx, tr.nodes)
    a = codec.compress_traversals(gpu_ctx, d_nodes, tr.lengths, 16, 8, 2, 1)
    b = codec.compress_traversals(gpu_ctx, tr.nodes, tr.lengths, 16, 8, 2, 1)
    np.testing.assert_array_equal(a.encoded, b.encoded)
    np.testing.assert_array_equal(a.rules_first, b.rules_first)
    rb = codec.Rulebook(gpu_ctx, a.rules_first, a.rules_second, a.start_id)
    d_enc = codec.DeviceArray.from_host(gpu_ctx, a.encoded)
    out = codec.DeviceArray(gpu_ctx, tr.n_steps, np.int32)
    res, lens, mism = codec.decompress_traversals(gpu_ctx, rb, d_enc, a.encoded_lengths,
                                                  tr.lengths, 1, out=out)
    assert res is out and not mism
    np.testing.assert_array_equal(out.to_host(), tr.nodes)


def test_corrupt_rulebooks_fail_cleanly(gpu_ctx):
    say("corrupt rulebooks")
    # rule 10 = (11, 1), rule 11 = (10, 2): a cycle
    with pytest.raises(RuntimeError, match="cyclic"):
        codec.Rulebook(gpu_ctx, [11, 10], [1, 2], 10)
    # decoded lengths that disagree with the stored originals only flag a mismatch
    rb = codec.Rulebook(gpu_ctx, [1], [2], 10)
    dec, lens, mism = codec.decompress_traversals(gpu_ctx, rb, np.array([10, 3], np.int32),
                                                  np.array([2], np.uint32),
                                                  np.array([2], np.uint32), 0)
    assert mism and dec.tolist() == [1, 2, 3] and lens.tolist() == [3]
    # rule sizes: (1,2) -> 2
    assert rb.rule_size(0) == 2
    with pytest.raises(RuntimeError, match="truncated"):
        codec.decompress_traversals(gpu_ctx, rb, np.array([10], np.int32),
                                    np.array([5], np.uint32))


def test_rulebook_of_another_ctx_is_refused(gpu_ctx):
    # A rulebook's last kernels run on its own ctx's stream, so every call that decodes
    # with it must use that ctx.
    say("rulebook ctx check")
    other = codec.Context(0)
    try:
        rb = codec.Rulebook(other, [1], [2], 10)
        enc, lens = np.array([10, 3], np.int32), np.array([2], np.uint32)
        with pytest.raises(RuntimeError, match="another ctx"):
            codec.decompress_traversals(gpu_ctx, rb, enc, lens)  # plans through gfz_decoded_size
        with pytest.raises(RuntimeError, match="another ctx"):
            codec.decompress_traversals(gpu_ctx, rb, enc, lens, np.array([3], np.uint32), 0)
        with pytest.raises(RuntimeError, match="another ctx"):
            codec.decompress_stream(gpu_ctx, rb, enc, lens, 0, consumer=lambda *a: None)
        with pytest.raises(RuntimeError, match="another ctx"):
            codec.Resident(gpu_ctx, rb, enc, lens, delta_round=0)
        dec, got_lens, mism = codec.decompress_traversals(other, rb, enc, lens,
                                                          np.array([3], np.uint32), 0)
        assert not mism and dec.tolist() == [1, 2, 3] and got_lens.tolist() == [3]
        rb.close()
    finally:
        other.close()


def test_bodies_and_descent_paths_agree(gpu_ctx, monkeypatch):
    """Expansion through pre-expanded bodies (default) and by descending the rule
    tree (bodies disabled by a zero budget) both reproduce the input."""
    g = goldens.load("plain_20k_r8_t2_d1")
    nodes, lengths, _ = goldens.inputs(g)
    exp = goldens.expected(g)
    rb = codec.Rulebook(gpu_ctx, exp["rules_first"], exp["rules_second"], exp["start_id"])
    info = rb.info()
    say(f"bodies info {info}")
    assert info["has_bodies"] and info["max_depth"] == 7
    rb.close()
    monkeypatch.setenv("GFAZ_BODY_BUDGET_MB", "0")
    rb = codec.Rulebook(gpu_ctx, exp["rules_first"], exp["rules_second"], exp["start_id"])
    assert not rb.info()["has_bodies"]
    dec, _, mism = codec.decompress_traversals(gpu_ctx, rb, exp["encoded"], exp["encoded_lengths"],
                                               lengths, 1)
    rb.close()
    assert not mism
    np.testing.assert_array_equal(dec, nodes)


def test_non_topological_rulebook_uses_descent(gpu_ctx):
    """A child with a higher ID than its parent (legal for the recursive CPU
    decoder) makes depth non-monotone in rule ID: expansion still succeeds."""
    say("non-topological rulebook")
    rb = codec.Rulebook(gpu_ctx, [11, 1], [1, 2], 10)  # 10 = (11, 1), 11 = (1, 2)
    assert not rb.info()["has_bodies"]
    dec, lens, _ = codec.decompress_traversals(gpu_ctx, rb, np.array([10, -10, 3], np.int32),
                                               np.array([3], np.uint32), None, 0)
    assert dec.tolist() == [1, 2, 1, -1, -2, -1, 3]
    rb.close()


def _size_class(n):
    c = 0
    while n >= 4:
        n >>= 2
        c += 1
    return c


@pytest.mark.parametrize("delta", [0, 1, 2])
def test_every_expansion_tile_size_agrees(gpu_ctx, delta):
    """Each tile-size candidate of the autotuner decodes bit-exactly (the
    autotuner may pick any of them)."""
    say(f"tile sizes delta={delta}")
    tr = synth.generate(synth.CONFIGS["plain_20k"])
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, delta, threads=8)
    rb = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
    gpu_ctx.set_autotune(False)
    try:
        for items in (8, 16, 32):
            gpu_ctx.set_tuned(f"expand_items/{_size_class(tr.n_steps)}", items)
            dec, _, mism = codec.decompress_traversals(gpu_ctx, rb, ref["encoded"],
                                                       ref["encoded_lengths"], tr.lengths, delta)
            assert not mism
            np.testing.assert_array_equal(dec, tr.nodes)
    finally:
        gpu_ctx.set_autotune(True)
        rb.close()


def _expand_py(sym, first, second, min_id):
    """Recursive definition of rule expansion (decompression_workflow.cpp:70-106)."""
    r = abs(sym) - min_id
    if not (0 <= r < len(first)):
        return [sym]
    if sym > 0:
        return _expand_py(first[r], first, second, min_id) + _expand_py(second[r], first, second, min_id)
    return _expand_py(-second[r], first, second, min_id) + _expand_py(-first[r], first, second, min_id)


def test_deep_chain_rulebook(gpu_ctx):
    """A 300-level chain (rule i = (rule i-1, node)): more levels than one scalar
    readback holds, both orientations, with and without an inverse delta."""
    say("deep chain rulebook")
    import sys as _sys
    _sys.setrecursionlimit(10000)
    min_id, depth = 1000, 300
    first = [5] + [min_id + i - 1 for i in range(1, depth)]
    second = [7] + [(-1) ** i * (i % 13 + 1) for i in range(1, depth)]
    rb = codec.Rulebook(gpu_ctx, first, second, min_id)
    info = rb.info()
    assert info["max_depth"] == depth and info["has_bodies"]
    top = min_id + depth - 1
    enc = np.array([3, top, -top, min_id + 7, -(min_id + 100)], np.int32)
    want = []
    for s in enc.tolist():
        want += _expand_py(s, first, second, min_id)
    for delta in (0, 1):
        dec, lens, _ = codec.decompress_traversals(gpu_ctx, rb, enc, np.array([enc.size], np.uint32),
                                                   None, delta)
        exp = np.array(want, np.int64)
        if delta:
            exp = np.cumsum(exp)
        np.testing.assert_array_equal(dec.astype(np.int64), (exp + 2**31) % 2**32 - 2**31)
    rb.close()


def test_shuffled_rule_ids_decode_by_descent(gpu_ctx):
    """A grammar whose rule IDs are shuffled (depth no longer grows with ID, legal for
    the recursive CPU decoder) over tens of thousands of rules: the rule-size passes,
    which start at the first rule the previous pass left unresolved, still resolve every
    rule across many blocks, the rulebook falls back to descent, and the traversals
    decode to the input and to the oracle's decode of the same rulebook."""
    cfg = synth.CONFIGS["plain_500k"].scaled(backbone=60000, n_haps=24)
    tr = synth.generate(cfg)
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=8)
    R, lo = ref["rules_first"].size, ref["start_id"]
    say(f"shuffled rule IDs: {R} rules")
    assert R > 20000
    perm = np.random.default_rng(5).permutation(R).astype(np.int64)

    def remap(x):
        x = x.astype(np.int64)
        a = np.abs(x)
        rule = (a >= lo) & (a < lo + R)
        out = x.copy()
        out[rule] = np.sign(x[rule]) * (lo + perm[a[rule] - lo])
        return out.astype(np.int32)

    rf = np.empty(R, np.int32)
    rs = np.empty(R, np.int32)
    rf[perm] = remap(ref["rules_first"])
    rs[perm] = remap(ref["rules_second"])
    enc = remap(ref["encoded"])
    rb0 = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], lo)
    depth = rb0.info()["max_depth"]
    rb0.close()
    rb = codec.Rulebook(gpu_ctx, rf, rs, lo)
    info = rb.info()
    assert not info["has_bodies"] and info["max_depth"] == depth >= 4
    dec, _, mism = codec.decompress_traversals(gpu_ctx, rb, enc, ref["encoded_lengths"],
                                               tr.lengths, 1)
    rb.close()
    assert not mism
    np.testing.assert_array_equal(dec, tr.nodes)
    want = oracle.decompress(enc, ref["encoded_lengths"], rf, rs, lo, 1, tr.n_steps, threads=8)
    np.testing.assert_array_equal(dec, want)
