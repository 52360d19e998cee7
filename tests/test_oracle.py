"""CPU tests: the oracle against the reference's golden outputs, a brute-force
restatement, and invariants of the grammar. No GPU needed."""
import itertools
import os
import random
import subprocess

import numpy as np
import pytest

import goldens
from oracle import oracle


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


@pytest.mark.parametrize("name", goldens.names())
def test_oracle_matches_reference_golden(name):
    g = goldens.load(name)
    nodes, lengths, _ = goldens.inputs(g)
    m = g["meta"]
    exp = goldens.expected(g)
    r = oracle.compress(nodes, lengths, m["rounds"], m["threshold"], m["delta_round"], threads=4)
    assert r["start_id"] == exp["start_id"] and r["end_id"] == exp["end_id"]
    np.testing.assert_array_equal(r["encoded_lengths"], exp["encoded_lengths"])
    np.testing.assert_array_equal(r["encoded"], exp["encoded"])
    np.testing.assert_array_equal(r["rules_first"], exp["rules_first"])
    np.testing.assert_array_equal(r["rules_second"], exp["rules_second"])
    np.testing.assert_array_equal(exp["original_lengths"], lengths)


@pytest.mark.parametrize("name", goldens.names())
def test_oracle_decodes_reference_golden(name):
    g = goldens.load(name)
    nodes, lengths, _ = goldens.inputs(g)
    exp = goldens.expected(g)
    dec = oracle.decompress(exp["encoded"], exp["encoded_lengths"], exp["rules_first"],
                            exp["rules_second"], exp["start_id"], exp["delta_round"], nodes.size)
    np.testing.assert_array_equal(dec, nodes)


# ---- brute force: Appendix-style restatement with the scan formulation ------------

def _canon(a, b):
    k = (a << 32) | (b & 0xFFFFFFFF)
    r = ((-b) << 32) | ((-a) & 0xFFFFFFFF)
    # signed 64-bit compare
    def s64(v):
        v &= (1 << 64) - 1
        return v - (1 << 64) if v >> 63 else v
    return min(s64(k), s64(r)), s64(k)


def brute_compress(seqs, rounds, delta_round):
    seqs = [list(s) for s in seqs]
    mx = 0
    if delta_round == 0:
        mx = max([abs(v) for s in seqs for v in s], default=0)
    for _ in range(delta_round):
        for s in seqs:
            if s:
                mx = max(mx, abs(s[0]))
            for i in range(len(s) - 1, 0, -1):
                s[i] = s[i] - s[i - 1]
                mx = max(mx, abs(s[i]))
    nxt = mx + 1
    start = nxt
    rules = []
    for _ in range(rounds):
        count = {}
        for s in seqs:
            for i in range(len(s) - 1):
                c, _k = _canon(s[i], s[i + 1])
                count[c] = count.get(c, 0) + 1
        K = {c for c, n in count.items() if n >= 2}
        if not K:
            break
        picks_all = []
        used = set()
        for s in seqs:
            pick = []
            prev = False
            for i in range(len(s)):
                cand = i + 1 < len(s) and _canon(s[i], s[i + 1])[0] in K
                p = cand and not prev
                pick.append(p)
                prev = p
                if p:
                    used.add(_canon(s[i], s[i + 1])[0])
            picks_all.append(pick)
        U = sorted(used)
        fid = {k: nxt + r for r, k in enumerate(U)}
        new = []
        for s, pick in zip(seqs, picks_all):
            out, i = [], 0
            while i < len(s):
                if pick[i]:
                    c, k = _canon(s[i], s[i + 1])
                    out.append(fid[c] if k == c else -fid[c])
                    i += 2
                else:
                    out.append(s[i])
                    i += 1
            new.append(out)
        seqs = new
        for k in U:
            rules.append(((k >> 32), ((k & 0xFFFFFFFF) ^ 0x80000000) - 0x80000000))
        nxt += len(U)
    return seqs, rules, start, nxt


def brute_expand(sym, rules, start):
    R = len(rules)
    if start <= abs(sym) < start + R:
        a, b = rules[abs(sym) - start]
        if sym > 0:
            return brute_expand(a, rules, start) + brute_expand(b, rules, start)
        return brute_expand(-b, rules, start) + brute_expand(-a, rules, start)
    return [sym]


@pytest.mark.parametrize("seed", range(40))
def test_oracle_matches_bruteforce_scan_formulation(seed):
    rng = random.Random(seed)
    seqs = []
    for _ in range(rng.randint(1, 5)):
        n = rng.randint(0, 14)
        alphabet = rng.randint(1, 4)
        seqs.append([rng.choice([1, -1]) * rng.randint(1, alphabet) for _ in range(n)])
    rounds = rng.choice([1, 2, 3, 8])
    delta = rng.choice([0, 1, 2])
    bs, brules, bstart, bend = brute_compress(seqs, rounds, delta)
    nodes = np.array([v for s in seqs for v in s], dtype=np.int32)
    lengths = np.array([len(s) for s in seqs], dtype=np.uint32)
    r = oracle.compress(nodes, lengths, rounds, 2, delta, threads=2)
    assert (r["start_id"], r["end_id"]) == (bstart, bend)
    assert r["encoded"].tolist() == [v for s in bs for v in s]
    assert r["encoded_lengths"].tolist() == [len(s) for s in bs]
    assert list(zip(r["rules_first"].tolist(), r["rules_second"].tolist())) == brules
    # expansion by the recursive definition, then inverse delta, gives the input back
    dec = []
    for s in bs:
        e = [x for v in s for x in brute_expand(v, brules, bstart)]
        for _ in range(delta):
            for i in range(1, len(e)):
                e[i] += e[i - 1]
        dec.extend(e)
    assert dec == nodes.tolist()


@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_bruteforce_large_symbols(seed):
    """Node symbols of magnitude 2^30 .. 2^31 - 1 (rule IDs start above them): the
    oracle's packed keys, signed order and dense IDs agree with the scan formulation.
    Deltas are kept inside int32 so that the exact arithmetic here needs no wrap."""
    rng = random.Random(1000 + seed)
    top = (1 << 31) - 5000
    big = [top - rng.randint(0, 3000) for _ in range(3)] + [(1 << 30) + rng.randint(0, 9)]
    seqs = []
    delta = seed % 2
    for _ in range(rng.randint(1, 4)):
        n = rng.randint(2, 16)
        if delta:
            seqs.append([rng.choice(big[-1:] + [1, 2]) for _ in range(n)])
        else:
            seqs.append([rng.choice([1, -1]) * rng.choice(big + [1, 2]) for _ in range(n)])
    bs, brules, bstart, bend = brute_compress(seqs, 8, delta)
    nodes = np.array([v for s in seqs for v in s], dtype=np.int32)
    lengths = np.array([len(s) for s in seqs], dtype=np.uint32)
    r = oracle.compress(nodes, lengths, 8, 2, delta, threads=2)
    assert (r["start_id"], r["end_id"]) == (bstart, bend)
    assert r["encoded"].tolist() == [v for s in bs for v in s]
    assert list(zip(r["rules_first"].tolist(), r["rules_second"].tolist())) == brules


def test_bruteforce_exhaustive_tiny():
    """All length-<=6 single traversals over {+-1, +-2}: oracle == scan formulation."""
    syms = [1, -1, 2, -2]
    for n in range(0, 7):
        for seq in itertools.product(syms, repeat=n):
            bs, brules, bstart, bend = brute_compress([list(seq)], 8, 0)
            r = oracle.compress(np.array(seq, dtype=np.int32), np.array([n], dtype=np.uint32), 8,
                                2, 0, threads=1)
            assert r["encoded"].tolist() == bs[0], seq
            assert list(zip(r["rules_first"].tolist(), r["rules_second"].tolist())) == brules


def test_grammar_invariants_plain_20k():
    g = goldens.load("plain_20k_r8_t2_d1")
    nodes, lengths, _ = goldens.inputs(g)
    r = oracle.compress(nodes, lengths, 8, 2, 1, threads=4)
    start, end = r["start_id"], r["end_id"]
    f = r["rules_first"].astype(np.int64)
    s = r["rules_second"].astype(np.int64)
    key = (f << 32) | (s & 0xFFFFFFFF)
    rev = ((-s) << 32) | ((-f) & 0xFFFFFFFF)
    assert np.all(key <= rev), "every rule key is canonical (signed min of key and reverse)"
    # rules sorted ascending by signed key within each round, depth == creating round
    depth = np.zeros(end - start, dtype=np.int64)
    o = 0
    for rnd, cnt in enumerate(r["rules_per_round"], start=1):
        seg = key[o:o + cnt]
        assert np.all(np.diff(seg) > 0), f"round {rnd} rules sorted by signed key"
        for i in range(o, o + cnt):
            ch = [abs(int(v)) for v in (f[i], s[i])]
            d = [depth[c - start] if start <= c < end else 0 for c in ch]
            assert all(c < start + i or not (start <= c < end) for c in ch)
            depth[i] = 1 + max(d)
            assert depth[i] == rnd, "a rule's depth equals the round that created it"
        o += cnt
    enc = np.abs(r["encoded"].astype(np.int64))
    assert np.all((enc < start) | ((enc >= start) & (enc < end)))


def test_encode_existing_matches_reference_add_haplotypes():
    g = goldens.load("plain_20k_add6")
    base = goldens.load("plain_20k_r8_t2_d1")
    exp = goldens.expected(base)
    enc, lens = oracle.encode_existing(
        g["appended_nodes"], g["appended_lengths"], exp["rules_first"], exp["rules_second"],
        [(exp["start_id"], exp["end_id"], 0)], 1)
    np.testing.assert_array_equal(lens, g["appended_encoded_lengths"])
    np.testing.assert_array_equal(enc, g["appended_encoded"])
    assert bool(g["prefix_unchanged"])


def test_bits_cpp():
    """The bit-parallel greedy pick (csrc/bits.hpp) against the sequential recurrence."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = os.path.join(root, "tests", "cpp", "test_bits.cpp")
    exe = os.path.join("/tmp", f"gfaz_test_bits_{os.getpid()}")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(root, "gfaz_amd", "csrc"), src,
                    "-o", exe], check=True)
    try:
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
        assert "bad=0" in out
    finally:
        os.unlink(exe)


def test_invariant_checker_catches_planted_mistakes():
    """tests/invariants.check_grammar (used on full-size GPU outputs) accepts the
    oracle's grammar and rejects plausible mistakes planted in it."""
    from invariants import check_grammar
    g = goldens.load("plain_20k_r8_t2_d1")
    nodes, lengths, _ = goldens.inputs(g)
    r = oracle.compress(nodes, lengths, 8, 2, 1, threads=4)
    args = [r["encoded"], r["rules_first"], r["rules_second"], r["start_id"], r["end_id"],
            list(r["rules_per_round"])]
    check_grammar(*args)
    swapped = list(args)
    swapped[1], swapped[2] = args[2], args[1]            # transposed rule columns
    shifted = list(args)
    rp = list(args[5])
    rp[0], rp[1] = rp[0] - 1, rp[1] + 1                  # a rule counted in the wrong round
    shifted[5] = rp
    negated = list(args)
    negated[1] = -args[1]                                # a dropped sign
    beyond = list(args)
    beyond[0] = args[0].copy()
    beyond[0][7] = r["end_id"]                           # an ID past the rule window
    for bad in (swapped, shifted, negated, beyond):
        with pytest.raises(AssertionError):
            check_grammar(*bad)
