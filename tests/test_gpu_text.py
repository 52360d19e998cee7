"""P/W line step text on the GPU (gfz_format_steps / gfz_parse_steps): the lines the
reference wrote (tests/golden/*_lines.txt, edge_text_lines.json) are reproduced byte for
byte and parsed back to the reference's traversals; random cases across tile borders
match the CPU restatement (oracle/text_oracle.py); a full synthetic pangenome's P-lines
equal what the generator writes in the reference writer's format."""
import json
import os
import sys
import tempfile

import numpy as np
import pytest

import goldens
from gfaz_amd import codec, synth
from oracle import text_oracle as to

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_lines(ctx, lines):
    for kind in (to.P, to.W):
        sel = [l for l in lines if l[0] == kind]
        if not sel:
            continue
        steps = [l[2] for l in sel]
        loff = np.concatenate([[0], np.cumsum([len(s) for s in steps])])
        want_nodes, want_len = to.parse_steps(kind, b"".join(steps), loff)
        nodes, lens = codec.parse_steps(ctx, kind, b"".join(steps), loff)
        np.testing.assert_array_equal(nodes, want_nodes)
        np.testing.assert_array_equal(lens, want_len)
        text = codec.format_steps(ctx, kind, nodes, lens, [l[1] for l in sel], [l[3] for l in sel])
        assert text == to.format_steps(kind, nodes, lens, [l[1] for l in sel], [l[3] for l in sel])


@pytest.mark.parametrize("name", ["example_lines.txt", "cli_fixture_lines.txt",
                                  "plain_3k_lines.txt"])
def test_reference_lines_reproduced(gpu_ctx, name):
    say(f"text {name}")
    with open(os.path.join(goldens.GOLDEN_DIR, name), "rb") as f:
        text = f.read()
    lines = to.split_lines(text)
    _check_lines(gpu_ctx, lines)
    # the reference's lines themselves, byte for byte
    for kind in (to.P, to.W):
        sel = [l for l in lines if l[0] == kind]
        if sel:
            steps = [l[2] for l in sel]
            loff = np.concatenate([[0], np.cumsum([len(s) for s in steps])])
            nodes, lens = codec.parse_steps(gpu_ctx, kind, b"".join(steps), loff)
            assert codec.format_steps(gpu_ctx, kind, nodes, lens, [l[1] for l in sel],
                                      [l[3] for l in sel]) == b"".join(l[1] + l[2] + l[3] for l in sel)


def test_edge_step_lists_and_errors(gpu_ctx):
    say("text edge cases")
    with open(os.path.join(goldens.GOLDEN_DIR, "edge_text_lines.json")) as f:
        fx = json.load(f)
    src = to.split_lines(fx["input"].encode())
    ref = to.split_lines(fx["reference_output"].encode())
    for kind in (to.P, to.W):
        a = [l[2] for l in src if l[0] == kind]
        b = [l for l in ref if l[0] == kind]
        nodes, lens = codec.parse_steps(gpu_ctx, kind, b"".join(a),
                                        np.concatenate([[0], np.cumsum([len(s) for s in a])]))
        got = codec.format_steps(gpu_ctx, kind, nodes, lens, [l[1] for l in b], [l[3] for l in b])
        assert got == b"".join(l[1] + l[2] + l[3] for l in b)
    for label, text, kind in (("no_digits", b"1+,x+", to.P), ("overflow", b">1>4294967296", to.W)):
        msg = fx["errors"][label]["stderr"][-1].split("what():")[-1].strip()
        with pytest.raises(RuntimeError, match=msg):
            codec.parse_steps(gpu_ctx, kind, text, [0, len(text)])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_lines_match_the_restatement(gpu_ctx, seed):
    say(f"text random seed={seed}")
    rng = np.random.default_rng(seed)
    # lengths around the 2048-step tile: empty, short, several tiles, many per tile
    lens = rng.choice([0, 1, 2, 7, 300, 2047, 2048, 2049, 6000], size=40).astype(np.uint32)
    n = int(lens.sum())
    mags = rng.choice([1, 9, 10, 99, 12345, 2 ** 31 - 1], size=n) * rng.integers(1, 3, n)
    nodes = (np.minimum(mags, 2 ** 31 - 1) * rng.choice([-1, 1], size=n)).astype(np.int64)
    nodes[rng.integers(0, n, 5)] = -2 ** 31
    nodes = nodes.astype(np.int32)
    heads = [b"P\t" + b"x" * int(k) + b"\t" if k % 3 else b"" for k in rng.integers(0, 9, lens.size)]
    tails = [b"\t*\n" if k % 2 else b"\n" for k in rng.integers(0, 4, lens.size)]
    for kind in (to.P, to.W):
        got = codec.format_steps(gpu_ctx, kind, nodes, lens, heads, tails)
        want = to.format_steps(kind, nodes, lens, heads, tails)
        assert got == want
        # the step fields parse back (P names of 2^31 give INT32_MIN again)
        offs = np.concatenate([[0], np.cumsum(lens, dtype=np.int64)])
        fields = [to.format_steps(kind, nodes[offs[i]:offs[i + 1]], [lens[i]], [b""], [b""])
                  for i in range(lens.size)]
        loff = np.concatenate([[0], np.cumsum([len(f) for f in fields])])
        pn, pl = codec.parse_steps(gpu_ctx, kind, b"".join(fields), loff)
        np.testing.assert_array_equal(pn, nodes)
        np.testing.assert_array_equal(pl, lens)


def test_device_buffers_and_sizes(gpu_ctx):
    tr = synth.generate(synth.CONFIGS["plain_20k"])
    heads = [f"P\thap{h}\t".encode() for h in range(tr.lengths.size)]
    tails = [b"\t*\n"] * tr.lengths.size
    d_nodes = codec.DeviceArray.from_host(gpu_ctx, tr.nodes)
    size = len(to.format_steps(to.P, tr.nodes, tr.lengths, heads, tails))
    out = codec.DeviceArray(gpu_ctx, size + 17, np.uint8)
    assert codec.format_steps(gpu_ctx, to.P, d_nodes, tr.lengths, heads, tails, out=out) == size
    text = out.to_host(size).tobytes()
    assert text == to.format_steps(to.P, tr.nodes, tr.lengths, heads, tails)
    lines = to.split_lines(text)
    steps = b"".join(l[2] for l in lines)
    loff = np.concatenate([[0], np.cumsum([len(l[2]) for l in lines])])
    d_txt = codec.DeviceArray.from_host(gpu_ctx, np.frombuffer(steps, np.uint8).copy())
    d_back = codec.DeviceArray(gpu_ctx, tr.n_steps, np.int32)
    n, lens = codec.parse_steps(gpu_ctx, to.P, d_txt, loff, out=d_back)
    assert n == tr.n_steps
    np.testing.assert_array_equal(lens, tr.lengths)
    np.testing.assert_array_equal(d_back.to_host(), tr.nodes)
    with pytest.raises(RuntimeError, match="capacity"):
        codec.format_steps(gpu_ctx, to.P, d_nodes, tr.lengths, heads, tails,
                           out=codec.DeviceArray(gpu_ctx, size - 1, np.uint8))


def test_full_pangenome_p_lines(gpu_ctx):
    """plain_500k (39.6 M steps): the P-lines the generator writes in the reference
    writer's format (header, S and L lines skipped) equal the GPU's, and parse back."""
    cfg = synth.CONFIGS["plain_500k"]
    tr = synth.generate(cfg)
    say(f"text plain_500k ({tr.n_steps} steps)")
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "g.gfa")
        synth.write_gfa(cfg, tr, path)
        with open(path, "rb") as f:
            text = f.read()
    start = text.index(b"\nP\t") + 1
    want = text[start:]
    heads = [f"P\thap{h}\t".encode() for h in range(tr.lengths.size)]
    tails = [b"\t*\n"] * tr.lengths.size
    got = codec.format_steps(gpu_ctx, to.P, tr.nodes, tr.lengths, heads, tails)
    assert got == want
    lines = to.split_lines(want)
    loff = np.concatenate([[0], np.cumsum([len(l[2]) for l in lines])])
    nodes, lens = codec.parse_steps(gpu_ctx, to.P, b"".join(l[2] for l in lines), loff)
    np.testing.assert_array_equal(lens, tr.lengths)
    assert np.array_equal(nodes, tr.nodes)


def test_parse_fields_inside_whole_lines(gpu_ctx):
    """Step fields parsed in place inside whole P-lines (heads and tails skipped), with
    the line starts that formatting reports."""
    tr = synth.generate(synth.CONFIGS["plain_20k"].scaled(backbone=4000, n_haps=5))
    lens = np.concatenate([tr.lengths, [0]]).astype(np.uint32)  # and an empty line
    heads = [f"P\thap{h}\t".encode() for h in range(lens.size)]
    tails = [b"\t3M,2M\n" if h % 2 else b"\t*\n" for h in range(lens.size)]
    text, line_off = codec.format_steps(gpu_ctx, to.P, tr.nodes, lens, heads, tails,
                                        with_offsets=True)
    assert text == to.format_steps(to.P, tr.nodes, lens, heads, tails)
    hl = np.array([len(h) for h in heads], np.uint64)
    tl = np.array([len(t) for t in tails], np.uint64)
    begin, end = line_off[:-1] + hl, line_off[1:] - tl
    nodes, got_lens = codec.parse_steps(gpu_ctx, to.P, text, fields=(begin, end))
    np.testing.assert_array_equal(got_lens, lens)
    np.testing.assert_array_equal(nodes, tr.nodes)


def test_edge_tokens_inside_long_fields(gpu_ctx):
    """Edge tokens (10 and 11 digits, leading zeros, trailing letters, empty names,
    zero, the largest name) placed among thousands of ordinary ones, so that tokens
    of every length from 1 to 14 bytes sit at every offset of the parse tiles' LDS
    windows: the GPU parse equals the CPU restatement, and malformed names deep in a
    field raise the reference's errors."""
    say("text edge tokens in long fields")
    rng = np.random.default_rng(11)
    p_toks = [b"1+", b"9-", b"4294967295+", b"4294967295-", b"0+", b"+", b"12ab+", b"007-",
              b"2147483648-", b"1234567890+", b"01234567890-", b"001234567890+",
              b"0001234567890-", b"55x9+", b"3"]
    w_toks = [b">1", b"<9", b">4294967295", b"<4294967295", b"<0", b">", b"<12ab", b">007",
              b"<2147483648", b">1234567890", b"<01234567890", b">001234567890",
              b"<0001234567890", b">55x9"]
    for kind, toks, sep in ((to.P, p_toks, b","), (to.W, w_toks, b"")):
        fields = []
        for n in (3000, 1, 0, 777, 5000):
            pick = rng.integers(0, len(toks), n)
            fields.append(sep.join(toks[i] for i in pick))
        loff = np.concatenate([[0], np.cumsum([len(f) for f in fields])])
        want_nodes, want_len = to.parse_steps(kind, b"".join(fields), loff)
        nodes, lens = codec.parse_steps(gpu_ctx, kind, b"".join(fields), loff)
        np.testing.assert_array_equal(lens, want_len)
        np.testing.assert_array_equal(nodes, want_nodes)
    for kind, bad, msg in ((to.P, b"x+", "no digits"), (to.P, b"4294967296+", "overflow"),
                           (to.P, b"99999999999-", "overflow"), (to.W, b">x", "no digits"),
                           (to.W, b"<4294967296", "overflow"), (to.W, b">99999999999", "overflow")):
        ok = b"12+," * 700 if kind == to.P else b">12" * 700
        text = ok + bad + (b"," if kind == to.P else b"") + ok
        with pytest.raises(to.ParseError, match=msg):
            to.parse_steps(kind, text, [0, len(text)])
        with pytest.raises(RuntimeError, match=msg):
            codec.parse_steps(gpu_ctx, kind, text, [0, len(text)])
