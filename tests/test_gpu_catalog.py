"""Extraction front-end on the GPU: gfz_extract_path_lines / gfz_extract_walk_lines /
gfz_extract_walk_lines_by_name over a resident store and a catalog.

Pinned by the reference's own extract output for the CLI fixture
(tests/golden/cli_fixture_kat.json: extract_path_lines_by_name and extract_walk_line
through the compiled reference) and, at larger sizes, by the text oracle's
format_steps over the input traversals (itself pinned by the reference's lines)."""
import json
import os

import numpy as np
import pytest

import goldens
import gfaz_py
from gfaz_amd import codec, synth
from oracle import oracle, text_oracle

pytestmark = pytest.mark.gpu


def _strings(blob: bytes, lengths) -> list[bytes]:
    out, o = [], 0
    for n in np.asarray(lengths, dtype=np.int64):
        out.append(blob[o:o + int(n)])
        o += int(n)
    return out


def _varint_i64(data: bytes, n: int) -> list[int]:
    """decompress_varint_int64 (src/codec/codec.cpp:460-526): LEB128, then zigzag."""
    out, i = [], 0
    while i < len(data) and len(out) < n:
        v, shift = 0, 0
        while i < len(data):
            b = data[i]
            i += 1
            v |= (b & 0x7F) << shift
            if not b & 0x80:
                break
            shift += 7
        out.append((v >> 1) ^ -(v & 1))
    return out


def _split(nodes, lengths):
    offs = np.concatenate([[0], np.cumsum(lengths, dtype=np.int64)])
    return [nodes[offs[i]:offs[i + 1]] for i in range(len(lengths))]


def test_cli_fixture_lines_equal_reference_extract(gpu_ctx):
    g = gfaz_py.read_gfaz(os.path.join(goldens.GOLDEN_DIR, "cli_fixture_r8_t2_d1.gfaz"))
    with open(os.path.join(goldens.GOLDEN_DIR, "cli_fixture_kat.json")) as f:
        kat = json.load(f)
    b = g.blocks
    names = _strings(b["names"].decode(), b["name_lengths"].uint32())
    overlaps = _strings(b["overlaps"].decode(), b["overlap_lengths"].uint32())
    nw = int(g.walk_lengths.size)
    samples = _strings(b["walk_sample_ids"].decode(), b["walk_sample_id_lengths"].uint32())
    haps = b["walk_hap_indices"].uint32()
    seqs = _strings(b["walk_seq_ids"].decode(), b["walk_seq_id_lengths"].uint32())
    starts = _varint_i64(b["walk_seq_starts"].decode(), nw)
    ends = _varint_i64(b["walk_seq_ends"].decode(), nw)
    rf, rs = g.rules()
    enc = np.concatenate([g.paths.int32(), g.walks.int32()])
    enc_len = np.concatenate([g.sequence_lengths, g.walk_lengths]).astype(np.uint32)
    rb = codec.Rulebook(gpu_ctx, rf, rs, g.min_rule_id())
    res = codec.Resident(gpu_ctx, rb, enc, enc_len, g.delta_round)
    cat = codec.Catalog()
    cat.add_paths(names, 0, overlaps)
    cat.add_walks(list(zip(samples, haps, seqs, starts, ends)), len(names))
    try:
        got = codec.extract_path_lines(gpu_ctx, res, cat, ["pathB", "pathA"])
        assert b"".join(got).decode() == kat["extract_path"]
        got = codec.extract_walk_lines(gpu_ctx, res, cat, [("sampleB", 1, "chr2", -1, -1)])
        assert b"".join(got).decode() == kat["extract_walk"]
        by_name = codec.extract_walk_lines_by_name(gpu_ctx, res, cat, ["sampleB"])
        assert b"".join(by_name).decode() == kat["extract_walk"]
        with pytest.raises(RuntimeError, match="path not found: pathZ"):
            codec.extract_path_lines(gpu_ctx, res, cat, ["pathZ"])
        assert codec.extract_path_lines(gpu_ctx, res, cat, []) == []
    finally:
        res.close()
        rb.close()


@pytest.mark.parametrize("delta", [0, 1, 2])
def test_lines_and_appends_equal_text_oracle(gpu_ctx, delta):
    """11 synthetic haplotypes stored as 7 named paths and 4 walks; then 3 new paths and
    2 new walks encoded with the stored rules and appended (no re-framing): every
    line, old and new, in any order, equals the text oracle over the input traversals."""
    print(f"catalog lines delta={delta}", flush=True)
    tr = synth.generate(synth.CONFIGS["plain_20k"])
    trav = _split(tr.nodes, tr.lengths)
    n_base = 11
    base_len = tr.lengths[:n_base]
    ref = oracle.compress(np.concatenate(trav[:n_base]), base_len, 8, 2, delta, threads=8)
    rb = codec.Rulebook(gpu_ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
    res = codec.Resident(gpu_ctx, rb, ref["encoded"], ref["encoded_lengths"], delta)
    cat = codec.Catalog()
    paths = {f"hap{i}#0#chr1": i for i in range(7)}
    walks = {("sample%d" % i, i % 3, "chr%d" % (i % 2), 10 * i if i % 2 else -1,
              10 * i + 5 if i % 2 else -1): 7 + k for k, i in enumerate(range(4))}
    cat.add_paths(list(paths), 0, ["*" if i % 2 else f"{i}M" for i in range(7)])
    cat.add_walks(list(walks), 7)
    # append haplotypes 11.. (3 paths, 2 walks) encoded with the stored rules
    new = trav[n_base:n_base + 5]
    enc, enc_len = codec.encode_with_rules(
        gpu_ctx, np.concatenate(new), np.array([t.size for t in new], np.uint32),
        ref["rules_first"], ref["rules_second"],
        [(ref["start_id"], ref["end_id"], 0)], delta)
    first = res.append(enc, enc_len)
    assert first == n_base
    new_paths = {f"new{i}": first + i for i in range(3)}
    new_walks = {("newsample%d" % i, 0, "chrX", -1, -1): first + 3 + i for i in range(2)}
    cat.add_paths(list(new_paths), first)
    cat.add_walks(list(new_walks), first + 3)
    overlap = {f"hap{i}#0#chr1": ("*" if i % 2 else f"{i}M") for i in range(7)}
    try:
        rng = np.random.default_rng(delta)
        all_paths = {**paths, **new_paths}
        names = list(all_paths)
        req = [names[i] for i in rng.integers(0, len(names), 12)]
        got = codec.extract_path_lines(gpu_ctx, res, cat, req)
        for line, name in zip(got, req):
            t = all_paths[name]
            want = text_oracle.format_steps(text_oracle.P, trav[t], [trav[t].size],
                                            [f"P\t{name}\t".encode()],
                                            [f"\t{overlap.get(name, '*')}\n".encode()])
            assert line == want, name
        all_walks = {**walks, **new_walks}
        keys = list(all_walks)
        req = [keys[i] for i in rng.integers(0, len(keys), 9)]
        got = codec.extract_walk_lines(gpu_ctx, res, cat, req)
        by_name = codec.extract_walk_lines_by_name(gpu_ctx, res, cat, [k[0] for k in req])
        assert got == by_name
        for line, k in zip(got, req):
            t = all_walks[k]
            head = "W\t%s\t%d\t%s\t%s\t%s\t" % (k[0], k[1], k[2], k[3] if k[3] >= 0 else "*",
                                                 k[4] if k[4] >= 0 else "*")
            want = text_oracle.format_steps(text_oracle.W, trav[t], [trav[t].size],
                                            [head.encode()], [b"\n"])
            assert line == want, k
    finally:
        res.close()
        rb.close()
