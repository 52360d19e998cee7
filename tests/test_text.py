"""CPU checks of the P/W step-text restatement (oracle/text_oracle.py) against the
lines the reference wrote and parsed (tests/golden/*_lines.txt, edge_text_lines.json)."""
import json
import os

import numpy as np
import pytest

import goldens
from oracle import text_oracle as to

GD = goldens.GOLDEN_DIR


def _fixture(name):
    with open(os.path.join(GD, name), "rb") as f:
        return f.read()


@pytest.mark.parametrize("name,golden", [("example_lines.txt", "example_r8_t2_d1"),
                                         ("cli_fixture_lines.txt", "cli_fixture_r8_t2_d1")])
def test_format_and_parse_reproduce_reference_lines(name, golden):
    text = _fixture(name)
    g = goldens.load(golden)
    nodes, lengths, n_paths = goldens.inputs(g)
    lines = to.split_lines(text)
    assert len(lines) == lengths.size
    offs = np.concatenate([[0], np.cumsum(lengths, dtype=np.int64)])
    for kind in (to.P, to.W):
        idx = [i for i, l in enumerate(lines) if l[0] == kind]
        if not idx:
            continue
        sub_nodes = np.concatenate([nodes[offs[i]:offs[i + 1]] for i in idx])
        got = to.format_steps(kind, sub_nodes, lengths[idx], [lines[i][1] for i in idx],
                              [lines[i][3] for i in idx])
        want = b"".join(lines[i][1] + lines[i][2] + lines[i][3] for i in idx)
        assert got == want
        steps = [lines[i][2] for i in idx]
        loff = np.concatenate([[0], np.cumsum([len(s) for s in steps])])
        pn, pl = to.parse_steps(kind, b"".join(steps), loff)
        np.testing.assert_array_equal(pn, sub_nodes)
        np.testing.assert_array_equal(pl, lengths[idx])


def test_plain_lines_round_trip():
    lines = to.split_lines(_fixture("plain_3k_lines.txt"))
    steps = [l[2] for l in lines]
    loff = np.concatenate([[0], np.cumsum([len(s) for s in steps])])
    nodes, lengths = to.parse_steps(to.P, b"".join(steps), loff)
    assert lengths.size == 6 and nodes.size > 15000
    assert to.format_steps(to.P, nodes, lengths, [l[1] for l in lines],
                           [l[3] for l in lines]) == b"".join(l[1] + l[2] + l[3] for l in lines)


def test_edge_step_lists_match_the_reference_parser():
    fx = json.loads(_fixture("edge_text_lines.json"))
    src = to.split_lines(fx["input"].encode())
    ref = to.split_lines(fx["reference_output"].encode())
    for kind in (to.P, to.W):
        a = [l for l in src if l[0] == kind]
        b = [l for l in ref if l[0] == kind]
        sa = [l[2] for l in a]
        n1, l1 = to.parse_steps(kind, b"".join(sa), np.concatenate([[0], np.cumsum([len(s) for s in sa])]))
        sb = [l[2] for l in b]
        n2, l2 = to.parse_steps(kind, b"".join(sb), np.concatenate([[0], np.cumsum([len(s) for s in sb])]))
        np.testing.assert_array_equal(n1, n2)
        np.testing.assert_array_equal(l1, l2)
        assert to.format_steps(kind, n1, l1, [l[1] for l in b], [l[3] for l in b]) == \
            b"".join(l[1] + l[2] + l[3] for l in b)
    for label, e in fx["errors"].items():
        text = {"no_digits": b"1+,x+", "overflow": b">1>4294967296"}[label]
        kind = to.P if label == "no_digits" else to.W
        with pytest.raises(to.ParseError) as info:
            to.parse_steps(kind, text, [0, len(text)])
        assert e["rc"] != 0 and str(info.value) in e["stderr"][-1]
