"""TEST INFRASTRUCTURE ONLY - CPU restatement of the reference's P/W line step text.

Only tests/ and bench.py's CPU-baseline leg import this module; the product formats and
parses step text on the GPU (gfaz_amd/csrc/text.hip).

* format_steps follows format_path_line_numeric / format_walk_line_numeric
  (src/io/gfa_write_utils.cpp:238-290, append_numeric_node_name :167-169): a P step is
  "," (not before the first), the decimal |id| (as uint32) and '+' or '-'; a W step is
  '>' or '<' and the decimal |id|.
* parse_steps follows parse_p_line / parse_w_line (src/io/gfa_parser.cpp:676-790) with
  resolve_node_id's numeric-name path (:1210-1219) and parse_uint32 (:69-91).

Pinned by the reference's own output (tests/golden/*_lines.txt, edge_text_lines.json,
written by tests/golden/make_goldens.py through the compiled reference).
"""
from __future__ import annotations

import numpy as np

P, W = 0, 1


def _u32(v: int) -> int:
    return v & 0xFFFFFFFF if v >= 0 else (-v) & 0xFFFFFFFF


def format_steps(kind: int, nodes, lengths, heads: list[bytes], tails: list[bytes]) -> bytes:
    out = []
    o = 0
    nodes = [int(v) for v in np.asarray(nodes, dtype=np.int32)]
    for t, n in enumerate(np.asarray(lengths, dtype=np.uint64)):
        n = int(n)
        seq = nodes[o:o + n]
        o += n
        if kind == P:
            steps = ",".join(f"{_u32(v)}{'-' if v < 0 else '+'}" for v in seq)
        else:
            steps = "".join(f"{'<' if v < 0 else '>'}{_u32(v)}" for v in seq)
        out.append(heads[t] + steps.encode() + tails[t])
    return b"".join(out)


class ParseError(ValueError):
    pass


def _parse_uint32(name: bytes) -> int:
    val, any_digit = 0, False
    for c in name:
        if not 48 <= c <= 57:
            break
        any_digit = True
        val = val * 10 + (c - 48)
        if val > 0xFFFFFFFF:
            raise ParseError("parse_uint32: overflow")
    if not any_digit:
        raise ParseError("parse_uint32: no digits")
    return val


def _node(val: int, reverse: bool) -> int:
    """uint32 id -> int32 NodeId, negated for reverse (the reference's conversions)."""
    v = (-val if reverse else val) & 0xFFFFFFFF
    return v - (1 << 32) if v >= 1 << 31 else v


def parse_steps(kind: int, text: bytes, line_off) -> tuple[np.ndarray, np.ndarray]:
    nodes, lengths = [], []
    for l in range(len(line_off) - 1):
        s = text[int(line_off[l]):int(line_off[l + 1])]
        seq = []
        if kind == P:
            for tok in s.split(b","):
                if not tok:
                    continue
                name, orient = tok[:-1], tok[-1:]
                if not name:
                    continue
                val = _parse_uint32(name)
                if val:
                    seq.append(_node(val, orient == b"-"))
        else:
            i = 0
            while i < len(s):
                if s[i] not in b"<>":
                    i += 1
                    continue
                j = i + 1
                while j < len(s) and s[j] not in b"<>":
                    j += 1
                name = s[i + 1:j]
                if name:
                    val = _parse_uint32(name)
                    if val:
                        seq.append(_node(val, s[i] == 60))  # '<'
                i = j
        nodes.extend(seq)
        lengths.append(len(seq))
    return np.array(nodes, dtype=np.int32), np.array(lengths, dtype=np.uint32)


def split_lines(text: bytes):
    """P/W lines of GFA text -> (kind, heads, step fields, tails) per line, in order: the
    head is every byte before the step field, the tail every byte after it."""
    out = []
    for line in text.splitlines(keepends=True):
        if line[:2] == b"P\t":
            f = line.split(b"\t")
            head = b"\t".join(f[:2]) + b"\t"
            steps = f[2]
            tail = line[len(head) + len(steps):]
            out.append((P, head, steps, tail))
        elif line[:2] == b"W\t":
            f = line.split(b"\t")
            head = b"\t".join(f[:6]) + b"\t"
            steps = f[6].rstrip(b"\n")
            tail = line[len(head) + len(steps):]
            out.append((W, head, steps, tail))
    return out
