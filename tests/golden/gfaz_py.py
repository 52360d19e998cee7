"""Test-side reader for the `.gfaz` v5 container (src/codec/serialization.cpp:137-346).

Pure Python over ``struct`` plus the system libzstd through ctypes. It reads the traversal
fields that the grammar codec owns and keeps every other block as opaque bytes, so tests can
compare payloads written by the reference and by this project field by field.
"""
from __future__ import annotations

import ctypes
import struct
from dataclasses import dataclass, field

import numpy as np

_zstd = ctypes.CDLL("libzstd.so.1")
_zstd.ZSTD_decompress.restype = ctypes.c_size_t
_zstd.ZSTD_decompress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
_zstd.ZSTD_isError.restype = ctypes.c_uint
_zstd.ZSTD_isError.argtypes = [ctypes.c_size_t]

MAGIC = 0x5A414647
VERSION = 5


def zstd_decompress(payload: bytes, original_size: int) -> bytes:
    if original_size == 0:
        return b""
    out = ctypes.create_string_buffer(original_size)
    src = ctypes.create_string_buffer(payload, len(payload))
    n = _zstd.ZSTD_decompress(out, original_size, src, len(payload))
    if _zstd.ZSTD_isError(n) or n != original_size:
        raise ValueError("zstd payload does not decode to its original size")
    return out.raw


@dataclass
class Block:
    original_size: int
    payload: bytes

    def decode(self) -> bytes:
        return zstd_decompress(self.payload, self.original_size)

    def int32(self) -> np.ndarray:
        return np.frombuffer(self.decode(), dtype=np.int32).copy()

    def uint32(self) -> np.ndarray:
        return np.frombuffer(self.decode(), dtype=np.uint32).copy()


@dataclass
class Gfaz:
    header: str
    ranges: list  # (k, start, end, flattened_offset, element_count)
    sequence_lengths: np.ndarray
    original_path_lengths: np.ndarray
    rules_first: Block
    rules_second: Block
    paths: Block
    delta_round: int
    walk_lengths: np.ndarray
    original_walk_lengths: np.ndarray
    walks: Block
    blocks: dict = field(default_factory=dict)  # every other block, by name

    def min_rule_id(self) -> int:
        return self.ranges[0][1] if self.ranges else 0

    def rules(self):
        """Rule columns after the whole-column prefix sum (codec.cpp:88-95)."""
        f = np.cumsum(self.rules_first.int32(), dtype=np.int64).astype(np.int32)
        s = np.cumsum(self.rules_second.int32(), dtype=np.int64).astype(np.int32)
        return f, s


class _Reader:
    def __init__(self, data: bytes):
        self.d = data
        self.o = 0

    def val(self, fmt: str):
        v = struct.unpack_from("<" + fmt, self.d, self.o)[0]
        self.o += struct.calcsize("<" + fmt)
        return v

    def raw(self, n: int) -> bytes:
        v = self.d[self.o:self.o + n]
        if len(v) != n:
            raise ValueError("truncated .gfaz")
        self.o += n
        return v

    def string(self) -> str:
        return self.raw(self.val("Q")).decode()

    def u32vec(self) -> np.ndarray:
        n = self.val("Q")
        return np.frombuffer(self.raw(4 * n), dtype=np.uint32).copy()

    def block(self) -> Block:
        orig = self.val("Q")
        n = self.val("Q")
        return Block(orig, self.raw(n))


def read_gfaz(path: str) -> Gfaz:
    with open(path, "rb") as f:
        r = _Reader(f.read())
    if r.val("I") != MAGIC:
        raise ValueError("invalid file format (magic mismatch)")
    version = r.val("I")
    if version != VERSION:
        raise ValueError(f"unsupported file version {version}")
    header = r.string()
    ranges = [(r.val("i"), r.val("I"), r.val("I"), r.val("Q"), r.val("Q")) for _ in range(r.val("Q"))]
    seq_len = r.u32vec()
    orig_path = r.u32vec()
    rf, rs, paths = r.block(), r.block(), r.block()
    delta = max(0, r.val("i"))
    blocks = {}
    for name in ("names", "name_lengths", "overlaps", "overlap_lengths", "segment_sequences",
                 "segment_seq_lengths"):
        blocks[name] = r.block()

    def opt_cols(prefix):
        cols = []
        for _ in range(r.val("Q")):
            tag = r.string()
            typ = r.raw(1)
            n = r.val("Q")
            cols.append((tag, typ, n, [r.block() for _ in range(8)]))
        blocks[prefix] = cols

    opt_cols("segment_optional")
    for name in ("link_from_ids", "link_to_ids", "link_from_orients", "link_to_orients",
                 "link_overlap_nums", "link_overlap_ops"):
        blocks[name] = r.block()
    blocks["num_links"] = r.val("Q")
    opt_cols("link_optional")
    blocks["num_jumps"] = r.val("Q")
    blocks["jumps"] = [r.block() for _ in range(8)]
    blocks["num_containments"] = r.val("Q")
    blocks["containments"] = [r.block() for _ in range(9)]
    wl = r.u32vec()
    owl = r.u32vec()
    walks = r.block()
    for name in ("walk_sample_ids", "walk_sample_id_lengths", "walk_hap_indices", "walk_seq_ids",
                 "walk_seq_id_lengths", "walk_seq_starts", "walk_seq_ends"):
        blocks[name] = r.block()
    if r.o != len(r.d):
        raise ValueError("trailing bytes after .gfaz fields")
    return Gfaz(header, ranges, seq_len, orig_path, rf, rs, paths, delta, wl, owl, walks, blocks)


def parse_traversals(gfa_path: str):
    """P- and W-line traversals of a GFA whose segment names are 1..n in file order (the
    reference parser's numeric fast path, src/io/gfa_parser.cpp:605-615,1210-1219)."""
    paths, walks = [], []
    with open(gfa_path) as f:
        for line in f:
            if line.startswith("P\t"):
                steps = line.rstrip("\n").split("\t")[2]
                seq = [int(s[:-1]) if s[-1] == "+" else -int(s[:-1]) for s in steps.split(",") if s]
                paths.append(seq)
            elif line.startswith("W\t"):
                body = line.rstrip("\n").split("\t")[6]
                seq, i = [], 0
                while i < len(body):
                    sign = 1 if body[i] == ">" else -1
                    j = i + 1
                    while j < len(body) and body[j] not in "<>":
                        j += 1
                    seq.append(sign * int(body[i + 1:j]))
                    i = j
                walks.append(seq)
    seqs = paths + walks
    lengths = np.array([len(s) for s in seqs], dtype=np.uint32)
    nodes = np.array([v for s in seqs for v in s], dtype=np.int32)
    return nodes, lengths, len(paths)
