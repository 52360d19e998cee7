"""TEST INFRASTRUCTURE ONLY - ctypes binding of the CPU oracle (gfaz_oracle.cpp).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this
module; it is the checker and the timed CPU baseline, never the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_oracle", "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "gfaz_ref")

_lib = None


class _Result(ctypes.Structure):
    _fields_ = [
        ("encoded", ctypes.POINTER(ctypes.c_int32)),
        ("n_encoded", ctypes.c_uint64),
        ("encoded_lengths", ctypes.POINTER(ctypes.c_uint32)),
        ("n_trav", ctypes.c_uint64),
        ("rules_first", ctypes.POINTER(ctypes.c_int32)),
        ("rules_second", ctypes.POINTER(ctypes.c_int32)),
        ("n_rules", ctypes.c_uint64),
        ("start_id", ctypes.c_uint32),
        ("end_id", ctypes.c_uint32),
        ("rounds_run", ctypes.c_int32),
        ("rules_per_round", ctypes.c_uint64 * 64),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def build_reference() -> bool:
    """Compile the reference CPU path from /root/reference (when present)."""
    if not os.path.isdir("/root/reference/src"):
        return False
    subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.oracle_compress.argtypes = [vp, ctypes.c_uint64, vp, ctypes.c_uint64, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(_Result)]
        L.oracle_compress.restype = ctypes.c_int
        L.oracle_free_result.argtypes = [ctypes.POINTER(_Result)]
        L.oracle_decompress.argtypes = [vp, ctypes.c_uint64, vp, ctypes.c_uint64, vp, vp,
                                        ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                                        ctypes.c_int, vp, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_decompress.restype = ctypes.c_int
        L.oracle_encode_existing.argtypes = [vp, ctypes.c_uint64, vp, ctypes.c_uint64, vp, vp,
                                             ctypes.c_uint64, vp, vp, vp, ctypes.c_uint64,
                                             ctypes.c_int, vp, vp,
                                             ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_encode_existing.restype = ctypes.c_int
        _lib = L
    return _lib


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def compress(nodes, lengths, rounds=8, threshold=2, delta_round=1, threads=0) -> dict:
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    res = _Result()
    rc = lib().oracle_compress(nodes.ctypes.data, nodes.size, lengths.ctypes.data, lengths.size,
                               rounds, threshold, delta_round, threads, ctypes.byref(res))
    if rc != 0:
        raise RuntimeError(f"oracle_compress failed ({rc})")
    try:
        out = {
            "encoded": _arr(res.encoded, res.n_encoded, np.int32),
            "encoded_lengths": _arr(res.encoded_lengths, res.n_trav, np.uint32),
            "rules_first": _arr(res.rules_first, res.n_rules, np.int32),
            "rules_second": _arr(res.rules_second, res.n_rules, np.int32),
            "start_id": int(res.start_id),
            "end_id": int(res.end_id),
            "rounds_run": int(res.rounds_run),
            "rules_per_round": [int(res.rules_per_round[i]) for i in range(min(64, res.rounds_run))],
        }
    finally:
        lib().oracle_free_result(ctypes.byref(res))
    return out


def decompress(encoded, encoded_lengths, rules_first, rules_second, min_rule_id, delta_round,
               total, threads=0) -> np.ndarray:
    encoded = np.ascontiguousarray(encoded, dtype=np.int32)
    encoded_lengths = np.ascontiguousarray(encoded_lengths, dtype=np.uint32)
    rf = np.ascontiguousarray(rules_first, dtype=np.int32)
    rs = np.ascontiguousarray(rules_second, dtype=np.int32)
    out = np.empty(max(1, total), dtype=np.int32)
    n = ctypes.c_uint64()
    rc = lib().oracle_decompress(encoded.ctypes.data, encoded.size, encoded_lengths.ctypes.data,
                                 encoded_lengths.size, rf.ctypes.data, rs.ctypes.data, rf.size,
                                 min_rule_id, delta_round, threads, out.ctypes.data, total,
                                 ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(f"oracle_decompress failed ({rc}); decoded {n.value} > {total}")
    return out[: n.value].copy()


def encode_existing(nodes, lengths, rules_first, rules_second, layers, delta_round):
    """layers: list of (start_id, end_id, row_offset)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    rf = np.ascontiguousarray(rules_first, dtype=np.int32)
    rs = np.ascontiguousarray(rules_second, dtype=np.int32)
    ls = np.array([l[0] for l in layers], dtype=np.uint32)
    le = np.array([l[1] for l in layers], dtype=np.uint32)
    lo = np.array([l[2] for l in layers], dtype=np.uint64)
    out = np.empty(max(1, nodes.size), dtype=np.int32)
    out_len = np.empty(max(1, lengths.size), dtype=np.uint32)
    n = ctypes.c_uint64()
    rc = lib().oracle_encode_existing(nodes.ctypes.data, nodes.size, lengths.ctypes.data,
                                      lengths.size, rf.ctypes.data, rs.ctypes.data, rf.size,
                                      ls.ctypes.data, le.ctypes.data, lo.ctypes.data, len(layers),
                                      delta_round, out.ctypes.data, out_len.ctypes.data,
                                      ctypes.byref(n))
    if rc == -1:
        raise RuntimeError("appended sequence collides with the stored rule region")
    if rc != 0:
        raise RuntimeError(f"oracle_encode_existing failed ({rc})")
    return out[: n.value].copy(), out_len[: lengths.size].copy()


def ref_available() -> bool:
    return os.path.exists(REF_BIN)
