"""Python host side of the gfx950 grammar codec, over the C ABI of include/gfaz_amd.h.

Names follow the reference's operator interface for this path:

* ``compress_traversals``  - the traversal grammar of ``run_path_compression_gpu``
  (include/gpu/compression/compression_workflow_gpu.hpp:75-78) with the reference CPU
  compressor's exact output (src/workflows/compression_workflow.cpp:235-279,548-700).
* ``Rulebook`` / ``decompress_traversals`` - ``prepare_gpu_traversal_rulebook`` and
  ``decompress_gpu_traversal_materialized``
  (src/gpu/decompression/decompression_primitives_gpu.cu:18-39,
  src/gpu/decompression/traversal_decode_gpu.cu:76-120).

Errors raise ``RuntimeError`` carrying the library's message, as the reference's
pybind module does. There is no CPU fallback: without the built HIP library, or
without a GPU, these calls fail.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native

GFZ_INPUT_DEVICE = 1
GFZ_OUTPUT_DEVICE = 2

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32


class _Block(ctypes.Structure):
    _fields_ = [("original_size", _u64), ("payload", ctypes.POINTER(ctypes.c_uint8)),
                ("payload_size", _u64)]


class _Framed(ctypes.Structure):
    _fields_ = [("start_id", _u32), ("end_id", _u32), ("n_rules", _u64),
                ("rounds_run", ctypes.c_int32), ("delta_round", ctypes.c_int32),
                ("n_trav", _u64), ("n_paths", _u32),
                ("encoded_lengths", ctypes.POINTER(ctypes.c_uint32)),
                ("rules_first", _Block), ("rules_second", _Block), ("paths", _Block),
                ("walks", _Block), ("gpu_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


class _Strings(ctypes.Structure):
    _fields_ = [("bytes", ctypes.c_char_p), ("off", ctypes.POINTER(_u64)), ("n", _u64)]


class _WalkKeys(ctypes.Structure):
    _fields_ = [("sample_ids", _Strings), ("hap_indices", ctypes.POINTER(ctypes.c_uint32)),
                ("seq_ids", _Strings), ("seq_starts", ctypes.POINTER(ctypes.c_int64)),
                ("seq_ends", ctypes.POINTER(ctypes.c_int64))]


class _Lines(ctypes.Structure):
    _fields_ = [("text", ctypes.POINTER(ctypes.c_uint8)), ("n_bytes", _u64),
                ("line_off", ctypes.POINTER(_u64)), ("n_lines", _u64)]


class _Layer(ctypes.Structure):
    _fields_ = [("start_id", ctypes.c_uint32), ("end_id", ctypes.c_uint32),
                ("row_offset", _u64)]


class _EncodeOut(ctypes.Structure):
    _fields_ = [("encoded", ctypes.POINTER(ctypes.c_int32)), ("n_encoded", _u64),
                ("encoded_lengths", ctypes.POINTER(ctypes.c_uint32)), ("n_trav", _u64),
                ("flags", ctypes.c_uint32), ("internal", _vp)]


class _CompressOut(ctypes.Structure):
    _fields_ = [
        ("encoded", ctypes.POINTER(ctypes.c_int32)),
        ("n_encoded", _u64),
        ("encoded_lengths", ctypes.POINTER(ctypes.c_uint32)),
        ("n_trav", _u64),
        ("rules_first", ctypes.POINTER(ctypes.c_int32)),
        ("rules_second", ctypes.POINTER(ctypes.c_int32)),
        ("n_rules", _u64),
        ("start_id", _u32),
        ("end_id", _u32),
        ("rounds_run", ctypes.c_int32),
        ("flags", _u32),
        ("round_rules", _u64 * 32),
        ("round_lengths", _u64 * 32),
        ("round_distinct", _u64 * 32),
        ("round_slots", _u64 * 32),
        ("internal", _vp),
    ]


_SIGNATURES = {
    "gfz_last_error": ([], ctypes.c_char_p),
    "gfz_version": ([], ctypes.c_char_p),
    "gfz_ctx_create": ([ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "gfz_ctx_destroy": ([_vp], None),
    "gfz_ctx_sync": ([_vp], ctypes.c_int),
    "gfz_ctx_set_memory_limit": ([_vp, _u64], ctypes.c_int),
    "gfz_ctx_set_profiling": ([_vp, ctypes.c_int], ctypes.c_int),
    "gfz_debug_sort_u64": ([_vp, _vp, _u64, ctypes.c_uint32], ctypes.c_int),
    "gfz_ctx_set_profile_groups": ([_vp, ctypes.c_char_p], ctypes.c_int),
    "gfz_ctx_reset_stats": ([_vp], ctypes.c_int),
    "gfz_ctx_kernel_stat": ([_vp, ctypes.c_char_p, ctypes.POINTER(_u64),
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
                            ctypes.c_int),
    "gfz_ctx_num_cus": ([_vp], ctypes.c_int),
    "gfz_ctx_set_autotune": ([_vp, ctypes.c_int], ctypes.c_int),
    "gfz_ctx_tuned": ([_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "gfz_ctx_set_tuned": ([_vp, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "gfz_ctx_graph_stats": ([_vp, ctypes.POINTER(_u64)], ctypes.c_int),
    "gfz_autotune_compress": ([_vp, _vp, _u64, _vp, _u64, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)],
                              ctypes.c_int),
    "gfz_ctx_save_tuning": ([_vp, ctypes.c_char_p], ctypes.c_int),
    "gfz_ctx_load_tuning": ([_vp, ctypes.c_char_p], ctypes.c_int),
    "gfz_dev_alloc": ([_vp, _u64, ctypes.POINTER(_vp)], ctypes.c_int),
    "gfz_dev_free": ([_vp, _vp], ctypes.c_int),
    "gfz_memcpy_h2d": ([_vp, _vp, _vp, _u64], ctypes.c_int),
    "gfz_memcpy_d2h": ([_vp, _vp, _vp, _u64], ctypes.c_int),
    "gfz_frame_int32": ([_vp, _u64, ctypes.c_int, _vp], ctypes.c_int),
    "gfz_frame_rule_columns": ([_vp, _vp, _u64, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "gfz_unframe_int32": ([_vp, _u64, _u64, _vp, _u64], ctypes.c_int),
    "gfz_unframe_rule_columns": ([_vp, _u64, _u64, _vp, _u64, _u64, _vp, _vp, _u64],
                                 ctypes.c_int),
    "gfz_block_free": ([_vp], None),
    "gfz_catalog_create": ([ctypes.POINTER(_vp)], ctypes.c_int),
    "gfz_catalog_free": ([_vp], None),
    "gfz_catalog_add_paths": ([_vp, ctypes.POINTER(_Strings), ctypes.POINTER(_Strings), _u64],
                              ctypes.c_int),
    "gfz_catalog_add_walks": ([_vp, ctypes.POINTER(_WalkKeys), _u64], ctypes.c_int),
    "gfz_catalog_info": ([_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64)], ctypes.c_int),
    "gfz_catalog_find_paths": ([_vp, ctypes.POINTER(_Strings), _vp], ctypes.c_int),
    "gfz_catalog_find_walks": ([_vp, ctypes.POINTER(_WalkKeys), _vp], ctypes.c_int),
    "gfz_catalog_find_walks_by_name": ([_vp, ctypes.POINTER(_Strings), _vp], ctypes.c_int),
    "gfz_extract_path_lines": ([_vp, _vp, _vp, ctypes.POINTER(_Strings), ctypes.POINTER(_Lines)],
                               ctypes.c_int),
    "gfz_extract_walk_lines": ([_vp, _vp, _vp, ctypes.POINTER(_WalkKeys), ctypes.POINTER(_Lines)],
                               ctypes.c_int),
    "gfz_extract_walk_lines_by_name": ([_vp, _vp, _vp, ctypes.POINTER(_Strings),
                                        ctypes.POINTER(_Lines)], ctypes.c_int),
    "gfz_lines_free": ([ctypes.POINTER(_Lines)], None),
    "gfz_framer_begin": ([_vp, _u64, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "gfz_framer_ready": ([_vp, _u64], ctypes.c_int),
    "gfz_framer_end": ([_vp, _vp], ctypes.c_int),
    "gfz_framer_free": ([_vp], None),
    "gfz_compress_framed": ([_vp, _vp, _u64, _vp, _u64, _u32, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int, _u32, ctypes.c_int, ctypes.POINTER(_Framed)],
                            ctypes.c_int),
    "gfz_framed_free": ([ctypes.POINTER(_Framed)], None),
    "gfz_decompress_framed": ([_vp, ctypes.POINTER(_Framed), _vp, _vp, _u64, _u32,
                               ctypes.POINTER(_u64), _vp, ctypes.POINTER(ctypes.c_int), _vp],
                              ctypes.c_int),
    "gfz_resident_load": ([_vp, _vp, _vp, _u64, _vp, _u64, ctypes.c_int, ctypes.c_uint32,
                           ctypes.POINTER(_vp)], ctypes.c_int),
    "gfz_resident_free": ([_vp], None),
    "gfz_format_steps": ([_vp, ctypes.c_int, _vp, _u64, _vp, _u64, _vp, _vp, _vp, _vp,
                          ctypes.c_uint32, _vp, _u64, ctypes.POINTER(_u64), _vp], ctypes.c_int),
    "gfz_parse_steps": ([_vp, ctypes.c_int, _vp, _u64, _vp, _vp, _u64, ctypes.c_uint32, _vp,
                         _u64, _vp, ctypes.POINTER(_u64)], ctypes.c_int),
    "gfz_extract_noalloc": ([_vp, _vp, _vp, _u64, _vp, _u64, _vp, ctypes.POINTER(_u64)],
                            ctypes.c_int),
    "gfz_resident_append": ([_vp, _vp, _vp, _u64, _vp, _u64, ctypes.c_uint32, ctypes.POINTER(_u64)],
                            ctypes.c_int),
    "gfz_resident_info": ([_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64)], ctypes.c_int),
    "gfz_resident_lengths": ([_vp, _vp], ctypes.c_int),
    "gfz_extract": ([_vp, _vp, _vp, _u64, _vp, _u64, ctypes.c_uint32, _vp,
                     ctypes.POINTER(_u64)], ctypes.c_int),
    "gfz_encode_with_rules": ([_vp, _vp, _u64, _vp, _u64, _vp, _vp, _u64, _vp, _u64, ctypes.c_int,
                               ctypes.c_uint32, _vp], ctypes.c_int),
    "gfz_encode_free": ([_vp], None),
    "gfz_decompress_stream": ([_vp, _vp, _vp, _u64, _vp, _u64, ctypes.c_int, _u64,
                               ctypes.c_uint32, _vp, _vp], ctypes.c_int),
    "gfz_calibration_copy": ([_vp, _vp, _vp, _u64, ctypes.POINTER(ctypes.c_double)],
                             ctypes.c_int),
    "gfz_compress_traversals": ([_vp, _vp, _u64, _vp, _u64, _u32, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, _u32, ctypes.POINTER(_CompressOut)], ctypes.c_int),
    "gfz_compress_free": ([ctypes.POINTER(_CompressOut)], None),
    "gfz_rulebook_load": ([_vp, _vp, _vp, _u64, _u32, _u32, ctypes.POINTER(_vp)], ctypes.c_int),
    "gfz_rulebook_free": ([_vp], None),
    "gfz_rulebook_rule_size": ([_vp, _u64, ctypes.POINTER(_u32)], ctypes.c_int),
    "gfz_rulebook_info": ([_vp, ctypes.POINTER(_u32), ctypes.POINTER(ctypes.c_int),
                           ctypes.POINTER(_u64)], ctypes.c_int),
    "gfz_decoded_size": ([_vp, _vp, _vp, _u64, _vp, _u64, _u32, ctypes.POINTER(_u64)],
                         ctypes.c_int),
    "gfz_decompress_traversals": ([_vp, _vp, _vp, _u64, _vp, _vp, _u64, ctypes.c_int, _vp, _u64,
                                   _u32, ctypes.POINTER(_u64), _vp, ctypes.POINTER(ctypes.c_int)],
                                  ctypes.c_int),
}


def lib() -> ctypes.CDLL:
    L = _native.codec_lib()
    if not getattr(L, "_gfz_typed", False):
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        L._gfz_typed = True
    return L


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().gfz_last_error()
        raise RuntimeError((msg.decode() if msg else "") or f"gfaz_amd error {rc}")


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


class DeviceArray:
    """A pool-backed device buffer of a numpy dtype (no torch needed)."""

    def __init__(self, ctx: "Context", n: int, dtype):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = _vp()
        _check(lib().gfz_dev_alloc(ctx.handle, max(1, self.n) * self.dtype.itemsize, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx: "Context", a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        if a.size:
            _check(lib().gfz_memcpy_h2d(ctx.handle, d.ptr, a.ctypes.data, a.nbytes))
        return d

    def to_host(self, n: int | None = None) -> np.ndarray:
        n = self.n if n is None else int(n)
        out = np.empty(n, dtype=self.dtype)
        if n:
            _check(lib().gfz_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr is not None and self.ctx.handle:
            lib().gfz_dev_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device context: stream, caching device pool, optional kernel timing."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().gfz_ctx_create(device, ctypes.byref(h)))
        self.handle = h.value
        self.device = device

    def close(self) -> None:
        if self.handle:
            lib().gfz_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self) -> None:
        _check(lib().gfz_ctx_sync(self.handle))

    def set_memory_limit(self, nbytes: int) -> None:
        _check(lib().gfz_ctx_set_memory_limit(self.handle, int(nbytes)))

    def set_profiling(self, on: bool, groups=None) -> None:
        """Per-group kernel timing. `groups` limits it to those kernel groups
        (None = all): each timed group drains the stream around its launches."""
        _check(lib().gfz_ctx_set_profile_groups(self.handle,
                                                ",".join(groups or ()).encode()))
        _check(lib().gfz_ctx_set_profiling(self.handle, int(bool(on))))

    def reset_stats(self) -> None:
        _check(lib().gfz_ctx_reset_stats(self.handle))

    def kernel_stat(self, name: str) -> dict:
        n, ms, b = _u64(), ctypes.c_double(), ctypes.c_double()
        _check(lib().gfz_ctx_kernel_stat(self.handle, name.encode(), ctypes.byref(n),
                                         ctypes.byref(ms), ctypes.byref(b)))
        return {"launches": int(n.value), "total_ms": float(ms.value), "bytes": float(b.value)}

    def num_cus(self) -> int:
        return int(lib().gfz_ctx_num_cus(self.handle))

    def set_autotune(self, on: bool) -> None:
        _check(lib().gfz_ctx_set_autotune(self.handle, int(bool(on))))

    def tuned(self, key: str) -> int:
        v = ctypes.c_int()
        _check(lib().gfz_ctx_tuned(self.handle, key.encode(), ctypes.byref(v)))
        return v.value

    def set_tuned(self, key: str, value: int) -> None:
        _check(lib().gfz_ctx_set_tuned(self.handle, key.encode(), int(value)))

    def graph_stats(self) -> dict:
        """Round graphs of compression (gfz_ctx_graph_stats): launches, graphs
        instantiated and graphs updated in place over the ctx's life."""
        out = (_u64 * 3)()
        _check(lib().gfz_ctx_graph_stats(self.handle, out))
        return {"launches": int(out[0]), "instantiated": int(out[1]), "updated": int(out[2])}

    def autotune_compress(self, nodes, lengths, rounds: int = 8, threshold: int = 2,
                          delta_round: int = 1) -> float:
        """Tune the pair-counting knobs on this input (gfz_autotune_compress); returns
        the fastest setting's compress time in ms."""
        lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
        flags = 0
        if isinstance(nodes, DeviceArray):
            flags |= GFZ_INPUT_DEVICE
            nptr, nn = nodes.ptr, nodes.n
        else:
            nodes = np.ascontiguousarray(nodes, dtype=np.int32)
            nptr, nn = nodes.ctypes.data, nodes.size
        ms = ctypes.c_double()
        _check(lib().gfz_autotune_compress(self.handle, nptr, nn, lengths.ctypes.data,
                                           lengths.size, rounds, threshold, delta_round, flags,
                                           ctypes.byref(ms)))
        return float(ms.value)

    def save_tuning(self, path: str) -> None:
        _check(lib().gfz_ctx_save_tuning(self.handle, path.encode()))

    def load_tuning(self, path: str) -> None:
        _check(lib().gfz_ctx_load_tuning(self.handle, path.encode()))

    def calibration_copy(self, dst: "DeviceArray", src: "DeviceArray", nbytes: int) -> float:
        """Plain device copy of nbytes; returns its duration in ms."""
        ms = ctypes.c_double()
        _check(lib().gfz_calibration_copy(self.handle, dst.ptr, src.ptr, int(nbytes),
                                          ctypes.byref(ms)))
        return float(ms.value)


@dataclass
class Compressed:
    """Traversal grammar: what CompressedData's traversal fields hold before zstd."""

    encoded: np.ndarray | None        # int32 (host) or None when kept on the device
    encoded_lengths: np.ndarray       # uint32 per traversal
    rules_first: np.ndarray | None    # int32 raw, rule-ID order
    rules_second: np.ndarray | None
    start_id: int
    end_id: int
    rounds_run: int
    rules_per_round: list = field(default_factory=list)
    lengths_per_round: list = field(default_factory=list)
    n_encoded: int = 0
    n_rules: int = 0


def compress_traversals(ctx: Context, nodes, lengths, n_paths: int = 0, rounds: int = 8,
                        threshold: int = 2, delta_round: int = 1,
                        keep_on_device: bool = False) -> Compressed:
    """Grammar-compress flat traversals. ``nodes`` is a host int32 array or a DeviceArray."""
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    flags = 0
    if isinstance(nodes, DeviceArray):
        flags |= GFZ_INPUT_DEVICE
        nptr, nn = nodes.ptr, nodes.n
    else:
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        nptr, nn = nodes.ctypes.data, nodes.size
    if keep_on_device:
        flags |= GFZ_OUTPUT_DEVICE
    out = _CompressOut()
    _check(lib().gfz_compress_traversals(ctx.handle, nptr, nn, lengths.ctypes.data, lengths.size,
                                         n_paths, rounds, threshold, delta_round, flags,
                                         ctypes.byref(out)))
    try:
        def arr(p, n, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)

        host = not keep_on_device
        res = Compressed(
            encoded=arr(out.encoded, out.n_encoded, np.int32) if host else None,
            encoded_lengths=arr(out.encoded_lengths, out.n_trav, np.uint32),
            rules_first=arr(out.rules_first, out.n_rules, np.int32) if host else None,
            rules_second=arr(out.rules_second, out.n_rules, np.int32) if host else None,
            start_id=int(out.start_id), end_id=int(out.end_id), rounds_run=int(out.rounds_run),
            rules_per_round=[int(out.round_rules[i]) for i in range(min(32, out.rounds_run))],
            lengths_per_round=[int(out.round_lengths[i]) for i in range(min(32, out.rounds_run))],
            n_encoded=int(out.n_encoded), n_rules=int(out.n_rules))
    finally:
        lib().gfz_compress_free(ctypes.byref(out))
    return res


class Rulebook:
    """Device-resident rulebook: rule columns (raw, rule-ID order) and the ID window."""

    def __init__(self, ctx: Context, rules_first, rules_second, min_rule_id: int):
        self.ctx = ctx
        rf = np.ascontiguousarray(rules_first, dtype=np.int32)
        rs = np.ascontiguousarray(rules_second, dtype=np.int32)
        if rf.size != rs.size:
            raise ValueError("rule columns differ in length")
        h = _vp()
        _check(lib().gfz_rulebook_load(ctx.handle, rf.ctypes.data, rs.ctypes.data, rf.size,
                                       int(min_rule_id), 0, ctypes.byref(h)))
        self.handle = h.value
        self.n_rules = int(rf.size)
        self.min_rule_id = int(min_rule_id)

    def info(self) -> dict:
        d, hb, be = _u32(), ctypes.c_int(), _u64()
        _check(lib().gfz_rulebook_info(self.handle, ctypes.byref(d), ctypes.byref(hb),
                                       ctypes.byref(be)))
        return {"max_depth": int(d.value), "has_bodies": bool(hb.value),
                "body_elems": int(be.value)}

    def rule_size(self, index: int) -> int:
        s = _u32()
        _check(lib().gfz_rulebook_rule_size(self.handle, index, ctypes.byref(s)))
        return int(s.value)

    def close(self) -> None:
        if self.handle:
            lib().gfz_rulebook_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def debug_sort_u64(ctx: Context, keys: np.ndarray, local_cap: int = 0) -> np.ndarray:
    """Sorts distinct uint64 keys on the device with the rule-ordering sort
    (a test entry; local_cap > 0 sends larger buckets down its in-memory path)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    d = DeviceArray.from_host(ctx, keys)
    try:
        _check(lib().gfz_debug_sort_u64(ctx.handle, d.ptr, keys.size, int(local_cap)))
        return d.to_host()
    finally:
        d.free()


def decompress_traversals(ctx: Context, rb: Rulebook, encoded, encoded_lengths,
                          orig_lengths=None, delta_round: int = 1, out: DeviceArray | None = None):
    """Expand + inverse delta. Returns (decoded int32 host array or DeviceArray, lengths,
    length_mismatch)."""
    enc_len = np.ascontiguousarray(encoded_lengths, dtype=np.uint32)
    orig = None if orig_lengths is None else np.ascontiguousarray(orig_lengths, dtype=np.uint32)
    flags = 0
    if isinstance(encoded, DeviceArray):
        flags |= GFZ_INPUT_DEVICE
        eptr, en = encoded.ptr, encoded.n
    else:
        encoded = np.ascontiguousarray(encoded, dtype=np.int32)
        eptr, en = encoded.ctypes.data, encoded.size
    n_out = _u64()
    lens = np.empty(max(1, enc_len.size), dtype=np.uint32)
    mism = ctypes.c_int()
    if out is not None:
        flags |= GFZ_OUTPUT_DEVICE
        result, optr, cap = out, out.ptr, out.n
    else:
        # The stored original lengths give the decoded size without a planning
        # pass; a disagreeing rulebook is caught by the library's capacity check.
        if orig is not None:
            total = int(orig.sum(dtype=np.uint64))
        else:
            _check(lib().gfz_decoded_size(ctx.handle, rb.handle, eptr, en, enc_len.ctypes.data,
                                          enc_len.size, flags, ctypes.byref(n_out)))
            total = int(n_out.value)
        result = np.empty(max(1, total), dtype=np.int32)
        optr, cap = result.ctypes.data, total
    args = (ctx.handle, rb.handle, eptr, en, enc_len.ctypes.data, _ptr(orig), enc_len.size,
            delta_round)
    rc = lib().gfz_decompress_traversals(*args, optr, cap, flags, ctypes.byref(n_out),
                                         lens.ctypes.data, ctypes.byref(mism))
    if rc != 0 and int(n_out.value) > cap:
        if out is not None:
            raise ValueError(f"device output holds {out.n} < {int(n_out.value)}")
        result = np.empty(max(1, int(n_out.value)), dtype=np.int32)
        optr, cap = result.ctypes.data, int(n_out.value)
        rc = lib().gfz_decompress_traversals(*args, optr, cap, flags, ctypes.byref(n_out),
                                             lens.ctypes.data, ctypes.byref(mism))
    _check(rc)
    total = int(n_out.value)
    if out is None:
        result = result[:total]
    return result, lens[: enc_len.size], bool(mism.value)


class DeviceCompressed:
    """Compression result kept in HBM (GFZ_OUTPUT_DEVICE): encoded stream and rule
    columns stay on the device until ``close``; lengths and counts are on the host."""

    def __init__(self, ctx: Context, nodes, lengths, n_paths: int = 0, rounds: int = 8,
                 threshold: int = 2, delta_round: int = 1):
        self.ctx = ctx
        self.lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
        flags = GFZ_OUTPUT_DEVICE
        if isinstance(nodes, DeviceArray):
            flags |= GFZ_INPUT_DEVICE
            nptr, nn = nodes.ptr, nodes.n
        else:
            nodes = np.ascontiguousarray(nodes, dtype=np.int32)
            nptr, nn = nodes.ctypes.data, nodes.size
        self._out = _CompressOut()
        _check(lib().gfz_compress_traversals(ctx.handle, nptr, nn, self.lengths.ctypes.data,
                                             self.lengths.size, n_paths, rounds, threshold,
                                             delta_round, flags, ctypes.byref(self._out)))
        o = self._out
        self.n_encoded = int(o.n_encoded)
        self.n_rules = int(o.n_rules)
        self.start_id = int(o.start_id)
        self.end_id = int(o.end_id)
        self.rounds_run = int(o.rounds_run)
        self.rules_per_round = [int(o.round_rules[i]) for i in range(min(32, o.rounds_run))]
        self.lengths_per_round = [int(o.round_lengths[i]) for i in range(min(32, o.rounds_run))]
        self.distinct_per_round = [int(o.round_distinct[i]) for i in range(min(32, o.rounds_run))]
        self.slots_per_round = [int(o.round_slots[i]) for i in range(min(32, o.rounds_run))]
        self.encoded_lengths = np.ctypeslib.as_array(o.encoded_lengths, shape=(max(1, o.n_trav),))[
            : o.n_trav].copy()

    @property
    def encoded_ptr(self) -> int:
        return ctypes.cast(self._out.encoded, _vp).value

    def rulebook(self) -> "Rulebook":
        rb = Rulebook.__new__(Rulebook)
        rb.ctx = self.ctx
        h = _vp()
        _check(lib().gfz_rulebook_load(self.ctx.handle, ctypes.cast(self._out.rules_first, _vp),
                                       ctypes.cast(self._out.rules_second, _vp), self.n_rules,
                                       self.start_id, GFZ_INPUT_DEVICE, ctypes.byref(h)))
        rb.handle = h.value
        rb.n_rules = self.n_rules
        rb.min_rule_id = self.start_id
        return rb

    def encoded_view(self) -> DeviceArray:
        """Non-owning DeviceArray over the encoded stream."""
        d = DeviceArray.__new__(DeviceArray)
        d.ctx, d.n, d.dtype, d.ptr = self.ctx, self.n_encoded, np.dtype(np.int32), self.encoded_ptr
        d.free = lambda: None
        return d

    def to_host(self) -> Compressed:
        o = self._out
        enc = np.empty(self.n_encoded, np.int32)
        rf = np.empty(self.n_rules, np.int32)
        rs = np.empty(self.n_rules, np.int32)
        for dst, src in ((enc, o.encoded), (rf, o.rules_first), (rs, o.rules_second)):
            if dst.size:
                _check(lib().gfz_memcpy_d2h(self.ctx.handle, dst.ctypes.data,
                                            ctypes.cast(src, _vp), dst.nbytes))
        return Compressed(enc, self.encoded_lengths.copy(), rf, rs, self.start_id, self.end_id,
                          self.rounds_run, list(self.rules_per_round),
                          list(self.lengths_per_round), self.n_encoded, self.n_rules)

    def close(self) -> None:
        if self._out is not None:
            lib().gfz_compress_free(ctypes.byref(self._out))
            self._out = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- traversal-column framing (host; gfz_frame_* / gfz_unframe_*) -------------

def _take_block(b: _Block) -> tuple[int, bytes]:
    data = ctypes.string_at(b.payload, b.payload_size) if b.payload_size else b""
    lib().gfz_block_free(ctypes.byref(b))
    return int(b.original_size), data


def frame_int32(values, threads: int = 0) -> tuple[int, bytes]:
    """(original_size, zstd payload) of an int32 column, as zstd_compress_int32_vector."""
    a = np.ascontiguousarray(values, dtype=np.int32)
    b = _Block()
    _check(lib().gfz_frame_int32(a.ctypes.data, a.size, int(threads), ctypes.byref(b)))
    return _take_block(b)


def frame_rule_columns(first, second, threads: int = 0):
    """Raw rule columns -> ((orig, payload), (orig, payload)) as process_rules + zstd."""
    f = np.ascontiguousarray(first, dtype=np.int32)
    s_ = np.ascontiguousarray(second, dtype=np.int32)
    if f.size != s_.size:
        raise ValueError("rule columns differ in length")
    bf, bs = _Block(), _Block()
    _check(lib().gfz_frame_rule_columns(f.ctypes.data, s_.ctypes.data, f.size, int(threads),
                                        ctypes.byref(bf), ctypes.byref(bs)))
    return _take_block(bf), _take_block(bs)


def unframe_int32(payload: bytes, original_size: int) -> np.ndarray:
    out = np.empty(max(1, original_size // 4), dtype=np.int32)
    buf = ctypes.create_string_buffer(payload, len(payload)) if payload else None
    _check(lib().gfz_unframe_int32(buf, len(payload), int(original_size), out.ctypes.data,
                                   out.size))
    return out[: original_size // 4]


def unframe_rule_columns(first: tuple[int, bytes], second: tuple[int, bytes], n_rules: int):
    f = np.empty(max(1, n_rules), dtype=np.int32)
    s_ = np.empty(max(1, n_rules), dtype=np.int32)
    pf = ctypes.create_string_buffer(first[1], len(first[1])) if first[1] else None
    ps = ctypes.create_string_buffer(second[1], len(second[1])) if second[1] else None
    _check(lib().gfz_unframe_rule_columns(pf, len(first[1]), int(first[0]), ps, len(second[1]),
                                          int(second[0]), f.ctypes.data, s_.ctypes.data,
                                          int(n_rules)))
    return f[:n_rules], s_[:n_rules]


class Framer:
    """Piecewise zstd framing of one column (gfz_framer_*): ``ready(n)`` declares
    bytes [0, n) of the column final; ``end()`` returns (original_size, payload),
    the bytes gfz_frame_int32 gives for the same data."""

    def __init__(self, column: np.ndarray, threads: int = 0):
        self.column = column  # the library reads it until end()
        h = _vp()
        _check(lib().gfz_framer_begin(column.ctypes.data if column.nbytes else None,
                                      column.nbytes, int(threads), ctypes.byref(h)))
        self.handle = h.value

    def ready(self, nbytes: int) -> None:
        _check(lib().gfz_framer_ready(self.handle, int(nbytes)))

    def end(self) -> tuple[int, bytes]:
        b = _Block()
        try:
            _check(lib().gfz_framer_end(self.handle, ctypes.byref(b)))
            return _take_block(b)
        finally:
            lib().gfz_framer_free(self.handle)
            self.handle = None

    def __del__(self):
        if getattr(self, "handle", None):
            lib().gfz_framer_free(self.handle)
            self.handle = None


@dataclass
class Framed:
    """The traversal fields of CompressedData, framed (gfz_compress_framed)."""

    start_id: int
    end_id: int
    n_rules: int
    rounds_run: int
    delta_round: int
    n_paths: int
    encoded_lengths: np.ndarray            # uint32 per traversal
    rules_first: tuple[int, bytes]         # (original_size, zstd payload)
    rules_second: tuple[int, bytes]
    paths: tuple[int, bytes]
    walks: tuple[int, bytes]
    gpu_ms: float = 0.0
    total_ms: float = 0.0


def compress_framed(ctx: Context, nodes, lengths, n_paths: int = 0, rounds: int = 8,
                    threshold: int = 2, delta_round: int = 1, threads: int = 0) -> Framed:
    """Grammar-compress and frame the columns, the framing overlapped with the
    copies out of HBM (gfz_compress_framed)."""
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    flags = 0
    if isinstance(nodes, DeviceArray):
        flags |= GFZ_INPUT_DEVICE
        nptr, nn = nodes.ptr, nodes.n
    else:
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        nptr, nn = nodes.ctypes.data, nodes.size
    out = _Framed()
    _check(lib().gfz_compress_framed(ctx.handle, nptr, nn, lengths.ctypes.data, lengths.size,
                                     int(n_paths), rounds, threshold, delta_round, flags,
                                     int(threads), ctypes.byref(out)))
    try:
        def blk(b: _Block):
            data = ctypes.string_at(b.payload, b.payload_size) if b.payload_size else b""
            return int(b.original_size), data

        el = (np.ctypeslib.as_array(out.encoded_lengths, shape=(lengths.size,)).copy()
              if lengths.size else np.zeros(0, dtype=np.uint32))
        return Framed(int(out.start_id), int(out.end_id), int(out.n_rules),
                      int(out.rounds_run), int(out.delta_round), int(out.n_paths), el,
                      blk(out.rules_first), blk(out.rules_second), blk(out.paths),
                      blk(out.walks), float(out.gpu_ms), float(out.total_ms))
    finally:
        lib().gfz_framed_free(ctypes.byref(out))


def decompress_framed(ctx: Context, f: Framed, orig_lengths=None):
    """Unframe and decode (gfz_decompress_framed). Returns (nodes, lengths,
    length_mismatch, timings_ms) with timings = (rulebook built, payloads decoded,
    done) on the host clock."""
    keep = []

    def blk(pair):
        b = _Block()
        b.original_size = int(pair[0])
        if pair[1]:
            buf = ctypes.create_string_buffer(pair[1], len(pair[1]))
            keep.append(buf)
            b.payload = ctypes.cast(buf, ctypes.POINTER(ctypes.c_uint8))
        b.payload_size = len(pair[1])
        return b

    el = np.ascontiguousarray(f.encoded_lengths, dtype=np.uint32)
    s = _Framed()
    s.start_id, s.end_id, s.n_rules = f.start_id, f.end_id, f.n_rules
    s.rounds_run, s.delta_round = f.rounds_run, f.delta_round
    s.n_trav, s.n_paths = el.size, f.n_paths
    s.encoded_lengths = el.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    s.rules_first, s.rules_second = blk(f.rules_first), blk(f.rules_second)
    s.paths, s.walks = blk(f.paths), blk(f.walks)
    ol = None if orig_lengths is None else np.ascontiguousarray(orig_lengths, dtype=np.uint32)
    n_out, mism = _u64(), ctypes.c_int()
    lens = np.zeros(max(1, el.size), dtype=np.uint32)
    tms = (ctypes.c_double * 3)()
    # Size first (out_cap 0 fails with the decoded size), then decode.
    cap = int(ol.sum()) if ol is not None else None
    if cap is None:
        rc = lib().gfz_decompress_framed(ctx.handle, ctypes.byref(s), None, None, 0, 0,
                                         ctypes.byref(n_out), None, None, None)
        if rc != 0 and n_out.value == 0:
            _check(rc)
        cap = int(n_out.value)
    out = np.empty(max(1, cap), dtype=np.int32)
    _check(lib().gfz_decompress_framed(ctx.handle, ctypes.byref(s), _ptr(ol), out.ctypes.data,
                                       cap, 0, ctypes.byref(n_out), lens.ctypes.data,
                                       ctypes.byref(mism), tms))
    return (out[: n_out.value], lens[: el.size], bool(mism.value), tuple(tms))


# ---- resident decoding / extract ----------------------------------------------

class Resident:
    """Encoded traversals kept decodable in HBM against a loaded Rulebook
    (gfz_resident_load); ``extract`` decodes any subset in one launch."""

    def __init__(self, ctx: Context, rb: "Rulebook", encoded, encoded_lengths,
                 delta_round: int = 1):
        self.ctx, self.rb = ctx, rb
        enc_len = np.ascontiguousarray(encoded_lengths, dtype=np.uint32)
        flags = 0
        if isinstance(encoded, DeviceArray):
            flags |= GFZ_INPUT_DEVICE
            eptr, en = encoded.ptr, encoded.n
        else:
            encoded = np.ascontiguousarray(encoded, dtype=np.int32)
            eptr, en = encoded.ctypes.data, encoded.size
        h = _vp()
        _check(lib().gfz_resident_load(ctx.handle, rb.handle, eptr, en, enc_len.ctypes.data,
                                       enc_len.size, int(delta_round), flags, ctypes.byref(h)))
        self.handle = h.value
        nt, nd = _u64(), _u64()
        _check(lib().gfz_resident_info(self.handle, ctypes.byref(nt), ctypes.byref(nd)))
        self.n_trav, self.n_decoded = int(nt.value), int(nd.value)

    def append(self, encoded, encoded_lengths) -> int:
        """Append encoded traversals (gfz_resident_append); returns the first new index."""
        enc_len = np.ascontiguousarray(encoded_lengths, dtype=np.uint32)
        flags = 0
        if isinstance(encoded, DeviceArray):
            flags |= GFZ_INPUT_DEVICE
            eptr, en = encoded.ptr, encoded.n
        else:
            encoded = np.ascontiguousarray(encoded, dtype=np.int32)
            eptr, en = encoded.ctypes.data, encoded.size
        first = _u64()
        _check(lib().gfz_resident_append(self.ctx.handle, self.handle, eptr, en,
                                         enc_len.ctypes.data, enc_len.size, flags,
                                         ctypes.byref(first)))
        nt, nd = _u64(), _u64()
        _check(lib().gfz_resident_info(self.handle, ctypes.byref(nt), ctypes.byref(nd)))
        self.n_trav, self.n_decoded = int(nt.value), int(nd.value)
        return int(first.value)

    def lengths(self) -> np.ndarray:
        out = np.empty(max(1, self.n_trav), dtype=np.uint32)
        _check(lib().gfz_resident_lengths(self.handle, out.ctypes.data))
        return out[: self.n_trav]

    def extract(self, ids, out: DeviceArray | None = None):
        """Decode traversals ``ids`` back to back: (values, offsets[n_ids + 1])."""
        ids = np.ascontiguousarray(ids, dtype=np.uint64)
        offs = np.empty(ids.size + 1, dtype=np.uint64)
        n = _u64()
        if out is not None:
            _check(lib().gfz_extract(self.ctx.handle, self.handle, ids.ctypes.data, ids.size,
                                     out.ptr, out.n, GFZ_OUTPUT_DEVICE, offs.ctypes.data,
                                     ctypes.byref(n)))
            return out, offs
        lens = self.lengths()
        total = int(lens[ids.astype(np.int64)].sum(dtype=np.uint64)) if ids.size and \
            int(ids.max()) < self.n_trav else 0
        res = np.empty(max(1, total), dtype=np.int32)
        _check(lib().gfz_extract(self.ctx.handle, self.handle, ids.ctypes.data, ids.size,
                                 res.ctypes.data, total, 0, offs.ctypes.data, ctypes.byref(n)))
        return res[: int(n.value)], offs

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().gfz_resident_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- add-haplotypes encoding ------------------------------------------------------

def encode_with_rules(ctx: Context, nodes, lengths, rules_first, rules_second, layers,
                      delta_round: int = 1):
    """Encode new traversals with stored rules (gfz_encode_with_rules).
    layers: [(start_id, end_id, row_offset)]. Returns (encoded, encoded_lengths)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    rf = np.ascontiguousarray(rules_first, dtype=np.int32)
    rs = np.ascontiguousarray(rules_second, dtype=np.int32)
    if rf.size != rs.size:
        raise ValueError("rule columns differ in length")
    arr = (_Layer * max(1, len(layers)))(*[_Layer(int(a), int(b), int(c)) for a, b, c in layers])
    out = _EncodeOut()
    _check(lib().gfz_encode_with_rules(ctx.handle, nodes.ctypes.data, nodes.size,
                                       lengths.ctypes.data, lengths.size, rf.ctypes.data,
                                       rs.ctypes.data, rf.size, arr, len(layers),
                                       int(delta_round), 0, ctypes.byref(out)))
    try:
        enc = (np.ctypeslib.as_array(out.encoded, shape=(out.n_encoded,)).copy()
               if out.n_encoded else np.zeros(0, np.int32))
        lens = (np.ctypeslib.as_array(out.encoded_lengths, shape=(out.n_trav,)).copy()
                if out.n_trav else np.zeros(0, np.uint32))
    finally:
        lib().gfz_encode_free(ctypes.byref(out))
    return enc, lens


# ---- streaming decompress -----------------------------------------------------------

class _Chunk(ctypes.Structure):
    _fields_ = [("values", ctypes.POINTER(ctypes.c_int32)), ("n_values", _u64),
                ("trav_begin", _u64), ("trav_end", _u64), ("expanded_begin", _u64),
                ("lengths", ctypes.POINTER(ctypes.c_uint32))]


_CHUNK_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(_Chunk), ctypes.c_void_p)


def decompress_stream(ctx: Context, rb: "Rulebook", encoded, encoded_lengths,
                      delta_round: int = 1, chunk_bytes: int = 0, consumer=None) -> None:
    """Decode every traversal in chunks of whole traversals (gfz_decompress_stream).
    consumer(values, trav_begin, trav_end, expanded_begin, lengths) gets views of pinned
    memory valid during the call; a truthy return value stops the stream."""
    enc_len = np.ascontiguousarray(encoded_lengths, dtype=np.uint32)
    flags = 0
    if isinstance(encoded, DeviceArray):
        flags |= GFZ_INPUT_DEVICE
        eptr, en = encoded.ptr, encoded.n
    else:
        encoded = np.ascontiguousarray(encoded, dtype=np.int32)
        eptr, en = encoded.ctypes.data, encoded.size
    errors = []

    def cb(ptr, _user):
        c = ptr.contents
        try:
            vals = (np.ctypeslib.as_array(c.values, shape=(c.n_values,)) if c.n_values
                    else np.zeros(0, np.int32))
            n_t = c.trav_end - c.trav_begin
            lens = (np.ctypeslib.as_array(c.lengths, shape=(n_t,)) if n_t
                    else np.zeros(0, np.uint32))
            return 1 if consumer(vals, int(c.trav_begin), int(c.trav_end),
                                 int(c.expanded_begin), lens) else 0
        except Exception as e:  # surfaced after the stream stops
            errors.append(e)
            return 1

    fn = _CHUNK_FN(cb)
    rc = lib().gfz_decompress_stream(ctx.handle, rb.handle, eptr, en, enc_len.ctypes.data,
                                     enc_len.size, int(delta_round), int(chunk_bytes), flags,
                                     ctypes.cast(fn, _vp), None)
    if errors:
        raise errors[0]
    _check(rc)


# ---- P/W line step text --------------------------------------------------------------

GFZ_TEXT_P, GFZ_TEXT_W = 0, 1


def _concat(parts) -> tuple[bytes, np.ndarray]:
    parts = [bytes(p) for p in parts]
    off = np.zeros(len(parts) + 1, dtype=np.uint64)
    if parts:
        np.cumsum([len(p) for p in parts], dtype=np.uint64, out=off[1:])
    return b"".join(parts), off


def format_steps(ctx: Context, kind: int, nodes, lengths, heads, tails, out=None,
                 with_offsets: bool = False):
    """Lines head_t + steps_t + tail_t (gfz_format_steps). nodes: host array or
    DeviceArray; heads/tails: one bytes object per traversal. Returns the text as bytes,
    or (with ``out`` a DeviceArray of uint8) the text size, written there; with
    ``with_offsets`` also the n_trav + 1 line starts."""
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    flags = 0
    if isinstance(nodes, DeviceArray):
        flags |= GFZ_INPUT_DEVICE
        nptr, nn = nodes.ptr, nodes.n
    else:
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        nptr, nn = nodes.ctypes.data, nodes.size
    hb, ho = _concat(heads)
    tb, tof = _concat(tails)
    hbuf = ctypes.create_string_buffer(hb, max(1, len(hb)))
    tbuf = ctypes.create_string_buffer(tb, max(1, len(tb)))
    n = _u64()
    line_off = np.zeros(lengths.size + 1, dtype=np.uint64)
    L = lib()
    args = (ctx.handle, int(kind), nptr, nn, lengths.ctypes.data, lengths.size, hbuf,
            ho.ctypes.data, tbuf, tof.ctypes.data)
    if out is not None:
        _check(L.gfz_format_steps(*args, flags | GFZ_OUTPUT_DEVICE, out.ptr, out.n,
                                  ctypes.byref(n), line_off.ctypes.data))
        return (int(n.value), line_off) if with_offsets else int(n.value)
    _check(L.gfz_format_steps(*args, flags, None, 0, ctypes.byref(n), None))
    res = np.empty(max(1, int(n.value)), dtype=np.uint8)
    _check(L.gfz_format_steps(*args, flags, res.ctypes.data, res.size, ctypes.byref(n),
                              line_off.ctypes.data))
    text = res[: int(n.value)].tobytes()
    return (text, line_off) if with_offsets else text


def format_size(ctx: Context, kind: int, nodes, lengths, heads, tails):
    """Text size and line starts of gfz_format_steps, without writing the text."""
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    flags = 0
    if isinstance(nodes, DeviceArray):
        flags |= GFZ_INPUT_DEVICE
        nptr, nn = nodes.ptr, nodes.n
    else:
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        nptr, nn = nodes.ctypes.data, nodes.size
    hb, ho = _concat(heads)
    tb, tof = _concat(tails)
    hbuf = ctypes.create_string_buffer(hb, max(1, len(hb)))
    tbuf = ctypes.create_string_buffer(tb, max(1, len(tb)))
    n = _u64()
    line_off = np.zeros(lengths.size + 1, dtype=np.uint64)
    _check(lib().gfz_format_steps(ctx.handle, int(kind), nptr, nn, lengths.ctypes.data,
                                  lengths.size, hbuf, ho.ctypes.data, tbuf, tof.ctypes.data,
                                  flags, None, 0, ctypes.byref(n), line_off.ctypes.data))
    return int(n.value), line_off


def parse_steps(ctx: Context, kind: int, text, line_off=None, out=None, fields=None):
    """Step fields -> (nodes, lengths) (gfz_parse_steps). The fields are either
    text[line_off[l]:line_off[l + 1]] (contiguous) or, with ``fields=(begin, end)``,
    text[begin[l]:end[l]]. text: bytes or a DeviceArray of uint8; with ``out`` (a
    DeviceArray of int32) the nodes stay on the device and (n_nodes, lengths) is
    returned."""
    if fields is None:
        line_off = np.ascontiguousarray(line_off, dtype=np.uint64)
        fb = np.ascontiguousarray(line_off[:-1])
        fe = np.ascontiguousarray(line_off[1:])
    else:
        fb = np.ascontiguousarray(fields[0], dtype=np.uint64)
        fe = np.ascontiguousarray(fields[1], dtype=np.uint64)
    n_lines = fb.size
    fbp = fb.ctypes.data if n_lines else None
    fep = fe.ctypes.data if n_lines else None
    flags = 0
    if isinstance(text, DeviceArray):
        flags |= GFZ_INPUT_DEVICE
        tptr, tn = text.ptr, text.n
    else:
        tbuf = np.frombuffer(bytes(text), dtype=np.uint8) if len(text) else np.zeros(1, np.uint8)
        tptr, tn = tbuf.ctypes.data, len(text)
    lengths = np.zeros(max(1, n_lines), dtype=np.uint32)
    n = _u64()
    L = lib()
    if out is not None:
        _check(L.gfz_parse_steps(ctx.handle, int(kind), tptr, tn, fbp, fep, n_lines,
                                 flags | GFZ_OUTPUT_DEVICE, out.ptr, out.n, lengths.ctypes.data,
                                 ctypes.byref(n)))
        return int(n.value), lengths[:n_lines]
    _check(L.gfz_parse_steps(ctx.handle, int(kind), tptr, tn, fbp, fep, n_lines, flags, None,
                             0, lengths.ctypes.data, ctypes.byref(n)))
    nodes = np.empty(max(1, int(n.value)), dtype=np.int32)
    _check(L.gfz_parse_steps(ctx.handle, int(kind), tptr, tn, fbp, fep, n_lines, flags,
                             nodes.ctypes.data, nodes.size, lengths.ctypes.data,
                             ctypes.byref(n)))
    return nodes[: int(n.value)], lengths[:n_lines]


# ---- extraction front-end: catalog + lines (gfz_catalog_*, gfz_extract_*_lines) ----

def _as_bytes(v) -> bytes:
    return v if isinstance(v, bytes) else str(v).encode()


class _StringList:
    """A list of strings as gfz_strings (keeps the buffers alive)."""

    def __init__(self, items):
        bs = [_as_bytes(v) for v in items]
        self.blob = b"".join(bs)
        off = np.zeros(len(bs) + 1, dtype=np.uint64)
        if bs:
            off[1:] = np.cumsum([len(b) for b in bs])
        self.off = off
        self.s = _Strings(self.blob, off.ctypes.data_as(ctypes.POINTER(_u64)), len(bs))


class _WalkKeyList:
    def __init__(self, keys):
        keys = list(keys)
        self.samples = _StringList([k[0] for k in keys])
        self.seqs = _StringList([k[2] for k in keys])
        self.hap = np.array([int(k[1]) for k in keys] or [0], dtype=np.uint32)
        self.start = np.array([int(k[3]) for k in keys] or [0], dtype=np.int64)
        self.end = np.array([int(k[4]) for k in keys] or [0], dtype=np.int64)
        self.s = _WalkKeys(self.samples.s, self.hap.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                           self.seqs.s, self.start.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                           self.end.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        self.n = len(keys)


def _take_lines(l: _Lines) -> list[bytes]:
    try:
        text = ctypes.string_at(l.text, l.n_bytes) if l.n_bytes else b""
        off = [int(l.line_off[i]) for i in range(l.n_lines + 1)] if l.n_lines else [0]
        return [text[off[i]:off[i + 1]] for i in range(l.n_lines)]
    finally:
        lib().gfz_lines_free(ctypes.byref(l))


class Catalog:
    """Path names and walk identifiers of resident traversals (gfz_catalog_*).
    Walk keys are (sample_id, hap_index, seq_id, seq_start, seq_end) tuples, -1 = '*'."""

    def __init__(self):
        h = _vp()
        _check(lib().gfz_catalog_create(ctypes.byref(h)))
        self.handle = h.value

    def close(self) -> None:
        if self.handle:
            lib().gfz_catalog_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_paths(self, names, first_trav_id: int, overlaps=None) -> None:
        n = _StringList(names)
        o = _StringList(overlaps) if overlaps is not None else None
        _check(lib().gfz_catalog_add_paths(self.handle, ctypes.byref(n.s),
                                           ctypes.byref(o.s) if o else None, int(first_trav_id)))

    def add_walks(self, keys, first_trav_id: int) -> None:
        k = _WalkKeyList(keys)
        _check(lib().gfz_catalog_add_walks(self.handle, ctypes.byref(k.s), int(first_trav_id)))

    def info(self) -> tuple[int, int]:
        a, b = _u64(), _u64()
        _check(lib().gfz_catalog_info(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return int(a.value), int(b.value)

    def find_paths(self, names) -> np.ndarray:
        n = _StringList(names)
        out = np.zeros(max(1, n.s.n), dtype=np.uint64)
        _check(lib().gfz_catalog_find_paths(self.handle, ctypes.byref(n.s), out.ctypes.data))
        return out[: n.s.n]

    def find_walks(self, keys) -> np.ndarray:
        k = _WalkKeyList(keys)
        out = np.zeros(max(1, k.n), dtype=np.uint64)
        _check(lib().gfz_catalog_find_walks(self.handle, ctypes.byref(k.s), out.ctypes.data))
        return out[: k.n]

    def find_walks_by_name(self, names) -> np.ndarray:
        n = _StringList(names)
        out = np.zeros(max(1, n.s.n), dtype=np.uint64)
        _check(lib().gfz_catalog_find_walks_by_name(self.handle, ctypes.byref(n.s),
                                                    out.ctypes.data))
        return out[: n.s.n]


def extract_path_lines(ctx: Context, res: "Resident", cat: Catalog, names) -> list[bytes]:
    """extract_path_lines_by_name on a resident store: P lines, one per name."""
    n = _StringList(names)
    out = _Lines()
    _check(lib().gfz_extract_path_lines(ctx.handle, res.handle, cat.handle, ctypes.byref(n.s),
                                        ctypes.byref(out)))
    return _take_lines(out)


def extract_walk_lines(ctx: Context, res: "Resident", cat: Catalog, keys) -> list[bytes]:
    """extract_walk_lines: W lines for (sample, hap, seq, start, end) keys."""
    k = _WalkKeyList(keys)
    out = _Lines()
    _check(lib().gfz_extract_walk_lines(ctx.handle, res.handle, cat.handle, ctypes.byref(k.s),
                                        ctypes.byref(out)))
    return _take_lines(out)


def extract_walk_lines_by_name(ctx: Context, res: "Resident", cat: Catalog, names) -> list[bytes]:
    """extract_walk_lines_by_name: W lines for sample IDs."""
    n = _StringList(names)
    out = _Lines()
    _check(lib().gfz_extract_walk_lines_by_name(ctx.handle, res.handle, cat.handle,
                                                ctypes.byref(n.s), ctypes.byref(out)))
    return _take_lines(out)
