"""Python mirror of the serving ABI (gfz_server_*): a rulebook and encoded traversals
resident in HBM, extract and add-haplotypes requests batched continuously.

    srv = Server(rules_first, rules_second, start_id, encoded, encoded_lengths,
                 delta_round=1, hbm_budget_bytes=1 << 30)
    t = srv.submit_extract([4, 17])
    r = srv.wait(t)          # {"values", "offsets", "latency_ms", ...}
    t = srv.submit_add(nodes, lengths)
    r = srv.wait(t)          # {"first_new_id", "encoded_lengths", "latency_ms"}
    srv.stats()              # p50/p99 latency, GB/s, batches
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import codec
from .codec import _check, _u64, _vp, lib

_dbl = ctypes.c_double


class _Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_budget_bytes", _u64),
                ("max_batch_requests", ctypes.c_uint32), ("max_queue", ctypes.c_uint32),
                ("wait_timeout_ms", ctypes.c_uint32)]


class _Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("is_add", ctypes.c_int), ("latency_ms", _dbl),
                ("values", ctypes.POINTER(ctypes.c_int32)), ("n_values", _u64),
                ("offsets", ctypes.POINTER(_u64)), ("n_items", _u64), ("first_new_id", _u64),
                ("encoded_lengths", ctypes.POINTER(ctypes.c_uint32))]


class _Stats(ctypes.Structure):
    _fields_ = [("completed", _u64), ("rejected", _u64), ("batches", _u64), ("bytes_out", _u64),
                ("n_traversals", _u64), ("p50_ms", _dbl), ("p99_ms", _dbl), ("mean_ms", _dbl),
                ("elapsed_s", _dbl), ("gbps", _dbl)]


_SIGS = {
    "gfz_server_create": ([_vp, _vp, _vp, _u64, ctypes.c_uint32, _vp, _u64, _vp, _u64, _vp, _u64,
                           ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "gfz_server_destroy": ([_vp], None),
    "gfz_submit_extract": ([_vp, _vp, _u64, ctypes.POINTER(_u64)], ctypes.c_int),
    "gfz_submit_add": ([_vp, _vp, _u64, _vp, _u64, ctypes.POINTER(_u64)], ctypes.c_int),
    "gfz_poll": ([_vp, _u64, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "gfz_wait": ([_vp, _u64, ctypes.POINTER(_Result)], ctypes.c_int),
    "gfz_result_free": ([ctypes.POINTER(_Result)], None),
    "gfz_server_stats": ([_vp, ctypes.POINTER(_Stats)], ctypes.c_int),
    "gfz_server_reset_stats": ([_vp], ctypes.c_int),
}


def _bind():
    L = lib()
    if not getattr(L, "_gfz_serving_bound", False):
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        L._gfz_serving_bound = True
    return L


class Server:
    def __init__(self, rules_first, rules_second, start_id: int, encoded, encoded_lengths,
                 delta_round: int = 1, hbm_budget_bytes: int = 1 << 30, device: int = 0,
                 layers=None, max_batch_requests: int = 0, max_queue: int = 0,
                 wait_timeout_ms: int = 0):
        L = _bind()
        rf = np.ascontiguousarray(rules_first, dtype=np.int32)
        rs = np.ascontiguousarray(rules_second, dtype=np.int32)
        enc = np.ascontiguousarray(encoded, dtype=np.int32)
        el = np.ascontiguousarray(encoded_lengths, dtype=np.uint32)
        if layers is None:
            layers = [(int(start_id), int(start_id) + int(rf.size), 0)] if rf.size else []
        arr = (codec._Layer * max(1, len(layers)))(*[codec._Layer(int(a), int(b), int(c))
                                                    for a, b, c in layers])
        cfg = _Config(device, int(hbm_budget_bytes), int(max_batch_requests), int(max_queue),
                      int(wait_timeout_ms))
        h = _vp()
        _check(L.gfz_server_create(ctypes.byref(cfg), rf.ctypes.data, rs.ctypes.data, rf.size,
                                   int(start_id), arr, len(layers), enc.ctypes.data, enc.size,
                                   el.ctypes.data, el.size, int(delta_round), ctypes.byref(h)))
        self.handle = h.value

    def submit_extract(self, ids) -> int:
        ids = np.ascontiguousarray(ids, dtype=np.uint64)
        t = _u64()
        _check(_bind().gfz_submit_extract(self.handle, ids.ctypes.data, ids.size, ctypes.byref(t)))
        return int(t.value)

    def submit_add(self, nodes, lengths) -> int:
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
        t = _u64()
        _check(_bind().gfz_submit_add(self.handle, nodes.ctypes.data, nodes.size,
                                      lengths.ctypes.data, lengths.size, ctypes.byref(t)))
        return int(t.value)

    def poll(self, ticket: int) -> bool:
        d = ctypes.c_int()
        _check(_bind().gfz_poll(self.handle, ticket, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int) -> dict:
        L = _bind()
        r = _Result()
        rc = L.gfz_wait(self.handle, ticket, ctypes.byref(r))
        try:
            _check(rc)
            out = {"latency_ms": r.latency_ms, "is_add": bool(r.is_add)}
            if r.is_add:
                out["first_new_id"] = int(r.first_new_id)
                out["encoded_lengths"] = (np.ctypeslib.as_array(r.encoded_lengths,
                                                                shape=(r.n_items,)).copy()
                                          if r.n_items else np.zeros(0, np.uint32))
            else:
                out["values"] = (np.ctypeslib.as_array(r.values, shape=(r.n_values,)).copy()
                                 if r.n_values else np.zeros(0, np.int32))
                out["offsets"] = np.ctypeslib.as_array(r.offsets, shape=(r.n_items + 1,)).copy()
            return out
        finally:
            L.gfz_result_free(ctypes.byref(r))

    def stats(self) -> dict:
        s = _Stats()
        _check(_bind().gfz_server_stats(self.handle, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _Stats._fields_}

    def reset_stats(self) -> None:
        _check(_bind().gfz_server_reset_stats(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            _bind().gfz_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
