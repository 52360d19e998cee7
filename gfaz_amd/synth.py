"""Seeded synthetic pangenomes (the host generator in csrc/synth.cpp).

Workloads, named by what they are:

* ``plain_20k``   - 20,000-site backbone, 16 P-lines (about 0.31M steps), CPU-runnable.
* ``plain_500k``  - 500,000-site backbone, 64 P-lines plus 0.1% tandem duplications of
  50-500 steps (about 32M steps).
* ``walks_2m``    - 2,000,000-site backbone, 2,000 W-lines of log-uniform length in
  [1e3, 2e6] steps (the batched, uneven-length form).
* ``plain_10m``   - 10,000,000-site backbone, 256 P-lines (about 2.56G steps).

Every site with ``i % 10 == 9`` is a 2-allele bubble; each step is reverse-oriented with
p = 0.02 and a deletion skips 1-5 sites with p = 0.01.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, replace

import numpy as np

from . import _native

SEED_BASE = 0x67666158


class _Params(ctypes.Structure):
    _fields_ = [
        ("backbone", ctypes.c_uint64),
        ("n_haps", ctypes.c_uint32),
        ("walks", ctypes.c_int32),
        ("p_inv", ctypes.c_double),
        ("p_del", ctypes.c_double),
        ("p_dup", ctypes.c_double),
        ("del_max", ctypes.c_uint32),
        ("dup_min", ctypes.c_uint32),
        ("dup_max", ctypes.c_uint32),
        ("len_min", ctypes.c_uint64),
        ("len_max", ctypes.c_uint64),
        ("seed", ctypes.c_uint64),
    ]


@dataclass(frozen=True)
class SynthConfig:
    name: str
    backbone: int
    n_haps: int
    walks: bool = False
    p_inv: float = 0.02
    p_del: float = 0.01
    p_dup: float = 0.0
    del_max: int = 5
    dup_min: int = 50
    dup_max: int = 500
    len_min: int = 1000
    len_max: int = 2_000_000
    seed: int = SEED_BASE + 1

    def scaled(self, backbone: int | None = None, n_haps: int | None = None,
               len_max: int | None = None, seed: int | None = None) -> "SynthConfig":
        kw = {}
        if backbone is not None:
            kw["backbone"] = backbone
        if n_haps is not None:
            kw["n_haps"] = n_haps
        if len_max is not None:
            kw["len_max"] = len_max
        if seed is not None:
            kw["seed"] = seed
        return replace(self, name=f"{self.name}_scaled", **kw)

    def _params(self) -> _Params:
        return _Params(self.backbone, self.n_haps, int(self.walks), self.p_inv, self.p_del,
                       self.p_dup, self.del_max, self.dup_min, self.dup_max, self.len_min,
                       self.len_max, self.seed)


CONFIGS = {
    "plain_20k": SynthConfig("plain_20k", 20_000, 16, seed=SEED_BASE + 1),
    "plain_500k": SynthConfig("plain_500k", 500_000, 64, p_dup=0.001, seed=SEED_BASE + 2),
    "walks_2m": SynthConfig("walks_2m", 2_000_000, 2_000, walks=True, p_dup=0.001,
                            seed=SEED_BASE + 3),
    "plain_10m": SynthConfig("plain_10m", 10_000_000, 256, p_dup=0.001, seed=SEED_BASE + 4),
}


@dataclass
class Traversals:
    """Flat traversals: paths first, then walks (the FlattenedPaths layout)."""

    nodes: np.ndarray     # int32, sign = orientation
    lengths: np.ndarray   # uint32, one per traversal
    n_paths: int

    @property
    def n_steps(self) -> int:
        return int(self.nodes.size)

    def offsets(self) -> np.ndarray:
        off = np.zeros(self.lengths.size + 1, dtype=np.uint64)
        np.cumsum(self.lengths, dtype=np.uint64, out=off[1:])
        return off


def _lib():
    lib = _native.synth_lib()
    if not getattr(lib, "_gfz_typed", False):
        P = ctypes.POINTER(_Params)
        lib.gfz_synth_traversals.argtypes = [
            P, ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)), ctypes.POINTER(ctypes.c_uint64),
            ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32)), ctypes.POINTER(ctypes.c_uint64)]
        lib.gfz_synth_traversals.restype = ctypes.c_int
        lib.gfz_synth_free.argtypes = [ctypes.c_void_p]
        lib.gfz_synth_gfa_bytes.argtypes = [P, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        lib.gfz_synth_gfa_bytes.restype = ctypes.c_int
        lib.gfz_synth_write_gfa.argtypes = [P, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_uint64, ctypes.c_char_p]
        lib.gfz_synth_write_gfa.restype = ctypes.c_int
        lib.gfz_synth_num_segments.argtypes = [P]
        lib.gfz_synth_num_segments.restype = ctypes.c_uint64
        lib._gfz_typed = True
    return lib


def generate(cfg: SynthConfig) -> Traversals:
    lib = _lib()
    p = cfg._params()
    nodes_p = ctypes.POINTER(ctypes.c_int32)()
    lens_p = ctypes.POINTER(ctypes.c_uint32)()
    n_nodes = ctypes.c_uint64()
    n_trav = ctypes.c_uint64()
    rc = lib.gfz_synth_traversals(ctypes.byref(p), ctypes.byref(nodes_p), ctypes.byref(n_nodes),
                                  ctypes.byref(lens_p), ctypes.byref(n_trav))
    if rc != 0:
        raise RuntimeError(f"synthetic generator failed ({rc})")
    try:
        nodes = np.ctypeslib.as_array(nodes_p, shape=(max(1, n_nodes.value),))[: n_nodes.value].copy()
        lengths = np.ctypeslib.as_array(lens_p, shape=(max(1, n_trav.value),))[: n_trav.value].copy()
    finally:
        lib.gfz_synth_free(ctypes.cast(nodes_p, ctypes.c_void_p))
        lib.gfz_synth_free(ctypes.cast(lens_p, ctypes.c_void_p))
    n_paths = 0 if cfg.walks else int(n_trav.value)
    return Traversals(nodes, lengths, n_paths)


def gfa_bytes(cfg: SynthConfig, tr: Traversals) -> int:
    out = ctypes.c_uint64()
    rc = _lib().gfz_synth_gfa_bytes(ctypes.byref(cfg._params()), tr.nodes.ctypes.data,
                                    tr.lengths.ctypes.data, tr.lengths.size, ctypes.byref(out))
    if rc != 0:
        raise RuntimeError("gfa byte count failed")
    return int(out.value)


def write_gfa(cfg: SynthConfig, tr: Traversals, path: str) -> None:
    rc = _lib().gfz_synth_write_gfa(ctypes.byref(cfg._params()), tr.nodes.ctypes.data,
                                    tr.lengths.ctypes.data, tr.lengths.size, path.encode())
    if rc != 0:
        raise RuntimeError(f"writing {path} failed ({rc})")


def num_segments(cfg: SynthConfig) -> int:
    return int(_lib().gfz_synth_num_segments(ctypes.byref(cfg._params())))
