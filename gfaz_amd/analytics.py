"""Pangenome growth and presence/absence (PAV) on decoded traversals: the Python mirror
of compute_growth (src/workflows/growth_workflow.cpp:261-503) and compute_pav
(src/workflows/pav_workflow.cpp:475-760).

The host steps restate the reference's: traversal grouping keys (PanSN names of P-lines,
sample/hap/seq of W-lines, per mode), the BED reader, and the window walk along each
reference traversal (gfz_pav_window_nodes, in the C library). The per-node coverage
counts and the per-window, per-group sums run on the GPU (gfz_coverage_*); the growth
curve is the reference's closed form (gfz_growth_curve).

Grouping modes (include/workflows/growth_workflow.hpp:12-28): 0 PerPathWalk,
1 SampleHapSeq, 2 SampleHap, 3 Sample.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import codec
from .codec import GFZ_INPUT_DEVICE, DeviceArray, _check, _u64, _vp, lib

PER_PATH_WALK, SAMPLE_HAP_SEQ, SAMPLE_HAP, SAMPLE = 0, 1, 2, 3

_SIGS = {
    "gfz_coverage_build": ([_vp, _vp, _u64, _vp, _u64, _vp, ctypes.c_uint32, ctypes.c_uint32,
                            ctypes.c_uint32, ctypes.POINTER(_vp)], ctypes.c_int),
    "gfz_coverage_free": ([_vp], None),
    "gfz_coverage_hist": ([_vp, _vp], ctypes.c_int),
    "gfz_coverage_windows": ([_vp, _vp, _vp, _u64, _vp, _vp, _vp], ctypes.c_int),
    "gfz_growth_curve": ([_vp, ctypes.c_uint32, _vp], ctypes.c_int),
    "gfz_pav_window_nodes": ([_vp, _u64, _vp, ctypes.c_uint32, _vp, _vp, _u64, _vp, _vp, _u64],
                             ctypes.c_int),
}


def _bind():
    L = lib()
    if not getattr(L, "_gfz_analytics_bound", False):
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        L._gfz_analytics_bound = True
    return L


# ---- grouping keys ---------------------------------------------------------------

def _strip_coords(s: str) -> str:
    """strip_pansn_coords_inplace: drop a trailing ':<digits>-<digits>'."""
    colon = s.rfind(":")
    if colon <= 0:
        return s
    dash = s.find("-", colon + 1)
    if dash < 0 or dash == colon + 1 or dash + 1 == len(s):
        return s
    if not (s[colon + 1:dash].isdigit() and s[dash + 1:].isdigit()):
        return s
    if not all(c.isascii() for c in s[colon + 1:]):
        return s
    return s[:colon]


def _growth_pansn(name: str):
    """growth_workflow.cpp parse_pansn_path_name -> (sample, hap, seq, has_hap, has_seq)."""
    h1 = name.find("#")
    if h1 < 0:
        return _strip_coords(name), "", "", False, False
    sample = name[:h1]
    h2 = name.find("#", h1 + 1)

    def two(hap_raw):
        if not hap_raw:
            return sample, "", "", False, False
        return sample, _strip_coords(hap_raw), "", True, False

    if h2 < 0:
        return two(name[h1 + 1:])
    if h2 == h1 + 1:
        return sample, "", "", False, False
    hap_raw = name[h1 + 1:h2]
    seq_raw = name[h2 + 1:]
    if not seq_raw or seq_raw[0] == "#":
        return two(hap_raw)
    return sample, hap_raw, _strip_coords(seq_raw), True, True


def growth_keys(path_names, walks, mode: int) -> list[str] | None:
    """Group keys of compute_growth (growth_workflow.cpp path_group_key / walk_group_key),
    paths first then walks (walks: (sample, hap, seq) per W-line); None for PerPathWalk,
    where every traversal is its own group."""
    if mode == PER_PATH_WALK:
        return None
    keys = []
    for n in path_names:
        sample, hap, seq, has_hap, has_seq = _growth_pansn(n)
        if mode == SAMPLE:
            keys.append(sample)
        elif mode == SAMPLE_HAP:
            keys.append(sample + "#" + (hap if has_hap else ""))
        elif has_hap:
            keys.append(f"{sample}#{hap}#{seq}" if has_seq else f"{sample}#{hap}")
        elif has_seq:
            keys.append(f"{sample}#*#{seq}")
        else:
            keys.append(sample)
    for sample, hap, seq in walks:
        keys.append(sample if mode == SAMPLE else f"{sample}#{hap}" if mode == SAMPLE_HAP
                    else f"{sample}#{hap}#{seq}")
    return keys


def _pav_pansn(name: str):
    """pav_workflow.cpp parse_pansn_path_name."""
    h1 = name.find("#")
    if h1 < 0:
        return _strip_coords(name), "", "", False, False
    sample = name[:h1]
    h2 = name.find("#", h1 + 1)
    if h2 < 0:
        hap = name[h1 + 1:]
        return sample, _strip_coords(hap), "", bool(hap), False
    hap, seq = name[h1 + 1:h2], name[h2 + 1:]
    return sample, hap, _strip_coords(seq), bool(hap), bool(seq)


def pav_keys(path_names, walks, mode: int) -> list[str]:
    """Group keys of compute_pav (pav_workflow.cpp path_group_key / walk_group_key)."""
    keys = []
    for n in path_names:
        sample, hap, seq, has_hap, has_seq = _pav_pansn(n)
        if mode == SAMPLE:
            keys.append(sample)
        elif mode == SAMPLE_HAP:
            keys.append(f"{sample}#{hap}" if has_hap else sample)
        elif mode == SAMPLE_HAP_SEQ:
            keys.append(f"{sample}#{hap}#{seq}" if has_hap and has_seq
                        else f"{sample}#{hap}" if has_hap else sample)
        else:
            keys.append(n)
    for sample, hap, seq, *_ in walks:
        keys.append(sample if mode == SAMPLE else f"{sample}#{hap}" if mode == SAMPLE_HAP
                    else f"{sample}#{hap}#{seq}")
    return keys


def group_ids(keys: list[str]):
    """First-seen order of distinct keys: (group_of per traversal, group names)."""
    ids, names = {}, []
    out = np.empty(len(keys), dtype=np.uint32)
    for i, k in enumerate(keys):
        if k not in ids:
            ids[k] = len(names)
            names.append(k)
        out[i] = ids[k]
    return out, names


def walk_reference_name(sample: str, hap: int, seq: str, start: int, end: int) -> str:
    """pav_workflow.cpp walk_reference_name."""
    name = f"{sample}#{hap}#{seq}"
    if start != -1 or end != -1:
        name += f":{start}-{end}"
    return name


def read_bed(path: str):
    """pav_workflow.cpp read_bed: (chrom, start, end, name) per non-comment line."""
    ranges = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line[0] == "#":
                continue
            parts = line.split()
            if len(parts) < 3:
                raise RuntimeError("pav: malformed BED line: " + line)
            chrom, start, end = parts[0], int(parts[1]), int(parts[2])
            name = parts[3] if len(parts) > 3 else f"{chrom}:{start}-{end}"
            if end < start:
                raise RuntimeError("pav: BED end before start: " + line)
            ranges.append((chrom, start, end, name))
    if not ranges:
        raise RuntimeError("pav: BED file contains no ranges")
    return ranges


# ---- GPU coverage ------------------------------------------------------------------

class Coverage:
    """Group membership of nodes 1..num_nodes (gfz_coverage_build) for decoded traversals
    (host array or DeviceArray), one group id per traversal."""

    def __init__(self, ctx: codec.Context, nodes, lengths, group_of, n_groups: int,
                 num_nodes: int):
        L = _bind()
        self.ctx = ctx
        lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
        group_of = np.ascontiguousarray(group_of, dtype=np.uint32)
        flags = 0
        if isinstance(nodes, DeviceArray):
            flags |= GFZ_INPUT_DEVICE
            nptr, nn = nodes.ptr, nodes.n
        else:
            nodes = np.ascontiguousarray(nodes, dtype=np.int32)
            nptr, nn = nodes.ctypes.data, nodes.size
        h = _vp()
        _check(L.gfz_coverage_build(ctx.handle, nptr, nn, lengths.ctypes.data, lengths.size,
                                    group_of.ctypes.data, int(n_groups), int(num_nodes), flags,
                                    ctypes.byref(h)))
        self.handle = h.value
        self.n_groups = int(n_groups)

    def hist(self) -> np.ndarray:
        out = np.zeros(self.n_groups + 1, dtype=np.uint64)
        _check(_bind().gfz_coverage_hist(self.handle, out.ctypes.data))
        return out

    def windows(self, window_off, window_nodes, segment_lengths):
        window_off = np.ascontiguousarray(window_off, dtype=np.uint64)
        window_nodes = np.ascontiguousarray(window_nodes, dtype=np.uint32)
        seg = np.ascontiguousarray(segment_lengths, dtype=np.uint32)
        nw = window_off.size - 1
        den = np.zeros(max(1, nw), dtype=np.uint64)
        num = np.zeros(max(1, nw * self.n_groups), dtype=np.uint64)
        _check(_bind().gfz_coverage_windows(self.handle, window_off.ctypes.data,
                                            window_nodes.ctypes.data if window_nodes.size
                                            else None, nw, seg.ctypes.data if seg.size
                                            else None, den.ctypes.data, num.ctypes.data))
        return den[:nw], num[:nw * self.n_groups].reshape(nw, self.n_groups)

    def close(self):
        if getattr(self, "handle", None):
            _bind().gfz_coverage_free(self.handle)
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


def growth_curve(hist) -> np.ndarray:
    hist = np.ascontiguousarray(hist, dtype=np.uint64)
    n = hist.size - 1
    out = np.zeros(n + 1, dtype=np.float64)
    _check(_bind().gfz_growth_curve(hist.ctypes.data, n, out.ctypes.data))
    return out


def window_nodes(ref_nodes, segment_lengths, starts, ends):
    """gfz_pav_window_nodes: (window_off, window_nodes) of ranges on one reference
    traversal."""
    L = _bind()
    ref = np.ascontiguousarray(ref_nodes, dtype=np.int32)
    seg = np.ascontiguousarray(segment_lengths, dtype=np.uint32)
    st = np.ascontiguousarray(starts, dtype=np.uint64)
    en = np.ascontiguousarray(ends, dtype=np.uint64)
    off = np.zeros(st.size + 1, dtype=np.uint64)
    args = (ref.ctypes.data if ref.size else None, ref.size,
            seg.ctypes.data if seg.size else None, seg.size,
            st.ctypes.data if st.size else None, en.ctypes.data if en.size else None, st.size,
            off.ctypes.data)
    _check(L.gfz_pav_window_nodes(*args, None, 0))
    nodes = np.zeros(max(1, int(off[-1])), dtype=np.uint32)
    _check(L.gfz_pav_window_nodes(*args, nodes.ctypes.data, nodes.size))
    return off, nodes[: int(off[-1])]


def growth(ctx, nodes, lengths, num_nodes: int, path_names=(), walks=(), mode=PER_PATH_WALK):
    """compute_growth on decoded traversals (paths first, then walks): (hist, growth,
    number of groups)."""
    keys = growth_keys(list(path_names), list(walks), mode)
    n = np.asarray(lengths).size
    if keys is None:
        group_of, n_groups = np.arange(n, dtype=np.uint32), n
    else:
        group_of, names = group_ids(keys)
        n_groups = len(names)
    if n_groups == 0:
        return np.zeros(1, np.uint64), np.zeros(1), 0
    with Coverage(ctx, nodes, lengths, group_of, n_groups, num_nodes) as cov:
        hist = cov.hist()
    return hist, growth_curve(hist), n_groups


def pav(ctx, nodes, lengths, segment_lengths, ranges, path_names=(), walks=(),
        mode=PER_PATH_WALK):
    """compute_pav on decoded traversals (paths first, then walks; walks as (sample, hap,
    seq, start, end)): (group names, denominators, numerators[window][group])."""
    path_names, walks = list(path_names), list(walks)
    keys = pav_keys(path_names, walks, mode)
    group_of, names = group_ids(keys)
    if not names:
        return names, np.zeros(len(ranges), np.uint64), np.zeros((len(ranges), 0), np.uint64)
    slice_of = {}
    for i, n in enumerate(path_names):
        slice_of.setdefault(n, i)
    for i, w in enumerate(walks):
        slice_of.setdefault(walk_reference_name(*w), len(path_names) + i)
    offs = np.concatenate([[0], np.cumsum(lengths, dtype=np.int64)])
    host_nodes = nodes.to_host() if isinstance(nodes, DeviceArray) else np.asarray(nodes)
    seg = np.ascontiguousarray(segment_lengths, dtype=np.uint32)
    by_chrom = {}
    for r, (chrom, *_rest) in enumerate(ranges):
        by_chrom.setdefault(chrom, []).append(r)
    win_nodes = [None] * len(ranges)
    for chrom, rids in by_chrom.items():
        if chrom not in slice_of:
            raise RuntimeError("pav: BED reference path not found: " + chrom)
        s = slice_of[chrom]
        off, wn = window_nodes(host_nodes[offs[s]:offs[s + 1]], seg,
                               [ranges[r][1] for r in rids], [ranges[r][2] for r in rids])
        for j, r in enumerate(rids):
            win_nodes[r] = wn[off[j]:off[j + 1]]
    woff = np.concatenate([[0], np.cumsum([w.size for w in win_nodes], dtype=np.uint64)])
    wn_all = np.concatenate(win_nodes) if win_nodes else np.zeros(0, np.uint32)
    with Coverage(ctx, nodes, lengths, group_of, len(names), seg.size) as cov:
        den, num = cov.windows(woff, wn_all, seg)
    return names, den, num
