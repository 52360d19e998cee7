"""TEST INFRASTRUCTURE ONLY - brute-force restatement of the coverage counts behind
compute_growth (src/workflows/growth_workflow.cpp:261-503) and compute_pav
(src/workflows/pav_workflow.cpp:475-760), from decoded traversals and a group id per
traversal. Pinned by the reference's outputs (tests/golden/analytics.json, written
through the compiled reference by tests/golden/make_goldens.py)."""
from __future__ import annotations

import numpy as np


def group_sets(nodes, lengths, group_of, n_groups, num_nodes):
    offs = np.concatenate([[0], np.cumsum(lengths, dtype=np.int64)])
    sets = [set() for _ in range(n_groups)]
    a = np.abs(np.asarray(nodes, dtype=np.int64))
    for t in range(len(lengths)):
        seg = a[offs[t]:offs[t + 1]]
        seg = seg[(seg >= 1) & (seg <= num_nodes)]
        sets[int(group_of[t])].update(np.unique(seg).tolist())
    return sets


def coverage_hist(nodes, lengths, group_of, n_groups, num_nodes):
    cov = np.zeros(num_nodes + 1, dtype=np.int64)
    for s in group_sets(nodes, lengths, group_of, n_groups, num_nodes):
        if s:
            cov[np.fromiter(s, dtype=np.int64)] += 1
    return np.bincount(cov[1:], minlength=n_groups + 1).astype(np.uint64)


def growth_curve(hist):
    """The closed form of growth_workflow.cpp:486-500 (frac_covered_union), in doubles."""
    n = len(hist) - 1

    def frac(N, c, k):
        if c == 0 or k == 0:
            return 0.0
        if c >= N or N - c < k:
            return 1.0
        ratio = 1.0
        for i in range(c):
            ratio *= float(N - k - i) / float(N - i)
        return 1.0 - ratio

    out = [0.0] * (n + 1)
    for k in range(1, n + 1):
        s = 0.0
        for c in range(1, n + 1):
            if hist[c]:
                s += float(hist[c]) * frac(n, c, k)
        out[k] = s
    return np.array(out)


def window_nodes(ref_nodes, segment_lengths, ranges):
    """Distinct nodes of each range [start, end) overlapping it on the reference
    traversal (brute force over every node occurrence)."""
    out = []
    for start, end in ranges:
        sel, off = set(), 0
        for v in ref_nodes:
            a = abs(int(v))
            if a == 0 or a > len(segment_lengths):
                continue
            ln = int(segment_lengths[a - 1])
            if end > off and start < off + ln:
                sel.add(a)
            off += ln
        out.append(sorted(sel))
    return out


def pav_sums(win_nodes, sets, segment_lengths):
    den = np.array([sum(int(segment_lengths[a - 1]) for a in w) for w in win_nodes], np.uint64)
    num = np.array([[sum(int(segment_lengths[a - 1]) for a in w if a in s) for s in sets]
                    for w in win_nodes], np.uint64).reshape(len(win_nodes), len(sets))
    return den, num
