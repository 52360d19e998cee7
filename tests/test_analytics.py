"""CPU checks of growth and PAV: the grouping keys, the host steps of the library
(gfz_growth_curve, gfz_pav_window_nodes) and the brute-force restatement
(oracle/analytics_oracle.py) against what the reference computed
(tests/golden/analytics.json)."""
import os

import numpy as np
import pytest

import goldens
from analytics_cases import case, groups
from gfaz_amd import analytics as an
from oracle import analytics_oracle as ao


@pytest.mark.parametrize("label", ["cli_fixture", "pansn_3k"])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_growth_matches_reference(label, mode):
    c = case(label)
    ref = c["ref"]["growth"][str(mode)]
    g, n = groups(c, mode, "growth")
    assert n == ref["num_haplotypes"]
    hist = ao.coverage_hist(c["nodes"], c["lengths"], g, n, c["seg"].size)
    assert hist.tolist() == ref["hist"]
    assert ao.growth_curve(hist).tolist() == ref["growth"]
    assert an.growth_curve(hist).tolist() == ref["growth"]   # the library's host step


@pytest.mark.parametrize("label", ["cli_fixture", "pansn_3k"])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_pav_matches_reference(label, mode):
    c = case(label)
    offs = np.concatenate([[0], np.cumsum(c["lengths"], dtype=np.int64)])
    g, names = groups(c, mode, "pav")
    sets = ao.group_sets(c["nodes"], c["lengths"], g, len(names), c["seg"].size)
    slice_of = {}
    for i, n in enumerate(c["names"]):
        slice_of.setdefault(n, i)
    for i, w in enumerate(c["walks"]):
        slice_of.setdefault(an.walk_reference_name(*w), len(c["names"]) + i)
    for bed in c["beds"]:
        ref = c["ref"]["pav"][f"{bed}/{mode}"]
        assert names == ref["groups"]
        ranges = an.read_bed(os.path.join(goldens.GOLDEN_DIR, bed))
        assert [list(r) for r in ranges] == ref["ranges"]
        wins = []
        for chrom, start, end, _ in ranges:
            s = slice_of[chrom]
            ref_nodes = c["nodes"][offs[s]:offs[s + 1]]
            brute = ao.window_nodes(ref_nodes, c["seg"], [(start, end)])[0]
            off, wn = an.window_nodes(ref_nodes, c["seg"], [start], [end])
            assert wn.tolist() == brute                     # the library's host step
            wins.append(brute)
        den, num = ao.pav_sums(wins, sets, c["seg"])
        assert den.tolist() == ref["denominators"]
        assert num.reshape(-1).tolist() == ref["numerators"]
