"""Growth and PAV with the GPU coverage counts (gfz_coverage_*): the reference's results
(tests/golden/analytics.json) for every grouping mode, and a 64-haplotype pangenome of
39.6 M steps against a numpy count."""
import os
import sys

import numpy as np
import pytest

import goldens
from analytics_cases import case
from gfaz_amd import analytics as an
from gfaz_amd import codec, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("label", ["cli_fixture", "pansn_3k"])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_growth_and_pav_match_reference(gpu_ctx, label, mode):
    say(f"analytics {label} mode={mode}")
    c = case(label)
    hist, growth, n = an.growth(gpu_ctx, c["nodes"], c["lengths"], c["seg"].size, c["names"],
                                [w[:3] for w in c["walks"]], mode)
    ref = c["ref"]["growth"][str(mode)]
    assert n == ref["num_haplotypes"]
    assert hist.tolist() == ref["hist"]
    assert growth.tolist() == ref["growth"]
    d_nodes = codec.DeviceArray.from_host(gpu_ctx, c["nodes"])
    for bed in c["beds"]:
        ranges = an.read_bed(os.path.join(goldens.GOLDEN_DIR, bed))
        names, den, num = an.pav(gpu_ctx, d_nodes, c["lengths"], c["seg"], ranges, c["names"],
                                 c["walks"], mode)
        r = c["ref"]["pav"][f"{bed}/{mode}"]
        assert names == r["groups"]
        assert den.tolist() == r["denominators"]
        assert num.reshape(-1).tolist() == r["numerators"]


def test_large_pangenome_coverage(gpu_ctx):
    cfg = synth.CONFIGS["plain_500k"]
    tr = synth.generate(cfg)
    num_nodes = synth.num_segments(cfg)
    say(f"analytics plain_500k ({tr.n_steps} steps, {num_nodes} nodes)")
    group_of = (np.arange(tr.lengths.size) // 8).astype(np.uint32)   # 8 samples
    d_nodes = codec.DeviceArray.from_host(gpu_ctx, tr.nodes)
    with an.Coverage(gpu_ctx, d_nodes, tr.lengths, group_of, 8, num_nodes) as cov:
        hist = cov.hist()
        offs = np.concatenate([[0], np.cumsum(tr.lengths, dtype=np.int64)])
        members = []
        cnt = np.zeros(num_nodes + 1, np.int64)
        for g in range(8):
            a = np.abs(tr.nodes[offs[8 * g]:offs[8 * g + 8]].astype(np.int64))
            u = np.unique(a[(a >= 1) & (a <= num_nodes)])
            members.append(u)
            cnt[u] += 1
        want = np.bincount(cnt[1:], minlength=9)
        np.testing.assert_array_equal(hist, want)
        # windows: 200 random node sets
        rng = np.random.default_rng(3)
        seg = rng.integers(1, 65, num_nodes).astype(np.uint32)
        wins = [np.unique(rng.integers(1, num_nodes + 1, int(rng.integers(0, 5000))))
                for _ in range(200)]
        woff = np.concatenate([[0], np.cumsum([w.size for w in wins])]).astype(np.uint64)
        den, num = cov.windows(woff, np.concatenate(wins).astype(np.uint32), seg)
        for w, win in enumerate(wins):
            assert int(den[w]) == int(seg[win - 1].sum(dtype=np.int64))
            for g in range(8):
                sel = win[np.isin(win, members[g], assume_unique=True)]
                assert int(num[w, g]) == int(seg[sel - 1].sum(dtype=np.int64))
