"""The autotuner on the GPU: every pair-counting setting it tries gives the oracle's
grammar (gfz_autotune_compress checks them against each other; the result is checked
against the oracle here), the choice lands in the ctx's tuned table, survives a
save/load through the JSON file, and a file for another device is refused."""
import json
import os
import sys
import tempfile

import numpy as np
import pytest

from gfaz_amd import codec, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_autotune_compress_and_tuning_file():
    cfg = synth.CONFIGS["plain_500k"].scaled(backbone=60_000, n_haps=16)
    tr = synth.generate(cfg)
    say(f"autotune compress ({tr.n_steps} steps)")
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=16)
    cls = 0
    n = tr.n_steps
    while n >= 4:
        n >>= 2
        cls += 1
    with codec.Context(0) as ctx:
        ms = ctx.autotune_compress(tr.nodes, tr.lengths)
        assert ms > 0
        scale = ctx.tuned(f"count_bucket_scale/{cls}")
        slots = ctx.tuned(f"count_agg_slots/{cls}")
        assert scale in (1, 2, 4) and slots in (2048, 4096)
        got = codec.compress_traversals(ctx, tr.nodes, tr.lengths, tr.n_paths, 8, 2, 1)
        np.testing.assert_array_equal(got.encoded, ref["encoded"])
        np.testing.assert_array_equal(got.rules_first, ref["rules_first"])
        np.testing.assert_array_equal(got.rules_second, ref["rules_second"])
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "tuning.json")
            ctx.save_tuning(path)
            with open(path) as f:
                saved = json.load(f)
            assert saved["tuned"][f"count_bucket_scale/{cls}"] == scale
            assert saved["arch"].startswith("gfx950")
            with codec.Context(0) as fresh:
                fresh.load_tuning(path)
                assert fresh.tuned(f"count_bucket_scale/{cls}") == scale
                assert fresh.tuned(f"count_agg_slots/{cls}") == slots
                other = dict(saved, arch="gfx000")
                bad = os.path.join(tmp, "other.json")
                with open(bad, "w") as f:
                    json.dump(other, f)
                fresh.set_tuned("probe", 7)
                with pytest.raises(RuntimeError, match="tuning file is for gfx000"):
                    fresh.load_tuning(bad)
                assert fresh.tuned("probe") == 7
                with open(bad, "w") as f:
                    f.write("{not json")
                with pytest.raises(RuntimeError, match="malformed"):
                    fresh.load_tuning(bad)
