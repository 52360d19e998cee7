"""Pair counting layouts on the GPU: every bucket count (one-level up to 4096
buckets, two-level beyond), and the fall back to the HBM counting table when a
bucket overflows (forced with a tiny repeated-key cap, also after a two-level
partition has reused the aggregated tiles' memory), give the oracle's grammar bit
for bit. So does tile aggregation in rounds of few keys with and without the known
repeated pairs dropped first (their keys then enter as an extra tile), also when
the round falls back to the table, and with rule-table regions from barely above
their longest list (long probe chains that wrap within a region) to four times it.
Candidate marking in the cross-traversal tile order (24 traversals of similar
length: the order's grid pads the tiles by under a quarter, so it is taken) and in
the launch order gives the same grammar too.
The knobs are the ctx tuned values "count_buckets", "count_rep_cap",
"count_hot_drop", "rule_region_pct" and "cand_order"."""
import sys

import numpy as np
import pytest

from gfaz_amd import codec, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def case():
    cfg = synth.CONFIGS["plain_500k"].scaled(backbone=120_000, n_haps=24)
    tr = synth.generate(cfg)
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=16)
    return tr, ref


def _check(ctx, tr, ref):
    got = codec.compress_traversals(ctx, tr.nodes, tr.lengths, tr.n_paths, 8, 2, 1)
    assert (got.start_id, got.end_id) == (ref["start_id"], ref["end_id"])
    np.testing.assert_array_equal(got.encoded_lengths, ref["encoded_lengths"])
    np.testing.assert_array_equal(got.encoded, ref["encoded"])
    np.testing.assert_array_equal(got.rules_first, ref["rules_first"])
    np.testing.assert_array_equal(got.rules_second, ref["rules_second"])


@pytest.mark.parametrize("buckets", [64, 1024, 4096, 8192, 32768])
def test_every_bucket_count_gives_the_oracle_grammar(gpu_ctx, case, buckets):
    tr, ref = case
    say(f"count buckets={buckets} ({tr.n_steps} steps)")
    gpu_ctx.set_tuned("count_buckets", buckets)
    try:
        _check(gpu_ctx, tr, ref)
    finally:
        gpu_ctx.set_tuned("count_buckets", 0)


@pytest.mark.parametrize("buckets", [1024, 16384])
def test_bucket_overflow_falls_back_to_the_table(gpu_ctx, case, buckets):
    tr, ref = case
    say(f"count fallback buckets={buckets}")
    gpu_ctx.set_tuned("count_buckets", buckets)
    gpu_ctx.set_tuned("count_rep_cap", 1)
    try:
        _check(gpu_ctx, tr, ref)
    finally:
        gpu_ctx.set_tuned("count_buckets", 0)
        gpu_ctx.set_tuned("count_rep_cap", 0)


def test_walks_two_level(gpu_ctx):
    """Uneven walks (rounds of mostly distinct pairs partition the raw stream)."""
    cfg = synth.CONFIGS["walks_2m"].scaled(n_haps=60, len_max=200_000)
    tr = synth.generate(cfg)
    say(f"count walks two-level ({tr.n_steps} steps)")
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=16)
    gpu_ctx.set_tuned("count_buckets", 16384)
    try:
        _check(gpu_ctx, tr, ref)
    finally:
        gpu_ctx.set_tuned("count_buckets", 0)


@pytest.mark.parametrize("hot_drop", [0, 1, -1])
@pytest.mark.parametrize("rep_cap", [0, 1])
def test_known_hot_drop_gives_the_oracle_grammar(gpu_ctx, case, hot_drop, rep_cap):
    tr, ref = case
    say(f"count hot_drop={hot_drop} rep_cap={rep_cap}")
    gpu_ctx.set_tuned("count_hot_drop", hot_drop)
    gpu_ctx.set_tuned("count_rep_cap", rep_cap)
    try:
        _check(gpu_ctx, tr, ref)
    finally:
        gpu_ctx.set_tuned("count_hot_drop", -1)
        gpu_ctx.set_tuned("count_rep_cap", 0)


@pytest.mark.parametrize("pct", [101, 400])
def test_rule_region_sizes_give_the_oracle_grammar(gpu_ctx, case, pct):
    tr, ref = case
    say(f"rule regions {pct}% of the longest list")
    gpu_ctx.set_tuned("rule_region_pct", pct)
    try:
        _check(gpu_ctx, tr, ref)
    finally:
        gpu_ctx.set_tuned("rule_region_pct", -1)


@pytest.mark.parametrize("order", [0, 1])
def test_candidate_tile_orders_give_the_oracle_grammar(gpu_ctx, case, order):
    tr, ref = case
    say(f"candidate tile order {order}")
    gpu_ctx.set_tuned("cand_order", order)
    try:
        _check(gpu_ctx, tr, ref)
    finally:
        gpu_ctx.set_tuned("cand_order", -1)
