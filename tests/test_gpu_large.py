"""Full-size configurations on the GPU.

* plain_10m (10,000,000-site backbone, 256 P-lines, 3,168,123,299 steps: past 2^31
  elements): device-resident compress, the grammar's size-independent invariants, an
  exact decode into HBM, resident extraction of traversals that end past 2^31 decoded
  positions, and the streamed decode through pinned host buffers.
* walks_2m (2,000 W-lines of log-uniform length, 550 M steps) and a plain_10m-shaped
  graph at 1/10 of its backbone with all 256 P-lines (317 M steps): bit for bit against
  the CPU oracle.
"""
import sys

import numpy as np
import pytest

from gfaz_amd import codec, synth
from invariants import check_grammar
from oracle import oracle

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_plain_10m_round_trip_and_invariants(gpu_ctx):
    say("plain_10m: generate")
    tr = synth.generate(synth.CONFIGS["plain_10m"])
    assert tr.n_steps > 2 ** 31 and tr.lengths.size == 256
    d_nodes = codec.DeviceArray.from_host(gpu_ctx, tr.nodes)
    say("plain_10m: compress")
    comp = codec.DeviceCompressed(gpu_ctx, d_nodes, tr.lengths, tr.n_paths, 8, 2, 1)
    d_nodes.free()
    host = comp.to_host()
    say(f"plain_10m: {host.n_encoded} symbols, {host.n_rules} rules; invariants")
    assert host.rounds_run == 8
    check_grammar(host.encoded, host.rules_first, host.rules_second, host.start_id,
                  host.end_id, host.rules_per_round)
    rb = comp.rulebook()
    say("plain_10m: decompress into HBM")
    out = codec.DeviceArray(gpu_ctx, tr.n_steps, np.int32)
    _, lens, mism = codec.decompress_traversals(gpu_ctx, rb, comp.encoded_view(),
                                                comp.encoded_lengths, tr.lengths, 1, out=out)
    assert not mism
    np.testing.assert_array_equal(lens, tr.lengths)
    got = out.to_host()
    assert np.array_equal(got, tr.nodes)
    del got
    out.free()
    offs = np.concatenate([[0], np.cumsum(tr.lengths, dtype=np.int64)])
    assert offs[200] > 2 ** 31
    say("plain_10m: resident extract")
    res = codec.Resident(gpu_ctx, rb, comp.encoded_view(), comp.encoded_lengths, 1)
    for ids in ([255, 0, 200], [173]):
        vals, o = res.extract(ids)
        want = np.concatenate([tr.nodes[offs[i]:offs[i + 1]] for i in ids])
        assert np.array_equal(vals, want)
    res.close()
    say("plain_10m: streamed decode")
    state = {"chunks": 0, "ok": True, "next": 0}

    def verify(values, b, e, expanded_begin, lengths):
        state["chunks"] += 1
        state["ok"] = bool(state["ok"] and b == state["next"]
                           and int(expanded_begin) == int(offs[b])
                           and np.array_equal(values, tr.nodes[offs[b]:offs[e]]))
        state["next"] = e
        return False

    codec.decompress_stream(gpu_ctx, rb, comp.encoded_view(), comp.encoded_lengths, 1, 1 << 30,
                            verify)
    assert state["ok"] and state["next"] == 256 and state["chunks"] >= 12
    rb.close()
    comp.close()


def _against_oracle(ctx, cfg, label):
    say(f"{label}: generate")
    tr = synth.generate(cfg)
    say(f"{label}: {tr.n_steps} steps; oracle")
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=ORACLE_THREADS)
    say(f"{label}: GPU")
    got = codec.compress_traversals(ctx, tr.nodes, tr.lengths, tr.n_paths, 8, 2, 1)
    assert (got.start_id, got.end_id) == (ref["start_id"], ref["end_id"])
    assert np.array_equal(got.encoded_lengths, ref["encoded_lengths"])
    assert np.array_equal(got.encoded, ref["encoded"])
    assert np.array_equal(got.rules_first, ref["rules_first"])
    assert np.array_equal(got.rules_second, ref["rules_second"])
    return tr, got


def test_full_size_walks_2m_matches_oracle(gpu_ctx):
    tr, got = _against_oracle(gpu_ctx, synth.CONFIGS["walks_2m"], "walks_2m")
    assert tr.lengths.size == 2000 and tr.n_steps > 500_000_000


def test_plain_10m_shape_at_tenth_backbone_matches_oracle(gpu_ctx):
    cfg = synth.CONFIGS["plain_10m"].scaled(backbone=1_000_000)
    tr, got = _against_oracle(gpu_ctx, cfg, "plain_10m/10")
    assert tr.lengths.size == 256 and tr.n_steps > 300_000_000
