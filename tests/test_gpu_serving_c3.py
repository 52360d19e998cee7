"""The serving loop on the batching-form graph held resident: walks_2m (2,000
walks of 1k-2M steps, about 550 M steps) compressed on the GPU, its rulebook and
encoded walks loaded into a server, then a fixed mix of 160 requests - 80 %
extracts of 1-3 random walks, 20 % adds of 1-4 new walks - run one at a time and
all submitted at once. Every extract equals the input walks, every result is the
same alone and in the mix, and added walks extract back as added. (The grammar of
this graph is checked against the oracle in test_gpu_large.py; the add encoding
against the oracle's in test_gpu_serving.py.)"""
import sys

import numpy as np
import pytest

from gfaz_amd import codec, serving, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_serving_on_resident_walks_2m(gpu_ctx):
    cfg = synth.CONFIGS["walks_2m"]
    tr = synth.generate(cfg)
    say(f"serving C3: {tr.lengths.size} walks, {tr.n_steps} steps")
    got = codec.compress_traversals(gpu_ctx, tr.nodes, tr.lengths, tr.n_paths, 8, 2, 1)
    offs = np.concatenate([[0], np.cumsum(tr.lengths, dtype=np.int64)])
    walk = lambda i: tr.nodes[offs[i]:offs[i + 1]]
    rng = np.random.default_rng(23)
    small = cfg.scaled(backbone=50_000, n_haps=4, len_max=40_000)
    reqs = []
    for k in range(160):
        if rng.random() < 0.2:
            new = synth.generate(small.scaled(n_haps=int(rng.integers(1, 5)), seed=cfg.seed + 77 + k))
            reqs.append(("add", new.nodes, new.lengths))
        else:
            reqs.append(("extract", rng.integers(0, tr.lengths.size, int(rng.integers(1, 4))),
                         None))

    def run(mixed):
        with serving.Server(got.rules_first, got.rules_second, got.start_id, got.encoded,
                            got.encoded_lengths, 1, hbm_budget_bytes=8 << 30,
                            wait_timeout_ms=60000) as srv:
            if mixed:
                out = [srv.wait(t) for t in
                       [srv.submit_extract(a) if kind == "extract" else srv.submit_add(a, b)
                        for kind, a, b in reqs]]
            else:
                out = [srv.wait(srv.submit_extract(a) if kind == "extract" else
                                srv.submit_add(a, b)) for kind, a, b in reqs]
            back = []
            if mixed:  # added walks extract back as added
                for r, (kind, nodes, lengths) in zip(out, reqs):
                    if kind == "add":
                        ids = np.arange(r["first_new_id"], r["first_new_id"] + lengths.size)
                        back.append((srv.wait(srv.submit_extract(ids))["values"], nodes))
            return out, back, srv.stats()

    say("serving C3 alone")
    alone, _, st_alone = run(False)
    say("serving C3 mixed")
    mixed, back, st = run(True)
    say(f"serving C3 stats alone={st_alone} mixed={st}")
    assert st["completed"] == len(reqs) + len(back) and st["rejected"] == 0
    assert st["batches"] <= st_alone["batches"]
    for a, m, (kind, ids, lengths) in zip(alone, mixed, reqs):
        if kind == "extract":
            np.testing.assert_array_equal(a["values"], m["values"])
            np.testing.assert_array_equal(a["offsets"], m["offsets"])
            np.testing.assert_array_equal(m["values"], np.concatenate([walk(i) for i in ids]))
        else:
            assert a["first_new_id"] == m["first_new_id"]
            np.testing.assert_array_equal(a["encoded_lengths"], m["encoded_lengths"])
    assert back
    for values, nodes in back:
        np.testing.assert_array_equal(values, nodes)
