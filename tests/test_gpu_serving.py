"""Serving loop on the GPU: continuous batching of extract and add-haplotypes
requests against a resident rulebook. Every result equals the input traversals,
and is the same when the request runs alone as when it runs in a mix."""
import sys
import threading

import numpy as np
import pytest

from gfaz_amd import serving, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def base():
    cfg = synth.CONFIGS["plain_20k"]
    tr = synth.generate(cfg)
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=8)
    offs = np.concatenate([[0], np.cumsum(tr.lengths, dtype=np.int64)])
    trav = [tr.nodes[offs[i]:offs[i + 1]] for i in range(tr.lengths.size)]
    return cfg, tr, ref, trav


def _requests(cfg, n_base, seed):
    """A fixed mix: 80% extracts of 1-5 random stored traversals, 20% adds of 1-3
    new synthetic haplotypes."""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(40):
        if rng.random() < 0.2:
            new = synth.generate(cfg.scaled(n_haps=int(rng.integers(1, 4)),
                                            seed=cfg.seed + 1000 + k))
            reqs.append(("add", new.nodes, new.lengths))
        else:
            reqs.append(("extract", rng.integers(0, n_base, int(rng.integers(1, 6))), None))
    return reqs


def _server(ref, budget=64 << 20):
    # a request unfinished after 30 s fails its wait, naming its stage, instead of blocking
    return serving.Server(ref["rules_first"], ref["rules_second"], ref["start_id"],
                          ref["encoded"], ref["encoded_lengths"], 1, hbm_budget_bytes=budget,
                          wait_timeout_ms=30000)


def _run(srv, reqs, mixed):
    out = []
    if mixed:
        tickets = [srv.submit_extract(a) if kind == "extract" else srv.submit_add(a, b)
                   for kind, a, b in reqs]
        out = [srv.wait(t) for t in tickets]
    else:
        for kind, a, b in reqs:
            t = srv.submit_extract(a) if kind == "extract" else srv.submit_add(a, b)
            out.append(srv.wait(t))
    return out


def test_serving_alone_equals_mixed_and_input(base):
    cfg, tr, ref, trav = base
    reqs = _requests(cfg, len(trav), 7)
    say("serving alone")
    with _server(ref) as srv:
        alone = _run(srv, reqs, mixed=False)
        st_alone = srv.stats()
    say("serving mixed")
    with _server(ref) as srv:
        mixed = _run(srv, reqs, mixed=True)
        st = srv.stats()
        # added traversals are extractable and decode to what was added
        added = [(r, q) for r, q in zip(mixed, reqs) if q[0] == "add"]
        for r, (_, nodes, lengths) in added:
            ids = np.arange(r["first_new_id"], r["first_new_id"] + lengths.size)
            got = srv.wait(srv.submit_extract(ids))
            np.testing.assert_array_equal(got["values"], nodes)
    say(f"stats alone={st_alone} mixed={st}")
    assert st["completed"] == len(reqs) and st["rejected"] == 0
    assert st["batches"] <= st_alone["batches"]
    assert st["p99_ms"] >= st["p50_ms"] > 0
    for a, m, (kind, ids, lengths) in zip(alone, mixed, reqs):
        if kind == "extract":
            np.testing.assert_array_equal(a["values"], m["values"])
            np.testing.assert_array_equal(a["offsets"], m["offsets"])
            want = np.concatenate([trav[i] for i in ids])
            np.testing.assert_array_equal(m["values"], want)
        else:
            assert a["first_new_id"] == m["first_new_id"]
            np.testing.assert_array_equal(a["encoded_lengths"], m["encoded_lengths"])
            want_enc, want_len = oracle.encode_existing(
                ids, lengths, ref["rules_first"], ref["rules_second"],
                [(ref["start_id"], ref["end_id"], 0)], 1)
            np.testing.assert_array_equal(m["encoded_lengths"], want_len)


def test_serving_concurrent_submitters(base):
    cfg, tr, ref, trav = base
    say("serving threads")
    errors = []
    with _server(ref, budget=8 << 20) as srv:
        def client(seed):
            rng = np.random.default_rng(seed)
            try:
                for _ in range(25):
                    ids = rng.integers(0, len(trav), int(rng.integers(1, 4)))
                    r = srv.wait(srv.submit_extract(ids))
                    np.testing.assert_array_equal(r["values"], np.concatenate([trav[i] for i in ids]))
            except Exception as e:  # reported below
                errors.append(e)
        th = [threading.Thread(target=client, args=(s,)) for s in range(6)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = srv.stats()
    assert not errors, errors[0]
    assert st["completed"] == 150


def test_serving_budget_and_bad_requests(base):
    cfg, tr, ref, trav = base
    say("serving budget")
    longest = int(np.argmax([t.size for t in trav]))
    budget = 2 * 4 * (trav[longest].size - 1)   # one slot is just too small for it
    with _server(ref, budget=budget) as srv:
        with pytest.raises(RuntimeError, match="HBM budget"):
            srv.wait(srv.submit_extract([longest]))
        with pytest.raises(RuntimeError, match="out of range"):
            srv.wait(srv.submit_extract([len(trav) + 5]))
        small = int(np.argmin([t.size for t in trav]))
        r = srv.wait(srv.submit_extract([small]))
        np.testing.assert_array_equal(r["values"], trav[small])
        r = srv.wait(srv.submit_extract([]))
        assert r["values"].size == 0 and r["offsets"].tolist() == [0]
        st = srv.stats()
        assert st["rejected"] == 2
