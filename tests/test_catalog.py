"""Traversal catalog lookups (gfz_catalog_*), host only: the reference's lookup
semantics and messages (src/workflows/extraction_workflow.cpp:206-424)."""
import threading

import numpy as np
import pytest

from gfaz_amd import codec

PREFIX = "Extraction workflow error: "


def fixture_catalog():
    c = codec.Catalog()
    c.add_paths(["pathA", "pathB"], 0, ["2M,1M", ""])
    c.add_walks([("sampleA", 0, "chr1", 0, 6), ("sampleB", 1, "chr2", -1, -1)], 2)
    return c


def test_paths_by_name_in_request_order():
    c = fixture_catalog()
    assert c.info() == (2, 2)
    assert list(c.find_paths(["pathB", "pathA", "pathB"])) == [1, 0, 1]
    assert c.find_paths([]).size == 0


def test_missing_names_and_tuples_name_the_request():
    c = fixture_catalog()
    with pytest.raises(RuntimeError, match=f"^{PREFIX}path not found: pathZ$"):
        c.find_paths(["pathA", "pathZ"])
    with pytest.raises(RuntimeError, match=f"^{PREFIX}walk not found: sampleZ$"):
        c.find_walks_by_name(["sampleZ"])
    with pytest.raises(RuntimeError, match=f"^{PREFIX}walk not found for the provided "
                                           "identifier: sampleB\t1\tchr2\t0\t-1$"):
        c.find_walks([("sampleB", 1, "chr2", 0, -1)])


def test_walk_tuples_match_all_five_fields_first_stored_wins():
    c = fixture_catalog()
    assert list(c.find_walks([("sampleB", 1, "chr2", -1, -1), ("sampleA", 0, "chr1", 0, 6)])) == [3, 2]
    for wrong in [("sampleA", 1, "chr1", 0, 6), ("sampleA", 0, "chr2", 0, 6),
                  ("sampleA", 0, "chr1", 1, 6), ("sampleA", 0, "chr1", 0, 7),
                  ("sample", 0, "Achr1", 0, 6)]:
        with pytest.raises(RuntimeError, match="walk not found for the provided identifier"):
            c.find_walks([wrong])
    # an identical tuple appended later: the first stored walk keeps answering
    c.add_walks([("sampleA", 0, "chr1", 0, 6)], 10)
    assert list(c.find_walks([("sampleA", 0, "chr1", 0, 6)])) == [2]


def test_any_repeated_stored_name_fails_every_lookup_by_name():
    """The reference builds its name map before looking anything up and throws at the
    first repeat in storage order, whatever was requested."""
    c = fixture_catalog()
    c.add_paths(["pathC", "pathB", "pathC"], 4)
    with pytest.raises(RuntimeError, match=f"^{PREFIX}path name is ambiguous: pathB$"):
        c.find_paths(["pathA"])
    assert c.find_paths([]).size == 0  # an empty request returns before the map
    c.add_walks([("sampleA", 3, "chr9", -1, -1)], 7)
    with pytest.raises(RuntimeError, match=f"^{PREFIX}walk name is ambiguous: sampleA$"):
        c.find_walks_by_name(["sampleB"])
    # tuple lookups do not use names
    assert list(c.find_walks([("sampleA", 3, "chr9", -1, -1)])) == [7]


def test_appends_take_the_given_traversal_ids():
    c = fixture_catalog()
    c.add_paths(["pathC"], 4, ["*"])
    c.add_walks([("sampleC", 0, "chr3", -1, -1)], 5)
    assert list(c.find_paths(["pathC", "pathA"])) == [4, 0]
    assert list(c.find_walks_by_name(["sampleC", "sampleA"])) == [5, 2]


def test_bytes_names_and_bad_arguments():
    c = codec.Catalog()
    c.add_paths([b"p\xff\x00x", "plain"], 0)
    assert list(c.find_paths([b"p\xff\x00x"])) == [0]
    with pytest.raises(RuntimeError, match="overlaps must be empty or one per path"):
        c.add_paths(["a", "b"], 2, ["x"])


def test_lookups_run_concurrently_with_appends():
    c = codec.Catalog()
    c.add_paths([f"p{i}" for i in range(1000)], 0)
    errors = []

    def reader(seed):
        rng = np.random.default_rng(seed)
        try:
            for _ in range(200):
                ids = rng.integers(0, 1000, 16)
                got = c.find_paths([f"p{i}" for i in ids])
                assert np.array_equal(got, ids)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    def writer():
        for k in range(50):
            c.add_paths([f"q{k}_{j}" for j in range(20)], 1000 + 20 * k)

    ts = [threading.Thread(target=reader, args=(s,)) for s in range(4)] + [threading.Thread(target=writer)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors
    assert c.info() == (2000, 0)
    assert list(c.find_paths(["q49_19", "p999"])) == [1999, 999]
