"""gfz_extract_noalloc (extraction with the plan in the library's fixed device
scratch array, no device allocation) against gfz_extract: the same values and
offsets on two calls in a row, the second smaller; clean errors beyond the array
and for delta_round 2. Through the C ABI."""
import ctypes
import sys

import numpy as np
import pytest

from gfaz_amd import codec, synth
from gfaz_amd.codec import GFZ_OUTPUT_DEVICE, _u64, lib
from oracle import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _split(nodes, lengths):
    offs = np.concatenate([[0], np.cumsum(lengths, dtype=np.int64)])
    return [nodes[offs[i]:offs[i + 1]] for i in range(lengths.size)]


def _noalloc(ctx, res, ids, out):
    ids = np.ascontiguousarray(ids, dtype=np.uint64)
    offs = np.zeros(ids.size + 1, np.uint64)
    n = _u64()
    rc = lib().gfz_extract_noalloc(ctx.handle, res.handle, ids.ctypes.data, ids.size, out.ptr,
                                   out.n, offs.ctypes.data, ctypes.byref(n))
    return rc, offs, int(n.value)


def _with_alloc(ctx, res, ids, out):
    ids = np.ascontiguousarray(ids, dtype=np.uint64)
    offs = np.zeros(ids.size + 1, np.uint64)
    n = _u64()
    codec._check(lib().gfz_extract(ctx.handle, res.handle, ids.ctypes.data, ids.size, out.ptr,
                                   out.n, GFZ_OUTPUT_DEVICE, offs.ctypes.data, ctypes.byref(n)))
    return offs, int(n.value)


def test_noalloc_extract_equals_extract():
    say("no-alloc extract: two calls against gfz_extract")
    tr = synth.generate(synth.CONFIGS["plain_20k"])
    ref = oracle.compress(tr.nodes, tr.lengths, 8, 2, 1, threads=8)
    trav = _split(tr.nodes, tr.lengths)
    with codec.Context(0) as ctx:
        rb = codec.Rulebook(ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
        res = codec.Resident(ctx, rb, ref["encoded"], ref["encoded_lengths"], 1)
        cap = 3 * tr.n_steps
        a = codec.DeviceArray(ctx, cap, np.int32)
        b = codec.DeviceArray(ctx, cap, np.int32)
        rng = np.random.default_rng(5)
        for ids in (np.concatenate([np.arange(tr.lengths.size), rng.integers(0, 16, 20)]),
                    np.array([9, 2, 9])):  # the second call is smaller
            rc, offs, n = _noalloc(ctx, res, ids, a)
            assert rc == 0, lib().gfz_last_error()
            offs2, n2 = _with_alloc(ctx, res, ids, b)
            assert n == n2 == int(offs[-1])
            np.testing.assert_array_equal(offs, offs2)
            got = a.to_host(n)
            np.testing.assert_array_equal(got, b.to_host(n2))
            np.testing.assert_array_equal(got, np.concatenate([trav[i] for i in ids]))
        rc, offs, n = _noalloc(ctx, res, [], a)
        assert rc == 0 and n == 0 and offs.tolist() == [0]
        rc, _, _ = _noalloc(ctx, res, [0], codec.DeviceArray(ctx, 1, np.int32))
        assert rc == -1 and b"output capacity" in lib().gfz_last_error()
        res.close()
        rb.close()


def test_noalloc_rejects_beyond_scratch_and_delta2():
    say("no-alloc extract: limits")
    nodes = np.array([5, 6, 7, 5, 6, 7], np.int32)
    lengths = np.array([1, 2, 3], np.uint32)
    with codec.Context(0) as ctx:
        for delta in (1, 2):
            ref = oracle.compress(nodes, lengths, 8, 2, delta, threads=1)
            rb = codec.Rulebook(ctx, ref["rules_first"], ref["rules_second"], ref["start_id"])
            res = codec.Resident(ctx, rb, ref["encoded"], ref["encoded_lengths"], delta)
            out = codec.DeviceArray(ctx, 20_000, np.int32)
            if delta == 1:
                # 20,000 one-step segments: a plan of about 1.1 MB, above the 1 MiB array
                rc, _, _ = _noalloc(ctx, res, np.zeros(20_000, np.uint64), out)
                assert rc == -1 and b"more than the fixed" in lib().gfz_last_error()
                rc, offs, n = _noalloc(ctx, res, [2, 0], out)
                assert rc == 0 and n == 4
                np.testing.assert_array_equal(out.to_host(4), [5, 6, 7, 5])
            else:
                rc, _, _ = _noalloc(ctx, res, [0], out)
                assert rc == -1 and b"delta_round 0 and 1" in lib().gfz_last_error()
            res.close()
            rb.close()
