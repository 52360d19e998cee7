"""The sort that orders each round's rules (ascending distinct 64-bit keys) against
numpy's sort: every size boundary of its paths (one block, sample sort, radix sort
beyond 512 K keys), skewed and presorted inputs, and buckets forced past the
register limit onto the in-memory radix path."""
import sys

import numpy as np
import pytest

from gfaz_amd import codec

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def distinct_keys(rng, n, hi=1 << 62):
    k = np.unique(rng.integers(0, hi, size=int(n * 1.1) + 16, dtype=np.uint64))
    rng.shuffle(k)
    return k[:n]


def check(ctx, keys, local_cap=0):
    say(f"  sort n={keys.size} local_cap={local_cap}")
    got = codec.debug_sort_u64(ctx, keys, local_cap)
    np.testing.assert_array_equal(got, np.sort(keys))


@pytest.mark.parametrize("n", [0, 1, 2, 3, 100, 1024, 1025, 2487, 4095, 4096, 4097, 8192, 8193,
                               50_000, 515_000, 524_288, 524_289, 2_000_000])
def test_sort_sizes(gpu_ctx, n):
    check(gpu_ctx, distinct_keys(np.random.default_rng(n), n))


def test_sort_skewed_and_presorted(gpu_ctx):
    rng = np.random.default_rng(7)
    # Most keys in one narrow range, a few spread far: uneven buckets.
    dense = np.arange(200_000, dtype=np.uint64) * 3 + (1 << 40)
    sparse = distinct_keys(rng, 5_000)
    keys = np.unique(np.concatenate([dense, sparse]))
    rng.shuffle(keys)
    check(gpu_ctx, keys)
    check(gpu_ctx, np.sort(keys))
    check(gpu_ctx, np.sort(keys)[::-1].copy())


@pytest.mark.parametrize("n,cap", [(50, 16), (200, 16), (3000, 64), (6000, 64), (20_000, 64),
                                   (100_000, 128)])
def test_sort_in_memory_path(gpu_ctx, n, cap):
    check(gpu_ctx, distinct_keys(np.random.default_rng(n + cap), n), cap)
