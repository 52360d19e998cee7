"""One ctx per thread: several host threads compress, frame and decode different
inputs on their own contexts of the same GPU at once (ctypes releases the GIL for
the calls, so the library really runs them concurrently), and every result equals
the same call made alone."""
import threading

import numpy as np
import pytest

from gfaz_amd import codec, container, synth

pytestmark = pytest.mark.gpu

N_THREADS = 4


def _inputs():
    out = []
    for k in range(N_THREADS):
        cfg = synth.CONFIGS["plain_500k"].scaled(backbone=150_000 + 20_000 * k, n_haps=16 + 4 * k,
                                                 len_max=200_000)
        tr = synth.generate(cfg)
        out.append((tr.nodes, tr.lengths, tr.lengths.size // 2))
    return out


def _work(ctx, nodes, lengths, n_paths):
    c = codec.compress_traversals(ctx, nodes, lengths, n_paths, 8, 2, 1)
    rb = codec.Rulebook(ctx, c.rules_first, c.rules_second, c.start_id)
    try:
        dec, lens, mism = codec.decompress_traversals(ctx, rb, c.encoded, c.encoded_lengths,
                                                      lengths, 1)
        res = codec.Resident(ctx, rb, c.encoded, c.encoded_lengths, 1)
        try:
            ext, _ = res.extract(np.arange(lengths.size)[::-1])
        finally:
            res.close()
    finally:
        rb.close()
    cols = container.compress_traversals_gpu(ctx, nodes, lengths, n_paths, threads=2)
    framed = tuple(getattr(cols, f).pair() for f in
                   ("rules_first_zstd", "rules_second_zstd", "paths_zstd", "walks_zstd"))
    return (c.encoded, c.rules_first, c.rules_second, dec, lens, mism, ext, framed)


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            np.testing.assert_array_equal(x, y)
        else:
            assert x == y


def test_contexts_on_threads_match_alone():
    inputs = _inputs()
    alone = []
    with codec.Context(0) as ctx:
        for nodes, lengths, n_paths in inputs:
            alone.append(_work(ctx, nodes, lengths, n_paths))
    results = [None] * N_THREADS
    errors = []
    start = threading.Barrier(N_THREADS)

    def run(k):
        try:
            with codec.Context(0) as ctx:
                start.wait()
                for _ in range(3):  # repeated, so calls of different threads interleave
                    results[k] = _work(ctx, *inputs[k])
        except Exception as e:  # noqa: BLE001
            errors.append((k, e))

    print(f"{N_THREADS} threads, one ctx each", flush=True)
    ts = [threading.Thread(target=run, args=(k,)) for k in range(N_THREADS)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for k in range(N_THREADS):
        _same(results[k], alone[k])
        nodes, lengths, _ = inputs[k]
        np.testing.assert_array_equal(results[k][3], nodes)


def test_a_ctx_refuses_another_ctx_objects():
    """Objects belong to the ctx that made them (the reference has one global CUDA
    context; here a rulebook or resident store used with another ctx fails cleanly)."""
    nodes, lengths, _ = _inputs()[0]
    with codec.Context(0) as a, codec.Context(0) as b:
        c = codec.compress_traversals(a, nodes, lengths, 0, 8, 2, 1)
        rb = codec.Rulebook(a, c.rules_first, c.rules_second, c.start_id)
        res = codec.Resident(a, rb, c.encoded, c.encoded_lengths, 1)
        try:
            with pytest.raises(RuntimeError, match="another ctx"):
                codec.decompress_traversals(b, rb, c.encoded, c.encoded_lengths, lengths, 1)
            res.ctx = b
            with pytest.raises(RuntimeError, match="another ctx"):
                res.extract([0])
        finally:
            res.ctx = a
            res.close()
            rb.close()
