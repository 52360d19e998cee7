"""Piecewise framing (gfz_framer_*) and the framing behind gfz_frame_int32.

The reference frames a column with one ZSTD_compress2 call (src/codec/codec.cpp:
180-216). The library feeds zstd whole 128 KiB blocks of the column as they become
final; these tests hold its payloads to a ZSTD_compress2 call made here through
ctypes on the same libzstd, at sizes of many blocks (the .gfaz goldens are all
smaller than one block), at several levels and for any sequence of ready() calls.
"""
import ctypes

import numpy as np
import pytest

from gfaz_amd import codec

_Z = ctypes.CDLL("libzstd.so.1")
_Z.ZSTD_createCCtx.restype = ctypes.c_void_p
_Z.ZSTD_freeCCtx.argtypes = [ctypes.c_void_p]
_Z.ZSTD_CCtx_setParameter.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
_Z.ZSTD_CCtx_setParameter.restype = ctypes.c_size_t
_Z.ZSTD_compress2.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                              ctypes.c_void_p, ctypes.c_size_t]
_Z.ZSTD_compress2.restype = ctypes.c_size_t
_Z.ZSTD_compressBound.argtypes = [ctypes.c_size_t]
_Z.ZSTD_compressBound.restype = ctypes.c_size_t
_Z.ZSTD_isError.argtypes = [ctypes.c_size_t]
_Z.ZSTD_isError.restype = ctypes.c_uint
_ZSTD_C_LEVEL, _ZSTD_C_WORKERS = 100, 400  # zstd.h ZSTD_cParameter values


def zstd_compress2(data: bytes, level: int, workers: int) -> bytes:
    """One ZSTD_compress2 call configured as configure_zstd_context
    (src/codec/codec.cpp:148-177): level, then the worker count (refusal ignored)."""
    c = _Z.ZSTD_createCCtx()
    try:
        assert not _Z.ZSTD_isError(_Z.ZSTD_CCtx_setParameter(c, _ZSTD_C_LEVEL, level))
        _Z.ZSTD_CCtx_setParameter(c, _ZSTD_C_WORKERS, workers)
        bound = _Z.ZSTD_compressBound(len(data))
        buf = ctypes.create_string_buffer(bound)
        n = _Z.ZSTD_compress2(c, buf, bound, data, len(data))
        assert not _Z.ZSTD_isError(n)
        return buf.raw[:n]
    finally:
        _Z.ZSTD_freeCCtx(c)


def column(n: int, seed: int = 0) -> np.ndarray:
    """Grammar-like int32 column: node IDs along a drifting walk plus rule IDs."""
    rng = np.random.default_rng(seed)
    steps = rng.integers(-3, 5, size=n).cumsum() % 200_000
    rules = rng.integers(1 << 20, (1 << 20) + 5000, size=n)
    pick = rng.random(n) < 0.3
    sign = np.where(rng.random(n) < 0.1, -1, 1)
    return (np.where(pick, rules, steps) * sign).astype(np.int32)


SIZES = [1, 1000, 32768, 32769, 300_000, 2_500_000]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("level", [1, 9, 19])
def test_frame_matches_zstd_compress2(monkeypatch, n, level):
    monkeypatch.setenv("GFA_COMPRESSION_ZSTD_LEVEL", str(level))
    a = column(n, seed=n + level)
    orig, payload = codec.frame_int32(a, threads=4)
    assert orig == a.nbytes
    assert payload == zstd_compress2(a.tobytes(), level, 4)
    np.testing.assert_array_equal(codec.unframe_int32(payload, orig), a)


def test_default_level_is_nine():
    a = column(400_000, seed=3)
    assert codec.frame_int32(a)[1] == zstd_compress2(a.tobytes(), 9, 8)


@pytest.mark.parametrize("seed", range(4))
def test_pieces_in_any_order_of_readiness(seed):
    """ready() at arbitrary byte counts (block multiples or not, repeated, shrinking)
    never changes the frame."""
    rng = np.random.default_rng(seed)
    a = column(int(rng.integers(200_000, 1_500_000)), seed=seed)
    want = codec.frame_int32(a)
    f = codec.Framer(a)
    done = 0
    while done < a.nbytes:
        step = int(rng.choice([1, 4, 131072, 131072 * 3, 500_000, 4 << 20]))
        done = min(a.nbytes, done + step)
        f.ready(done)
        if rng.random() < 0.2:
            f.ready(done // 2)  # a smaller count than before is a no-op
    assert f.end() == want


def test_end_without_ready_and_empty_column():
    a = column(100_000, seed=9)
    assert codec.Framer(a).end() == codec.frame_int32(a)
    assert codec.Framer(np.zeros(0, np.int32)).end() == (0, b"")


def test_ready_beyond_the_column_is_refused():
    f = codec.Framer(column(10))
    with pytest.raises(RuntimeError, match="exceeds"):
        f.ready(41)
    f.end()


def test_unframe_accepts_concatenated_frames():
    """ZSTD_decompress (src/codec/codec.cpp:224-244) decodes a payload of several
    frames; so does the library (it decodes a single frame block by block)."""
    a, b = column(50_000, 1), column(70_000, 2)
    pa, pb = codec.frame_int32(a)[1], codec.frame_int32(b)[1]
    got = codec.unframe_int32(pa + pb, a.nbytes + b.nbytes)
    np.testing.assert_array_equal(got, np.concatenate([a, b]))


def test_unframe_rejects_truncated_and_oversized():
    a = column(300_000, 4)
    orig, p = codec.frame_int32(a)
    with pytest.raises(RuntimeError):
        codec.unframe_int32(p[: len(p) // 2], orig)
    with pytest.raises(RuntimeError):
        codec.unframe_int32(p, orig + 4)  # frame holds fewer bytes than the block says
    with pytest.raises(RuntimeError):
        codec.unframe_int32(p, orig - 4)
