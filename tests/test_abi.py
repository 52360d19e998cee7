"""CPU tests of the boundary: the built library loads and exports every entry point that
include/*.h declares; host-side argument checks fail cleanly without a GPU."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(gfz_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    for must in ("gfz_compress_traversals", "gfz_rulebook_load", "gfz_decompress_traversals",
                 "gfz_last_error", "gfz_ctx_create"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from gfaz_amd import _native
    lib = ctypes.CDLL(_native.CODEC_LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_are_status_codes_not_exceptions():
    from gfaz_amd import codec
    L = codec.lib()
    assert L.gfz_compress_traversals(None, None, 0, None, 0, 0, 8, 2, 1, 0, None) == -1
    assert b"null argument" in L.gfz_last_error()
    assert L.gfz_rulebook_load(None, None, None, 0, 0, 0, None) == -1
    assert L.gfz_ctx_set_profiling(None, 1) == -1
    assert L.gfz_version().startswith(b"gfaz_amd")


def test_no_gpu_here_fails_loudly():
    from gfaz_amd import codec
    import torch  # noqa: F401  (only to learn whether a GPU is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        codec.Context(0)


def test_synthetic_generator_is_deterministic():
    from gfaz_amd import synth
    cfg = synth.CONFIGS["plain_20k"].scaled(backbone=3000, n_haps=4)
    a, b = synth.generate(cfg), synth.generate(cfg)
    assert np.array_equal(a.nodes, b.nodes) and np.array_equal(a.lengths, b.lengths)
    assert a.n_steps > 3000 * 4 * 0.9


def test_serving_argument_checks_need_no_gpu():
    """gfz_server_create rejects a missing config or a tiny HBM budget before it
    touches a device."""
    from gfaz_amd import serving
    L = serving._bind()
    out = ctypes.c_void_p()
    assert L.gfz_server_create(None, None, None, 0, 1, None, 0, None, 0, None, 0, 1,
                               ctypes.byref(out)) == -1
    assert b"null argument" in L.gfz_last_error()
    with pytest.raises(RuntimeError, match="HBM budget too small"):
        serving.Server(np.zeros(0, np.int32), np.zeros(0, np.int32), 1,
                       np.zeros(0, np.int32), np.zeros(0, np.uint32), 1, hbm_budget_bytes=1024)
