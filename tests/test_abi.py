"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and host-side helpers (tiling, tile layout) behave."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    src = open(os.path.join(ROOT, "include", "mpqc_ccsd.h")).read()
    return sorted(set(re.findall(r"\b(mpqc_ccsd_\w+)\s*\(", src)))


def test_library_exports_every_header_symbol():
    from mpqc_amd import _lib
    lib = _lib.load()
    syms = _header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.EXPORTS)
    assert b"gfx950" in lib.mpqc_ccsd_version()


def test_last_error_without_context():
    from mpqc_amd import _lib
    lib = _lib.load()
    st = lib.mpqc_ccsd_ctx_create(0, 0, None, 0, None, 0, 0, None)
    assert st == _lib.ERR_ARG
    assert b"null" in lib.mpqc_ccsd_last_error(None)


def test_trange1_matches_reference_rule():
    from mpqc_amd.ccsd import compute_trange1
    assert compute_trange1(90, 24) == [0, 24, 48, 72, 90]
    assert compute_trange1(10, 24) == [0, 10]
    assert compute_trange1(19, 4) == [0, 4, 8, 12, 16, 19]
    assert compute_trange1(24, 24) == [0, 24]


def test_ta_tile_layout_roundtrip():
    from mpqc_amd.ccsd import from_ta_tiles, to_ta_tiles
    rng = np.random.default_rng(0)
    x = rng.normal(size=(7, 5, 3, 4))
    hm = [[0, 3, 7], [0, 5], [0, 1, 3], [0, 2, 4]]
    buf = to_ta_tiles(x, hm)
    assert buf.size == x.size
    # first tile is x[0:3, 0:5, 0:1, 0:2] row-major
    assert np.array_equal(buf[:30], x[0:3, 0:5, 0:1, 0:2].ravel())
    assert np.array_equal(from_ta_tiles(buf, x.shape, hm), x)
