"""Host-side checks of libspecfab_gpu that need no GPU: the library loads,
exports every entry point include/specfab_gpu.h declares, and its host-only
helpers (indexing) behave."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "specfab_gpu.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sfg_[a-z0-9_A-Z]+)\s*\(", txt)))


def test_library_exports_header():
    from specfab_amd import _lib
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    # and the Python binding declares a signature for every one of them
    assert set(syms) <= set(_lib.SIGNATURES), sorted(set(syms) - set(_lib.SIGNATURES))


def test_indexing_helpers():
    from specfab_amd import _lib
    L = _lib.lib()
    for Lc, n in ((0, 1), (2, 6), (4, 15), (8, 45), (12, 91), (20, 231)):
        assert L.sfg_nlm_len(Lc) == n
    assert L.sfg_nlm_len(3) < 0
    lm = (ctypes.c_int * 30)()
    assert L.sfg_lm(4, lm) == 0
    pairs = [(lm[2 * j], lm[2 * j + 1]) for j in range(15)]
    assert pairs == [(l, m) for l in (0, 2, 4) for m in range(-l, l + 1)]


def test_no_device_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("has a GPU")
    from specfab_amd import _lib
    with pytest.raises(_lib.SFGError):
        _lib.Context(0)
