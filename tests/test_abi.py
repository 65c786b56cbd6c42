"""CPU-side checks of the C-ABI library: it loads and exports every symbol
include/ars.h declares (no compute calls: this container has no GPU)."""
import os
import re

import pytest

import _pkg

ROOT = _pkg.ROOT


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ars.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ars_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("ars_create", "ars_encode", "ars_recover", "ars_recompute_data"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    ars = _pkg.load()
    if not os.path.exists(ars.LIB_PATH):
        pytest.skip("libars.so not built")
    lib = ars.lib()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert set(declared_symbols()) <= set(ars.SIGNATURES), "python bindings miss a symbol"
    assert b"gfx950" in lib.ars_version()


def test_validation_errors_without_gpu():
    """Argument checks run before any HIP call, so they work on a CPU host."""
    ars = _pkg.load()
    if not os.path.exists(ars.LIB_PATH):
        pytest.skip("libars.so not built")
    lib = ars.lib()
    # null context -> ARS_E_NULL
    assert lib.ars_encode(None, 64, 32, 1, 1, 64, 0, None) == ars.ARS_E_NULL
    assert lib.ars_recover(None, 64, 32, 1, 64, None, 0, None) == ars.ARS_E_NULL
