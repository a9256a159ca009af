"""The C-ABI library loads and exports every symbol include/mosaic_amd.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"^\s*int\s+(mosaic_\w+)\s*\(", text, flags=re.M)
    return names


def test_header_declares_entry_points():
    names = declared()
    assert "mosaic_kin_sum" in names and len(names) == len(set(names))


def test_library_exports_all_declared():
    from mosaic_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libmosaic_amd.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for n in declared():
        assert hasattr(L, n), n
    assert set(declared()) == set(_lib.EXPORTS)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from mosaic_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_LIB", None)
    with pytest.raises(_lib.MosaicLibraryError):
        _lib.lib()


def test_every_export_has_a_signature():
    """ctypes converts Python floats only with declared argument types."""
    from mosaic_amd import _lib
    assert set(_lib._SIGS) == set(_lib.EXPORTS)
