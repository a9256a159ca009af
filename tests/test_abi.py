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


def test_pixel_tables_cache_key():
    """The pixel prologue is reused only for the same coordinate tensor,
    unchanged: another tensor at the same address (here a view), or an
    in-place change, recomputes it."""
    import torch
    from mosaic_amd import kinematic as K
    spc = K.species_table(__import__("mosaic_amd.tables", fromlist=["x"]).species_constants(
        ["Au"], 10000.0))
    opts = K.kin_opts(5e10)
    c = torch.zeros(3, 16, dtype=torch.float64)
    key = K._pixels_key(spc, opts, c)
    assert K._same_pixels(key, spc, opts, c)
    assert not K._same_pixels(key, spc, opts, c.view(3, 16))     # same address, other tensor
    assert not K._same_pixels(key, spc, K.kin_opts(5e10, apply_pol=True), c)
    c.add_(1.0)
    assert not K._same_pixels(key, spc, opts, c)
    assert not K._same_pixels(None, spc, opts, c)


def test_staged_sum_size_limits():
    """mosaic_kin_workspace (called first by every staged-sum entry) rejects
    sizes past the kernels' grids: more than 65535 poses, more than 2^30
    pixels (a 40000^2 detector would overflow 32-bit indices), more than
    2^31 - 1 pixel-poses, 2^32 atoms."""
    import ctypes
    from mosaic_amd import _lib
    lib = _lib.lib()
    n = ctypes.c_size_t(0)
    counts = (ctypes.c_int64 * 1)(1000)

    def ws(atoms, frames, ny, nz):
        counts[0] = atoms
        return lib.mosaic_kin_workspace(atoms, 1, counts, frames, ny, nz, ctypes.byref(n))

    assert ws(1000, 64, 256, 256) == 0 and n.value > 0
    assert ws(1000, 65535, 8, 8) == 0
    assert ws(1000, 65536, 8, 8) == _lib.MOSAIC_ERR_ARGS
    assert ws(1000, 1, 32768, 32768) == 0                   # 2^30 pixels
    assert ws(1000, 1, 40000, 40000) == _lib.MOSAIC_ERR_ARGS
    assert ws(1 << 32, 1, 8, 8) == _lib.MOSAIC_ERR_ARGS
    assert ws(1000, 65535, 256, 256) == _lib.MOSAIC_ERR_ARGS    # 4.3e9 pixel-poses
